"""CPU oracle of the FDTD time loop.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never calls it.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_REF64 = os.path.join(HERE, "_build", "cpu_ref64")
CPU_REF32 = os.path.join(HERE, "_build", "cpu_ref32")
CPU_REF64_FAST = os.path.join(HERE, "_build", "cpu_ref64_fast")
CPU_REF32_FAST = os.path.join(HERE, "_build", "cpu_ref32_fast")
REF_BIN = os.path.join(HERE, "_ref", "vulture_ref")


def run_mesh(mesh, outdir, precision="fp64", full_precision=True, steps=-1, threads=0, timeout=3600):
    """Run the oracle on a mesh; returns the time-loop seconds it reports."""
    exe = CPU_REF64 if precision == "fp64" else CPU_REF32
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: run make -C oracle")
    os.makedirs(outdir, exist_ok=True)
    cmd = [exe, "-o", outdir]
    if full_precision:
        cmd.append("-f")
    if steps >= 0:
        cmd += ["-n", str(steps)]
    if threads:
        cmd += ["-t", str(threads)]
    cmd.append(os.path.abspath(mesh))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"oracle failed ({r.returncode}): {r.stderr.strip()}")
    for line in r.stdout.splitlines():
        if line.startswith("loop_seconds"):
            return float(line.split()[1])
    return float("nan")


def time_mesh(mesh, steps, warmup, threads, precision="fp64", outdir=None, timeout=600):
    """Time `steps` oracle steps after `warmup` untimed ones on `threads`
    OpenMP threads (the -O3 timing build of the same sources).  Returns
    (timed_seconds, gobox_cells, timed_steps)."""
    exe = CPU_REF64_FAST if precision == "fp64" else CPU_REF32_FAST
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: run make -C oracle")
    cmd = [exe, "-n", str(steps + warmup), "-w", str(warmup), "-t", str(threads)]
    if outdir:
        os.makedirs(outdir, exist_ok=True)
        cmd += ["-o", outdir]
    cmd.append(os.path.abspath(mesh))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout,
                       cwd=outdir or os.path.dirname(os.path.abspath(mesh)))
    if r.returncode != 0:
        raise RuntimeError(f"oracle failed ({r.returncode}): {r.stderr.strip()}")
    vals = dict(line.split() for line in r.stdout.splitlines() if line.count(" ") == 1)
    return float(vals["loop_seconds"]), float(vals["cells"]), steps


def run_reference(mesh_dir, mesh_name, timeout=3600):
    """Run the reference float build (only where it was built from the sources)."""
    if not os.path.exists(REF_BIN):
        raise RuntimeError("reference build not present")
    return subprocess.run([REF_BIN, mesh_name], cwd=mesh_dir, capture_output=True, text=True, timeout=timeout)
