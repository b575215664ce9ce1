#!/usr/bin/env python3
"""FP64-oracle fixtures for the device parity tests whose oracle run is too
slow to repeat inside a test.

The FP64 oracle (oracle/_build/cpu_ref64: real = double, -O2 without
contraction, observer values with 17 significant digits) is pinned byte for
byte to the reference built in double precision from its own sources
(tests/test_oracle_reference.py).  Its observer files for these meshes are
kept here as arrays, so the device can be compared with them on a machine
that has neither the reference nor the time to run the oracle:

  the reference's ABC-reflection tests sgref_mur, sgref_pml6, sgref_pml12,
  dgref_ref, dgrefmb_ref (140^3 and 150^3 cells with 12-17 observers);
  the full configurations: the graded shielded enclosure (200x160x120, 4000
  steps, 512-frequency DFT), the Debye-sphere volume (256^3, 3000 steps) and
  the SIBC + Mur box (384x384x256, 6000 steps), as vulture_amd.meshgen writes
  them.

tests/golden/fp64/<case>.npz holds, for every eh_*_{td,fd}.asc file of the
oracle run, td/<file> as float64 [steps][2 + ncomp] (step, time, values) and
fd/<file> as float64 [nfreq][1 + 2 ncomp]; for the full configurations only
the first `keep_td` time-domain files are kept (every frequency-domain file
is).  mesh_sha256 records the mesh text the arrays belong to.  For the
reference meshes, whose fields are cancellation-dominated, float_td and
float_fd_db record the error of the reference's own float arithmetic (the
float oracle, byte-identical to the reference's float build) against FP64,
the yardstick of the FP32 device run there.

For a generated case, `--float-error case ...` runs the float oracle on the
same mesh and adds float_td and float_fd_db against the stored FP64 arrays to
its fixture (for configurations long and resonant enough that float rounding
alone exceeds the FP32 bars).

Usage: python tests/golden/make_fp64_fixtures.py [case ...]   (default: all)
       python tests/golden/make_fp64_fixtures.py --float-error case ...
"""
import hashlib
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from vulture_amd import meshgen  # noqa: E402
from vulture_amd.obsfiles import observer_files, read_fd, read_td  # noqa: E402

OUT = os.path.join(HERE, "fp64")
REFERENCE = ["sgref_mur", "sgref_pml6", "sgref_pml12", "dgref_ref", "dgrefmb_ref"]
# generated full configurations: (case, meshgen name, kwargs, time-domain files kept)
GENERATED = [("enclosure_full", "enclosure", {}, 8), ("debye_full", "debye", {}, 2),
             ("sibc_box_full", "sibc_box", {}, 2)]


def stage(case, work):
    """Write the case's mesh (and its data files) under `work`; return the mesh path."""
    if case in REFERENCE:
        d = os.path.join(work, "in")
        shutil.copytree(os.path.join(HERE, "ref_inputs", case), d)
        return os.path.join(d, [f for f in os.listdir(d) if f.endswith(".mesh")][0])
    for c, name, kw, _ in GENERATED:
        if c == case:
            return meshgen.write(name, os.path.join(work, "in"), **kw)
    raise KeyError(case)


def keep_td(case):
    for c, _, _, k in GENERATED:
        if c == case:
            return k
    return None


def mesh_sha256(mesh):
    return hashlib.sha256(open(mesh, "rb").read()).hexdigest()


def make(case):
    work = tempfile.mkdtemp(prefix=f"fp64_{case}_")
    try:
        mesh = stage(case, work)
        out = os.path.join(work, "oracle")
        t0 = time.time()
        oracle.run_mesh(mesh, out, precision="fp64", full_precision=True, threads=os.cpu_count() or 1,
                        timeout=24 * 3600)
        arrays = {"mesh_sha256": np.array(mesh_sha256(mesh))}
        td = observer_files(out, "eh_*_td.asc")
        k = keep_td(case)
        for f in td[:k] if k is not None else td:
            n, t, v = read_td(f)
            arrays["td/" + os.path.basename(f)] = np.column_stack([n.astype(np.float64), t, v])
        for f in observer_files(out, "eh_*_fd.asc"):
            fr, v = read_fd(f)
            cols = [fr]
            for c in range(v.shape[1]):
                cols += [v[:, c].real, v[:, c].imag]
            arrays["fd/" + os.path.basename(f)] = np.column_stack(cols)
        if case in REFERENCE:
            from vulture_amd.obsfiles import compare_outputs
            out32 = os.path.join(work, "oracle32")
            oracle.run_mesh(mesh, out32, precision="fp32", full_precision=True, threads=os.cpu_count() or 1)
            r = compare_outputs(out32, out)
            arrays["float_td"] = np.array(r["td"])
            arrays["float_fd_db"] = np.array(r["fd_db"])
        os.makedirs(OUT, exist_ok=True)
        np.savez_compressed(os.path.join(OUT, f"{case}.npz"), **arrays)
        print(f"{case}: {len(arrays) - 1} files, oracle {time.time() - t0:.0f} s", flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def write_fixture_files(fx, dst, wf_src):
    """The fixture's arrays as observer files under dst, beside the waveform
    spectrum files of the run in wf_src (host-computed, the same in every run)."""
    import glob
    os.makedirs(dst, exist_ok=True)
    for key in fx.files:
        if key.startswith(("td/", "fd/")):
            with open(os.path.join(dst, key[3:]), "w") as f:
                for row in fx[key]:
                    head = f"{int(row[0])} " if key.startswith("td/") else f"{float(row[0])!r} "
                    f.write(head + " ".join(repr(float(x)) for x in row[1:]) + "\n")
    for w in glob.glob(os.path.join(wf_src, "wf_*_fd.asc")):
        shutil.copy(w, dst)


def add_float_error(case):
    from vulture_amd.obsfiles import compare_outputs
    path = os.path.join(OUT, f"{case}.npz")
    fx = np.load(path)
    arrays = {k: fx[k] for k in fx.files}
    work = tempfile.mkdtemp(prefix=f"f32_{case}_")
    try:
        mesh = stage(case, work)
        if mesh_sha256(mesh) != str(fx["mesh_sha256"]):
            raise SystemExit(f"{case}: the mesh generator changed since the fixture was made")
        t0 = time.time()
        out32 = os.path.join(work, "oracle32")
        oracle.run_mesh(mesh, out32, precision="fp32", full_precision=True, threads=os.cpu_count() or 1,
                        timeout=24 * 3600)
        ref = os.path.join(work, "fixture")
        write_fixture_files(fx, ref, out32)
        r = compare_outputs(out32, ref)
        arrays["float_td"] = np.array(r["td"])
        arrays["float_fd_db"] = np.array(r["fd_db"])
        np.savez_compressed(path, **arrays)
        print(f"{case}: float oracle against FP64: td {r['td']:.3e}, fd_db {r['fd_db']:.3e} "
              f"({r['files']} files), {time.time() - t0:.0f} s", flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    if sys.argv[1:2] == ["--float-error"]:
        for c in sys.argv[2:]:
            add_float_error(c)
    else:
        cases = sys.argv[1:] or REFERENCE + [c for c, _, _, _ in GENERATED]
        for c in cases:
            make(c)
