#!/usr/bin/env python3
"""Throughput of the FDTD time loop on one MI355X.

Workload: the Debye-sphere volume -- a 256^3 cubic mesh of 2 mm cells with an
8-cell UPML on all faces (272^3 = 20,123,648 updated cells), 200 seeded
single-pole Debye spheres, a TF/SF plane wave at seeded angles on the
auxiliary-grid path and 8 point observers with 64-frequency running DFTs.
The mesh file is generated (vulture_amd.meshgen, seed 3) and fed to the same
parser and set-up as any Vulture mesh.

A step is one full leapfrog update (every kernel of the time loop) of every
cell.  Both storage forms run: FP64 (the plain form) and FP32 fields with
FP64 DFT/SIBC state; `value` is the FP32 form's Mcell-updates/s and the FP64
figures sit in `fp64`.  Fields stay resident in HBM; W untimed steps run
first, then K steps are timed with hipEvents on the solve's stream.

`roofline` prices the dominant kernel (the fast E or fast H sweep over the
inner grid, whichever takes longer) by its own compulsory bytes per launch
(kernel_bytes below: 9 field words and the 2-byte coefficient class per cell,
plus the fused Debye sites for E).  Its average duration comes from hipEvents
around every launch in the timed steps.  `step_roofline` prices the whole
step with the compulsory bytes of one pass in the same representation
(step_bytes: 12 field words and two 2-byte classes per cell, 18 more words
per UPML cell, the Debye state), over ms_per_step.

`cpu_baseline` times the reference itself (its float build from its own
sources with its OpenMP option, built by oracle/Makefile) on a bounded number
of steps of the same mesh on this host, when that build is present; the CPU
restatement of the loop is timed beside it (`cpu_port`).

Further records of the same run: `fp32_error` (the full 3000 steps of both
forms against the FP64 oracle's series), `sibc_box` (the SIBC + Mur box),
`packed` (16 small requests one at a time and as one packed group) and
`serving` (the solve server on a seeded stream of the contract's shape).
"""
import argparse
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mcells/s FP64 vs FP32 at 256^3 + %HBM peak; p50/p99 sweep-solve latency"
HBM_PEAK_GBS = 8000.0


def _traffic(kernel, prec):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, or None."""
    path = os.path.join(HERE, "profiles", f"traffic_{kernel}_debye256_{prec}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split()[1])
    return None


def debye_site_words(st):
    """Compulsory words per Debye site: lastE read and written, J read and
    written per pole (two words per pole when complex)."""
    return 2 + 2 * max(st["debye_poles"], 1) * (2 if st["debye_complex"] else 1)


def kernel_bytes(st, esz):
    """Compulsory bytes per launch of the two volume kernels of the inner grid.

    Fused sweep (k_leap) and its fix-up (k_h_fix): the fused sweep reads old E
    and H (6 words) and both coefficient classes (2 + 2 B; 9 words of
    alpha/beta/gamma without classes) and writes new E and H (6 words) of
    every fast-box cell; the fix-up reads old H and the cell's new E (6 words)
    and its H class (2 B; 3 words of gamma) and writes new H (3) of each H
    cell updated again after the E corrections.

    Separate sweeps (k_e_fast, k_h_fast): fast E reads E and H (6 words) and
    the class (2 B; 6 words of alpha/beta without classes), writes E (3),
    plus the fused Debye sites; fast H reads H and E (6), the class (2 B; 3
    words of gamma), writes H (3)."""
    ce, ch = st["fast_cells"]
    if st.get("leap"):
        coef, coef_h = (4, 2) if st["indexed_media"] else (9 * esz, 3 * esz)
        return ce * (12 * esz + coef), st["fix_cells"] * (9 * esz + coef_h)
    coef_e, coef_h = (2, 2) if st["indexed_media"] else (6 * esz, 3 * esz)
    e = ce * (9 * esz + coef_e) + st["debye_fused_sites"] * debye_site_words(st) * esz
    h = ch * (9 * esz + coef_h)
    return e, h


def step_bytes(st, esz):
    """Compulsory bytes of one step in the representation the kernels use:
    every field word read and written once (12 words per cell), the
    coefficients once (two 2-byte classes, or 9 words), 9 UPML auxiliaries
    read and written (18 words per UPML cell), and the Debye state."""
    cells, upml = st["cells"], st["cells"] - st["mesh_cells"]
    coef = 4 if st["indexed_media"] else 9 * esz
    return cells * (12 * esz + coef) + upml * 18 * esz + st["debye_sites"] * debye_site_words(st) * esz


def measure(mesh, prec, steps, warmup):
    from vulture_amd import DeviceSolve

    t0 = time.time()
    with DeviceSolve(mesh, precision=prec, steps=warmup + steps, kernel_events=steps) as ds:
        setup_s = time.time() - t0
        st = ds.stats()
        ds.run(0, warmup)
        ds.synchronize()
        ds.run(warmup, steps)
        total_ms = ds.synchronize()
        e_ms, h_ms = ds.kernel_ms()
    esz = 8 if prec == "fp64" else 4
    eb, hb = kernel_bytes(st, esz)
    names = ("k_leap", "k_h_fix") if st.get("leap") else ("k_e_fast", "k_h_fast")
    if e_ms >= h_ms:
        name, kb, kms = names[0], eb, e_ms
    else:
        name, kb, kms = names[1], hb, h_ms
    ms_step = total_ms / steps
    achieved = kb / (kms * 1e-3) / 1e9
    sb = step_bytes(st, esz)
    step_gbs = sb / (ms_step * 1e-3) / 1e9
    traffic = _traffic(name, prec)
    return {
        "prec": prec, "setup_s": round(setup_s, 2), "stats": st,
        "value": st["cells"] * steps / (total_ms * 1e-3) / 1e6,
        "ms_per_step": ms_step,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": kb, "avg_ms": round(kms, 5),
                     "e_ms": round(e_ms, 5), "h_ms": round(h_ms, 5)},
        "step_roofline": {"achieved": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                          "bytes_per_step": sb, "unit": "GB/s"},
    }


def sibc_box(work, steps, warmup=5):
    """The SIBC + Mur box (384x384x256 cubic 1.5 mm, Mur faces, six fitted
    conductor walls, 64 dielectric blocks, current-moment dipole) in FP32:
    throughput and the split between the volume sweeps and the rest."""
    from vulture_amd import DeviceSolve, meshgen
    mesh = meshgen.write("sibc_box", os.path.join(work, "sibc"), steps=warmup + steps)
    t0 = time.time()
    with DeviceSolve(mesh, precision="fp32", steps=warmup + steps, kernel_events=steps) as ds:
        setup_s = time.time() - t0
        st = ds.stats()
        ds.run(0, warmup)
        ds.synchronize()
        ds.run(warmup, steps)
        total_ms = ds.synchronize()
        e_ms, h_ms = ds.kernel_ms()
    return {"workload": "SIBC + Mur box 384x384x256 cubic 1.5 mm, six fitted conductor walls, 64 dielectric "
                        "blocks, current-moment dipole (seed 4)", "dtype": "fp32", "steps": steps,
            "cells": st["cells"], "sibc_faces": st["sibc_faces"],
            "sibc_state": ["complex per filter", "real per filter", "real per input"][st["sibc_state"]],
            "value": round(st["cells"] * steps / (total_ms * 1e-3) / 1e6, 1),
            "ms_per_step": round(total_ms / steps, 4), "fast_e_ms": round(e_ms, 5), "fast_h_ms": round(h_ms, 5),
            "setup_s": round(setup_s, 2)}


def packed(work, count, steps=2000):
    """Small serving requests (32-72 cells per axis before the 8-cell UPML,
    `steps` steps, FP32) run one at a time, in chunks of 200 steps replayed
    from captured graphs, and then all together as one packed group
    (vgpu_run_group, chunks of 200 steps replayed from a captured graph).
    Solo throughput counts the device time of each solve's loop; the group's
    counts the host's wall time of its loop, so the comparison favours solo.
    Every solve's files must be byte-identical in the two runs."""
    import filecmp
    import numpy as np
    import vulture_amd
    from vulture_amd import meshgen
    rng = np.random.default_rng(21)
    d = os.path.join(work, "packed")
    os.makedirs(d, exist_ok=True)
    meshes, cells = [], 0
    for q in range(count):
        text, n, _ = meshgen.serving_request(rng, nmin=32, nmax=72, steps_min=steps, steps_max=steps)
        meshes.append(os.path.join(d, f"r{q}.mesh"))
        with open(meshes[-1], "w") as f:
            f.write(text)
        cells += (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    solo_ms = 0.0
    for q, m in enumerate(meshes):
        solo_ms += vulture_amd.run_mesh(m, os.path.join(d, f"solo{q}"), precision="fp32", chunk=200)
    group_ms = vulture_amd.run_meshes_group(meshes, [os.path.join(d, f"group{q}") for q in range(count)],
                                            precision="fp32", chunk=200)
    same = True
    for q in range(count):
        a, b = os.path.join(d, f"solo{q}"), os.path.join(d, f"group{q}")
        fa = sorted(os.listdir(a))
        same = same and fa == sorted(os.listdir(b)) and all(
            filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False) for f in fa)
    work_cu = cells * steps
    return {"workload": f"{count} serving requests, cells per axis log-uniform [32, 72] + 8-cell UPML, graded "
                        f"lines, SIMPLE/DEBYE blocks, TF/SF, {steps} steps (seed 21)", "dtype": "fp32",
            "requests": count, "mean_cells": round(cells / count), "solo_loop_s": round(solo_ms * 1e-3, 4),
            "group_loop_s": round(group_ms * 1e-3, 4),
            "solo_device_mcells_per_s": round(work_cu / (solo_ms * 1e-3) / 1e6, 1),
            "group_mcells_per_s": round(work_cu / (group_ms * 1e-3) / 1e6, 1),
            "group_over_solo": round(solo_ms / group_ms, 3), "files_identical": same}


def _fixture_dir(npz, dst, wf_src):
    """Observer files under dst from an FP64-oracle fixture (arrays written by
    tests/golden/make_fp64_fixtures.py), beside wf_src's waveform spectrum."""
    import glob
    import shutil
    import numpy as np
    fx = np.load(npz)
    os.makedirs(dst, exist_ok=True)
    for key in fx.files:
        if key.startswith(("td/", "fd/")):
            with open(os.path.join(dst, key[3:]), "w") as f:
                for row in fx[key]:
                    head = f"{int(row[0])} " if key.startswith("td/") else f"{float(row[0])!r} "
                    f.write(head + " ".join(repr(float(x)) for x in row[1:]) + "\n")
    for w in glob.glob(os.path.join(wf_src, "wf_*_fd.asc")):
        shutil.copy(w, dst)
    return str(fx["mesh_sha256"])


def accuracy(work):
    """The Debye-sphere volume for its full 3000 steps in both storage forms:
    the FP32 observers against the FP64 device run (time-series relative L2
    over all samples and components; max dB deviation of the spectra over the
    band where the waveform spectrum lies within 60 dB of its peak, bins
    below 1e-3 of the largest reference magnitude left out (fd_db) and kept
    (fd_db_all)), and both forms against the FP64 oracle's series when its
    fixture (tests/golden/fp64/debye_full.npz) is present."""
    import hashlib
    from vulture_amd import meshgen, run_mesh
    from vulture_amd.obsfiles import compare_outputs
    d = os.path.join(work, "acc")
    mesh = meshgen.write("debye", d)
    outs = {}
    for prec in ("fp64", "fp32"):
        outs[prec] = os.path.join(d, prec)
        run_mesh(mesh, outs[prec], precision=prec)
    keys = ("td", "fd_raw", "fd_db", "fd_db_all")
    r = compare_outputs(outs["fp32"], outs["fp64"])
    res = {"steps": 3000, "fp32_vs_fp64": {k: float(f"{r[k]:.3e}") for k in keys}}
    fx = os.path.join(HERE, "tests", "golden", "fp64", "debye_full.npz")
    if os.path.exists(fx):
        sha = _fixture_dir(fx, os.path.join(d, "oracle"), outs["fp64"])
        if sha == hashlib.sha256(open(mesh, "rb").read()).hexdigest():
            for prec in ("fp64", "fp32"):
                q = compare_outputs(outs[prec], os.path.join(d, "oracle"))
                res[f"{prec}_vs_fp64_oracle"] = {k: float(f"{q[k]:.3e}") for k in keys}
            res["oracle_observers_compared"] = q["files"]
    return res


def host_cpu():
    """The host CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def time_reference(exe, mesh, steps, cells, threads):
    """Loop seconds of a reference build on `steps` steps of `mesh`: wall time
    of the full run minus that of its set-up-only run (-p), on `threads`
    OpenMP threads.  Returns (Mcell-updates/s, full wall, set-up wall)."""
    import re
    import subprocess
    d = os.path.join(os.path.dirname(mesh), "ref_" + os.path.basename(exe))
    os.makedirs(d, exist_ok=True)
    text = re.sub(r"^NT \d+$", f"NT {steps}", open(mesh).read(), flags=re.M)
    rmesh = os.path.join(d, "ref.mesh")
    with open(rmesh, "w") as f:
        f.write(text)
    walls = []
    for opt in (["-p"], []):
        t0 = time.monotonic()
        r = subprocess.run([exe, "-n", str(threads)] + opt + [rmesh], cwd=d, capture_output=True, text=True,
                           timeout=600)
        walls.append(time.monotonic() - t0)
        print(f"# cpu baseline: {os.path.basename(exe)} {' '.join(opt) or 'full run'} {walls[-1]:.1f} s",
              file=sys.stderr, flush=True)
        if r.returncode != 0:
            return None
    return cells * steps / (walls[1] - walls[0]) / 1e6, walls[1], walls[0]


def cpu_baseline(mesh, steps, cells):
    """The reference itself on `steps` steps of the same mesh, on every core
    this process may use (OMP_NUM_THREADS when set: the job's share of the
    host): its float build from its own sources with its own Release flags
    (CMakeLists.txt:64,68: -Ofast -flto -funroll-loops -march=core2) and its
    OpenMP option (oracle/_ref/vulture_ref_rel).  Beside it: the same sources
    at -O3 -march=x86-64-v3 (vulture_ref_omp), giving the calibration ratio
    between the two builds, and the CPU restatement of the loop (cpu_port)."""
    import oracle
    model, avail = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or avail)
    sec, pcells, n = oracle.time_mesh(mesh, steps=steps, warmup=2, threads=threads, precision="fp32",
                                      outdir=os.path.join(os.path.dirname(mesh), "cpu"))
    port = {"value": round(pcells * n / sec / 1e6, 2), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} timed steps (after 2 untimed) of the same 272^3 Debye mesh, FP32 CPU restatement "
                      f"(-O3, OpenMP, {threads} threads): {sec:.2f} s"}
    rel = os.path.join(HERE, "oracle", "_ref", "vulture_ref_rel")
    o3 = os.path.join(HERE, "oracle", "_ref", "vulture_ref_omp")
    if not os.path.exists(rel):
        return port, None
    r = time_reference(rel, mesh, steps, cells, threads)
    if r is None:
        return port, None
    base = {"value": round(r[0], 2), "unit": "Mcell-updates/s", "cores": threads, "kind": "reference",
            "cpu": model, "host_cores": avail,
            "sample": f"the reference's own float build with its Release flags and OpenMP option, {threads} "
                      f"threads, on {steps} steps of the same 272^3 Debye mesh: full run {r[1]:.2f} s minus "
                      f"set-up-only run (-p) {r[2]:.2f} s = {r[1] - r[2]:.2f} s"}
    if os.path.exists(o3):
        q = time_reference(o3, mesh, steps, cells, threads)
        if q is not None:
            base["o3_build"] = round(q[0], 2)
            base["release_over_o3"] = round(r[0] / q[0], 3)
    return base, port


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="mesh cells per axis")
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--cpu-steps", type=int, default=100, help="timed CPU baseline steps (0 skips it)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the full-length FP32 vs FP64 accuracy run")
    ap.add_argument("--spheres", type=int, default=200, help="Debye spheres (workload variants for profiling)")
    ap.add_argument("--boundary", choices=["PML", "PEC", "MUR"], default="PML")
    ap.add_argument("--no-planewave", action="store_true", help="hard point source instead of the TF/SF plane wave")
    ap.add_argument("--serve", type=int, default=32, help="requests in the serving sweep (0 skips it)")
    ap.add_argument("--packed", type=int, default=16,
                    help="small requests run solo and as one packed group (0 skips it; at most 16)")
    ap.add_argument("--serve-active", type=int, default=64,
                    help="solves resident at once when serving (device memory permitting)")
    ap.add_argument("--serve-workers", type=int, default=12, help="host threads setting up and writing solves")
    ap.add_argument("--serve-nmax", type=int, default=320, help="largest cells per axis of a serving request")
    ap.add_argument("--serve-steps", type=int, nargs=2, default=[2000, 8000], help="step range of serving requests")
    ap.add_argument("--sibc-steps", type=int, default=30,
                    help="timed steps of the SIBC + Mur box at 384x384x256 in FP32 (0 skips it)")
    ap.add_argument("--sibc-only", action="store_true",
                    help="run only the SIBC + Mur box leg and print its record (for profiling that workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design, not built)", file=sys.stderr)
        return 2
    from vulture_amd import meshgen
    work = tempfile.mkdtemp(prefix="vgbench_")
    if args.sibc_only:
        print(json.dumps(sibc_box(work, max(args.sibc_steps, 1))), flush=True)
        return 0
    kw = dict(n=args.n, steps=args.warmup + args.steps, spheres=args.spheres, boundary=args.boundary,
              planewave=not args.no_planewave)
    # the device reads the spheres as one voxel object; the reference's CPU
    # build reads the same spheres as one MB card per run (same cells, media
    # and face masks: identical observer files)
    mesh = meshgen.write("debye", os.path.join(work, "vox"), voxel=True, **kw)
    mesh_cards = meshgen.write("debye", os.path.join(work, "cards"), **kw)
    res = {}
    for prec in (["fp64", "fp32"] if args.precision == "both" else [args.precision]):
        res[prec] = measure(mesh, prec, args.steps, args.warmup)
        rf = res[prec]["roofline"]
        print(f"# {prec}: {res[prec]['value']:.1f} Mcell-updates/s, {res[prec]['ms_per_step']:.3f} ms/step, "
              f"setup {res[prec]['setup_s']} s; volume kernels {rf['e_ms']:.4f} / {rf['h_ms']:.4f} ms, "
              f"fix cells {res[prec]['stats'].get('fix_cells', 0)}", file=sys.stderr, flush=True)
    head = res.get("fp32") or res["fp64"]
    st = head["stats"]
    out = {
        "metric": METRIC, "value": round(head["value"], 1), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": head["prec"],
        "data": "synthetic",
        "config": {"workload": f"debye spheres {args.n}^3 cubic 2 mm, 8-cell UPML, 200 single-pole Debye "
                               f"spheres (one voxel object), aux-grid TF/SF plane wave, 8 DFT observers (seed 3)",
                   "cells": st["cells"], "mesh_cells": st["mesh_cells"], "debye_sites": st["debye_sites"],
                   "device_bytes": st["bytes"], "coefficients": "16-bit classes" if st["indexed_media"] else
                   "per-cell arrays", "debye_fused_sites": st["debye_fused_sites"],
                   "sweep": "fused leapfrog" if st.get("leap") else "separate E and H sweeps",
                   "fix_cells": st.get("fix_cells", 0)},
        "roofline": head["roofline"], "step_roofline": head["step_roofline"],
    }
    if "fp64" in res and "fp32" in res:
        f64 = res["fp64"]
        out["fp64"] = {"value": round(f64["value"], 1), "ms_per_step": round(f64["ms_per_step"], 4),
                       "roofline": f64["roofline"], "step_roofline": f64["step_roofline"]}
        out["fp32_over_fp64"] = round(res["fp32"]["value"] / f64["value"], 3)
    if not args.no_accuracy and args.n == 256 and args.spheres == 200 and args.boundary == "PML" and \
            not args.no_planewave:
        out["fp32_error"] = accuracy(work)
        a = out["fp32_error"]["fp32_vs_fp64"]
        print(f"# fp32 vs fp64 over 3000 steps: td {a['td']:.2e}, {a['fd_db']:.2e} dB", file=sys.stderr, flush=True)
    out["cpu_baseline"] = None
    if args.cpu_steps > 0:
        base, port = cpu_baseline(mesh_cards, args.cpu_steps, st["cells"])
        out["cpu_baseline"] = base
        if port is not None:
            out["cpu_port"] = port
    if args.sibc_steps > 0:
        out["sibc_box"] = sibc_box(work, args.sibc_steps)
        print(f"# sibc_box: {out['sibc_box']['value']:.1f} Mcell-updates/s, {out['sibc_box']['ms_per_step']:.3f} "
              f"ms/step", file=sys.stderr, flush=True)
    if args.packed > 0:
        out["packed"] = packed(work, min(args.packed, 16))
        print(f"# packed: {out['packed']['group_mcells_per_s']:.0f} Mcell-updates/s as one group against "
              f"{out['packed']['solo_device_mcells_per_s']:.0f} one at a time, files identical "
              f"{out['packed']['files_identical']}", file=sys.stderr, flush=True)
    if args.serve > 0:
        from vulture_amd import serve
        rep = serve.sweep(os.path.join(work, "serve"), count=args.serve, load=0.7, precision="fp32",
                          max_active=args.serve_active, workers=args.serve_workers, chunk=200, nmin=32,
                          nmax=args.serve_nmax,
                          steps_min=args.serve_steps[0], steps_max=args.serve_steps[1], nfreq=64,
                          log=lambda m: print(f"# serving: {m}", file=sys.stderr, flush=True))
        rep["workload"] = (f"{args.serve} seeded requests: cells per axis log-uniform [32, {args.serve_nmax}], "
                           f"graded lines, 4-24 SIMPLE/DEBYE blocks, 8-cell UPML, TF/SF at random angles, "
                           f"{args.serve_steps[0]}-{args.serve_steps[1]} steps, 64-frequency DFT observers; "
                           f"Poisson arrivals at 70% of isolated capacity; FP32; up to {args.serve_active} solves "
                           f"resident, least work first, the two with least work left running two "
                           f"graph-captured chunks each (200 steps, fewer above 5 M cells: at most 1e9 "
                           f"cell-updates a chunk); a running solve under 1 M cells steps with up to 15 more "
                           f"such solves in one packed group (vgpu_run_group), and no larger solve runs "
                           f"beside such a group")
        out["serving"] = rep
        print(f"# serving: p50 {rep['p50_latency_s']:.3f} s p99 {rep['p99_latency_s']:.3f} s, "
              f"saturated {rep['saturated_mcells_per_s']:.0f} vs isolated {rep['isolated_mcells_per_s']:.0f} "
              f"Mcell-updates/s, p99 latency / isolated {rep['p99_latency_over_isolated']:.2f}", file=sys.stderr,
              flush=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
