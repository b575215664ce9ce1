"""The sweep planning on the CPU: tests/native/shell_check.cpp runs the host
set-up of each mesh against the stand-in device library and checks the
library's own planning code (fast boxes, shell pieces, the XCD launch order
of the shell blocks) cell by cell against the per-cell region classification
the sweep kernels apply (grid.c:709-888 ordinary updates, pml.c:859-1176
UPML slabs, mur.c normal-H boxes)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from vulture_amd import meshgen

HOST = os.path.join(ROOT, "vulture_amd", "host")
NATIVE = os.path.join(ROOT, "tests", "native")
SRC = ["vh_parse.c", "vh_setup.c", "vh_voxel.c", "vh_observe.c", "vh_gpu.c", "vh_serve.c"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None or not os.path.exists(HIPCC):
        pytest.skip("gcc or hipcc not available")
    d = tmp_path_factory.mktemp("shell")
    objs = []
    for f in SRC + ["vgpu_stub.c"]:
        src = os.path.join(NATIVE if f == "vgpu_stub.c" else HOST, f)
        o = str(d / f.replace(".c", ".o"))
        r = subprocess.run(["gcc", "-std=gnu99", "-O2", "-DVH_REAL=double", "-I" + HOST, "-w", "-c", src, "-o", o],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        objs.append(o)
    exe = str(d / "shell_check")
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "vulture_amd", "csrc"), "-w", "-x", "hip",
           os.path.join(NATIVE, "shell_check.cpp"), "-x", "none"] + objs + ["-lm", "-lpthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(checker, meshes, tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([checker, str(out)] + meshes, capture_output=True, text=True, timeout=900, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    assert len(lines) == len(meshes) and all(" OK " in ln for ln in lines), r.stdout[-3000:]
    return lines


def test_reference_meshes(checker, tmp_path):
    src = os.path.join(GOLDEN, "ref_inputs")
    meshes = []
    for name in sorted(os.listdir(src)):
        if name == "parser":  # deliberately malformed meshes
            continue
        shutil.copytree(os.path.join(src, name), tmp_path / name)
        meshes += [str(tmp_path / name / f) for f in sorted(os.listdir(tmp_path / name)) if f.endswith(".mesh")]
    assert len(meshes) > 90
    _run(checker, meshes, tmp_path)


def test_generated_workloads(checker, tmp_path):
    meshes = [meshgen.write("debye", str(tmp_path / "d96"), n=96, spheres=6, rmin=3, rmax=8, steps=10),
              meshgen.write("debye", str(tmp_path / "d48m"), n=48, spheres=3, rmin=3, rmax=6, steps=10,
                            boundary="MUR"),
              meshgen.write("debye", str(tmp_path / "d48p"), n=48, spheres=3, rmin=3, rmax=6, steps=10,
                            boundary="PEC"),
              meshgen.write("enclosure", str(tmp_path / "enc"), nx=48, ny=40, nz=32, pml=6, steps=10, nfreq=4),
              meshgen.write("sibc_box", str(tmp_path / "sb"), nx=40, ny=36, nz=32, blocks=6, steps=10,
                            observers=2, nfreq=4)]
    rng = np.random.default_rng(3)
    for q in range(6):
        text, _, _ = meshgen.serving_request(rng, nmin=12, nmax=40, steps_min=5, steps_max=10, nfreq=2)
        p = tmp_path / f"s{q}.mesh"
        p.write_text(text)
        meshes.append(str(p))
    lines = _run(checker, meshes, tmp_path)
    # a UPML volume cuts into a few dozen pieces, not one per cell row
    assert all(int(ln.split()[6]) < 100 for ln in lines), lines
