"""Update-once check on the device (needs an MI355X): vgpu_check_limits runs
one step of the field sweeps in their counting form and tallies the counts
against the outer field limits.  Against the reference's own limit-checking
build (tests/golden/limits_unvisited.json.gz, one-step runs):
  - no cell is updated twice and no cell outside the limits is updated;
  - the cells left un-updated are exactly the ones the reference leaves;
in both precisions, with the fast boxes exact or cut to whole lines, with
per-cell coefficient arrays, with the fused leapfrog sweep, and at the full
256^3 Debye size.  The check changes no field."""
import os
import shutil

import numpy as np
import pytest

import vulture_amd
from vulture_amd import meshgen

from conftest import GOLDEN
from test_limits import FIELDS, load_fixture

pytestmark = pytest.mark.gpu

FX = load_fixture() if os.path.exists(os.path.join(GOLDEN, "limits_unvisited.json.gz")) else {}
KEYS = sorted(FX)


def _stage(key, tmp_path):
    rec = FX[key]
    if "generator" in rec:
        return meshgen.write(rec["generator"], str(tmp_path / "gen"), **rec["kwargs"])
    d = tmp_path / "in"
    shutil.copytree(os.path.join(GOLDEN, "ref_inputs", os.path.dirname(key)), d)
    return str(d / os.path.basename(key))


def _check(mesh, rec, precision):
    print(f"case {os.path.basename(mesh)} {precision}", flush=True)
    with vulture_amd.DeviceSolve(mesh, precision=precision, steps=1) as ds:
        rep, counts = ds.check_limits()
        lim = ds.field_limits()
    assert rep["repeated"] == [0] * 6, rep
    assert rep["outside"] == [0] * 6, rep
    for c, f in enumerate(FIELDS):
        b = lim[c]
        box = counts[c, b[0]:b[1] + 1, b[2]:b[3] + 1, b[4]:b[5] + 1]
        got = {(int(i) + b[0], int(j) + b[2], int(k) + b[4]) for i, j, k in np.argwhere(box == 0)}
        exp = {tuple(x) for x in rec["unvisited"][f]}
        assert got == exp, (f, len(got), len(exp), sorted(got ^ exp)[:5])
        assert rep["missed"][c] == len(exp)
        assert (rep["first"][c] >= 0) == bool(exp)


@pytest.mark.parametrize("key", KEYS)
def test_update_once_fp64(key, tmp_path):
    _check(_stage(key, tmp_path), FX[key], "fp64")


@pytest.mark.parametrize("key", [k for k in KEYS if k.startswith("gen_") or "mur" in k or "pml" in k
                                 or "sibc" in k or "closedbox" in k])
def test_update_once_fp32(key, tmp_path):
    _check(_stage(key, tmp_path), FX[key], "fp32")


@pytest.mark.parametrize("env", [("VGPU_ALIGN_FAST", "0"), ("VGPU_NO_INDEXED_MEDIA", "1"), ("VGPU_LEAP", "1")])
@pytest.mark.parametrize("key", ["gen_debye_pml", "gen_sibc_box", "gen_enclosure",
                                 "freespace_mur/freespace_mur.mesh", "parplate1D_debyepml_zyx/parplate1D_debyepml_zyx.mesh"])
def test_update_once_variants(key, env, tmp_path, monkeypatch):
    monkeypatch.setenv(*env)
    _check(_stage(key, tmp_path), FX[key], "fp32")


@pytest.mark.parametrize("leap", ["0", "1"])
def test_full_size_update_once_and_fields_unchanged(leap, tmp_path, monkeypatch):
    """256^3 Debye volume with UPML (no Mur faces): every cell of the outer
    limits updated exactly once, in both precisions, with the separate and
    the fused sweeps; the fields after a check equal the fields before it."""
    monkeypatch.setenv("VGPU_LEAP", leap)
    mesh = tmp_path / "d.mesh"
    mesh.write_text(meshgen.debye(steps=20))
    for prec in ("fp32", "fp64"):
        print(f"case debye256 {prec} leap {leap}", flush=True)
        with vulture_amd.DeviceSolve(str(mesh), precision=prec) as ds:
            ds.run(0, 10)
            ds.synchronize()
            before = [ds.download(c) for c in (0, 4)]
            rep = ds.check_limits(counts=False)
            after = [ds.download(c) for c in (0, 4)]
        assert rep["missed"] == [0] * 6 and rep["repeated"] == [0] * 6 and rep["outside"] == [0] * 6, rep
        assert rep["first"] == [-1] * 6
        for a, b in zip(before, after):
            assert np.count_nonzero(a) > 1000
            assert np.array_equal(a, b)


@pytest.mark.parametrize("key,rc", [("freespace_mur/freespace_mur.mesh", 6), ("freespace_pml/freespace_pml.mesh", 0)])
def test_command_line_check(key, rc, tmp_path):
    """vgvulture -L: the command-line form of the check warns for exactly the
    reference's un-updated cells and exits 6 when there are any, as the
    reference's limit-checking build does."""
    import re
    import subprocess
    from conftest import ROOT
    mesh = _stage(key, tmp_path)
    print(f"case vgvulture -L {key}", flush=True)
    r = subprocess.run([os.path.join(ROOT, "vulture_amd", "bin", "vgvulture"), "-L", mesh], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == rc, r.stdout[-2000:] + r.stderr[-2000:]
    got = {(m.group(1), int(m.group(2)), int(m.group(3)), int(m.group(4)))
           for m in re.finditer(r"Warning: (\w+)\[(\d+)\]\[(\d+)\]\[(\d+)\] not updated", r.stdout)}
    exp = {(f, *c) for f in FIELDS for c in FX[key]["unvisited"][f]}
    assert got == exp
    assert "repeated 0 outside 0" in r.stdout
