"""Group runs (vgpu_run_group) against solo runs (needs an MI355X).

A group run advances several solves in lockstep with one packed launch per
step kernel over all of them.  Every block of a packed launch does what the
same block does in the solve's own launch, so each solve's observer files
must be byte-identical to those of the mesh run alone (vulture_amd.run_mesh).

Cases: 16 of the reference's meshes that between them hold every boundary,
medium and source kind (PEC, PMC and periodic ghosts, Mur, UPML, Debye, SIBC,
TF/SF with the aux grid and the analytic incident field, hard, soft, lumped
and surface sources, impulse.dat), in both precisions, in one group; and
eight serving requests (graded lines, SIMPLE and Debye blocks, TF/SF) of
different sizes and step counts in FP32, run in chunks of 100 steps, so the
group shrinks as its shorter solves finish.  Chunked runs repeat a group,
which then replays from a captured graph.
"""
import filecmp
import os
import shutil

import numpy as np
import pytest

import vulture_amd
from vulture_amd import meshgen

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

INP = os.path.join(GOLDEN, "ref_inputs")
MIXED = ["closedbox_pec", "closedbox_spar", "dgref_mur", "dgrefmb_pml6", "dipole_vry", "freespace_pml",
         "hardsrc_jmsxy", "parplate1D_debye_xyz", "parplate1D_debyepml_zxy", "parplate1D_planewave_yzx",
         "parplate1D_sibcgo04_xyz", "parplate1D_sibcspar_zyx", "planewave_pec_xyz", "planewave_empty_zyx",
         "softsrc_imdz", "waveform_ext"]


def _stage(name, dst):
    shutil.copytree(os.path.join(INP, name), dst)
    return str(dst / [f for f in os.listdir(dst) if f.endswith(".mesh")][0])


def _same_files(a, b):
    fa, fb = sorted(os.listdir(a)), sorted(os.listdir(b))
    assert fa == fb and fa, (fa, fb)
    return [f for f in fa if not filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False)]


def _solo_and_group(meshes, tmp_path, precision, chunk=0):
    solo = [str(tmp_path / f"solo{i}") for i in range(len(meshes))]
    grp = [str(tmp_path / f"group{i}") for i in range(len(meshes))]
    for m, o in zip(meshes, solo):
        print(f"case solo {os.path.basename(m)} {precision}", flush=True)
        vulture_amd.run_mesh(m, o, precision=precision, chunk=chunk)
    print(f"case group of {len(meshes)} {precision}", flush=True)
    vulture_amd.run_meshes_group(meshes, grp, precision=precision, chunk=chunk)
    return {os.path.basename(m): _same_files(s, g) for m, s, g in zip(meshes, solo, grp)}


@pytest.mark.parametrize("precision,chunk", [("fp64", 0), ("fp32", 0), ("fp64", 50), ("fp32", 50)])
def test_reference_meshes_grouped_equal_solo(precision, chunk, tmp_path):
    """chunk 0: one group run per membership; chunk 50: repeated runs of the
    same group, captured into a graph at the second and replayed after."""
    meshes = [_stage(n, tmp_path / "in" / n) for n in MIXED]
    differ = _solo_and_group(meshes, tmp_path, precision, chunk=chunk)
    assert not any(differ.values()), {k: v for k, v in differ.items() if v}


def test_serving_requests_grouped_equal_solo(tmp_path):
    rng = np.random.default_rng(11)
    meshes = []
    for q in range(8):
        text, n, steps = meshgen.serving_request(rng, nmin=16, nmax=56, steps_min=150, steps_max=650)
        p = tmp_path / "in" / f"req{q}.mesh"
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(text)
        meshes.append(str(p))
    differ = _solo_and_group(meshes, tmp_path, "fp32", chunk=100)
    assert not any(differ.values()), {k: v for k, v in differ.items() if v}


def test_group_limits():
    with pytest.raises(ValueError):
        vulture_amd.run_meshes_group([], [])
    with pytest.raises(ValueError):
        vulture_amd.run_meshes_group(["a"] * (vulture_amd.GROUP_MAX + 1), ["b"] * (vulture_amd.GROUP_MAX + 1))


def test_fused_sweep_solves_are_not_grouped(tmp_path, monkeypatch):
    """Contexts with the fused leapfrog sweep (VGPU_LEAP=1) have no packed
    form: a group run of them is refused with a message, and the solve server
    runs such small solves one by one, with the files of each run alone."""
    from vulture_amd import serve
    monkeypatch.setenv("VGPU_LEAP", "1")
    meshes = [_stage(n, tmp_path / "in" / n) for n in ("freespace_pml", "dipole_vry", "softsrc_imdz")]
    print("case group of fused-sweep contexts", flush=True)
    with pytest.raises(vulture_amd.VgpuError, match="separate sweeps"):
        vulture_amd.run_meshes_group(meshes, [str(tmp_path / f"g{i}") for i in range(3)], precision="fp32")
    solo = [str(tmp_path / f"solo{i}") for i in range(3)]
    for m, o in zip(meshes, solo):
        vulture_amd.run_mesh(m, o, precision="fp32", chunk=100)
    print("case server with fused-sweep contexts", flush=True)
    with serve.Server(precision="fp32", max_active=4, workers=4, chunk=100) as sv:
        rids = [sv.submit(m, str(tmp_path / f"srv{i}")) for i, m in enumerate(meshes)]
        infos = [sv.wait(r) for r in rids]
    assert all(i["status"] == 1 for i in infos), infos
    for i in range(3):
        assert not _same_files(solo[i], str(tmp_path / f"srv{i}")), i
