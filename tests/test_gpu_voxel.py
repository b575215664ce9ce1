"""Device runs of the voxel object loader at full size (needs an MI355X).

The Debye-sphere volume at 256^3 (200 spheres, 81,972 k-runs) written as one
voxel object must give the device the same observer files, byte for byte, as
the same spheres written as one MB card per run: the loader yields the same
cells, media and face masks, and the device's arithmetic does not depend on
the order in which the Debye blocks are listed.  Both precisions, 200 steps.
"""
import filecmp
import os

import pytest

import vulture_amd
from vulture_amd import meshgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_full_size_sphere_voxels_equal_mb_cards_on_device(precision, tmp_path):
    outs = []
    for v in (False, True):
        d = tmp_path / f"v{int(v)}"
        print(f"case debye 256^3 voxel={v} {precision}", flush=True)
        mesh = meshgen.write("debye", str(d), steps=200, voxel=v)
        outs.append(d / "out")
        vulture_amd.run_mesh(mesh, str(outs[-1]), precision=precision)
    fa = sorted(f for f in os.listdir(outs[0]) if f.endswith(".asc"))
    assert fa == sorted(f for f in os.listdir(outs[1]) if f.endswith(".asc")) and len(fa) >= 16
    differ = [f for f in fa if not filecmp.cmp(outs[0] / f, outs[1] / f, shallow=False)]
    assert not differ, differ


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rotated_voxel_object_against_oracle(precision, tmp_path):
    """A voxel object with SIMPLE and Debye keys, turned by theta 30, phi 40,
    on a 24^3 mesh: the device against the FP64 oracle (1e-12 in FP64;
    1e-4 / 0.05 dB in FP32)."""
    import numpy as np
    import oracle
    from vulture_amd.obsfiles import compare_outputs
    d = tmp_path / "in"
    d.mkdir()
    A = np.random.default_rng(7).integers(0, 4, size=(7, 6, 5)).astype(np.uint8)
    A.tofile(d / "obj.keys")
    (d / "obj.vox").write_text('"obj.keys"\n1\n7 6 5\n1e-3 1e-3 1e-3\n3.5e-3 3e-3 2.5e-3\n'
                               "1 2.5 0.01 1.0\n2 dd\n3 7.0 0.0 1.0\n")
    lines = ["VM 1.0.0", "CE rotated voxel object", "DM 24 24 24", "GS"]
    lines += [f"BT {f} PML 4" for f in ("XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI")]
    lines += ["MT dd DEBYE 4.0 1.0e-2 1.0 1.0e+11 -5.0e+10", 'MT obj VOXEL "obj.vox"', "MB 12 12 12 12 12 12 obj 30 40",
              meshgen._gauss_card("wf1", 60e9), "EX 6 6 7 7 6 7 src EZ wf1 1.0",
              "OP 12 12 12 12 12 12 o1 TDOM_ASCII", "OP 15 15 8 8 13 13 o2 TDOM_ASCII",
              "OP 9 9 14 14 10 10 o3 FDOM_ASCII", "GE", "NT 300", "OF 1e9 6e10 16", "MS 1e-3", "EN"]
    (d / "t.mesh").write_text("\n".join(lines) + "\n")
    print(f"case rotated voxel object {precision}", flush=True)
    g, o = tmp_path / "gpu", tmp_path / "oracle"
    vulture_amd.run_mesh(str(d / "t.mesh"), str(g), precision=precision)
    oracle.run_mesh(str(d / "t.mesh"), str(o), precision="fp64")
    r = compare_outputs(str(g), str(o))
    assert r["files"] == 3, r
    if precision == "fp64":
        assert r["td"] <= 1e-12 and r["fd_raw"] <= 1e-12, r
    else:
        assert r["td"] <= 1e-4 and r["fd_db"] <= 0.05, r
