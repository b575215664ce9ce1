"""The entry point that allocates nothing (needs an MI355X).

vgpu_check_limits_noalloc runs the update-once check from the library's fixed
per-device scratch area instead of allocating.  Through the C-ABI it must give
exactly what vgpu_check_limits gives (the report and every per-cell count),
on two calls in a row whose second grid is smaller than the first (so the
second call sees the first call's leftovers in the shared scratch), and it
must refuse a grid whose check needs more scratch than it holds.
"""
import os
import shutil

import numpy as np
import pytest

import vulture_amd
from vulture_amd import meshgen

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _stage(name, dst):
    shutil.copytree(os.path.join(GOLDEN, "ref_inputs", name), dst)
    return str(dst / [f for f in os.listdir(dst) if f.endswith(".mesh")][0])


def _both(mesh, precision):
    print(f"case {os.path.basename(mesh)} {precision}", flush=True)
    with vulture_amd.DeviceSolve(mesh, precision=precision, steps=1) as ds:
        rep_a, cnt_a = ds.check_limits()
        rep_b, cnt_b = ds.check_limits(noalloc=True)
    return rep_a, cnt_a, rep_b, cnt_b


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_check_matches_allocating_check(precision, tmp_path):
    # a larger grid with un-updated Mur edges, then a smaller one
    big = meshgen.write("sibc_box", str(tmp_path / "big"), nx=40, ny=36, nz=32, blocks=4, steps=1, observers=1,
                        nfreq=4)
    small = _stage("freespace_mur", tmp_path / "small")
    shapes = []
    for mesh in (big, small):
        rep_a, cnt_a, rep_b, cnt_b = _both(mesh, precision)
        shapes.append(cnt_a.size)
        assert rep_a == rep_b
        assert np.array_equal(cnt_a, cnt_b)
        assert rep_a["repeated"] == [0] * 6 and rep_a["outside"] == [0] * 6
        assert np.count_nonzero(cnt_b) > 1000
    assert shapes[1] < shapes[0]


def test_noalloc_refuses_grid_beyond_its_scratch(tmp_path):
    from vulture_amd import VgpuError, load_vgpu
    limit = load_vgpu().vgpu_noalloc_scratch_bytes()
    assert limit == 64 << 20
    # 154^3 allocation (160 padded): 4 bytes per component and cell of visit counts alone exceed 64 MiB
    mesh = meshgen.write("freespace", str(tmp_path / "large"), n=136, steps=1)
    print("case noalloc refusal", flush=True)
    with vulture_amd.DeviceSolve(mesh, precision="fp32", steps=1) as ds:
        with pytest.raises(VgpuError, match="needs .* bytes of scratch"):
            ds.check_limits(noalloc=True)
        # the allocating check still runs on it
        rep = ds.check_limits(counts=False)
    assert rep["repeated"] == [0] * 6
