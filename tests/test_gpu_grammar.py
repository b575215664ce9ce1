"""Device runs of the generated grammar meshes (BR and BE cards, all 84
source tags; tests/golden/grammar) against the FP64 oracle, which
tests/test_grammar.py pins to the reference builds.  Needs an MI355X.
Bars as for the reference meshes: FP64 within 1e-12 relative L2 (time series
and raw DFTs); FP32 within 1e-4 / 0.05 dB, or 1.5 times the reference's own
float arithmetic's error against FP64 where that is larger.
"""
import os
import shutil

import pytest

import oracle
import vulture_amd
from vulture_amd.obsfiles import compare_outputs

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

G = os.path.join(GOLDEN, "grammar")
NAMES = sorted(os.listdir(G)) if os.path.isdir(G) else []


def _stage(name, dst):
    os.makedirs(dst)
    shutil.copy(os.path.join(G, name, f"{name}.mesh"), dst)
    return os.path.join(dst, f"{name}.mesh")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", NAMES)
def test_grammar_mesh_on_device(name, precision, tmp_path):
    mesh = _stage(name, str(tmp_path / "in"))
    print(f"case {name} {precision}", flush=True)
    g, o = tmp_path / "gpu", tmp_path / "oracle"
    vulture_amd.run_mesh(mesh, str(g), precision=precision)
    oracle.run_mesh(mesh, str(o), precision="fp64")
    r = compare_outputs(str(g), str(o))
    assert r["files"] == 4, r
    if precision == "fp64":
        assert r["td"] <= 1e-12 and r["fd_raw"] <= 1e-12, r
    else:
        oracle.run_mesh(mesh, str(tmp_path / "oracle32"), precision="fp32")
        ref = compare_outputs(str(tmp_path / "oracle32"), str(o))
        assert r["td"] <= max(1e-4, 1.5 * ref["td"]), (r, ref)
        assert r["fd_db"] <= max(0.05, 1.5 * ref["fd_db"]), (r, ref)
