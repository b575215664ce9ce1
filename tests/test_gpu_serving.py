"""The solve server on the device: continuous batching gives every request
the same observer files it gets alone, and those match the oracle."""
import os

import pytest

import oracle
from vulture_amd import serve
from vulture_amd.obsfiles import compare_outputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sweep_alone_equals_mix(tmp_path, precision):
    print(f"case serving sweep {precision}", flush=True)
    rep = serve.sweep(str(tmp_path), count=8, load=0.7, precision=precision, max_active=4, chunk=40,
                      nmin=16, nmax=40, steps_min=60, steps_max=160, nfreq=16)
    assert rep["alone_equals_mix"], rep["mismatched"]
    assert rep["p99_latency_s"] > 0 and rep["saturated_mcells_per_s"] > 0
    if precision == "fp64":
        mesh = os.path.join(str(tmp_path), "meshes", "req0000.mesh")
        odir = str(tmp_path / "oracle0")
        oracle.run_mesh(mesh, odir, precision="fp64")
        r = compare_outputs(os.path.join(str(tmp_path), "mix", "0000"), odir)
        assert r["files"] > 0 and r["fd_raw"] <= 1e-12, r
