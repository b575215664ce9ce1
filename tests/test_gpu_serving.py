"""The solve server on the device: continuous batching gives every request
the same observer files it gets alone, and those match the oracle."""
import os

import pytest

import oracle
from vulture_amd import serve
from vulture_amd.obsfiles import compare_outputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graphs", ["0", "1"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sweep_alone_equals_mix(tmp_path, precision, graphs, monkeypatch):
    """With graphs (the default) every solve replays its runs from a
    captured HIP graph while the server's other threads set up, read back
    and free other solves; VGPU_GRAPHS=0 launches every run directly."""
    monkeypatch.setenv("VGPU_GRAPHS", graphs)
    print(f"case serving sweep {precision} graphs {graphs}", flush=True)
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


@pytest.mark.parametrize("name,chunk", [("parplate1D_planewave_xyz", 500), ("closedbox_go04", 100),
                                        ("parplate1D_debyepml_xyz", 700), ("dgref_mur", 90)])
def test_graph_replayed_runs(name, chunk, tmp_path, monkeypatch):
    """Runs replayed from a captured graph give the oracle's results in FP64,
    and the replays do happen."""
    import shutil
    import vulture_amd
    from conftest import GOLDEN
    monkeypatch.setenv("VGPU_GRAPHS", "1")
    d = tmp_path / "in"
    shutil.copytree(os.path.join(GOLDEN, "ref_inputs", name), d)
    mesh = str(d / [f for f in os.listdir(d) if f.endswith(".mesh")][0])
    print(f"case graphs {name} chunk {chunk}", flush=True)
    with vulture_amd.DeviceSolve(mesh, precision="fp64", chunk=chunk) as ds:
        total = ds.steps
        for first in range(0, total, chunk):
            ds.run(first, min(chunk, total - first))
        ds.synchronize()
        replays = ds.stats()["graph_runs"]
    assert replays >= total // chunk - 1
    g, o = tmp_path / "gpu", tmp_path / "oracle"
    vulture_amd.run_mesh(mesh, str(g), precision="fp64", chunk=chunk)
    oracle.run_mesh(mesh, str(o), precision="fp64")
    r = compare_outputs(str(g), str(o))
    assert r["files"] > 0 and r["td"] <= 1e-12 and r["fd_raw"] <= 1e-12, r
