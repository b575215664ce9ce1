"""The rocprof summary helpers (vulture_amd/profiling.py) on synthetic counter
files: a kernel's dispatches are told apart from kernels whose names start
with the same letters, and the traffic per launch is each counter scaled by
what it reads for a copy of known size."""
import csv
import subprocess
import sys

from conftest import ROOT
from vulture_amd import profiling


def test_kernel_names_match_exactly():
    k = "void vgpu::k_corr<float>(vgpu::CorrArgs<float>, double)"
    assert profiling._is_kernel(k, "k_corr")
    assert not profiling._is_kernel("void vgpu::k_corr_gather<float>(long long)", "k_corr")
    assert not profiling._is_kernel("void vgpu::k_corr_aux<float>(x)", "k_corr")
    assert profiling._is_kernel("void vgpu::k_corr_aux<float>(x)", "k_corr_aux")
    assert profiling._is_kernel("vgpu::k_advance_step(unsigned long long*)", "k_advance_step")
    assert not profiling._is_kernel("void vgpu::k_debye_gather<double>(x)", "k_debye")


def _pmc(d, rows):
    d.mkdir()
    with open(d / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], r)))


def test_traffic_scales_by_the_copy(tmp_path):
    k, g = "void vgpu::k_corr<float>(a)", "void vgpu::k_corr_gather<float>(b)"
    copy = "void at::native::vectorized_elementwise_kernel<4>(c)"
    # two k_corr dispatches (per-dispatch values summed over XCD rows), one gather that must not count
    _pmc(tmp_path / "f", [(1, k, "FETCH_SIZE", 100), (1, k, "FETCH_SIZE", 100), (2, k, "FETCH_SIZE", 400),
                          (3, g, "FETCH_SIZE", 9999)])
    _pmc(tmp_path / "w", [(1, k, "WRITE_SIZE", 50), (2, k, "WRITE_SIZE", 150), (3, g, "WRITE_SIZE", 9999)])
    _pmc(tmp_path / "cf", [(1, copy, "FETCH_SIZE", 512)])     # 1 MiB copy read reported as 512 KiB
    _pmc(tmp_path / "cw", [(1, copy, "WRITE_SIZE", 1024)])    # and written as 1024 KiB
    r = subprocess.run([sys.executable, "-m", "vulture_amd.profiling", "traffic", "--kernel", "k_corr",
                        "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"),
                        "--copy-fetch", str(tmp_path / "cf"), "--copy-write", str(tmp_path / "cw"), "--mib", "1"],
                       capture_output=True, text=True, cwd=ROOT, check=True)
    out = dict(line.split() for line in r.stdout.splitlines() if len(line.split()) == 2)
    assert out["dispatches_fetch"] == "2" and out["dispatches_write"] == "2"
    assert float(out["scale_fetch"]) == 2.0 and float(out["scale_write"]) == 1.0
    # fetch: mean 300 KiB x 2; write: mean 100 KiB x 1
    assert int(out["traffic_bytes_per_launch"]) == 300 * 1024 * 2 + 100 * 1024
