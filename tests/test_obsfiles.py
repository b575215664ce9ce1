"""Observer-file comparison (vulture_amd/obsfiles.py) on hand-made files.

The statistics behind every parity bar must fail on a broken run: a file
holding NaN or Inf, a missing file and a misshapen file all give infinity.
"""
import math
import os

import numpy as np

from vulture_amd.obsfiles import compare_outputs


def _write_td(path, vals):
    with open(path, "w") as f:
        f.write("# time-domain observer\n")
        for n, row in enumerate(vals):
            f.write(f"{n:8d} {n * 1e-11:16.8e} " + " ".join(f"{v:16.8e}" for v in row) + "\n")


def _write_fd(path, freqs, vals):
    with open(path, "w") as f:
        f.write("# frequency-domain observer\n")
        for fr, row in zip(freqs, vals):
            f.write(f"{fr:16.8e} " + " ".join(f"{v.real:16.8e} {v.imag:16.8e}" for v in row) + "\n")


def _pair(tmp_path, td_test, fd_test=None):
    ref, test = tmp_path / "ref", tmp_path / "test"
    ref.mkdir()
    test.mkdir()
    rng = np.random.default_rng(7)
    td = rng.standard_normal((50, 6))
    _write_td(ref / "eh_o1_td.asc", td)
    _write_td(test / "eh_o1_td.asc", td_test(td))
    freqs = np.linspace(1e8, 1e9, 16)
    fd = rng.standard_normal((16, 6)) + 1j * rng.standard_normal((16, 6)) + 3.0
    _write_fd(ref / "eh_o1_fd.asc", freqs, fd)
    _write_fd(test / "eh_o1_fd.asc", freqs, fd_test(fd) if fd_test else fd)
    return str(test), str(ref)


def test_identical_files_compare_to_zero(tmp_path):
    r = compare_outputs(*_pair(tmp_path, lambda v: v))
    assert r["files"] == 2
    assert r["td"] == 0.0 and r["fd_raw"] == 0.0 and r["fd_db"] == 0.0 and r["fd_db_all"] == 0.0


def test_small_error_is_measured(tmp_path):
    r = compare_outputs(*_pair(tmp_path, lambda v: v * (1 + 1e-6), lambda v: v * 1.01))
    assert 5e-7 < r["td"] < 2e-6
    assert abs(r["fd_db"] - 20 * math.log10(1.01)) < 1e-4
    assert abs(r["fd_db_all"] - 20 * math.log10(1.01)) < 1e-4


def test_nan_in_time_series_fails(tmp_path):
    def bad(v):
        w = v.copy()
        w[10, 2] = np.nan
        return w
    r = compare_outputs(*_pair(tmp_path, bad))
    assert math.isinf(r["td"])
    assert not r["td"] <= 1e-4


def test_all_nan_time_series_fails(tmp_path):
    r = compare_outputs(*_pair(tmp_path, lambda v: np.full_like(v, np.nan)))
    assert math.isinf(r["td"])


def test_inf_in_time_series_fails(tmp_path):
    def bad(v):
        w = v.copy()
        w[3, 0] = np.inf
        return w
    assert math.isinf(compare_outputs(*_pair(tmp_path, bad))["td"])


def test_nan_in_spectrum_fails(tmp_path):
    def bad(v):
        w = v.copy()
        w[4, 1] = complex(np.nan, 0.0)
        return w
    r = compare_outputs(*_pair(tmp_path, lambda v: v, bad))
    assert math.isinf(r["fd_raw"]) and math.isinf(r["fd_db"]) and math.isinf(r["fd_db_all"])


def test_missing_file_fails(tmp_path):
    test, ref = _pair(tmp_path, lambda v: v)
    os.remove(os.path.join(test, "eh_o1_td.asc"))
    assert math.isinf(compare_outputs(test, ref)["td"])


def test_misshapen_file_fails(tmp_path):
    assert math.isinf(compare_outputs(*_pair(tmp_path, lambda v: v[:-1]))["td"])
