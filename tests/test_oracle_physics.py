"""Closed forms and invariants the FP64 oracle must satisfy.

These complement byte identity with the reference build: they would catch a
wrong sign, offset or face in the TF/SF corrections, a wrong time label or
sign in the running DFT, or a scale error in the waveform, independently of
the reference.
"""
import math
import os

import numpy as np
import pytest

import oracle

PML = "\n".join(f"BT {f} PML 8" for f in ("XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI"))


def _mesh(tmp_path, body, steps, extra=""):
    text = f"VM 1.0.0\nCE test\nDM 30 30 30\nGS\n{PML}\n{body}\nGE\nNT {steps}\n{extra}MS 0.01\nEN\n"
    p = tmp_path / "m.mesh"
    p.write_text(text)
    return str(p)


def _td(d, name):
    return np.loadtxt(os.path.join(d, f"eh_{name}_td.asc"))


@pytest.mark.parametrize("angles,comp", [("90.0 0.0 90.0", 2), ("90.0 90.0 90.0", 2), ("0.0 0.0 90.0", 0)])
def test_tfsf_normal_incidence_has_no_scattered_field(tmp_path, angles, comp):
    """planewave.c:576-1007 with the 1-D auxiliary grid (cubic mesh): at normal
    incidence the auxiliary grid reproduces the 3-D grid's propagation exactly,
    so outside the TF/SF box the field is zero to rounding, and inside it the
    incident pulse arrives with unit peak amplitude in the polarised component."""
    body = (f"WF wf1 GAUSSIAN_PULSE 1.0\nPW 8 22 8 22 8 22 pw1 wf1 {angles} 111111\n"
            "OP 15 15 15 15 15 15 in TDOM_ASCII\nOP 4 4 15 15 15 15 ox TDOM_ASCII\n"
            "OP 15 15 4 4 15 15 oy TDOM_ASCII\nOP 15 15 15 15 26 26 oz TDOM_ASCII")
    mesh = _mesh(tmp_path, body, 300)
    oracle.run_mesh(mesh, str(tmp_path), precision="fp64")
    inside = _td(str(tmp_path), "in")[:, 2:]
    peak = np.abs(inside).max()
    assert abs(np.abs(inside[:, comp]).max() - 1.0) < 2e-3
    for o in ("ox", "oy", "oz"):
        assert np.abs(_td(str(tmp_path), o)[:, 2:]).max() < 1e-12 * peak, o


def test_waveform_dft_matches_the_gaussian_transform(tmp_path):
    """observer.c:857-899 accumulates X(f) = sum_n g(n dt) exp(-i 2 pi f n dt);
    for a Gaussian g(t) = exp(-((t - t0)/w)^2 / 2) sampled finely and fully
    inside the window, X(f) = (w sqrt(2 pi) / dt) exp(-(2 pi f w)^2 / 2)
    exp(-i 2 pi f t0) to within the truncation of the tails."""
    w, t0 = 2.0e-10, 1.6e-9
    body = (f"WF wf1 GAUSSIAN_PULSE 1.0 {t0} {w}\nEX 15 15 15 15 15 16 src =EZ wf1 1.0\n"
            "OP 16 16 15 15 15 15 op FDOM_ASCII")
    mesh = _mesh(tmp_path, body, 1200, extra="OF 1e8 2e9 40\n")
    oracle.run_mesh(mesh, str(tmp_path), precision="fp64")
    d = np.loadtxt(os.path.join(str(tmp_path), "wf_wf1_fd.asc"))
    f, x = d[:, 0], d[:, 1] + 1j * d[:, 2]
    dt = float(np.loadtxt(os.path.join(str(tmp_path), "wf_wf1_td.asc"))[1, 1])
    ref = (w * math.sqrt(2 * math.pi) / dt) * np.exp(-0.5 * (2 * math.pi * f * w) ** 2) * np.exp(-2j * math.pi * f * t0)
    assert np.abs(x - ref).max() < 1e-6 * np.abs(ref).max()


def test_plane_wave_amplitude_scales_exactly(tmp_path):
    """Every update is linear in the fields, and doubling is exact in binary
    floating point: doubling the plane-wave size doubles every sample."""
    out = []
    for size in ("1.0", "2.0"):
        d = tmp_path / size
        d.mkdir()
        body = (f"WF wf1 GAUSSIAN_PULSE 1.0\nPW 8 22 8 22 8 22 pw1 wf1 63.0 27.0 41.0 111111 {size}\n"
                "MT d DEBYE 3.0 0.01 1.0 1e9 -1e9\nMB 12 18 12 18 12 18 d\n"
                "OP 15 15 15 15 15 15 in TDOM_ASCII\nOP 4 4 15 15 15 15 ox TDOM_ASCII")
        mesh = _mesh(d, body, 150)
        oracle.run_mesh(mesh, str(d), precision="fp64")
        out.append([_td(str(d), n)[:, 2:] for n in ("in", "ox")])
    for a, b in zip(*out):
        assert np.abs(a).max() > 0
        assert np.array_equal(b, 2.0 * a)
