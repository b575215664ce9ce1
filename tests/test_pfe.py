"""The SIBC slab fitter (vulture_amd/pfe.py) on the CPU: the slab impedances
against their closed forms, the fits' accuracy and passivity over the
conductor ranges of the SIBC + Mur workload, the .pfe text round trip, and a
1-D FDTD run of the CPU oracle whose transmission through a fitted sheet
matches the slab's analytic transmission coefficient."""
import math
import os

import numpy as np
import pytest

import oracle
from vulture_amd import pfe
from vulture_amd.obsfiles import read_fd


def test_slab_closed_forms():
    sig, d = 3e6, 5e-4
    # DC: both entries tend to the sheet resistance 1/(sigma d)
    z11, z12 = pfe.slab_z([1e-2], sig, d)
    assert abs(z11[0] - 1 / (sig * d)) < 1e-6 / (sig * d)
    assert abs(z12[0] - 1 / (sig * d)) < 1e-6 / (sig * d)
    # thick slab, high frequency: Z11 is the surface impedance (1+j) sqrt(w mu / 2 sigma), Z12 vanishes
    f = np.array([5e9, 2e10])
    z11, z12 = pfe.slab_z(f, 6e7, 2e-3)
    w = 2 * math.pi * f
    zs = (1 + 1j) * np.sqrt(w * pfe.MU0 / (2 * 6e7))
    assert np.all(np.abs(z11 - zs) / np.abs(zs) < 1e-6)
    assert np.all(np.abs(z12) < 1e-30)
    # even and odd impedances: Zc coth(gamma d / 2), Zc tanh(gamma d / 2)
    f = np.logspace(6, 10, 9)
    s = 2j * math.pi * f
    for sig, d in [(1e5, 1e-4), (1e6, 2e-4)]:
        z11, z12 = pfe.slab_z(f, sig, d)
        y = sig + s * pfe.EPS0
        gam, zc = np.sqrt(s * pfe.MU0 * y), np.sqrt(s * pfe.MU0 / y)
        assert np.allclose(z11 + z12, zc / np.tanh(gam * d / 2), rtol=1e-10)
        assert np.allclose(z11 - z12, zc * np.tanh(gam * d / 2), rtol=1e-10)
    # thin sheet Z11 = Z12 = R: transmission 2R / (eta0 + 2R)
    r = 0.7
    assert abs(pfe.slab_s21(r, r) - 2 * r / (pfe.ETA0 + 2 * r)) < 1e-15
    # the transmission is the S21 whose S-to-Z conversion (sibc.c:790-803,
    # Z = eta0 (1 + S)(1 - S)^-1) returns the two-port: check the round trip
    z11, z12 = pfe.slab_z([2e9], 1e5, 1e-4)
    z = np.array([[z11[0], z12[0]], [z12[0], z11[0]]])
    s_mat = (z - pfe.ETA0 * np.eye(2)) @ np.linalg.inv(z + pfe.ETA0 * np.eye(2))
    assert abs(s_mat[1, 0] - pfe.slab_s21(z11[0], z12[0])) < 1e-12
    assert np.allclose(pfe.ETA0 * (np.eye(2) + s_mat) @ np.linalg.inv(np.eye(2) - s_mat), z, rtol=1e-10)


@pytest.mark.parametrize("seed", range(6))
def test_fit_accuracy_and_passivity(seed):
    """Conductivity logU[1e5, 6e7] S/m, thickness U[0.1, 2] mm: Z11 within 2 %
    over 10 MHz - 20 GHz; every term positive (passive); transmission within
    0.5 dB wherever |Z12| > 20 % of |Z11| (below that, Z12 = (Ze - Zo)/2
    carries the fits' error amplified by |Z11/Z12|, and for electrically
    thick slabs the transmission is only represented as negligible)."""
    rng = np.random.default_rng(seed)
    sig = 10 ** rng.uniform(5, math.log10(6e7))
    d = rng.uniform(1e-4, 2e-3)
    fit = pfe.fit_slab(sig, d)
    assert fit.ae >= 0 and fit.ao >= 0 and np.all(fit.we >= 0) and np.all(fit.wo >= 0)
    f = np.logspace(7, math.log10(2e10), 300)
    z11, z12 = pfe.slab_z(f, sig, d)
    g11, g12 = fit.z(f)
    assert np.max(np.abs(g11 - z11) / np.abs(z11)) < 0.02
    # passivity on the axis: Re(Z11) >= |Re(Z12)| (Re Ze, Re Zo >= 0)
    assert np.all(g11.real + 1e-15 >= np.abs(g12.real))
    t, tg = np.abs(pfe.slab_s21(z11, z12)), np.abs(pfe.slab_s21(g11, g12))
    m = np.abs(z12) > 0.2 * np.abs(z11)
    if m.any():
        assert np.max(np.abs(20 * np.log10(tg[m] / t[m]))) < 0.5


def test_pfe_text_round_trip():
    fit = pfe.fit_slab(2e6, 3e-4)
    ent = pfe.read_pfe(pfe.pfe_text(fit))
    f = np.logspace(7, 10, 50)
    g11, g12 = fit.z(f)
    for (p, q), e in ent.items():
        v = pfe.pfe_eval(e, f)
        if p // 2 != q // 2:
            assert np.all(v == 0)
        else:
            ref = g11 if p % 2 == q % 2 else g12
            assert np.allclose(v, ref, rtol=1e-12, atol=1e-12 * np.abs(g11).max())
        assert np.all(e[1].real < 0) and np.all(e[1].imag == 0)  # stable real poles


def _parplate(with_sheet):
    cards = ["VM 1.0.0", "CE 1-D parallel plate: transmission through a fitted slab", "DM 200 1 1", "GS",
             "BT XLO PML 8", "BT XHI PML 8", "BT YLO PERIODIC", "BT YHI PERIODIC", "BT ZLO PERIODIC",
             "BT ZHI PERIODIC"]
    if with_sheet:
        cards += ['BT sheet SIBC "slab.pfe"', "TB 100 100 0 1 0 1 sheet 1 0.0"]
    cards += ["WF wf1 GAUSSIAN_PULSE 1.0 8.3391e-10 1.6678e-10", "EX 5 5 0 1 0 1 ex1 EY wf1 1.0 0.0",
              "OP 175 175 1 1 1 1 op4 FDOM_ASCII", "GE", "NT 5000", "MS 0.01", "EN"]
    return "\n".join(cards) + "\n"


def test_sheet_transmission_in_fdtd(tmp_path):
    """A plane wave in a 1-D parallel-plate line (10 mm cells) crosses an SIBC
    sheet carrying the fitted slab (sigma 1e5 S/m, 0.1 mm).  The ratio of the
    transmitted field with and without the sheet is the fitted model's S21:
    within 0.1 % at DC, and within 1 % in magnitude and 0.005 rad in phase
    up to 1 GHz (30 cells per wavelength); the exact slab's within 2 % and
    0.02 rad (the fit's own error).  Phases compare once the half time step
    by which the sheet's E lags the H it is computed from (sibc.c:444-586)
    is taken out.
    The default Courant number sqrt(3)/2 on cubic 10 mm cells gives
    dt = 0.01 / (2 c0)."""
    fit = pfe.fit_slab(1e5, 1e-4)
    out = []
    for w in (0, 1):
        d = tmp_path / str(w)
        d.mkdir()
        (d / "m.mesh").write_text(_parplate(w))
        (d / "slab.pfe").write_text(pfe.pfe_text(fit))
        oracle.run_mesh(str(d / "m.mesh"), str(d / "out"), precision="fp64")
        out.append(read_fd(str(d / "out" / "eh_op4_fd.asc")))
    f, free = out[0]
    _, sheet = out[1]
    dt = 0.01 / (2 * pfe.C0)
    t_sim = sheet[:, 1] / free[:, 1] * np.exp(1j * math.pi * f * dt)
    g11, g12 = fit.z(np.maximum(f, 1.0))
    t_fit = pfe.slab_s21(g11, g12)
    assert abs(t_sim[0] / t_fit[0] - 1) < 1e-3
    z11, z12 = pfe.slab_z(np.maximum(f, 1.0), 1e5, 1e-4)
    t_exact = pfe.slab_s21(z11, z12)
    m = (f > 0) & (f <= 1e9)
    assert m.sum() >= 2
    for ref, tol_mag, tol_ph in ((t_fit, 0.01, 0.005), (t_exact, 0.02, 0.02)):
        assert np.max(np.abs(np.abs(t_sim[m]) / np.abs(ref[m]) - 1)) < tol_mag
        assert np.max(np.abs(np.angle(t_sim[m] / ref[m]))) < tol_ph
