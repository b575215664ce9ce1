"""Pole-residue models of conducting slabs for SIBC boundaries (.pfe files).

A surface-impedance boundary (`BT name SIBC "file.pfe"`, sibc.c:126-190)
reads a 4x4 matrix of pole-residue filters in the format of
yfReadPoleResidueM (filter.c:142-188):

    4 4
    <for each row, then each column:>
    numPoles asymp
    pole_re pole_im residue_re residue_im     (numPoles lines)

Entry (p, q) is Z_pq(s) = asymp + sum_k r_k / (s - p_k), taken as its real
part after recursive convolution (filter.c:309-368).  The matrix is the
two-port impedance matrix of the sheet, TM block in rows/columns 0-1 and TE
block in 2-3 (sibc.c:142-151), relating E and H on the two sides as the
reference's S-to-Z conversion does (tportS2Z, sibc.c:790-803:
Z = eta0 (1 + S)(1 - S)^-1).

The slab model: a homogeneous layer of conductivity sigma, thickness d,
permittivity eps_r eps0 and permeability mu_r mu0 is a transmission-line
section with gamma = sqrt(s mu (sigma + s eps)), Zc = sqrt(s mu / (sigma +
s eps)):

    Z11 = Z22 = Zc coth(gamma d),   Z12 = Z21 = Zc / sinh(gamma d).

Its even and odd impedances, Ze = Z11 + Z12 = Zc coth(gamma d / 2) and
Zo = Z11 - Z12 = Zc tanh(gamma d / 2), are diffusive: without displacement
current their poles lie on the negative real axis.  Each is fitted as

    Z(s) ~ a0 + sum_n w_n s / (s + s_n),   a0 >= 0, w_n >= 0,

over fixed poles -s_n spaced logarithmically (non-negative least squares).
Every term is a resistor, or a resistor in parallel with an inductor, so
both fits are positive-real and the fitted two-port is passive; the
recursive convolution of a passive model stays bounded.  Then Z11 =
(Ze + Zo)/2 and Z12 = (Ze - Zo)/2, and s/(s + s_n) = 1 - s_n/(s + s_n) gives
the pole-residue form (real poles, real residues).

The reference ships no fitter; this is the generator the SIBC + Mur
workload needs.  tests/test_pfe.py pins it against the closed forms above
(DC and skin-effect limits, the exact identities Ze, Zo) and against a 1-D
FDTD run of the CPU oracle whose transmission through the fitted sheet
matches the slab's analytic transmission coefficient.
"""
import math

import numpy as np
from scipy.optimize import nnls

EPS0 = 8.854187817e-12
MU0 = 4e-7 * math.pi
C0 = 299792458.0
ETA0 = math.sqrt(MU0 / EPS0)


def slab_z(f, sigma, thickness, eps_r=1.0, mu_r=1.0):
    """Z11 and Z12 of a conducting slab at frequencies f (Hz)."""
    s = 2j * math.pi * np.asarray(f, dtype=float)
    mu = mu_r * MU0
    y = sigma + s * eps_r * EPS0
    gam = np.sqrt(s * mu * y)
    zc = np.sqrt(s * mu / y)
    e = np.exp(-gam * thickness)  # |e| < 1: coth and csch without overflow
    q = e * e
    return zc * (1 + q) / (1 - q), zc * 2 * e / (1 - q)


def slab_s21(z11, z12, eta=ETA0):
    """Transmission coefficient of a symmetric two-port between two
    half-spaces of impedance eta: 2 Z12 eta / ((Z11 + eta)^2 - Z12^2)."""
    return 2 * z12 * eta / ((z11 + eta) ** 2 - z12 ** 2)


class SlabFit:
    """Fitted even/odd impedances on the common poles -s_n (rad/s)."""

    def __init__(self, sn, ae, we, ao, wo):
        self.sn, self.ae, self.we, self.ao, self.wo = sn, ae, we, ao, wo

    def z(self, f):
        s = 2j * math.pi * np.asarray(f, dtype=float)
        t = s[:, None] / (s[:, None] + self.sn[None, :])
        ze = self.ae + t @ self.we
        zo = self.ao + t @ self.wo
        return 0.5 * (ze + zo), 0.5 * (ze - zo)

    def pole_residue(self):
        """(poles, Z11 asymp, Z11 residues, Z12 asymp, Z12 residues), poles
        and residues real, terms with zero weight in both fits dropped."""
        keep = (self.we > 0) | (self.wo > 0)
        sn, we, wo = self.sn[keep], self.we[keep], self.wo[keep]
        a11 = 0.5 * (self.ae + self.ao + we.sum() + wo.sum())
        a12 = 0.5 * (self.ae - self.ao + we.sum() - wo.sum())
        r11 = -0.5 * (we + wo) * sn
        r12 = -0.5 * (we - wo) * sn
        return -sn, a11, r11, a12, r12


def _nnls_fit(s, z, sn):
    """z(s) ~ a0 + sum_n w_n s/(s + s_n), a0 and w >= 0, minimising the
    relative error; columns scaled to unit norm (NNLS is invariant to a
    positive column scaling, and the scaled problem is far better
    conditioned)."""
    basis = np.concatenate([np.ones((s.size, 1), dtype=complex), s[:, None] / (s[:, None] + sn[None, :])], axis=1)
    a = basis / np.abs(z)[:, None]
    r = z / np.abs(z)
    A = np.concatenate([a.real, a.imag])
    norm = np.linalg.norm(A, axis=0)
    x, _ = nnls(A / norm, np.concatenate([r.real, r.imag]), maxiter=200 * basis.shape[1])
    x = x / norm
    return x[0], x[1:]


def fit_slab(sigma, thickness, fmin=1e7, fmax=20e9, eps_r=1.0, mu_r=1.0, per_decade=4, nfreq=400):
    """Fit the slab's even and odd impedances over [fmin, fmax], each with
    relative accuracy, on poles from fmin/30 to 30 fmax.  Z12 = (Ze - Zo)/2
    inherits an absolute error of about the fits' relative error times |Ze|,
    so transmission far below that level (thick good conductors at high
    frequency) is represented only as "negligible"."""
    f = np.logspace(math.log10(fmin), math.log10(fmax), nfreq)
    s = 2j * math.pi * f
    z11, z12 = slab_z(f, sigma, thickness, eps_r, mu_r)
    lo, hi = 2 * math.pi * fmin / 30, 2 * math.pi * fmax * 30
    npole = int(math.ceil(per_decade * math.log10(hi / lo))) + 1
    sn = np.logspace(math.log10(lo), math.log10(hi), npole)
    ae, we = _nnls_fit(s, z11 + z12, sn)
    ao, wo = _nnls_fit(s, z11 - z12, sn)
    return SlabFit(sn, ae, we, ao, wo)


def pfe_text(fit):
    """The 4x4 .pfe text of a slab fit: TM and TE blocks [[Z11, Z12], [Z12,
    Z11]] (normal incidence: both polarisations alike), zero blocks off the
    diagonal."""
    poles, a11, r11, a12, r12 = fit.pole_residue()
    ent = {(0, 0): (a11, r11), (1, 1): (a11, r11), (0, 1): (a12, r12), (1, 0): (a12, r12)}
    lines = ["4 4"]
    for p in range(4):
        for q in range(4):
            blk = (p // 2 == q // 2)
            a, r = ent[(p % 2, q % 2)] if blk else (0.0, None)
            if r is None:
                lines.append(f"0 {0.0:.16e}")
                continue
            lines.append(f"{poles.size} {a:.16e}")
            for pk, rk in zip(poles, r):
                lines.append(f" {pk:.16e} {0.0:.16e} {rk:.16e} {0.0:.16e}")
    return "\n".join(lines) + "\n"


def read_pfe(text):
    """Parse .pfe text into {(p, q): (asymp, poles[], residues[])}."""
    t = text.split()
    m, n = int(t[0]), int(t[1])
    k = 2
    out = {}
    for p in range(m):
        for q in range(n):
            npl, a = int(t[k]), float(t[k + 1])
            k += 2
            pl, rs = [], []
            for _ in range(npl):
                pr, pi, rr, ri = (float(v) for v in t[k:k + 4])
                k += 4
                pl.append(complex(pr, pi))
                rs.append(complex(rr, ri))
            out[(p, q)] = (a, np.array(pl), np.array(rs))
    return out


def pfe_eval(entry, f):
    """Evaluate one parsed entry asymp + sum r/(s - p) at frequencies f."""
    a, pl, rs = entry
    s = 2j * math.pi * np.asarray(f, dtype=float)
    return a + (rs[None, :] / (s[:, None] - pl[None, :])).sum(axis=1) if pl.size else np.full(s.shape, a, complex)
