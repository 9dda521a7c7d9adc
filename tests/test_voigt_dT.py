"""Temperature derivatives of the tabulated Voigt profile (update_hvoigt_dT,
update_phi_dT), which carry the adjoint's gradients through the setup to T.

Hvoigt is a bin average of the piecewise-linear Voigt tables at x = Δν/ΔνD and
a, both ∝ T^-1/2.  Its derivative is checked against a central difference of
the same tabulated function evaluated in 40-digit arithmetic (mpmath) at a
relative step of 1e-25, where neither truncation nor rounding reaches 1e-12.
A float64 difference cannot serve: in the far wings the bin average subtracts
two values near √π/2, and a 1e-6 step turns their rounding into a 1e-2 error.
ϕ_ν's derivative is checked against a float64 difference of update_phi along
Hvoigt + ε dHvoigt (sums of positive terms, no cancellation) and against the
invariant Σ_core Δν_k dϕ_k = 0 of the core normalisation.
"""
import mpmath as mp
import numpy as np
import pytest

from lydion_amd import constants as K
from lydion_amd import workloads as W
from lydion_amd.lya_arrays import (VoigtLUT, compute_a, compute_dnuD, update_hvoigt, update_hvoigt_dT,
                                   update_phi, update_phi_dT)

DPS = 40


@pytest.fixture(scope="module")
def lut():
    return VoigtLUT()


@pytest.fixture(scope="module")
def grid(lut):
    c = W.expanding_cloud(8, 8, 128, tau0=1e7, u_max=200.0, sigma_ln=0.5, lut=lut)
    return c.lya


class _Tabulated:
    """The tabulated H, ∫H and bin average in mpmath arithmetic, on the float64 table nodes."""

    def __init__(self, lut):
        self.lut = lut
        self.xmin = mp.mpf(lut.xmin)
        self.inv_dx = mp.mpf(lut.inv_dx)

    def fn(self, which, x):
        lut = self.lut
        if abs(x) < lut.xmax:
            xt = (x - self.xmin) * self.inv_dx
            i = int(mp.floor(xt))
            if i + 2 > lut.nx:
                i = lut.nx - 2
            d = (i + 1) - xt
            return d * mp.mpf(lut.t[which][i]) + (1 - d) * mp.mpf(lut.t[which][i + 1])
        if which == 0:
            return mp.sign(x)
        if which == 1:
            return mp.mpf(0.5) / x + mp.mpf(0.25) / x ** 3
        return mp.mpf(0)

    def H(self, x, a):
        c = mp.mpf(K.inv_Sqrt_pi_over_2)
        return self.fn(2, x) + c * a * (2 * x * self.fn(1, x) - 1) + a * a * self.fn(4, x)

    def I(self, x, a):
        return (mp.mpf(K.Sqrt_pi_over_2) * self.fn(0, x) - mp.mpf(K.inv_Sqrt_pi_over_2) * a * self.fn(1, x)
                + a * a * self.fn(3, x))

    def hvoigt(self, T, dnu_lo, dnu_hi):
        b = mp.mpf(K.b_PREFACT) * mp.sqrt(T)
        a = mp.mpf(K.a_PREFACT) / b
        iu = 1 / (b * mp.mpf(K.nu_Lya_over_c))
        xM, xm = mp.mpf(dnu_hi) * iu, mp.mpf(dnu_lo) * iu
        return (self.I(xM, a) - self.I(xm, a)) / (xM - xm)


def test_hvoigt_dT_exact_for_the_tables(lut, grid):
    tab = _Tabulated(lut)
    rng = np.random.Generator(np.random.PCG64(5))
    Ts = np.concatenate([[3.0, 30.0, 1e4, 3e5], 10.0 ** rng.uniform(0.5, 5.5, size=8)])
    d = grid["dnu_rel"]
    assert d.max() / compute_dnuD(3.0) > lut.xmax          # the asymptotic branch is exercised
    dH = update_hvoigt_dT(lut, d, grid["dnuC_rel"], Ts[:, None])[:, 0, :]
    Hv = update_hvoigt(lut, d, grid["dnuC_rel"], compute_a(Ts)[:, None], 1.0 / compute_dnuD(Ts)[:, None])[:, 0, :]
    worst = 0.0
    with mp.workdps(DPS):
        for t, T in enumerate(Ts):
            Tm = mp.mpf(T)
            h = Tm * mp.mpf("1e-25")
            for k in range(0, grid.n_nu, 4):
                ref = tab.hvoigt(Tm, d[k], d[k + 1])
                assert abs(float(ref) - Hv[t, k]) <= 1e-6 * abs(float(ref))   # same function as the setup
                fd = (tab.hvoigt(Tm + h, d[k], d[k + 1]) - tab.hvoigt(Tm - h, d[k], d[k + 1])) / (2 * h)
                worst = max(worst, abs(float(fd) - dH[t, k]) / abs(float(fd)))
    assert worst < 1e-7, worst


def test_H_scale_exact_for_the_tables(lut):
    """The narrow-bin branch of Htophat evaluates H at the bin centre."""
    tab = _Tabulated(lut)
    rng = np.random.Generator(np.random.PCG64(11))
    # fixed points off the table nodes, where the piecewise-linear tables have one slope
    xs = np.concatenate([rng.uniform(-1200.0, 1200.0, size=24), [0.3037, -2.7113, 999.372, -1000.5]])
    aa = 10.0 ** rng.uniform(-4, -1, size=xs.size)
    D = lut.H_scale(xs, aa)
    with mp.workdps(DPS):
        for x, a, got in zip(xs, aa, D):
            s = mp.mpf("1e-25")
            xm, am = mp.mpf(x), mp.mpf(a)
            fd = (tab.H(xm * (1 + s), am * (1 + s)) - tab.H(xm * (1 - s), am * (1 - s))) / (2 * s)
            assert abs(float(fd) - got) <= 1e-9 * max(abs(float(fd)), 1e-12), (x, a)


def test_phi_dT_quotient_rule_and_normalisation(lut, grid):
    rng = np.random.Generator(np.random.PCG64(3))
    T = 10.0 ** rng.uniform(1, 5, size=(4, 5))
    Hv = update_hvoigt(lut, grid["dnu_rel"], grid["dnuC_rel"], compute_a(T), 1.0 / compute_dnuD(T))
    dHv = update_hvoigt_dT(lut, grid["dnu_rel"], grid["dnuC_rel"], T)
    kmin, kmax = grid.kcore
    dnu = grid["dnu"]
    dphi = update_phi_dT(dnu, Hv, dHv, kmin, kmax)
    # outside the core ϕ is 0 at every T
    assert np.all(dphi[..., :kmin] == 0) and np.all(dphi[..., kmax + 1:] == 0)
    # Σ_core Δν_k ϕ_k = 1 at every T, so its derivative vanishes
    core = slice(kmin, kmax + 1)
    scale = np.sum(dnu[core] * np.abs(dphi[..., core]), axis=-1)
    assert np.all(np.abs(np.sum(dnu[core] * dphi[..., core], axis=-1)) <= 1e-12 * scale)
    # directional difference of update_phi along dHvoigt
    eps = 1e-6 * np.max(np.abs(Hv[..., core]) / np.max(np.abs(dHv[..., core]), axis=-1, keepdims=True), axis=-1)
    e = eps[..., None]
    fd = (update_phi(dnu, Hv + e * dHv, kmin, kmax) - update_phi(dnu, Hv - e * dHv, kmin, kmax)) / (2 * e)
    ref = np.max(np.abs(dphi[..., core]), axis=-1, keepdims=True)
    assert np.all(np.abs(fd - dphi) <= 1e-6 * ref)
