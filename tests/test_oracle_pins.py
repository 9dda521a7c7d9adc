"""Pins of the CPU oracle (oracle/m1lya_oracle.c) against what mathematics and
library routines fix: closed forms, limits, invariants, LAPACK dgtsv (scipy)
and the Voigt function (scipy).  None of these checks retypes the oracle's
formulas; each is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign or index, a transposed operand) fails it.

The reference ships no bit-level golden vectors for this path (its validation
is run-and-plot against analytic and Monte-Carlo tables); the physics-level
tables are checked in test_physics.py.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.linalg.lapack import dgtsv
from scipy.special import voigt_profile

import lydion_amd as L
from lydion_amd import constants as K
from oracle.oracle import OracleState, load, ptr
from tests.helpers import randomize, rel, small_cloud

lib = load()


def _arr(n):
    return (C.c_double * n)()


# ---------------------------------------------------------------- exp
def test_exp_within_two_ulp_of_libm():
    x = np.concatenate([np.linspace(-45, 45, 20001), [-40.0, 40.0, 0.0, 1e-3, -1e-3, 700.0, -744.0]])
    got = np.array([lib.orc_exp(v) for v in x])
    ref = np.exp(x)
    ulp = np.spacing(ref)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0
    assert lib.orc_exp(0.0) == 1.0


# ---------------------------------------------------------------- closure
def test_levermore_closure_matches_boosted_isotropic_radiation():
    """Levermore (1984): χ(f) is the Eddington factor of radiation isotropic in a frame moving
    with speed β: f = 4β/(3+β²), χ = (1+3β²)/(3+β²)."""
    beta = np.linspace(0.0, 0.99999, 2001)
    f = 4 * beta / (3 + beta * beta)
    chi = (1 + 3 * beta * beta) / (3 + beta * beta)
    got = np.array([lib.orc_chi_L84(v) for v in f])
    assert np.max(np.abs(got - chi)) < 1e-14
    assert lib.orc_chi_L84(0.0) == pytest.approx(1 / 3, abs=1e-16)
    assert lib.orc_chi_L84(1.0) == 1.0


def test_eddington_tensor_eigenstructure():
    rng = np.random.default_rng(0)
    out = _arr(3)
    for _ in range(500):
        J = math.exp(rng.normal(0, 5))
        f = rng.uniform(0.01, 0.999)
        th = rng.uniform(0, 2 * np.pi)
        Hr, Hz = f * J * math.cos(th), f * J * math.sin(th)
        lib.orc_D_edd(Hr, Hz, J, out)
        D = np.array([[out[0], out[2]], [out[2], out[1]]])
        n = np.array([math.cos(th), math.sin(th)])
        chi = (1 + 3 * 0) if False else lib.orc_chi_L84(f)
        # D n = χ n ;  D n⊥ = (1-χ)/2 n⊥ ; trace over (r, Z, ϕ) = 1
        assert np.allclose(D @ n, chi * n, rtol=1e-12, atol=1e-14)
        nperp = np.array([-n[1], n[0]])
        assert np.allclose(D @ nperp, 0.5 * (1 - chi) * nperp, rtol=1e-12, atol=1e-14)
        assert out[0] + out[1] + 0.5 * (1 - chi) == pytest.approx(1.0, abs=1e-14)
    # isotropic branches: J <= Jeps, |H| = 0, |H| <= fmin J
    for args in ((1.0, 0.0, 1e-61), (0.0, 0.0, 1.0), (1e-11, 0.0, 1.0)):
        lib.orc_D_edd(*args, out)
        assert list(out) == [1 / 3, 1 / 3, 0.0]
    # free streaming along Z: D_ZZ -> 1, D_rr -> 0
    lib.orc_D_edd(0.0, 1.0, 1.0, out)
    assert out[1] == pytest.approx(1.0, abs=1e-9) and out[0] == pytest.approx(0.0, abs=1e-9)


def test_hll_eigenvalues_limits():
    out = _arr(2)
    lib.orc_hll_eigenvalues(0.0, 0.3, out)        # isotropic: ±1/√3
    assert out[0] == pytest.approx(-1 / math.sqrt(3), rel=1e-9)
    assert out[1] == pytest.approx(1 / math.sqrt(3), rel=1e-9)
    lib.orc_hll_eigenvalues(1.0, 1.0, out)        # free streaming along the normal: both -> 1
    assert out[0] == pytest.approx(1.0, abs=1e-4) and out[1] == pytest.approx(1.0, abs=1e-9)
    lib.orc_hll_eigenvalues(1.0, -1.0, out)
    assert out[0] == pytest.approx(-1.0, abs=1e-9) and out[1] == pytest.approx(-1.0, abs=1e-4)


def test_minmod():
    assert lib.orc_minmod(1.0, 2.0) == 1.0
    assert lib.orc_minmod(-3.0, -2.0) == -2.0
    assert lib.orc_minmod(1.0, -2.0) == 0.0
    assert lib.orc_minmod(0.0, 5.0) == 0.0


# ---------------------------------------------------------------- fluxes
def _glf(FJ, FHr, FHz, J, Hr, Hz, tau, Ji, Jip1):
    inp = _arr(16)
    vals = [FJ[0], FJ[1], FHr[0], FHr[1], FHz[0], FHz[1], J[0], J[1], Hr[0], Hr[1], Hz[0], Hz[1],
            tau[0], tau[1], Ji, Jip1]
    for m, v in enumerate(vals):
        inp[m] = v
    out = _arr(3)
    lib.orc_flux_glf(inp, out)
    return list(out)


def test_glf_thin_limit_is_lax_friedrichs():
    # τ = 0 -> ϵ = 1: F = ½(F_L + F_R) - ½ (U_R - U_L), for J and H
    F = _glf((0.3, 0.1), (0.2, 0.4), (0.05, -0.05), (1.0, 2.0), (0.3, 0.1), (0.1, 0.2), (0.0, 0.0), 1.0, 2.0)
    assert F[0] == pytest.approx(0.5 * 0.4 - 0.5 * 1.0, abs=1e-15)
    assert F[1] == pytest.approx(0.5 * 0.6 - 0.5 * (0.1 - 0.3), abs=1e-15)
    assert F[2] == pytest.approx(0.0 - 0.5 * (0.2 - 0.1), abs=1e-15)


def test_glf_thick_limit_is_diffusion():
    tau = 1e7
    # continuous face state, no flux at the face: F_J -> -(J_{i+1} - J_i)/(3τ) (Fick, D = 1/(3α))
    F = _glf((0.0, 0.0), (0.2, 0.2), (0.0, 0.0), (1.5, 1.5), (0.0, 0.0), (0.0, 0.0), (tau / 2, tau / 2), 1.0, 2.0)
    assert F[0] == pytest.approx(-(2.0 - 1.0) / (3 * tau), rel=1e-6)
    # the Lax-Friedrichs dissipation is weighted by ϵ ~ 1/τ: τ F_J -> -½ (J_R - J_L)
    F = _glf((0.0, 0.0), (0.2, 0.2), (0.0, 0.0), (1.0, 2.0), (0.0, 0.0), (0.0, 0.0), (tau / 2, tau / 2), 1.0, 1.0)
    assert tau * F[0] == pytest.approx(-0.5, rel=1e-6)
    # and the H fluxes keep their physical part ½(K_L + K_R) undamped
    assert F[1] == pytest.approx(0.2, rel=1e-12)


def test_reconstruction_is_exact_for_linear_data():
    inp = _arr(19)
    x = np.array([0.0, 1.0, 2.0, -1.0])          # cells i, i+1, i+2, i-1 at unit spacing
    J = 2.0 + 0.5 * x
    f = 0.1 * np.ones(4)
    vals = list(J) + list(f * J) + list(0.0 * J) + [0.5, 0.5, 1.0, 1.0, 1.0, 3.0, 4.0]
    for m, v in enumerate(vals):
        inp[m] = v
    out = _arr(14)
    lib.orc_reconstruct_interface(inp, out)
    assert out[0] == pytest.approx(2.25, abs=1e-15) and out[1] == pytest.approx(2.25, abs=1e-15)
    assert out[2] == pytest.approx(0.225, abs=1e-15) and out[3] == pytest.approx(0.225, abs=1e-15)
    assert out[6] == 1.5 and out[7] == 2.0                      # τ_L = α_i ΔL, τ_R = α_{i+1} ΔR
    assert out[8] == pytest.approx(0.1) and out[10] == pytest.approx(1.0)  # |f| and n_r


def test_kphiphi_on_axis():
    # isotropic field on the axis: K_ϕϕ = (1 - D_rr - D_ZZ) J = J/3
    assert lib.orc_Kphiphi_inner_axis(3.0, 0.0) == pytest.approx(1.0, rel=1e-15)
    # free streaming along Z: K_ϕϕ -> 0
    assert lib.orc_Kphiphi_inner_axis(1.0, 1.0) == pytest.approx(0.0, abs=1e-9)


# ---------------------------------------------------------------- tridiagonal
@pytest.mark.parametrize("n", [1, 2, 3, 7, 126])
def test_gtsv_matches_lapack(n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        dl = rng.normal(size=max(n - 1, 0))
        d = rng.normal(size=n) * (1.0 if trial % 2 else 5.0)   # odd trials force row interchanges
        du = rng.normal(size=max(n - 1, 0))
        b = rng.normal(size=n)
        x_ref = b / d if n == 1 else dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())[3]
        a1, a2, a3, a4 = dl.copy(), d.copy(), du.copy(), b.copy()
        info = lib.orc_gtsv(n, ptr(a1) if n > 1 else ptr(np.zeros(1)), ptr(a2), ptr(a3) if n > 1 else ptr(np.zeros(1)),
                            ptr(a4))
        assert info == 0
        assert np.max(np.abs(a4 - x_ref)) <= 1e-10 * np.max(np.abs(x_ref))


def test_gtsv_reports_singular():
    dl, d, du, b = np.zeros(2), np.zeros(3), np.zeros(2), np.ones(3)
    assert lib.orc_gtsv(3, ptr(dl), ptr(d), ptr(du), ptr(b)) > 0


# ---------------------------------------------------------------- frequency factors
def _freq_cloud(umax=0.0, recoil=True):
    c = small_cloud(n_r=6, n_z=5, n_nu=64, tau0=1e7, umax=umax)
    c.lya.recoil = recoil
    return randomize(c, seed=21)


def test_freq_factors_diffusion_limit():
    """No recoil, no velocity: P = 0 and F± reduce to the harmonic-mean central difference
    F⁺_k = 1/(Δν_k (δ_k/K_k + δ'_{k+1}/K_{k+1}))."""
    c = _freq_cloud(umax=0.0, recoil=False)
    o = OracleState(c)
    o.update_freq_factors()
    lya = c.lya
    nu, nuC, dnu = lya.nu, lya["nuC"], lya["dnu"]
    cell = (2, 3)
    pref = 0.5 * max(lya["alpha"][cell], lya.alpha_floor) * lya["dnuD"][cell] ** 2
    Kk = pref * lya["Hvoigt"][cell]
    for k in range(1, lya.n_nu - 2):
        Fp = 1.0 / (dnu[k] * ((nu[k + 1] - nuC[k]) / Kk[k] + (nuC[k + 1] - nu[k + 1]) / Kk[k + 1]))
        Fm = 1.0 / (dnu[k] * ((nu[k] - nuC[k - 1]) / Kk[k - 1] + (nu[k + 1] - nuC[k]) / Kk[k]))
        assert o["Fkp"][cell + (k,)] == pytest.approx(Fp, rel=1e-12)
        assert o["Fkm"][cell + (k,)] == pytest.approx(Fm, rel=1e-12)
        assert o["Fk"][cell + (k,)] == pytest.approx(Fp + Fm, rel=1e-12)


@pytest.mark.parametrize("umax", [0.0, 3200.0])
def test_freq_factors_conserve_photons(umax):
    """Exponential fitting makes the operator conservative: Δν_k F_k = Δν_{k+1} F⁻_{k+1} + Δν_{k-1} F⁺_{k-1},
    and all coefficients are positive (an M-matrix)."""
    c = _freq_cloud(umax=umax)
    o = OracleState(c)
    o.update_freq_factors()
    dnu = c.lya["dnu"]
    Fk, Fp, Fm = o["Fk"], o["Fkp"], o["Fkm"]
    inner = slice(1, c.lya.n_nu - 1)
    assert np.all(Fk[..., inner] > 0) and np.all(Fp[..., inner] > 0) and np.all(Fm[..., inner] > 0)
    k = np.arange(2, c.lya.n_nu - 2)
    lhs = dnu[k] * Fk[..., k]
    rhs = dnu[k + 1] * Fm[..., k + 1] + dnu[k - 1] * Fp[..., k - 1]
    # The half-bin widths δν are differences of absolute frequencies (~2.5e15 Hz), as in the
    # reference (src/FreqFactors.jl:207-209), so the identity holds to ~ε ν / (Δν/2) per bin.
    bound = 8 * np.finfo(float).eps * c.lya["nuC"][k] / (0.5 * np.minimum(dnu[k - 1], dnu[k + 1]))
    err = np.abs(lhs - rhs) / np.abs(rhs)
    assert np.all(err <= bound[None, None, :])


def test_frequency_solve_satisfies_system_and_conserves():
    c = _freq_cloud(umax=800.0)
    lya = c.lya
    # radiation confined to the middle bins, so no photons reach the absorbing edge bins
    mid = np.zeros(lya.n_nu)
    mid[20:44] = 1.0
    for k in ("J", "J_old"):
        lya[k] = lya[k] * mid + 1e-60
    o = OracleState(c)
    cdt = c.c_tilde * c.dt
    o.update_freq_factors()
    Jold = o["J_old"].copy()
    Fk, Fp, Fm = o["Fk"].copy(), o["Fkp"].copy(), o["Fkm"].copy()
    o.frequency_update_all_cells(cdt, update_ff=False)
    J = o["J"]
    dw = 1.0 / cdt
    k = np.arange(1, lya.n_nu - 1)
    res = Fm[..., k] * J[..., k - 1] - (Fk[..., k] + dw) * J[..., k] + Fp[..., k] * J[..., k + 1] + dw * Jold[..., k]
    # normalised per column (the solve output is floored at Jeps afterwards, which only moves
    # entries many orders of magnitude below the column's scale)
    scale = np.max(Fk[..., k] * J[..., k] + dw * Jold[..., k], axis=-1, keepdims=True)
    assert np.max(np.abs(res) / scale) < 1e-9
    e_old = (Jold * lya["dnu"]).sum(-1)
    e_new = (J * lya["dnu"]).sum(-1)
    assert rel(e_new, e_old) < 1e-9


# ---------------------------------------------------------------- advection invariants
def test_advection_well_balanced_uniform_isotropic():
    """Uniform isotropic J, H = 0, no opacity, no flow is a discrete steady state: the flux
    divergence of H_r, (J/3)(A_R - A_L)/ΔV, cancels the geometric term K_ϕϕ⟨1/r⟩ = (J/3)·2/(r_R + r_L)."""
    c = small_cloud(n_r=9, n_z=7, n_nu=8, tau0=1e6)
    lya = c.lya
    for k in ("J", "J_old"):
        lya[k] = np.full(lya[k].shape, 3.0)
    for k in ("H_r", "H_z", "H_r_old", "H_z_old", "alpha_F"):
        lya[k] = np.zeros(lya[k].shape)
    o = OracleState(c)
    o.advect_lya(c.c_tilde * c.dt)
    assert np.max(np.abs(o["J"] - 3.0)) < 1e-13 * 3
    assert np.max(np.abs(o["H_r"])) < 1e-13 * 3
    assert np.max(np.abs(o["H_z"])) < 1e-13 * 3


def test_advection_conserves_photons_with_reflecting_walls():
    c = randomize(small_cloud(n_r=10, n_z=8, n_nu=16, tau0=1e5, tau_abs=0.0), seed=31, scale=1.0)
    lya = c.lya
    J = lya["J"]
    lya["J"] = np.maximum(J, 1e-3)      # keep every cell far above the Jeps floor
    lya["J_old"] = lya["J"].copy()
    o = OracleState(c)
    dV = c.geometry.dV
    e0 = (o["J"] * dV[:, :, None]).sum(axis=(0, 1))
    o.advect_lya(c.c_tilde * c.dt, bc_r_right=0, bc_z_bot=0, bc_z_top=0)
    e1 = (o["J"] * dV[:, :, None]).sum(axis=(0, 1))
    assert rel(e1, e0) < 1e-12
    assert np.max(np.abs(o["J"] - lya["J"])) > 1e-6     # the state did move


# ---------------------------------------------------------------- emission / reductions
def test_emission_is_backward_euler():
    c = randomize(small_cloud(tau_abs=3.0), seed=41)
    lya = c.lya
    o = OracleState(c)
    J0, Hr0 = o["J_old"].copy(), o["H_r"].copy()
    o.emission_absorption(c.c_tilde, c.dt)
    ct, dt = c.c_tilde, c.dt
    absJ = lya["alpha_c"][..., None] + (lya["alpha"] * lya["pd"])[..., None] * lya["Hvoigt"]
    src = (lya["js"] * ct)[..., None] * lya["phi"] * dt
    # implicit equation: J_new (1 + α_abs c̃ Δt) = J_old + js c̃ ϕ Δt
    assert rel(o["J"] * (1 + absJ * ct * dt), J0 + src) < 1e-14
    Hn = o["H_r"]
    assert rel(Hn * (1 + lya["alpha_F"] * ct * dt), Hr0) < 1e-14
    # momentum: Δρu = (4π/c)(1/c̃) Σ Δν (H_old - H)
    mom = K.Four_pi_over_c / ct * ((Hr0 - Hn) * lya["dnu"]).sum(-1)
    assert rel(o["drhou_r"], mom) < 1e-12


def test_force_multiplier_closed_forms():
    c = small_cloud(n_r=8, n_z=8, n_nu=8)
    g = c.geometry
    o = OracleState(c)
    js0 = 2.5e-15
    o.set("js", js0)
    rc = g.rCoM[None, :]
    zc = g.ZCoM[:, None] - c.Zc
    R = np.sqrt(rc * rc + zc * zc)
    F0, dt = 3.0e-20, 7.0
    o.set("drhou_r", F0 * dt * rc / R)
    o.set("drhou_z", F0 * dt * zc / R)
    MF, net, out, Ltot = o.update_f_mf_l(dt, bc_z_bot=0)
    Vbox = math.pi * g.r[-1] ** 2 * g.Z[-1]
    Lsum = 4 * math.pi * js0 * Vbox
    assert Ltot == pytest.approx(2 * Lsum, rel=1e-12)                # reflecting midplane doubles L
    assert MF == pytest.approx(F0 * Vbox / (Lsum / K.c), rel=1e-12)  # |f| ΔV summed over the box
    assert net == pytest.approx(MF, rel=1e-12) and out == pytest.approx(MF, rel=1e-12)
    o.set("drhou_r", -F0 * dt * rc / R)                               # inward force
    o.set("drhou_z", -F0 * dt * zc / R)
    MF2, net2, out2, _ = o.update_f_mf_l(dt, bc_z_bot=0)
    assert MF2 == pytest.approx(MF, rel=1e-12) and net2 == pytest.approx(-MF, rel=1e-12) and out2 == 0.0
    o.set("js", 0.0)
    assert o.update_f_mf_l(dt, bc_z_bot=0)[:3] == (0.0, 0.0, 0.0)       # L <= 0 branch


def test_escape_fraction_closed_form():
    c = small_cloud(n_r=8, n_z=6, n_nu=10)
    g, lya = c.geometry, c.lya
    o = OracleState(c)
    h = 2.0e-3
    Hr = np.zeros(o["H_r"].shape)
    Hz = np.zeros(o["H_z"].shape)
    Hr[:, -1, :] = h          # outflow through the outer cylinder
    Hz[-1, :, :] = -h         # inward at the top: does not escape
    Hz[0, :, :] = -h          # downward at the bottom: escapes only with :Outflow
    o.set("H_r", Hr)
    o.set("H_z", Hz)
    span = lya.nu[-1] - lya.nu[0]
    A_side = 2 * math.pi * g.r[-1] * g.Z[-1]
    A_cap = math.pi * g.r[-1] ** 2
    Lall = 1e38
    assert o.f_esc(Lall, 0) == pytest.approx(4 * math.pi * h * span * A_side / (0.5 * Lall), rel=1e-12)
    assert o.f_esc(Lall, 1) == pytest.approx(4 * math.pi * h * span * (A_side + A_cap) / Lall, rel=1e-12)
    assert o.f_esc(0.0, 0) == 0.0


# ---------------------------------------------------------------- setup
def test_reference_freq_grid_bin_counts():
    """GenerateLyαArrays gives N_ν = 194 at q = 1.1 and 542 at q = 1.03 (Tmin = 30 K)."""
    for q, n in ((1.1, 194), (1.03, 542)):
        buf = np.zeros(4000)
        m = lib.orc_lya_freq_grid(30.0, q, q, ptr(buf), 4000)
        assert m - 1 == n
        nu = buf[:m]
        assert np.all(np.diff(nu) > 0)
        assert np.allclose(nu + nu[::-1], 2 * K.nu_Lya, rtol=1e-15)   # symmetric about line centre
        assert np.array_equal(nu, L.reference_freq_grid(30.0, q, q))  # host setup == oracle


def test_voigt_profile_against_scipy():
    lut = L.VoigtLUT()
    x = np.linspace(-60, 60, 4001)
    for a in (4.7e-4, 4.7e-3):
        H = lut.H(x, a)
        ref = math.sqrt(math.pi) * voigt_profile(x, 1 / math.sqrt(2), a)
        assert np.max(np.abs(H - ref) / ref) < 1e-4
    # ∫ H dx over the whole line = √π (telescoping bin averages)
    c = L.static_cloud_32x32x64(lut=lut)
    Hv = c.lya["Hvoigt"][0, 0]
    dx = c.lya["dnu"] / c.lya["dnuD"][0, 0]
    # grid spans ±x_max: the Lorentz wings beyond carry 2 ∫_{x_max}^∞ a/(√π x²) dx = 2a/(√π x_max)
    a = c.lya["a"][0, 0]
    x_max = (c.lya.nu[-1] - c.lya.nu[0]) / 2 / c.lya["dnuD"][0, 0]
    assert (Hv * dx).sum() == pytest.approx(math.sqrt(math.pi) - 2 * a / (math.sqrt(math.pi) * x_max), rel=2e-7)
    # normalised emission profile over the core bins
    k0, k1 = c.lya.kcore
    assert (c.lya["phi"][0, 0, k0:k1 + 1] * c.lya["dnu"][k0:k1 + 1]).sum() == pytest.approx(1.0, rel=1e-14)


def test_host_hvoigt_and_phi_match_oracle():
    lut = L.VoigtLUT()
    c = L.static_cloud_32x32x64(lut=lut)
    lya = c.lya
    n_cells = lya.n_r * lya.n_z
    Hv = np.zeros((lya.n_z, lya.n_r, lya.n_nu))
    lib.orc_update_hvoigt(ptr(Hv), lya.n_nu, n_cells, ptr(np.ascontiguousarray(lya["dnu_rel"])),
                          ptr(np.ascontiguousarray(lya["dnuC_rel"])), ptr(np.ascontiguousarray(lya["a"])),
                          ptr(np.ascontiguousarray(lya["inv_dnuD"])), ptr(np.ascontiguousarray(lut.t)), lut.nx,
                          lut.xmax)
    assert np.array_equal(Hv, lya["Hvoigt"])
    phi = np.zeros_like(Hv)
    lib.orc_update_phi(ptr(phi), lya.n_nu, n_cells, ptr(np.ascontiguousarray(lya["dnu"])), ptr(Hv), *lya.kcore)
    assert np.array_equal(phi, lya["phi"])


def test_host_geometry_matches_oracle_and_closed_forms():
    g = L.setup_cylindrical_grid(1.1 * K.pc, 1.1 * K.pc, 12, 9)
    from oracle.oracle import OrcGeom
    og = OrcGeom()
    og.n_r, og.n_z = 12, 9
    keep = {}
    shapes = dict(r=13, z=10, rcom=12, zcom=9, dV=108, inv_dV=108, A_r=117, A_z=120, inv_r_cell=12,
                  dr_L=12, dr_R=12, inv_dr_i_im1=12, inv_dr_ip1_i=12, inv_dr_ip2_ip1=12,
                  dz_L=9, dz_R=9, inv_dz_j_jm1=9, inv_dz_jp1_j=9, inv_dz_jp2_jp1=9)
    for n, s in shapes.items():
        keep[n] = np.zeros(s)
        setattr(og, n, ptr(keep[n]))
    lib.orc_cyl_grid_uniform(1.1 * K.pc, 1.1 * K.pc, C.byref(og))
    for n, h in (("r", g.r), ("z", g.Z), ("rcom", g.rCoM), ("zcom", g.ZCoM), ("dV", g.dV), ("A_r", g.A_r),
                 ("A_z", g.A_Z), ("inv_r_cell", g.inv_r_cell), ("dr_L", g.dr[0]), ("inv_dr_ip2_ip1", g.dr[4])):
        assert np.array_equal(keep[n], np.ravel(h)), n
    assert list(og.dr_inner) == list(g.dr[5]) and list(og.dz_outer) == list(g.dZ[6])
    R, Z = g.r[-1], g.Z[-1]
    assert g.dV.sum() == pytest.approx(math.pi * R * R * Z, rel=1e-13)
    assert g.rCoM[0] == pytest.approx(2 / 3 * g.r[1], rel=1e-15)
    # flux-form/geometric consistency used by the well-balanced property
    dA = (g.A_r[:, 1:] - g.A_r[:, :-1]) / g.dV
    assert rel(dA, np.broadcast_to(g.inv_r_cell[None, :], dA.shape)) < 1e-13
