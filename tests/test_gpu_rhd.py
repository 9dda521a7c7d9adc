"""The RHD-side entry points on the GPU against the oracle: GetRHDCFLTimeStep and
FindRSL_LyαRT as device reductions, and the Lyα absorption sub-step with its rejection rule.

Bars: the CFL step is a min of per-cell values computed with IEEE divisions and square roots,
so it is bit-identical; the RSL floor involves pow(aτ, 1/3) and hypot, whose last bit may
differ between the device and the host library (≤ 1e-14 relative); the sub-step's J, H_r,
H_Z are bit-identical, its per-cell sums (momentum shares, e_Lyα_new, Δe) within 1e-12
relative, and its accept/reject decisions identical.
"""
import ctypes as C

import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import _lib
from lydion_amd import constants as K
from lydion_amd import m1lya as M
from lydion_amd import workloads as W
from oracle import oracle as O
from oracle.oracle import OracleState
from tests.helpers import bitdiff, randomize, small_cloud

pytestmark = pytest.mark.gpu


def _rhd_inputs(cloud, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    s = (cloud.lya.n_z, cloud.lya.n_r)
    rho = 1e-22 * np.exp(rng.normal(0, 1.5, s))
    return dict(u_r=rng.normal(0, 3e6, s), u_z=rng.normal(0, 3e6, s), gamma_ad=np.full(s, 5 / 3),
                P=rho * 1e12 * np.exp(rng.normal(0, 0.5, s)), rho=rho, rho_d=rho[..., None] * 1e-3 * rng.uniform(0, 1, s + (3,)),
                xHII=rng.uniform(0, 1, s), xH2=rng.uniform(0, 0.2, s), T=10 ** rng.uniform(2, 4.3, s))


def test_cfl_time_step_bit_exact():
    cloud = L.static_cloud_32x32x64()
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    inp = _rhd_inputs(cloud)
    dt = C.c_double()
    ur, uz = np.ascontiguousarray(inp["u_r"]), np.ascontiguousarray(inp["u_z"])
    ctx._check(ctx.lib.m1lya_cfl_time_step(ctx.handle, K.c, _lib.ptr(ur), _lib.ptr(uz), 0.4, C.byref(dt)))
    ref = O.cfl_dt(cloud.geometry, K.c, ur, uz, 0.4)
    assert dt.value == ref


@pytest.mark.parametrize("seed", [3, 4])
def test_find_rsl_matches_oracle(seed):
    cloud = W.rhd_like_cloud(40, 56, 32)
    cloud.lya["pd"] = np.random.default_rng(seed).uniform(0, 1e-6, (cloud.lya.n_z, cloud.lya.n_r))
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    inp = {k: np.ascontiguousarray(v) for k, v in _rhd_inputs(cloud, seed).items()}
    kap = np.array([1.2e4, 3.1e4, 8.0e4])
    ri = _lib.RslInputs()
    for k in ("u_r", "u_z", "gamma_ad", "P", "rho", "rho_d", "xHII", "xH2", "T"):
        setattr(ri, k, _lib.ptr(inp[k]))
    ri.kappa_abs[:] = list(kap)
    ri.lc = _lib.LineConstants(K.nu_Lya, K.b_PREFACT, K.nu_Lya_over_c, K.a_PREFACT, K.sigma0_PREFACT)
    for c_old in (1e7, 3e8, 1e10):
        out = C.c_double()
        ctx._check(ctx.lib.m1lya_find_rsl(ctx.handle, c_old, 1e6, C.byref(ri), C.byref(out)))
        ref = O.find_rsl(c_old, 1e6, cloud.geometry, inp["u_r"], inp["u_z"], inp["gamma_ad"], inp["P"], inp["rho"],
                         inp["rho_d"], inp["xHII"], inp["xH2"], inp["T"], cloud.lya["pd"], kap)
        print(f"RSL c_old {c_old:.1e}: device {out.value:.17e} oracle {ref:.17e}")
        assert abs(out.value - ref) <= 1e-14 * ref


def _substep_case(seed, dt_scale, e_factor):
    cloud = randomize(small_cloud(n_r=20, n_z=14, n_nu=128, tau0=1e7, tau_abs=0.7), seed=seed)
    o = OracleState(cloud)
    o.emission_absorption(cloud.c_tilde, dt_scale * cloud.dt)    # the sub-step's J, never rejected
    e_new = K.Four_pi_over_c * np.sum(np.maximum(o["J"], K.Jeps) * cloud.lya["dnu"], axis=-1)
    rng = np.random.default_rng(seed)
    cloud.lya["e_lya"] = e_new * np.where(rng.uniform(size=e_new.shape) < 0.5, 1.0, e_factor)
    return cloud


@pytest.mark.parametrize("dt_scale,e_factor", [(1.0, 1.0), (30.0, 0.5), (1e4, 0.8)])
def test_absorption_substep_matches_oracle(dt_scale, e_factor):
    cloud = _substep_case(11, dt_scale, e_factor)
    nz, nr = cloud.lya.n_z, cloud.lya.n_r
    rng = np.random.default_rng(5)
    rho_d = np.ascontiguousarray(rng.uniform(0, 1e-24, (nz, nr, 3)))
    kF = np.array([2e5, 5e5, 9e5])
    dt_cell = np.ascontiguousarray(dt_scale * cloud.dt * rng.uniform(0.5, 1.5, (nz, nr)))
    o = OracleState(cloud)
    redo_o, out_o = o.absorption_substep(dt_cell, cloud.c_tilde, rho_d=rho_d, kappa_F=kF)
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    redo = np.zeros((nz, nr), dtype=np.int32)
    out = np.zeros((12, nz, nr))
    ctx._check(ctx.lib.m1lya_absorption_substep(ctx.handle, _lib.ptr(dt_cell), 0.0, cloud.c_tilde,
                                                _lib.ptr(rho_d), _lib.ptr(kF),
                                                redo.ctypes.data_as(C.POINTER(C.c_int)), _lib.ptr(out)))
    print(f"dt x{dt_scale}, e_old x{e_factor}: rejected {int(redo.sum())} of {redo.size} cells")
    assert np.array_equal(redo, redo_o)
    for k, n in (("J", "J"), ("H_r", "H_r"), ("H_z", "H_z")):
        assert bitdiff(ctx.download(k), o[n]) == 0, k
    scale = np.maximum(np.max(np.abs(out_o), axis=(1, 2)), 1e-300)[:, None, None]
    assert np.max(np.abs(out - out_o) / scale) <= 1e-12


@pytest.mark.parametrize("ots,destruction", [(True, True), (False, True), (True, False)])
def test_prepare_subcycling_chem_matches_oracle(ots, destruction):
    """Lyα source terms from the chemical state on the device against the oracle: js, α_c, p_d
    within 1e-13 (pow, exp and log10 of the device library against the host's); then the
    Voigt refresh (bit-identical to the host setup, test_gpu_refresh.py) with this α_c,F."""
    from lydion_amd.lya_arrays import compute_a, compute_dnuD, compute_sigma0, update_hvoigt
    lut = L.VoigtLUT()
    cloud = small_cloud(n_r=18, n_z=13, n_nu=64, tau0=1e7, lut=lut)
    lya = cloud.lya
    nz, nr = lya.n_z, lya.n_r
    rng = np.random.Generator(np.random.PCG64(43))
    chem = np.empty((8, nz, nr))
    chem[0] = 10.0 ** rng.uniform(1.0, 5.2, (nz, nr))                 # T across all the clamps
    for m, (lo, hi) in enumerate([(-2, 4), (-6, 2), (-3, 3), (-3, 2), (-4, 4), (-9, -3), (-9, -3)], start=1):
        chem[m] = 10.0 ** rng.uniform(lo, hi, (nz, nr))
    chem[1, 0, :3] = 1e-22                                             # n_HI below the H2-term cut
    rho_d = np.ascontiguousarray(10.0 ** rng.uniform(-27, -21, (nz, nr, 3)))
    kap = np.array([1.2e4, 3.1e4, 8.0e4, 1.5e4, 3.5e4, 9.0e4])
    ctx = M.M1LyaContext(lya, cloud.geometry)
    ctx.set_voigt_lut(lut)
    chem_c = np.ascontiguousarray(chem)
    ctx._check(ctx.lib.m1lya_prepare_subcycling_chem(ctx.handle, _lib.ptr(chem_c), _lib.ptr(rho_d), _lib.ptr(kap),
                                                     0.5, int(ots), int(destruction), 1))
    ref = O.lya_sources(chem, rho_d, kap, z=0.5, ots=ots, destruction=destruction)
    for m, key in ((0, "js"), (1, "alpha_c"), (3, "pd")):
        g = ctx.download(key)
        e = np.max(np.abs(g - ref[m]) / np.maximum(np.abs(ref[m]), 1e-300))
        print(f"ots={ots} destruction={destruction} {key}: max rel {e:.1e}")
        assert e <= 1e-13, key
    T = chem[0]
    Hv = update_hvoigt(lut, lya["dnu_rel"], lya["dnuC_rel"], compute_a(T), 1.0 / compute_dnuD(T))
    assert bitdiff(ctx.download("Hvoigt"), Hv) == 0
    aF = (chem[1] * compute_sigma0(T))[..., None] * Hv + ref[2][..., None]
    assert np.max(np.abs(ctx.download("alpha_F") - aF) / aF) <= 1e-13
    ctx.close()
