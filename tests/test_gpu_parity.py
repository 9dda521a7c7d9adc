"""Parity of the HIP FP64 path (libm1lya via its C ABI) against the CPU oracle.

The FP64 kernels evaluate every formula in the oracle's (= the reference's)
operation order with FP contraction off; division and sqrt are IEEE correctly
rounded on both sides and both use the same deterministic exp.  So J, H_r,
H_Z and every per-element field are required to be BIT-IDENTICAL to the
oracle.  Reductions (Δρu, M_F, L, f_esc, e_Lyα) are summed in a different
order on the device; they are required to agree to 1e-12 relative, far inside
the 1e-10 bar.
"""
import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import m1lya as M
from oracle.oracle import OracleState
from tests.helpers import STATE, bitdiff, randomize, rel, small_cloud

pytestmark = pytest.mark.gpu

RED_TOL = 1e-12


@pytest.fixture(scope="module")
def lut():
    return L.VoigtLUT()


def _pair(cloud, hll=False):
    cloud.lya.hll = hll
    return M.M1LyaContext(cloud.lya, cloud.geometry), OracleState(cloud, hll=hll)


def _assert_state_equal(ctx, o, keys=STATE):
    for k in keys:
        g = ctx.download(k)
        assert bitdiff(g, o[k]) == 0, f"{k}: {bitdiff(g, o[k])} elements differ, max rel {rel(g, o[k]):.3e}"


def test_selftest_math_bit_exact():
    rng = np.random.default_rng(3)
    a = np.exp(rng.uniform(-300, 300, 20000)) * rng.choice([-1, 1], 20000)
    b = np.exp(rng.uniform(-300, 300, 20000))
    assert bitdiff(M.selftest_math(0, a, b), a / b) == 0
    s = np.abs(a)
    assert bitdiff(M.selftest_math(1, s), np.sqrt(s)) == 0
    from oracle.oracle import load
    x = rng.uniform(-45, 45, 20000)
    ref = np.array([load().orc_exp(v) for v in x])
    assert bitdiff(M.selftest_math(2, x), ref) == 0


@pytest.mark.parametrize("hll", [False, True])
@pytest.mark.parametrize("bc_z_bot", [L.REFLECT, L.OUTFLOW])
def test_advect_lya_bit_exact(lut, hll, bc_z_bot):
    cloud = randomize(small_cloud(lut=lut, umax=300.0, theta_cone=np.pi / 4, bc_z_bot=bc_z_bot), seed=11)
    ctx, o = _pair(cloud, hll=hll)
    cdt = cloud.c_tilde * cloud.dt
    kr_o = o.find_active_bins()
    assert M.find_active_bins(ctx) == kr_o
    for bc_r_right, bc_z_top in ((L.OUTFLOW, L.OUTFLOW), (L.REFLECT, L.REFLECT)):
        o.advect_lya(cdt, bc_r_right=bc_r_right, bc_z_bot=bc_z_bot, bc_z_top=bc_z_top)
        M.advect_lya(cdt, ctx, bc_r_right=bc_r_right, bc_z_bot=bc_z_bot, bc_z_top=bc_z_top)
        _assert_state_equal(ctx, o, STATE + ("J_old", "H_r_old", "H_z_old"))


def test_emission_absorption_bit_exact(lut):
    cloud = randomize(small_cloud(lut=lut, tau_abs=2.0), seed=12)
    ctx, o = _pair(cloud)
    for _ in range(2):
        o.emission_absorption(cloud.c_tilde, cloud.dt)
        M.lya_emission_absorption_update(ctx, cloud.c_tilde, cloud.dt)
        _assert_state_equal(ctx, o, STATE + ("J_old", "H_r_old", "H_z_old"))
        for k in ("drhou_r", "drhou_z"):
            g = ctx.download(k)
            scale = np.max(np.abs(o[k]))
            assert np.max(np.abs(g - o[k])) <= RED_TOL * scale, k


@pytest.mark.parametrize("umax", [0.0, 800.0])
def test_frequency_update_bit_exact(lut, umax):
    cloud = randomize(small_cloud(lut=lut, n_nu=64, tau0=1e8, umax=umax), seed=13)
    ctx, o = _pair(cloud)
    cdt = cloud.c_tilde * cloud.dt
    o.update_freq_factors()
    M.update_freq_factors(ctx)
    for k in ("Fkp", "Fk", "Fkm"):
        assert bitdiff(ctx.download(k), o[k]) == 0, k
    o.frequency_update_all_cells(cdt, update_ff=True)
    M.frequency_update_j_lya_all_cells(cdt, ctx, update_freq_factors=True)
    _assert_state_equal(ctx, o, STATE + ("J_old", "H_r_old", "H_z_old"))
    # and with the stored factors (UpdateFreqFactors = false)
    o.frequency_update_all_cells(cdt, update_ff=False)
    M.frequency_update_j_lya_all_cells(cdt, ctx, update_freq_factors=False)
    _assert_state_equal(ctx, o, STATE)


def _interchanges(dl, d, du):
    """Row interchanges LAPACK dgtsv makes on one tridiagonal system (its pivot rule:
    swap rows i and i+1 when |d_i| < |dl_i| after the elimination of row i-1)."""
    d = d.copy()
    du = np.append(du.copy(), 0.0)
    swaps = 0
    for i in range(len(d) - 1):
        if abs(d[i]) >= abs(dl[i]):
            d[i + 1] -= dl[i] / d[i] * du[i]
        else:
            swaps += 1
            f = d[i] / dl[i]
            d[i + 1], du[i + 1] = du[i] - f * d[i + 1], -f * du[i + 1]
    return swaps


def test_frequency_solve_with_row_interchanges_bit_exact(lut):
    """Stored frequency factors that are no M-matrix, so dgtsv interchanges rows: the
    solve's interchange path, which physical factors never take, against the oracle."""
    cloud = randomize(small_cloud(lut=lut, n_nu=64, tau0=1e7), seed=21)
    ctx, o = _pair(cloud)
    cdt = cloud.c_tilde * cloud.dt
    dw = 1.0 / cdt
    shape = cloud.lya["J"].shape
    rng = np.random.default_rng(8)
    F = {"Fkp": dw * rng.uniform(-2.0, 2.0, shape), "Fk": dw * rng.uniform(-1.5, 0.5, shape),
         "Fkm": dw * rng.uniform(-2.0, 2.0, shape)}
    for k, v in F.items():
        ctx.upload(k, v)
        o.set(k, v)
    # rows 1..N_ν-2 of one cell's system, as the solve forms them
    Fkp, Fk, Fkm = (F[k][0, 0, 1:-1] for k in ("Fkp", "Fk", "Fkm"))
    assert _interchanges(Fkm[1:], -(Fk + dw), Fkp[:-1]) > 5
    o.frequency_update_all_cells(cdt, update_ff=False)
    M.frequency_update_j_lya_all_cells(cdt, ctx, update_freq_factors=False)
    _assert_state_equal(ctx, o, STATE)


def test_reductions_match(lut):
    cloud = randomize(small_cloud(lut=lut, n_nu=32), seed=14)
    ctx, o = _pair(cloud)
    o.emission_absorption(cloud.c_tilde, cloud.dt)
    M.lya_emission_absorption_update(ctx, cloud.c_tilde, cloud.dt)
    mo = o.update_f_mf_l(cloud.dt)
    mg = M.update_f_lya_mf_l_lya(ctx, cloud.dt, cloud.Zc, cloud.bc_z_bot)
    assert rel(mg, mo) < RED_TOL
    rng = np.random.default_rng(5)
    fg = rng.normal(size=(2, cloud.lya.n_z, cloud.lya.n_r))
    mo5 = o.update_f_mf_l_grav(cloud.dt, fg[0], fg[1])
    mg5 = M.update_f_lya_mf_l_lya(ctx, cloud.dt, cloud.Zc, cloud.bc_z_bot, fg[0], fg[1])
    assert rel(mg5, mo5) < RED_TOL
    for bc in (L.REFLECT, L.OUTFLOW):
        assert rel(M.compute_f_esc_lya_cylindrical(ctx, mo[3], bc), o.f_esc(mo[3], bc)) < RED_TOL
    assert M.compute_f_esc_lya_cylindrical(ctx, 0.0) == 0.0 == o.f_esc(0.0)
    assert rel(M.update_e_lya(ctx), o.update_e_lya()) < RED_TOL
    rg, ng = M.edge_spectra(ctx)
    ro, no = o.edge_spectra()
    assert bitdiff(rg, ro) == 0
    assert rel(ng, no) < RED_TOL


@pytest.mark.parametrize("case", ["static_cloud_32x32x64", "cone_dust_flow", "odd_grid_outflow",
                                  "reference_grid_194_bins", "reference_grid_542_bins"])
def test_fused_step_bit_exact(lut, case):
    if case == "static_cloud_32x32x64":
        cloud = L.static_cloud_32x32x64(lut=lut)
        nsteps = 40
    elif case == "reference_grid_542_bins":
        # the reference's q = 1.03 grid (542 bins, RunVelocityGradientSpectraTest.jl:25): nine bins
        # per lane, one cell column per wavefront
        from lydion_amd.static_cloud import setup_static_cloud
        cloud = randomize(setup_static_cloud(1e8, 1e4, tau_abs=0.5, rs=1.0, umax=400.0, recoil=False, n_r=10,
                                             n_z=8, nu=L.reference_freq_grid(30.0, 1.03, 1.03), lut=lut), seed=18)
        assert cloud.lya.n_nu == 542
        nsteps = 3
    elif case == "reference_grid_194_bins":
        # the reference's default frequency grid (q = 1.1, 194 bins; four bins per lane)
        from lydion_amd.static_cloud import setup_static_cloud
        cloud = randomize(setup_static_cloud(1e8, 1e4, tau_abs=0.5, rs=0.5 * (1.1 / 14), umax=50.0, n_r=14, n_z=12,
                                             nu=L.reference_freq_grid(), lut=lut), seed=17)
        nsteps = 4
    elif case == "odd_grid_outflow":
        # N_r and N_Z not multiples of the advection block (2 x 4 cells), open bottom
        cloud = randomize(small_cloud(n_r=13, n_z=11, n_nu=48, tau0=1e7, tau_abs=1.0, umax=100.0,
                                      bc_z_bot=L.OUTFLOW, lut=lut), seed=16)
        nsteps = 8
    else:
        cloud = randomize(small_cloud(n_r=24, n_z=20, n_nu=128, tau0=1e8, tau_abs=1.0, theta_cone=np.pi / 4,
                                      umax=200.0, lut=lut), seed=15)
        nsteps = 6
    ctx, o = _pair(cloud)
    for s in range(nsteps):
        so = o.step()
        sg = M.step(ctx, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
        assert (sg["kmin"], sg["kmax"]) == (so[5], so[6]), f"step {s}"
        for name, idx in (("M_F", 0), ("M_F_radial_net", 1), ("M_F_radial_out", 2), ("L_Lya", 3), ("f_esc", 4)):
            assert abs(sg[name] - so[idx]) <= RED_TOL * max(abs(so[idx]), 1e-300), (s, name, sg[name], so[idx])
    _assert_state_equal(ctx, o)
    for k in ("f_r", "f_z"):
        g = ctx.download(k)
        assert np.max(np.abs(g - o[k])) <= RED_TOL * np.max(np.abs(o[k])), k
    # after fused steps the *_old fields read back as the current state (J_old == J)
    assert bitdiff(ctx.download("J_old"), o["J_old"]) == 0


def test_fused_step_bit_exact_cone_cloud_256x256x128(lut):
    """The full-size cone cloud (τ0 = 1e8, τ_abs = 1, 128 bins) from a developed field: two fused
    steps bit-identical to the oracle, so the grid-stride column kernel and the advection blocks
    are compared at the benchmark's size."""
    from lydion_amd.static_cloud import developed_state
    cloud = developed_state(L.cone_dust_cloud_256x256x128(lut=lut))
    ctx, o = _pair(cloud)
    for s in range(2):
        so = o.step()
        sg = M.step(ctx, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
        assert (sg["kmin"], sg["kmax"]) == (so[5], so[6]), f"step {s}"
        for name, idx in (("M_F", 0), ("M_F_radial_net", 1), ("M_F_radial_out", 2), ("L_Lya", 3), ("f_esc", 4)):
            assert abs(sg[name] - so[idx]) <= RED_TOL * max(abs(so[idx]), 1e-300), (s, name, sg[name], so[idx])
    _assert_state_equal(ctx, o)
    ctx.close()


def test_run_steps_matches_single_steps(lut):
    cloud = L.static_cloud_32x32x64(lut=lut)
    c1 = M.M1LyaContext(cloud.lya, cloud.geometry)
    c2 = M.M1LyaContext(cloud.lya, cloud.geometry)
    hist = M.run_steps(c1, 12, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    rows = [M.step(c2, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot) for _ in range(12)]
    for s in range(12):
        assert hist[s, 0] == rows[s]["M_F"] and hist[s, 4] == rows[s]["f_esc"]
    assert bitdiff(c1.download("J"), c2.download("J")) == 0
