"""The mixed-precision batched step (m1lya_batch_* via M1LyaBatch) on the GPU.

Reference: the FP64 path (M1LyaContext), which is bit-identical to the CPU
oracle on these inputs (tests/test_gpu_parity.py), and the oracle itself on
the small cases.  Bar: M_F and f_esc within 1e-4 relative of the FP64 result
at every step compared, the stated target of the FP32 path; the frequency-
integrated J energy within 1e-4.  Structural properties are exact: a cloud's
state does not depend on which other clouds share its batch, on the tile
shape, on graph replay, or on an autotuning pass.
"""
import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import m1lya as M
from lydion_amd.batch import M1LyaBatch
from lydion_amd.static_cloud import developed_state
from oracle.oracle import OracleState
from tests import physics_cases as P
from tests.helpers import bitdiff, small_cloud

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4


@pytest.fixture(scope="module")
def lut():
    return L.VoigtLUT()


def _relerr(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _energy_err(cloud, J, Jref):
    w = cloud.lya["dnu"][None, None, :] * cloud.geometry.dV[:, :, None]
    return float(np.sum(np.abs(J - Jref) * w) / np.sum(np.abs(Jref) * w))


def _fp64_history(cloud, nsteps):
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    h = M.run_steps(ctx, nsteps, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    return h, ctx


def _check_history(hb, h64, what, start=0, fesc_min=0.0):
    """M_F at every step from `start`; f_esc once the FP64 value exceeds fesc_min."""
    for s in range(start, h64.shape[0]):
        for name, m in (("M_F", 0), ("f_esc", 4)):
            if m == 4 and not h64[s, m] > fesc_min:
                continue
            e = _relerr(hb[s, m], h64[s, m])
            assert e <= FP32_TOL, f"{what} step {s} {name}: fp32 {hb[s, m]:.9e} fp64 {h64[s, m]:.9e} rel {e:.2e}"


def test_static_cloud_32x32x64_vs_oracle(lut):
    """Plain static cloud (τ0 = 1e6, 64 bins) from its initial state, 200 steps against
    the oracle.  Until the radiation front reaches the box edge (about step 80), the
    escaping flux is at the level of the J floor (1e-30·J0 in FP32, Jeps = 1e-60 in
    FP64), so f_esc is compared from 1e-12 upwards; M_F from step 5."""
    cloud = L.static_cloud_32x32x64(lut=lut)
    o = OracleState(cloud)
    b = M1LyaBatch([cloud])
    hb = b.run_steps(200)[:, 0, :]
    ho = np.array([o.step() for _ in range(200)])
    assert ho[-1, 4] > 1e-6   # the front has reached the boundary
    _check_history(hb, ho, "static_cloud_32x32x64", start=5, fesc_min=1e-12)
    assert np.array_equal(hb[:, 5:7], ho[:, 5:7])   # same active bins every step
    assert _energy_err(cloud, b.download(0, "J"), o["J"]) <= FP32_TOL


@pytest.mark.parametrize("hll", [False, True])
def test_cone_flow_vs_fp64(lut, hll):
    """τ0 = 1e8, dust, cone source, Hubble outflow 200 km/s, 128 bins (two bins per lane)."""
    cloud = developed_state(small_cloud(n_r=24, n_z=20, n_nu=128, tau0=1e8, tau_abs=1.0, theta_cone=np.pi / 4,
                                        umax=200.0, lut=lut))
    cloud.lya.hll = hll
    h64, ctx = _fp64_history(cloud, 30)
    b = M1LyaBatch([cloud])
    hb = b.run_steps(30)[:, 0, :]
    _check_history(hb, h64, f"cone_flow hll={hll}")
    assert _energy_err(cloud, b.download(0, "J"), ctx.download("J")) <= FP32_TOL


def test_outflow_bottom_and_reflecting_right(lut):
    """Outflow Z boundary at the bottom and a reflecting outer r boundary."""
    cloud = developed_state(small_cloud(n_r=20, n_z=18, n_nu=64, tau0=1e7, tau_abs=0.5, bc_z_bot=L.OUTFLOW,
                                        lut=lut))
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    h64 = M.run_steps(ctx, 20, cloud.c_tilde, cloud.dt, cloud.Zc, bc_r_right=L.REFLECT, bc_z_bot=L.OUTFLOW)
    b = M1LyaBatch([cloud])
    b.set_step(0, cloud.c_tilde, cloud.dt, cloud.Zc, bc_r_right=L.REFLECT, bc_z_bot=L.OUTFLOW)
    hb = b.run_steps(20)[:, 0, :]
    _check_history(hb, h64, "outflow-bottom")


@pytest.mark.parametrize("hll,bc,n_nu,tiles", [(0, "reflect", 32, (16, 8)), (1, "outflow", 32, (7, 3)),
                                                (0, "outflow", 24, (5, 2)), (1, "reflect", 18, (1, 4))])
def test_pair_mode_vs_fp64(lut, hll, bc, n_nu, tiles):
    """Clouds of N_ν <= 32 run two cells per wavefront (k_step_pair): Hubble outflow of 200 km/s
    (upwind fluxes, D:∇u work), dust, a cone, an odd N_r and segment lengths that end a row
    with a half-filled wavefront; every step's M_F and f_esc, the J energy and the escaping
    spectrum against the FP64 step."""
    bcz = L.OUTFLOW if bc == "outflow" else L.REFLECT
    cloud = developed_state(small_cloud(n_r=19, n_z=13, n_nu=n_nu, tau0=1e7, tau_abs=1.0, theta_cone=np.pi / 4,
                                        umax=200.0, bc_z_bot=bcz, lut=lut))
    cloud.lya.hll = hll
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    h64 = M.run_steps(ctx, 25, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=bcz)
    b = M1LyaBatch([cloud])
    b.set_step(0, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=bcz)
    b.set_tiles(*tiles)
    b.set_profiling(1)
    hb = b.run_steps(25)[:, 0, :]
    assert "k_step_pair" in b.kernel_stats() and "k_step" not in b.kernel_stats()
    _check_history(hb, h64, f"pair hll={hll} bc={bc} n_nu={n_nu} tiles={tiles}")
    ee = _energy_err(cloud, b.download(0, "J"), ctx.download("J"))
    es = _spectrum_err(cloud, b, ctx)
    print(f"pair mode n_nu={n_nu} hll={hll} bc={bc} tiles={tiles}: J energy {ee:.2e}, spectrum {es:.2e}")
    assert ee <= FP32_TOL and es <= FP32_TOL
    ctx.close()
    b.close()


def _spectrum_err(cloud, b, ctx, idx=0):
    """Max relative error of the escaping spectrum (the per-bin terms of f_esc) over the bins
    above 1e-3 of its peak."""
    _, _, s32 = P.escaping_spectrum(cloud, b.download(idx, "H_r"), b.download(idx, "H_z"))
    _, _, s64 = P.escaping_spectrum(cloud, ctx.download("H_r"), ctx.download("H_z"))
    big = s64 > 1e-3 * s64.max()
    return float(np.max(np.abs(s32[big] - s64[big]) / s64[big]))


def test_cone_dust_cloud_256x256x128_vs_fp64(lut):
    """The full-size cloud: 8 steps from the developed state; M_F, f_esc, J energy and the
    escaping spectrum."""
    cloud = developed_state(L.cone_dust_cloud_256x256x128(lut=lut))
    h64, ctx = _fp64_history(cloud, 8)
    b = M1LyaBatch([cloud])
    hb = b.run_steps(8)[:, 0, :]
    _check_history(hb, h64, "cone_dust_256x256x128")
    assert _energy_err(cloud, b.download(0, "J"), ctx.download("J")) <= FP32_TOL
    es = _spectrum_err(cloud, b, ctx)
    print(f"cone 256x256x128: escaping spectrum max rel err {es:.2e}")
    assert es <= FP32_TOL


def _mixed_clouds(lut):
    return [
        developed_state(small_cloud(n_r=16, n_z=12, n_nu=32, tau0=1e5, tau_abs=0.0, lut=lut), seed=1),
        developed_state(small_cloud(n_r=9, n_z=21, n_nu=128, tau0=1e9, tau_abs=3.0, umax=800.0, lut=lut), seed=2),
        developed_state(small_cloud(n_r=30, n_z=7, n_nu=64, tau0=1e7, tau_abs=10.0, theta_cone=np.pi / 4, lut=lut),
                        seed=3),
        developed_state(small_cloud(n_r=5, n_z=40, n_nu=100, tau0=1e6, tau_abs=1.0, bc_z_bot=L.OUTFLOW, lut=lut),
                        seed=4),
    ]


def test_batch_equals_single_clouds(lut):
    """A cloud's state and scalars do not depend on the other clouds of its batch."""
    clouds = _mixed_clouds(lut)
    big = M1LyaBatch(clouds)
    hb = big.run_steps(6)
    for idx, cl in enumerate(clouds):
        one = M1LyaBatch([cl])
        h1 = one.run_steps(6)
        assert np.array_equal(hb[:, idx, :], h1[:, 0, :]), idx
        for k in ("J", "H_r", "H_z", "f_r", "f_z"):
            assert bitdiff(big.download(idx, k), one.download(0, k)) == 0, (idx, k)


def test_batch_with_sheared_cloud_equals_single_clouds(lut):
    """A batch holding a sheared-flow cloud runs the kernel variant that can build reference
    frequency rows per bin; every other cloud of that batch (static, cone flow, isotropic
    expansion with 128 bins) still gets bit for bit the result it gets alone."""
    from lydion_amd import workloads as W
    clouds = _mixed_clouds(lut) + [developed_state(W.rhd_like_cloud(24, 40, 32, lut=lut))]
    big = M1LyaBatch(clouds)
    hb = big.run_steps(6)
    for idx, cl in enumerate(clouds[:-1]):
        one = M1LyaBatch([cl])
        h1 = one.run_steps(6)
        assert np.array_equal(hb[:, idx, :], h1[:, 0, :]), idx
        for k in ("J", "H_r", "H_z", "f_r", "f_z"):
            assert bitdiff(big.download(idx, k), one.download(0, k)) == 0, (idx, k)
        one.close()
    big.close()


def test_mixed_batch_vs_fp64(lut):
    """Every cloud of an uneven batch (bins 32..128, τ0 1e5..1e9, outflow, dust) against its FP64 run."""
    clouds = _mixed_clouds(lut)
    b = M1LyaBatch(clouds)
    hb = b.run_steps(10)
    for idx, cl in enumerate(clouds):
        h64, _ = _fp64_history(cl, 10)
        _check_history(hb[:, idx, :], h64, f"cloud {idx}")


def test_tile_shape_does_not_change_state(lut):
    clouds = _mixed_clouds(lut)[:2]
    ref = M1LyaBatch(clouds)
    ref.set_tiles(16, 4)
    h_ref = ref.run_steps(4)
    for tiles in [(4, 1), (7, 2), (64, 3), (16, 8), (5, 6)]:
        b = M1LyaBatch(clouds)
        b.set_tiles(*tiles)
        h = b.run_steps(4)
        for idx in range(len(clouds)):
            assert bitdiff(b.download(idx, "J"), ref.download(idx, "J")) == 0, (tiles, idx)
        # reductions are summed in a different order: agreement to rounding
        assert np.max(np.abs(h[..., :5] - h_ref[..., :5]) / np.maximum(np.abs(h_ref[..., :5]), 1e-300)) < 1e-12


def test_graph_replay_matches_eager(lut):
    """hipGraph replays (one graph per buffer parity), with eager work and host
    synchronisation between replays, reproduce the eager steps exactly."""
    clouds = _mixed_clouds(lut)[:3]
    eager = M1LyaBatch(clouds)
    graph = M1LyaBatch(clouds)
    graph.set_graph(True)
    for rep in range(4):
        he = eager.run_steps(3)
        hg = graph.run_steps(3)
        assert np.array_equal(he, hg), rep
        graph.synchronize()
        _ = graph.scalars()
    for idx in range(len(clouds)):
        assert bitdiff(graph.download(idx, "J"), eager.download(idx, "J")) == 0


def test_autotune_restores_state(lut):
    clouds = _mixed_clouds(lut)[:2]
    a = M1LyaBatch(clouds)
    b = M1LyaBatch(clouds)
    s, w, ms = b.autotune(trial_steps=1)
    assert (s, w) == b.tiles() and ms > 0
    ha, hb = a.run_steps(3), b.run_steps(3)
    for idx in range(len(clouds)):
        assert bitdiff(a.download(idx, "J"), b.download(idx, "J")) == 0
    assert np.max(np.abs(ha[..., :5] - hb[..., :5]) / np.maximum(np.abs(ha[..., :5]), 1e-300)) < 1e-12


def test_scale_choice(lut):
    """J0 only sets the FP32 exponent range: results agree to FP32 rounding for J0 x 1e6."""
    cloud = developed_state(small_cloud(n_nu=64, tau0=1e7, lut=lut))
    a = M1LyaBatch([cloud])
    b = M1LyaBatch([cloud], scales=[a.scale(0) * 1e6])
    ha, hb = a.run_steps(10), b.run_steps(10)
    assert np.max(np.abs(ha[:, 0, 0] - hb[:, 0, 0]) / np.abs(ha[:, 0, 0])) < 1e-5
    assert _energy_err(cloud, b.download(0, "J"), a.download(0, "J")) < 1e-5


@pytest.fixture(scope="module")
def sweep(lut):
    from lydion_amd import workloads as W
    return W.sweep_clouds(64, lut=lut)


def test_sweep_of_64_clouds_vs_fp64(sweep):
    """The seeded 64-cloud sweep (sides 32..256, τ0 1e5..1e9, T 1e2..1e4 K, τ_abs 0..10,
    32..128 bins, point or uniform source) in one batch, from a developed state:
    every cloud's M_F and f_esc against its own FP64 run at each of 10 steps."""
    clouds = [developed_state(c) for c in sweep]
    b = M1LyaBatch(clouds)
    b.autotune(trial_steps=1)
    hb = b.run_steps(10)
    for idx, cl in enumerate(clouds):
        h64, ctx = _fp64_history(cl, 10)
        _check_history(hb[:, idx, :], h64, f"sweep cloud {idx}")
        assert _energy_err(cl, b.download(idx, "J"), ctx.download("J")) <= FP32_TOL, idx
        ctx.close()
    b.close()


def test_sweep_first_steps_from_rest(sweep):
    """The first 16 sweep clouds from their initial state (J = 0, sources on), 40 steps:
    M_F from step 5 on, f_esc once the FP64 value exceeds 1e-12."""
    from lydion_amd import workloads as W
    clouds = W.sweep_clouds(16)
    b = M1LyaBatch(clouds)
    hb = b.run_steps(40)
    for idx, cl in enumerate(clouds):
        h64, ctx = _fp64_history(cl, 40)
        _check_history(hb[:, idx, :], h64, f"sweep cloud {idx} from rest", start=5, fesc_min=1e-12)
        ctx.close()
    b.close()


def test_rhd_like_state_vs_fp64(lut):
    """A turbulent RHD-like state (48x96 r-Z, 64 bins: lognormal n_H, T from 1e2 to 1e4 K tied to
    density, Mach-2 velocity field, dust ∝ n_H, open Z ends) from a developed field: M_F and
    f_esc of every step of 20 against the FP64 step."""
    from lydion_amd import workloads as W
    cloud = developed_state(W.rhd_like_cloud(48, 96, 64, lut=lut))
    b = M1LyaBatch([cloud])
    hb = b.run_steps(20)
    h64, ctx = _fp64_history(cloud, 20)
    _check_history(hb[:, 0, :], h64, "rhd-like")
    assert _energy_err(cloud, b.download(0, "J"), ctx.download("J")) <= FP32_TOL
    ctx.close()
    b.close()


def test_rhd_like_state_full_size_vs_fp64(lut):
    """The RHD-like state at its full size (256x512 r-Z, 128 bins, FP32 state of 200 MB): 24
    substeps from a developed field, every step's M_F and f_esc, the J energy and the escaping
    spectrum against the FP64 step."""
    from lydion_amd import workloads as W
    cloud = developed_state(W.rhd_like_cloud(256, 512, 128, lut=lut))
    b = M1LyaBatch([cloud])
    hb = b.run_steps(24)
    h64, ctx = _fp64_history(cloud, 24)
    _check_history(hb[:, 0, :], h64, "rhd-like 256x512x128")
    ee = _energy_err(cloud, b.download(0, "J"), ctx.download("J"))
    es = _spectrum_err(cloud, b, ctx)
    print(f"rhd 256x512x128, 24 substeps: M_F rel {_relerr(hb[-1, 0, 0], h64[-1, 0]):.2e}, "
          f"f_esc rel {_relerr(hb[-1, 0, 4], h64[-1, 4]):.2e}, J energy {ee:.2e}, spectrum {es:.2e}")
    assert ee <= FP32_TOL and es <= FP32_TOL
    ctx.close()
    b.close()


def _reference_grid_cloud(lut, n_r=24, n_z=20, umax=200.0, tau0=1e8):
    """The reference's default frequency grid (q = 1.1, 194 bins; four bins per lane in the batch)."""
    from lydion_amd.static_cloud import setup_static_cloud
    return developed_state(setup_static_cloud(tau0, 1e4, tau_abs=1.0, rs=0.5 * (1.1 / n_r), umax=umax, n_r=n_r,
                                              n_z=n_z, nu=L.reference_freq_grid(), lut=lut))


def test_reference_grid_194_bins_vs_fp64(lut):
    """The 194-bin grid the reference's published timings use (RunStaticCloudTest.jl:46), in the
    four-bins-per-lane variant of the batched step: M_F, f_esc at every step and the J energy
    against the FP64 step, and against the oracle from the initial state."""
    cloud = _reference_grid_cloud(lut)
    assert cloud.lya.n_nu == 194
    h64, ctx = _fp64_history(cloud, 20)
    b = M1LyaBatch([cloud])
    hb = b.run_steps(20)[:, 0, :]
    _check_history(hb, h64, "reference grid 194 bins")
    assert np.array_equal(hb[:, 5:7], h64[:, 5:7])
    ee = _energy_err(cloud, b.download(0, "J"), ctx.download("J"))
    es = _spectrum_err(cloud, b, ctx)
    print(f"194 bins, 20 steps: M_F rel {_relerr(hb[-1, 0], h64[-1, 0]):.2e}, f_esc rel "
          f"{_relerr(hb[-1, 4], h64[-1, 4]):.2e}, J energy {ee:.2e}, spectrum {es:.2e}")
    assert ee <= FP32_TOL and es <= FP32_TOL
    ctx.close()
    b.close()


def test_wide_and_narrow_clouds_in_one_batch(lut):
    """A batch that holds a 194-bin cloud runs every cloud in the four-bins-per-lane variant; each
    cloud still matches its FP64 step, and each is bit-identical to its batch of one (the narrow
    one run there by the two-bins-per-lane variant)."""
    wide = _reference_grid_cloud(lut, n_r=13, n_z=17, umax=0.0, tau0=1e7)
    narrow = developed_state(small_cloud(n_r=16, n_z=12, n_nu=64, tau0=1e6, tau_abs=0.5, lut=lut), seed=3)
    b = M1LyaBatch([narrow, wide])
    hb = b.run_steps(10)
    for idx, cloud in enumerate((narrow, wide)):
        b1 = M1LyaBatch([cloud])
        h1 = b1.run_steps(10)
        for key in ("J", "H_r", "H_z"):
            assert bitdiff(b.download(idx, key), b1.download(0, key)) == 0, (idx, key)
        assert np.array_equal(hb[:, idx, :], h1[:, 0, :]), idx
        b1.close()
    for idx, cloud in enumerate((narrow, wide)):
        h64, ctx = _fp64_history(cloud, 10)
        _check_history(hb[:, idx, :], h64, f"mixed batch cloud {idx}")
        ctx.close()
    b.close()


def test_profiling_modes(lut):
    """Profiling level 2 brackets only the step kernels (two events per step); level 1 every
    launch.  A batch holding a cloud of N_ν <= 32 and a wider one launches both k_step_pair and
    k_step."""
    b = M1LyaBatch(_mixed_clouds(lut)[1:2])
    b.set_profiling(2)
    b.run_steps(3, history=False)
    st = b.kernel_stats()
    assert set(st) == {"k_step"} and st["k_step"][0] == 3
    b.reset_kernel_stats()
    b.set_profiling(1)
    b.run_steps(2, history=False)
    st = b.kernel_stats()
    assert {"k_step", "k_reduce", "k_final"} <= set(st) and all(v[0] == 2 for k, v in st.items() if k in ("k_step", "k_reduce", "k_final"))
    b.set_profiling(0)
    b.close()
    b = M1LyaBatch(_mixed_clouds(lut)[:2])
    b.set_profiling(2)
    b.run_steps(2, history=False)
    st = b.kernel_stats()
    assert set(st) == {"k_step", "k_step_pair"} and all(v[0] == 2 for v in st.values())
    b.close()
