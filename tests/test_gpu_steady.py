"""Steady-state stop and non-finite detection on the GPU.

The RT-only static cloud of 32×32 cells, 64 bins, τ0 = 1e6 (recoil on, point source) is run
until ΔJ/J = Σ Δν ΔV |J_new − J_start| / Σ Δν ΔV max(J_new − Jeps, 0) < 1e-8: the FP64 context
must stop at the same step as the oracle run to the same rule, with a bit-identical state;
the FP32 batch is run to its own stop and compared with the FP64 steady state (M_F, f_esc,
the escaping spectrum) at the 1e-4 bar.

Non-finite values are data, not faults: a NaN written into J must surface as
M1LYA_ERR_NONFINITE from the fused FP64 step and from the batch.
"""
import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import constants as K
from lydion_amd import m1lya as M
from lydion_amd._lib import ERR_NONFINITE, M1LyaError
from lydion_amd.batch import M1LyaBatch
from oracle.oracle import OracleState
from tests import physics_cases as P
from tests.helpers import bitdiff, small_cloud

pytestmark = pytest.mark.gpu

TOL = 1e-8


@pytest.fixture(scope="module")
def lut():
    return L.VoigtLUT()


def _rel_change(w, J, Jp):
    return float(np.sum(w * np.abs(J - Jp)) / np.sum(w * np.maximum(J - K.Jeps, 0.0)))


@pytest.fixture(scope="module")
def steady(lut):
    """(cloud, FP64 context history, oracle state, oracle step count)."""
    cloud = L.static_cloud_32x32x64(lut=lut)
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    h = M.run_until_steady(ctx, 40000, TOL, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    o = OracleState(cloud)
    w = cloud.geometry.dV[:, :, None] * cloud.lya["dnu"][None, None, :]
    Jp = o["J"].copy()
    n, rel = 0, 1.0
    while rel >= TOL and n < 40000:
        o.step()
        n += 1
        J = o["J"]
        rel = _rel_change(w, J, Jp)
        Jp = J.copy()
    return cloud, ctx, h, o, n, rel


def test_fp64_stops_with_the_oracle(steady):
    cloud, ctx, h, o, n, rel = steady
    print(f"steady after {h.shape[0]} steps (oracle {n}); ΔJ/J {h[-1, 7]:.3e} (oracle {rel:.3e}); "
          f"M_F {h[-1, 0]:.6e}, f_esc {h[-1, 4]:.6e}")
    assert h.shape[0] == n
    assert h[-1, 7] < TOL and h[-2, 7] >= TOL
    assert abs(h[-1, 7] - rel) <= 1e-10 * rel
    for key, name in (("J", "J"), ("H_r", "H_r"), ("H_z", "H_z")):
        assert bitdiff(ctx.download(key), o[name]) == 0, key


def test_fp64_steady_history_matches_run_steps(steady, lut):
    """The stop changes nothing before it: the steps it ran are m1lya_run_steps' steps."""
    cloud, ctx, h, _, _, _ = steady
    ctx2 = M.M1LyaContext(cloud.lya, cloud.geometry)
    n = 300
    h2 = M.run_steps(ctx2, n, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    assert np.array_equal(h[:n, :7], h2[:, :7])
    assert np.all(h2[:, 7] == 0.0)


def test_fp32_steady_state_vs_fp64(steady, lut):
    cloud, ctx, h, _, _, _ = steady
    b = M1LyaBatch([L.static_cloud_32x32x64(lut=lut)])
    hb, st = b.run_until_steady(20000, TOL)
    print(f"FP32 batch: ΔJ/J < {TOL} after {st[0]} steps (FP64: {h.shape[0]}); ΔJ/J of its last 5 steps "
          f"{hb[-5:, 0, 7]}")
    assert st[0] > 0 and st[0] == hb.shape[0], "the FP32 state's rounding kept ΔJ/J above the tolerance"
    # FP64 state at the same step count: continue or truncate the FP64 run there
    ref = M.M1LyaContext(cloud.lya, cloud.geometry)
    h64 = M.run_steps(ref, int(st[0]), cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    for name, m in (("M_F", 0), ("f_esc", 4)):
        e = abs(hb[-1, 0, m] - h64[-1, m]) / abs(h64[-1, m])
        print(f"{name}: fp32 {hb[-1, 0, m]:.9e} fp64 {h64[-1, m]:.9e} rel {e:.2e}")
        assert e <= 1e-4
    _, _, S32 = P.escaping_spectrum(cloud, b.download(0, "H_r"), b.download(0, "H_z"))
    _, _, S64 = P.escaping_spectrum(cloud, ref.download("H_r"), ref.download("H_z"))
    big = S64 > 1e-3 * S64.max()
    es = float(np.max(np.abs(S32[big] - S64[big]) / S64[big]))
    print(f"escaping spectrum (bins above 1e-3 of the peak): max rel err {es:.2e}")
    assert es <= 1e-4
    # the FP64 state is itself steady there: both stops agree on the solution
    e = abs(h64[-1, 0] - h[-1, 0]) / h[-1, 0]
    assert e <= 1e-4


def test_fp64_nonfinite_state_is_reported(lut):
    cloud = small_cloud(lut=lut)
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    M.step(ctx, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    J = ctx.download("J")
    J[3, 4, 10] = np.nan
    ctx.upload("J", J)
    with pytest.raises(M1LyaError) as ei:
        M.step(ctx, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    assert ei.value.code == ERR_NONFINITE and "non-finite" in str(ei.value)
    # a step queued without synchronisation reports at the next synchronising call
    J[3, 4, 10] = np.inf
    ctx.upload("J", J)
    rc = ctx.lib.m1lya_step(ctx.handle, cloud.c_tilde, cloud.dt, cloud.Zc, L.OUTFLOW, cloud.bc_z_bot, L.OUTFLOW,
                            None)
    assert rc == 0
    with pytest.raises(M1LyaError) as ei:
        ctx.synchronize()
    assert ei.value.code == ERR_NONFINITE
    # the flag is cleared once reported
    ctx.synchronize()


def test_batch_nonfinite_state_is_reported(lut):
    clouds = [small_cloud(lut=lut), small_cloud(n_r=16, n_z=12, n_nu=64, lut=lut)]
    b = M1LyaBatch(clouds)
    b.run_steps(2)
    Hr = b.download(1, "H_r")
    Hr[5, 7, 20] = np.nan
    b.upload(1, "H_r", Hr)
    with pytest.raises(M1LyaError) as ei:
        b.run_steps(3)
    assert ei.value.code == ERR_NONFINITE and "cloud 1" in str(ei.value)


def test_fp64_profiling_levels(lut):
    """Profiling level 1 brackets every launch (the fused step's four, and the active-bin
    search before the first step), level 2 only the fused step's column kernel; neither
    changes the state."""
    cloud = small_cloud(lut=lut)
    step_kernels = {"advect", "column", "reduce_cols", "step_final"}
    states = []
    for level in (0, 1, 2):
        ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
        ctx.set_profiling(level)
        M.run_steps(ctx, 3, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot, history=False)
        st = ctx.kernel_stats()
        if level == 0:
            assert st == {}
        elif level == 1:
            assert set(st) == step_kernels | {"active_partial", "active_final"}
            # reduce_cols also sums the active-bin search's partials
            assert {k: st[k][0] for k in step_kernels} == {"advect": 3, "column": 3, "reduce_cols": 4, "step_final": 3}
        else:
            assert set(st) == {"column"} and st["column"][0] == 3
        assert all(v[1] > 0.0 for v in st.values())
        ctx.set_profiling(0)
        states.append(ctx.download("J"))
        ctx.close()
    assert bitdiff(states[0], states[1]) == 0 and bitdiff(states[0], states[2]) == 0
