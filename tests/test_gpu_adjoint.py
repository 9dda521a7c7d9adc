"""The reverse sweep of the FP64 fused step (m1lya_adjoint_run) on the GPU.

Every gradient is checked against central finite differences of the forward
FP64 step (itself bit-identical to the oracle, tests/test_gpu_parity.py):

* directional derivatives along smooth directions of the initial state and of
  every coefficient field the step reads, over five steps with a checkpoint
  interval that splits the run into uneven segments, for GLF and HLL fluxes,
  both bottom boundaries, warm and cold gas, point and uniform sources;
* per-cell gradients with respect to the H I density and the temperature of
  an expanding lognormal cloud (128×128 r-Z, 128 bins, τ0 = 1e7, Hubble
  outflow) on 32 seeded random cells, through the setup's dependence of α,
  α_F, ΔνD, Hvoigt and ϕ_ν on n_H and T, and the source luminosity.

Structural properties are exact: the checkpoint interval does not change the
gradient, and the context ends in the state run_steps leaves.
"""
import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import adjoint as A
from lydion_amd import m1lya as M
from lydion_amd import workloads as W
from lydion_amd.static_cloud import config_grid, developed_state, setup_static_cloud
from tests.helpers import bitdiff

pytestmark = pytest.mark.gpu

FD_TOL = 1e-5


@pytest.fixture(scope="module")
def lut():
    return L.VoigtLUT()


def _lsum(h, cloud):
    return h[0, 3] / (2.0 if cloud.bc_z_bot == L.REFLECT else 1.0)


def _objective(cloud, n, mf, sp, fields=None):
    """Objective of n fused steps with some context fields replaced."""
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    for k, v in (fields or {}).items():
        ctx.upload(k, v)
    h = M.run_steps(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    S = A.escaping_spectrum(cloud.geometry, ctx.download("H_r"), ctx.download("H_z"), _lsum(h, cloud),
                            cloud.bc_z_bot)
    ctx.close()
    return A.objective(h, S, mf, sp)


def _adjoint(cloud, n, mf, sp, checkpoint_every=0):
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    res = A.adjoint_run(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, mf, sp, bc_z_bot=cloud.bc_z_bot,
                        checkpoint_every=checkpoint_every)
    return res, ctx


def _test_cloud(lut, hll=False, bc_z_bot=L.REFLECT, T0=1e4, source="point", rest=False, n_nu=32):
    """12x10 cloud, 32 bins (or n_nu), τ0 = 1e6, dust, Hubble outflow of 200 km/s, destruction
    probability p_d up to 1e-3; a developed radiation field, or J = H = 0 (rest)."""
    rs = 0.5 * (1.1 / 12) if source == "point" else 1.0
    cloud = setup_static_cloud(1e6, T0, tau_abs=0.5, rs=rs, umax=200.0, n_r=12, n_z=10, bc_z_bot=bc_z_bot,
                               nu=config_grid(n_nu, 1e6, T_ref=T0), lut=lut)
    if not rest:
        cloud = developed_state(cloud)
    cloud.lya.hll = hll
    rng = np.random.Generator(np.random.PCG64(11))
    cloud.lya["pd"] = np.where(cloud.lya["alpha"] > 0, rng.uniform(0, 1e-3, cloud.lya["alpha"].shape), 0.0)
    return cloud


def _smooth_direction(base):
    """base times a smooth positive modulation in (j, i) and k: a smooth direction moves the two
    one-sided slopes of every limiter together, so the perturbed runs stay on the limiter
    branches of the base run."""
    nz, nr = base.shape[:2]
    jj, ii = np.meshgrid(np.arange(nz), np.arange(nr), indexing="ij")
    w = 1.0 + 0.5 * np.sin(0.7 * ii + 0.3) * np.cos(0.5 * jj + 0.1)
    if base.ndim == 3:
        w = w[..., None] * (1.0 + 0.3 * np.cos(0.2 * np.arange(base.shape[2])))[None, None, :]
    return base * w


FAMILIES = [("J", "state"), ("H_r", "state"), ("H_z", "state"), ("alpha_F", "bin"), ("Hvoigt", "bin"),
            ("phi", "bin"), ("alpha", "cell"), ("alpha_c", "cell"), ("js", "cell"), ("dnuD", "cell"),
            ("T", "cell")]
# The temperature enters the step directly only through the recoil term, a weak dependence:
# its central difference needs a larger step to rise above the rounding of the objective.
FD_STEP = {"T": 1e-4}


def _directional_check(cloud, n, families, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    mf = rng.normal(size=(n, 3))
    sp = rng.normal(size=cloud.lya.n_nu) * cloud.lya["dnu"] / cloud.lya["dnu"].mean()
    res, ctx = _adjoint(cloud, n, mf, sp, checkpoint_every=2)
    ctx.close()
    assert res.objective == pytest.approx(_objective(cloud, n, mf, sp), rel=1e-12)
    for key, kind in families:
        base = np.asarray(cloud.T if key == "T" else cloud.lya[key], dtype=np.float64)
        delta = _smooth_direction(base)
        if not np.any(delta):
            continue
        grad = {"state": res.grad_state, "bin": res.grad_bin, "cell": res.grad_cell}[kind][key]
        ad = float(np.sum(grad * delta))
        h = FD_STEP.get(key, 1e-5)
        fd = (_objective(cloud, n, mf, sp, {key: base + h * delta})
              - _objective(cloud, n, mf, sp, {key: base - h * delta})) / (2 * h)
        err = abs(fd - ad) / max(abs(fd), 1e-300)
        print(f"[adjoint fd] {key:8s} fd {fd:+.10e} adjoint {ad:+.10e} (|adj/obj| {abs(ad / res.objective):.1e}) "
              f"rel {err:.2e}")
        assert err <= FD_TOL, (key, fd, ad)


@pytest.mark.parametrize("case", ["warm-point-glf-reflect", "cold-uniform-hll-outflow"])
def test_directional_derivatives_vs_finite_differences(lut, case):
    """Initial state and every coefficient field, 5 steps (checkpoints every 2), smooth
    directions.  ϕ_ν is checked from rest (next test): from a developed field one step's
    emission barely moves the objective.  The direct temperature term (recoil, ∝ 1/T) is
    checked in cold gas; at 1e4 K it moves the objective by < 1e-6 of its value, below what a
    central difference resolves."""
    if case.startswith("warm"):
        cloud = _test_cloud(lut, hll=False, bc_z_bot=L.REFLECT, T0=1e4, source="point")
        families = [f for f in FAMILIES if f[0] not in ("phi", "T")]
    else:
        cloud = _test_cloud(lut, hll=True, bc_z_bot=L.OUTFLOW, T0=1e2, source="uniform")
        families = [f for f in FAMILIES if f[0] != "phi"]
    _directional_check(cloud, 5, families)


def test_directional_derivatives_160_bins(lut):
    """More than 130 bins: the reverse frequency solve runs on one lane (Thomas) instead of the
    wavefront's cyclic reduction."""
    cloud = _test_cloud(lut, hll=False, bc_z_bot=L.REFLECT, T0=1e4, source="point", n_nu=160)
    _directional_check(cloud, 3, [("J", "state"), ("alpha_F", "bin"), ("Hvoigt", "bin"), ("alpha", "cell")])


def test_emission_profile_gradient_from_rest(lut):
    """From J = H = 0 every photon of the run comes from emission: ϕ_ν, ∫js and the opacities
    carry the objective."""
    cloud = _test_cloud(lut, hll=True, bc_z_bot=L.OUTFLOW, T0=1e2, source="uniform", rest=True)
    _directional_check(cloud, 5, [f for f in FAMILIES if f[0] in ("phi", "js", "alpha_F", "Hvoigt", "dnuD")])


@pytest.mark.parametrize("case", ["glf-reflect-bottom", "hll-open-bottom", "glf-reflecting-walls"])
def test_reverse_interfaces_match_dual_passes(lut, case):
    """The default adjoint differentiates each interface once in reverse mode (m1a::face_vjp)
    and gathers its input cotangents per cell through the ghost maps; mode 1 takes 16
    dual-number passes per cell instead.  The two give the same gradient to rounding, for
    GLF and HLL fluxes and every boundary map (axis mirror, open and reflecting walls)."""
    run = {"bc_z_bot": L.REFLECT}
    if case == "glf-reflect-bottom":
        cloud = _test_cloud(lut, hll=False, bc_z_bot=L.REFLECT, T0=1e4, source="point")
    elif case == "hll-open-bottom":
        cloud = _test_cloud(lut, hll=True, bc_z_bot=L.OUTFLOW, T0=1e2, source="uniform")
        run = {"bc_z_bot": L.OUTFLOW}
    else:
        cloud = _test_cloud(lut, hll=False, bc_z_bot=L.OUTFLOW, T0=1e3, source="uniform")
        run = {"bc_z_bot": L.OUTFLOW, "bc_r_right": L.REFLECT, "bc_z_top": L.REFLECT}
    n = 4
    rng = np.random.Generator(np.random.PCG64(21))
    mf = rng.normal(size=(n, 3))
    sp = rng.normal(size=cloud.lya.n_nu)
    out = []
    for mode in (0, 1):
        ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
        ctx.set_adjoint_mode(mode)
        out.append(A.adjoint_run(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, mf, sp, checkpoint_every=2, **run))
        ctx.close()
    rev, dual = out
    assert rev.objective == dual.objective
    for group in ("grad_state", "grad_bin", "grad_cell"):
        for k, v in getattr(dual, group).items():
            w = getattr(rev, group)[k]
            scale = np.max(np.abs(v))
            err = np.max(np.abs(w - v)) / scale if scale > 0 else np.max(np.abs(w))
            print(f"[reverse vs dual] {case} {group} {k}: {err:.1e}")
            # two roundings of the same derivative: a transcription slip shows as O(1)
            assert err <= 1e-8, (case, group, k, err)


def test_checkpoint_interval_and_final_state(lut):
    cloud = _test_cloud(lut)
    n = 5
    rng = np.random.Generator(np.random.PCG64(5))
    mf = rng.normal(size=(n, 3))
    sp = rng.normal(size=cloud.lya.n_nu)
    ref, ctx = _adjoint(cloud, n, mf, sp, checkpoint_every=1)
    fwd = M.M1LyaContext(cloud.lya, cloud.geometry)
    h = M.run_steps(fwd, n, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    assert np.array_equal(h, ref.history)
    for k in ("J", "H_r", "H_z"):
        assert bitdiff(ctx.download(k), fwd.download(k)) == 0, k
    ctx.close()
    for every in (2, 3, 5, 0):
        res, c2 = _adjoint(cloud, n, mf, sp, checkpoint_every=every)
        c2.close()
        for group in ("grad_state", "grad_bin", "grad_cell"):
            for k, v in getattr(ref, group).items():
                assert bitdiff(getattr(res, group)[k], v) == 0, (every, group, k)


@pytest.fixture(scope="module", params=[10.0, 200.0, 800.0, 3200.0])
def expanding(request, lut):
    u_max = request.param
    return developed_state(W.expanding_cloud(128, 128, 128, tau0=1e7, u_max=u_max, sigma_ln=0.5, lut=lut))


def _cells(cloud, count=32, seed=17):
    """Seeded random cells inside the cloud, away from the grid edges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    inside = np.argwhere(cloud.lya["alpha"][2:-2, 2:-2] > 0) + 2
    return inside[rng.choice(len(inside), size=count, replace=False)]


def _fd_pair(f, x0, steps):
    """Central differences of f at x0 with the given steps."""
    return [(f(x0 + h) - f(x0 - h)) / (2 * h) for h in steps]


def _per_cell_check(label, cells, grad, fd_of_cell, x0, rel_steps, cloud, obj=0.0):
    """Adjoint vs central differences at two steps.  Where the two differences disagree by more
    than the tolerance a kink (limiter switch, a clamp of the exponential-fitting exponent) lies
    within the larger step: such cells are reported and left out.  On the others the adjoint
    must agree with the smaller-step difference to FD_TOL, relative to max(|grad|, 1e-3 max|grad|,
    the resolution of the difference): a central difference with step h of an objective of size
    |obj| cannot resolve derivatives below about 8 ε |obj| / h."""
    scale = np.max(np.abs(grad))
    eps = np.finfo(float).eps
    worst, used = 0.0, 0
    for (j, i) in cells:
        ad = grad[j, i]
        resolution = 8 * eps * abs(obj) / (rel_steps[1] * abs(x0[j, i])) / FD_TOL
        floor = max(abs(ad), 1e-3 * scale, resolution)
        fds = fd_of_cell(j, i, [r * x0[j, i] for r in rel_steps])
        smooth = abs(fds[0] - fds[1]) <= FD_TOL * floor
        e = abs(fds[1] - ad) / floor
        print(f"[adjoint fd] {cloud.params.get('u_max', 'rhd')} {label} cell ({j},{i}) fd {fds[1]:+.8e} "
              f"(larger step {fds[0]:+.8e}) adj {ad:+.8e} rel {e:.1e}{'' if smooth else '  [kink: skipped]'}")
        if smooth:
            used += 1
            worst = max(worst, e)
    return worst, used


def test_expanding_cloud_density_and_temperature_gradients(expanding, lut):
    """dM_F/dn_H(cell) after 6 steps and d(spectrum)/dT(cell) after 64 steps (the escaping
    spectrum needs that long to feel the cloud's interior) on 32 random cells, and dObj/dL,
    against central finite differences through the setup."""
    cloud = expanding
    nn = cloud.lya.n_nu
    cells = _cells(cloud)
    nH0 = A.density(cloud)
    T0 = np.asarray(cloud.T, dtype=np.float64)

    n_mf = 6
    mf = np.zeros((n_mf, 3))
    mf[-1, 0] = 1.0                                       # M_F after the last step
    res_mf, ctx = _adjoint(cloud, n_mf, mf, None)
    ctx.close()
    g_mf = A.physical_gradients(cloud, res_mf, lut)

    def fd_nH(j, i, steps):
        def f(x):
            nH = nH0.copy()
            nH[j, i] = x
            return _objective(cloud, n_mf, mf, None, A.cloud_with(cloud, n_H=nH, lut=lut))
        return _fd_pair(f, nH0[j, i], steps)

    worst_nH, used_nH = _per_cell_check("dMF/dnH", cells, g_mf["n_H"], fd_nH, nH0, (1e-4, 1e-5), cloud,
                                        obj=res_mf.objective)

    n_sp = 64
    k = np.arange(nn)
    sp = cloud.lya["dnu"] * (1.0 + 0.5 * np.sin(0.3 * k))  # a weighted emergent spectrum
    res_sp, ctx = _adjoint(cloud, n_sp, None, sp)
    ctx.close()
    g_sp = A.physical_gradients(cloud, res_sp, lut)

    def fd_T(j, i, steps):
        def f(x):
            T = T0.copy()
            T[j, i] = x
            return _objective(cloud, n_sp, None, sp, A.cloud_with(cloud, T=T, lut=lut))
        return _fd_pair(f, T0[j, i], steps)

    worst_T, used_T = _per_cell_check("dS/dT", cells, g_sp["T"], fd_T, T0, (1e-4, 1e-5), cloud,
                                      obj=res_sp.objective)

    hL = 1e-5
    fd = (_objective(cloud, n_mf, mf, None, A.cloud_with(cloud, L_scale=1 + hL, lut=lut))
          - _objective(cloud, n_mf, mf, None, A.cloud_with(cloud, L_scale=1 - hL, lut=lut))) / (2 * hL)
    ad = g_mf["L"] * cloud.params["L_lya"]
    eL = abs(fd - ad) / max(abs(fd), 1e-300)
    print(f"[adjoint fd] u_max={cloud.params['u_max']} dMF/dlnL fd {fd:+.8e} adj {ad:+.8e} rel {eL:.1e}; "
          f"n_H: worst {worst_nH:.1e} over {used_nH} cells; T: worst {worst_T:.1e} over {used_T} cells")
    assert used_nH >= 24 and used_T >= 24, (used_nH, used_T)
    assert worst_nH <= FD_TOL and worst_T <= FD_TOL and eL <= FD_TOL, (worst_nH, worst_T, eL)


def test_rhd_like_state_radial_force_gradient(lut):
    """The integrated radial force Σ_s M_F_radial_net(s) over 8 steps of a turbulent RHD-like
    state (48x96 r-Z, 64 bins; open Z ends, T tied to density, Mach-2 velocity field, dust ∝ n_H
    inside α_c and α_F): d/dn_H on 12 random cells and d/d ln L against central differences."""
    cloud = developed_state(W.rhd_like_cloud(48, 96, 64, lut=lut))
    n = 8
    mf = np.zeros((n, 3))
    mf[:, 1] = 1.0
    res, ctx = _adjoint(cloud, n, mf, None, checkpoint_every=3)
    ctx.close()
    g = A.physical_gradients(cloud, res, lut, with_T=False)
    nH0 = A.density(cloud)
    rng = np.random.Generator(np.random.PCG64(23))
    cells = np.argwhere(np.ones_like(nH0, dtype=bool)[2:-2, 2:-2]) + 2
    cells = cells[rng.choice(len(cells), size=12, replace=False)]

    def fd_nH(j, i, steps):
        def f(x):
            nH = nH0.copy()
            nH[j, i] = x
            return _objective(cloud, n, mf, None, A.cloud_with(cloud, n_H=nH, lut=lut))
        return _fd_pair(f, nH0[j, i], steps)

    worst, used = _per_cell_check("dF/dnH", cells, g["n_H"], fd_nH, nH0, (1e-4, 1e-5), cloud, obj=res.objective)
    hL = 1e-5
    fd = (_objective(cloud, n, mf, None, A.cloud_with(cloud, L_scale=1 + hL, lut=lut))
          - _objective(cloud, n, mf, None, A.cloud_with(cloud, L_scale=1 - hL, lut=lut))) / (2 * hL)
    ad = g["L"] * cloud.params["L_lya"]
    eL = abs(fd - ad) / max(abs(fd), 1e-300)
    print(f"[adjoint fd] rhd dF/dlnL fd {fd:+.8e} adj {ad:+.8e} rel {eL:.1e}; n_H worst {worst:.1e} over {used}")
    assert used >= 9 and worst <= FD_TOL and eL <= FD_TOL, (worst, used, eL)


def test_rhd_like_state_full_size_radial_force_gradient(lut):
    """The integrated radial force Σ_s M_F_radial_net(s) over 4 substeps of the RHD-like state at
    its full size (256x512 r-Z, 128 bins), checkpointed every 2 substeps: d/dn_H on 4 random cells
    and d/d ln L against central differences."""
    cloud = developed_state(W.rhd_like_cloud(256, 512, 128, lut=lut))
    n = 4
    mf = np.zeros((n, 3))
    mf[:, 1] = 1.0
    res, ctx = _adjoint(cloud, n, mf, None, checkpoint_every=2)
    ctx.close()
    g = A.physical_gradients(cloud, res, lut, with_T=False)
    obj = res.objective
    del res
    nH0 = A.density(cloud)
    rng = np.random.Generator(np.random.PCG64(29))
    cells = np.argwhere(np.ones_like(nH0, dtype=bool)[2:-2, 2:-2]) + 2
    cells = cells[rng.choice(len(cells), size=4, replace=False)]

    def fd_nH(j, i, steps):
        def f(x):
            nH = nH0.copy()
            nH[j, i] = x
            return _objective(cloud, n, mf, None, A.cloud_with(cloud, n_H=nH, lut=lut))
        return _fd_pair(f, nH0[j, i], steps)

    worst, used = _per_cell_check("dF/dnH", cells, g["n_H"], fd_nH, nH0, (1e-4, 1e-5), cloud, obj=obj)
    hL = 1e-5
    fd = (_objective(cloud, n, mf, None, A.cloud_with(cloud, L_scale=1 + hL, lut=lut))
          - _objective(cloud, n, mf, None, A.cloud_with(cloud, L_scale=1 - hL, lut=lut))) / (2 * hL)
    ad = g["L"] * cloud.params["L_lya"]
    eL = abs(fd - ad) / max(abs(fd), 1e-300)
    print(f"[adjoint fd] rhd 256x512x128 dF/dlnL fd {fd:+.8e} adj {ad:+.8e} rel {eL:.1e}; "
          f"n_H worst {worst:.1e} over {used}")
    assert used >= 3 and worst <= FD_TOL and eL <= FD_TOL, (worst, used, eL)
