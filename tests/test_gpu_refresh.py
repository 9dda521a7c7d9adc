"""Device-side coefficient refresh (m1lya_prepare_subcycling) on the GPU.

The device evaluates the Voigt tables, the bin-averaged profile Htophat, ϕ_ν,
α = n_HI σ0(T) and α_F = α Hvoigt + α_c,F in the operation order of the host
setup (lydion_amd.lya_arrays, itself bit-identical to the oracle's
orc_update_hvoigt / orc_update_phi, tests/test_oracle_pins.py): the results
must agree bit for bit, for per-cell temperatures across 1e2..1e4 K and
both ϕ_ν normalisations (core bins, Update_ϕν!; all bins, the RHD form).
"""
import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import m1lya as M
from lydion_amd.lya_arrays import compute_a, compute_dnuD, compute_sigma0, update_hvoigt, update_phi
from tests.helpers import bitdiff, small_cloud

pytestmark = pytest.mark.gpu


def _phi_all(dnu, Hv):
    flat = Hv.reshape(-1, Hv.shape[-1])
    s = np.zeros(flat.shape[0])
    for k in range(Hv.shape[-1]):
        s = s + dnu[k] * flat[:, k]
    return (flat * (1.0 / s)[:, None]).reshape(Hv.shape)


@pytest.mark.parametrize("phi_all_bins", [False, True])
def test_refresh_matches_host_setup(phi_all_bins):
    lut = L.VoigtLUT()
    cloud = small_cloud(n_r=14, n_z=11, n_nu=64, tau0=1e7, lut=lut)
    lya = cloud.lya
    rng = np.random.Generator(np.random.PCG64(41))
    T = 10.0 ** rng.uniform(2.0, 4.0, (lya.n_z, lya.n_r))
    nHI = 10.0 ** rng.uniform(-2.0, 3.0, (lya.n_z, lya.n_r))
    acF = rng.uniform(0.0, 1e-19, (lya.n_z, lya.n_r))
    ac = rng.uniform(0.0, 1e-19, (lya.n_z, lya.n_r))
    ctx = M.M1LyaContext(lya, cloud.geometry)
    ctx.set_voigt_lut(lut)
    ctx.prepare_subcycling(T, nHI, alpha_c=ac, alpha_c_F=acF, phi_all_bins=phi_all_bins)
    dnuD = compute_dnuD(T)
    Hv = update_hvoigt(lut, lya["dnu_rel"], lya["dnuC_rel"], compute_a(T), 1.0 / dnuD)
    phi = _phi_all(lya["dnu"], Hv) if phi_all_bins else update_phi(lya["dnu"], Hv, *lya.kcore)
    alpha = nHI * compute_sigma0(T)
    aF = alpha[..., None] * Hv + acF[..., None]
    for key, ref in (("Hvoigt", Hv), ("phi", phi), ("alpha", alpha), ("alpha_F", aF), ("dnuD", dnuD), ("T", T),
                     ("alpha_c", ac)):
        assert bitdiff(ctx.download(key), ref) == 0, key
    ctx.close()


def test_refreshed_step_equals_host_setup_step():
    """A fused step after the device refresh equals one after uploading the host's tables."""
    lut = L.VoigtLUT()
    cloud = L.static_cloud.developed_state(small_cloud(n_r=12, n_z=10, n_nu=32, tau0=1e6, lut=lut))
    lya = cloud.lya
    T = np.full((lya.n_z, lya.n_r), 3e3)
    nHI = lya["alpha"] / compute_sigma0(cloud.T)
    a = M.M1LyaContext(lya, cloud.geometry)
    a.set_voigt_lut(lut)
    a.prepare_subcycling(T, nHI, phi_all_bins=False)
    b = M.M1LyaContext(lya, cloud.geometry)
    dnuD = compute_dnuD(T)
    Hv = update_hvoigt(lut, lya["dnu_rel"], lya["dnuC_rel"], compute_a(T), 1.0 / dnuD)
    alpha = nHI * compute_sigma0(T)
    for key, v in (("T", T), ("dnuD", dnuD), ("Hvoigt", Hv), ("phi", update_phi(lya["dnu"], Hv, *lya.kcore)),
                   ("alpha", alpha), ("alpha_F", alpha[..., None] * Hv)):
        b.upload(key, v)
    ha = M.run_steps(a, 3, cloud.c_tilde, cloud.dt, cloud.Zc)
    hb = M.run_steps(b, 3, cloud.c_tilde, cloud.dt, cloud.Zc)
    assert np.array_equal(ha, hb)
    for key in ("J", "H_r", "H_z"):
        assert bitdiff(a.download(key), b.download(key)) == 0, key


def test_batch_refresh_equals_host_uploads():
    """The batch's device refresh (two uneven clouds, RHD-form ϕ_ν, dust flux opacity) gives the
    same steps, bit for bit, as uploading the host's tables."""
    from lydion_amd.batch import M1LyaBatch
    lut = L.VoigtLUT()
    clouds = [L.static_cloud.developed_state(small_cloud(n_r=12, n_z=10, n_nu=32, tau0=1e6, lut=lut)),
              L.static_cloud.developed_state(small_cloud(n_r=9, n_z=14, n_nu=64, tau0=1e7, umax=200.0, lut=lut))]
    rng = np.random.Generator(np.random.PCG64(5))
    a = M1LyaBatch(clouds)
    b = M1LyaBatch(clouds)
    a.set_voigt_lut(lut)
    for idx, cl in enumerate(clouds):
        lya = cl.lya
        T = 10.0 ** rng.uniform(2.0, 4.0, (lya.n_z, lya.n_r))
        nHI = lya["alpha"] / compute_sigma0(cl.T)
        acF = rng.uniform(0.0, 1e-20, (lya.n_z, lya.n_r))
        a.prepare_subcycling(idx, T, nHI, alpha_c=acF, alpha_c_F=acF, phi_all_bins=True)
        dnuD = compute_dnuD(T)
        Hv = update_hvoigt(lut, lya["dnu_rel"], lya["dnuC_rel"], compute_a(T), 1.0 / dnuD)
        alpha = nHI * compute_sigma0(T)
        for key, v in (("T", T), ("dnuD", dnuD), ("Hvoigt", Hv), ("phi", _phi_all(lya["dnu"], Hv)),
                       ("alpha", alpha), ("alpha_F", alpha[..., None] * Hv + acF[..., None]), ("alpha_c", acF)):
            b.upload(idx, key, v)
    ha, hb = a.run_steps(4), b.run_steps(4)
    assert np.array_equal(ha, hb)
    for idx in range(len(clouds)):
        for key in ("J", "H_r", "H_z"):
            assert bitdiff(a.download(idx, key), b.download(idx, key)) == 0, (idx, key)
