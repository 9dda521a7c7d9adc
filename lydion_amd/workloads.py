"""Seeded multi-cloud workloads for the batched step.

The reference runs its parameter studies one cloud at a time, each through
SetupStaticCloudTest (src/StaticCloudTest.jl:92-322) in a host loop over
τ_cl, r_s or u_max (RunMFStaticCloudTest.jl, RunSourceSizeStaticCloudTest.jl,
RunVelocityGradientSpectraTest.jl).  The sweep here draws every parameter of
such a study at random, so that one batch mixes grid sizes, optical depths,
temperatures, dust and frequency resolutions:

    N_r, N_Z   uniform in {32, 64, 128, 256}, independently
    τ0         log-uniform in [1e5, 1e9]
    T          log-uniform in [1e2, 1e4] K (uniform within the cloud)
    τ_abs      uniform in [0, 10]
    N_ν        uniform in {32, 64, 128}, spanning ±4 (a(T) τ0)^{1/3} Doppler widths at T
    source     a central point source of radius 0.5 cells, or uniform in the cloud (r_s = 1)

Cloud i is drawn from its own generator PCG64(seed + i), so a cloud does not
change when the sweep grows or shrinks.
"""
from __future__ import annotations

from .lya_arrays import VoigtLUT
from .static_cloud import DEFAULT_SEED, StaticCloud, config_grid, setup_static_cloud

import numpy as np

SIDES = (32, 64, 128, 256)
BINS = (32, 64, 128)


def sweep_params(n_clouds: int = 64, seed: int = DEFAULT_SEED) -> list[dict]:
    """The drawn parameters of the first n_clouds clouds of the sweep."""
    out = []
    for i in range(n_clouds):
        rng = np.random.Generator(np.random.PCG64(seed + i))
        n_r = int(rng.choice(SIDES))
        n_z = int(rng.choice(SIDES))
        tau0 = float(10.0 ** rng.uniform(5.0, 9.0))
        T0 = float(10.0 ** rng.uniform(2.0, 4.0))
        tau_abs = float(rng.uniform(0.0, 10.0))
        n_nu = int(rng.choice(BINS))
        source = "point" if rng.integers(0, 2) == 0 else "uniform"
        out.append(dict(index=i, seed=seed + i, n_r=n_r, n_z=n_z, tau0=tau0, T=T0, tau_abs=tau_abs, n_nu=n_nu,
                        source=source))
    return out


def sweep_cloud(p: dict, lut: VoigtLUT | None = None) -> StaticCloud:
    """The static cloud of one parameter set of sweep_params."""
    rs = 0.5 * (1.1 / p["n_r"]) if p["source"] == "point" else 1.0
    return setup_static_cloud(p["tau0"], p["T"], tau_abs=p["tau_abs"], rs=rs, recoil=True, tmax_unitless=4.0,
                              n_r=p["n_r"], n_z=p["n_z"], nu=config_grid(p["n_nu"], p["tau0"], T_ref=p["T"]),
                              lut=lut)


def sweep_clouds(n_clouds: int = 64, seed: int = DEFAULT_SEED, lut: VoigtLUT | None = None) -> list[StaticCloud]:
    """The first n_clouds clouds of the seeded sweep, set up as SetupStaticCloudTest does."""
    lut = lut or VoigtLUT()
    return [sweep_cloud(p, lut) for p in sweep_params(n_clouds, seed)]


def updates_per_step(clouds) -> int:
    """Cell-frequency updates in one step of every cloud."""
    return int(sum(c.lya.n_nu * c.lya.n_r * c.lya.n_z for c in clouds))


def gaussian_field(n_z: int, n_r: int, slope: float = -11.0 / 3.0, seed: int = DEFAULT_SEED) -> np.ndarray:
    """Zero-mean, unit-variance Gaussian random field on the (N_Z, N_r) grid with power
    spectrum P(k) ∝ k^slope (white noise filtered by k^(slope/2) in Fourier space)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kz = np.fft.fftfreq(n_z)[:, None]
    kr = np.fft.fftfreq(n_r)[None, :]
    k = np.hypot(kz, kr)
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** (0.5 * slope)
    g = np.fft.ifft2(np.fft.fft2(rng.normal(size=(n_z, n_r))) * amp).real
    return (g - g.mean()) / g.std()


def lognormal_density(n_z: int, n_r: int, sigma_ln: float, slope: float = -11.0 / 3.0,
                      seed: int = DEFAULT_SEED) -> np.ndarray:
    """ρ/ρ̄ with ln ρ a Gaussian random field of standard deviation sigma_ln and P(k) ∝ k^slope,
    normalised to unit mean over the grid."""
    rho = np.exp(sigma_ln * gaussian_field(n_z, n_r, slope, seed))
    return rho / rho.mean()


def expanding_cloud(n_r: int = 128, n_z: int = 128, n_nu: int = 128, tau0: float = 1e7, T0: float = 1e4,
                    u_max: float = 200.0, sigma_ln: float = 0.5, seed: int = DEFAULT_SEED,
                    lut: VoigtLUT | None = None) -> StaticCloud:
    """Expanding cloud: a point-source sphere of line-centre optical depth τ0 in a Hubble flow of
    edge speed u_max km/s (GetHubbleFlow, src/StaticCloudTest.jl:34-59), its H I density
    multiplied by a lognormal field (σ_lnρ = sigma_ln, P(k) ∝ k^-11/3, unit mean)."""
    from .static_cloud import point_source_cloud
    cloud = point_source_cloud(n_r, n_z, n_nu, tau0, T0, umax=u_max, lut=lut)
    rho = lognormal_density(n_z, n_r, sigma_ln, seed=seed)
    L = cloud.lya
    L["alpha"] = L["alpha"] * rho
    L["alpha_F"] = L["alpha"][:, :, None] * L["Hvoigt"]
    cloud.params.update(sigma_ln=sigma_ln, seed=seed, u_max=u_max)
    return cloud
