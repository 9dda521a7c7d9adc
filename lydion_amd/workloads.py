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


# Dust absorption opacity at Lyα per gram of dust, a recorded constant of this generator:
# an extinction cross-section of 1.6e-21 cm² per H for Milky-Way dust at 1216 Å with a dust
# mass of 0.01 x 1.4 m_H per H (κ_ext ≈ 6.8e4 cm²/g), times an absorbed fraction of 0.35.
KAPPA_ABS_LYA_DUST = 2.4e4
MU_H = 1.4      # gas mass per H nucleus, in m_H
MU_GAS = 1.27   # mean particle mass of neutral gas, in m_H (isothermal sound speed)


def rhd_like_cloud(n_r: int = 256, n_z: int = 512, n_nu: int = 128, n_mean: float = 1e3, sigma_ln: float = 1.0,
                   mach: float = 2.0, s_T: float | None = None, dust_to_gas: float = 0.01, metallicity: float = 0.1,
                   seed: int = DEFAULT_SEED, lut: VoigtLUT | None = None) -> StaticCloud:
    """An RHD-like Lyα state on the r-Z grid of the static clouds (R_max = Z_max = 1.1 pc),
    open at both Z ends (the RHD runs' :Outflow bottom), filled with turbulent gas:

      n_H = n_mean ρ,  ln ρ Gaussian with σ_lnρ = sigma_ln and P(k) ∝ k^-11/3 (unit-mean ρ);
      T   = clip(1e4 K (n_H / min n_H)^-s_T, 1e2 K, 1e4 K), s_T (when not given) such that the
            densest cell sits at 1e2 K;
      u_r, u_Z Gaussian random fields (P(k) ∝ k^-11/3) with |u|_rms = mach × the isothermal
          sound speed at 1e4 K, u_r tapered to zero on the axis; u* on the faces by averaging
          the two cells, the velocity gradients by centred differences;
      α_c = κ_abs,Lyα × dust_to_gas × metallicity × μ m_H n_H, and α_F = α Hvoigt + α_c;
      a central point source of 1e5 L☉ in the cell nearest to (r, Z) = (0, Z_max / 2).

    The frequency grid spans ±4 (a τ0)^{1/3} Doppler widths at 1e4 K for the mean line-centre
    optical depth τ0 = n_mean σ0(1e4 K) R_max."""
    from . import constants as K
    from ._lib import OUTFLOW
    from .grid import setup_cylindrical_grid
    from .lya_arrays import compute_a, compute_sigma0, generate_lya_arrays
    from .static_cloud import cfl_time_step

    Rmax = 1.1 * K.pc
    Zmax = Rmax
    Zc = 0.5 * Zmax
    geom = setup_cylindrical_grid(Rmax, Zmax, n_r, n_z)
    rho = lognormal_density(n_z, n_r, sigma_ln, seed=seed)
    n_H = n_mean * rho
    if s_T is None:
        s_T = float(np.log(100.0) / np.log(n_H.max() / n_H.min()))
    T = np.clip(1e4 * (n_H / n_H.min()) ** (-s_T), 1e2, 1e4)
    tau0 = n_mean * compute_sigma0(1e4) * Rmax
    nu = config_grid(n_nu, tau0, T_ref=1e4)
    L = generate_lya_arrays(K.c, geom.r, geom.Z, T, nu, recoil=True, lut=lut)
    L["alpha"] = n_H * compute_sigma0(T)
    L["alpha_c"] = KAPPA_ABS_LYA_DUST * dust_to_gas * metallicity * MU_H * K.mH * n_H
    L["alpha_F"] = L["alpha"][:, :, None] * L["Hvoigt"] + L["alpha_c"][:, :, None]
    L["pd"] = np.zeros((n_z, n_r))
    L_lya = 1e5 * K.Lsun
    js = np.zeros((n_z, n_r))
    j0 = int(np.argmin(np.abs(geom.ZCoM - Zc)))
    js[j0, 0] = L_lya / (K.Four_pi * geom.dV[j0, 0])
    L["js"] = js
    # turbulent velocity
    c_s = np.sqrt(K.kB * 1e4 / (MU_GAS * K.mH))
    rc, zc = geom.rCoM, geom.ZCoM
    dr = geom.r[1] - geom.r[0]
    u_r = gaussian_field(n_z, n_r, seed=seed + 1) * (1.0 - np.exp(-rc[None, :] / (2.0 * dr)))
    u_Z = gaussian_field(n_z, n_r, seed=seed + 2)
    scale = mach * c_s / np.sqrt(np.mean(u_r * u_r + u_Z * u_Z))
    u_r, u_Z = u_r * scale, u_Z * scale
    durdz, durdr = np.gradient(u_r, zc, rc)
    duzdz, duzdr = np.gradient(u_Z, zc, rc)
    ur_star = np.zeros((n_z, n_r + 1))
    ur_star[:, 1:-1] = 0.5 * (u_r[:, :-1] + u_r[:, 1:])
    ur_star[:, -1] = u_r[:, -1]
    uz_star = np.empty((n_z + 1, n_r))
    uz_star[1:-1] = 0.5 * (u_Z[:-1] + u_Z[1:])
    uz_star[0], uz_star[-1] = u_Z[0], u_Z[-1]
    vel = dict(ur_star=ur_star, uz_star=uz_star, dur_dr=durdr, dur_dz=durdz, duz_dr=duzdr, duz_dz=duzdz,
               ur_over_r=u_r / rc[None, :], u_r=u_r, u_z=u_Z)
    for key in ("dur_dr", "dur_dz", "duz_dr", "duz_dz", "ur_over_r", "ur_star", "uz_star"):
        L[key] = np.ascontiguousarray(vel[key])
    L["T"] = T
    dt = cfl_time_step(geom, K.c, u_r, u_Z, 0.4)
    a_tau = compute_a(1e4) * tau0
    tdiff = max(a_tau ** (1 / 3), 3.0) * Rmax / K.c
    params = dict(tau_cl=tau0, T0=1e4, n_mean=n_mean, sigma_ln=sigma_ln, mach=mach, s_T=s_T,
                  dust_to_gas=dust_to_gas, metallicity=metallicity, kappa_abs=KAPPA_ABS_LYA_DUST, seed=seed,
                  n_r=n_r, n_z=n_z, n_nu=L.n_nu, L_lya=L_lya, Rcl=Rmax, c_s=c_s)
    return StaticCloud(geometry=geom, lya=L, T=T, vel=vel, c_tilde=K.c, dt=dt, tdiff=tdiff, tmax=4.0 * tdiff,
                       Zc=Zc, bc_z_bot=OUTFLOW, params=params)


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
