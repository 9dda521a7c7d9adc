"""RT-only static / expanding cloud: host mirror of SetupStaticCloudTest's
initialisation (src/StaticCloudTest.jl:34-322) and its time loop (:357-428)
driven through libm1lya, plus the seeded workloads the parity tests and the
benchmark use.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import constants as K
from .grid import GeometryCylindrical, setup_cylindrical_grid
from .lya_arrays import (LyaArrays, VoigtLUT, bins_freq_grid, compute_a, compute_sigma0, generate_lya_arrays,
                         reference_freq_grid)
from ._lib import OUTFLOW, REFLECT


def hubble_flow(r, Z, umax: float, Rcl: float, Zc: float):
    """GetHubbleFlow (src/StaticCloudTest.jl:34-59), vectorised.
    Returns (u_r, u_Z, ∂u_r/∂r, ∂u_r/∂Z, ∂u_Z/∂r, ∂u_Z/∂Z, u_r/r)."""
    r = np.asarray(r, dtype=np.float64)
    Z = np.asarray(Z, dtype=np.float64)
    R = np.hypot(r, Z - Zc)
    inside = ~(R > Rcl)
    H = umax / Rcl
    z = np.zeros(np.broadcast(r, Z).shape)
    u_r = np.where(inside, H * r, 0.0) + z
    u_Z = np.where(inside, H * (Z - Zc), 0.0) + z
    Hc = np.where(inside, H, 0.0) + z
    return u_r, u_Z, Hc, z.copy(), z.copy(), Hc.copy(), Hc.copy()


def cone_factor(mu, mucone: float, taucone: float, taucl: float, smooth: bool, dmu: float):
    """ConeOpticalDepthFactor (src/StaticCloudTest.jl:70-89)."""
    if mucone >= 1.0:
        return np.ones_like(mu)
    fc = taucone / taucl
    if smooth:
        return fc + 0.5 * (1.0 - fc) * (1.0 - np.tanh((mu - mucone) / dmu))
    return np.where(mu > mucone, fc, 1.0)


def cfl_time_step(geom: GeometryCylindrical, c_tilde: float, u_r, u_Z, cfl: float) -> float:
    """GetRHDCFLTimeStep (src/AdvectJandH.jl:29-68)."""
    dr = (geom.r[1:] - geom.r[:-1])[None, :]
    dz = (geom.Z[1:] - geom.Z[:-1])[:, None]
    u = np.hypot(u_r, u_Z)
    return cfl * float(np.min(0.5 * np.minimum(dr, dz) / (c_tilde + u)))


@dataclass
class StaticCloud:
    """Everything SetupStaticCloudTest builds before its time loop."""
    geometry: GeometryCylindrical
    lya: LyaArrays
    T: np.ndarray
    vel: dict
    c_tilde: float
    dt: float
    tdiff: float
    tmax: float
    Zc: float
    bc_z_bot: int
    params: dict = field(default_factory=dict)

    def velocity_args(self):
        v = self.vel
        return (v["ur_star"], v["uz_star"], v["dur_dr"], v["dur_dz"], v["duz_dr"], v["duz_dz"], v["ur_over_r"])


def setup_static_cloud(tau_cl: float, T0: float, tau_abs: float = 0.0, pd0: float = 0.0, rs: float = 1.0,
                       umax: float = 0.0, recoil: bool = True, theta_cone: float = 0.0, tau_cone: float | None = None,
                       smooth_cone: bool = True, dmu_cone: float = 0.1, ionized_source: bool = False,
                       tmax_unitless: float = 2.0, bc_z_bot: int = REFLECT, n_r: int = 64, n_z: int = 64,
                       nu: np.ndarray | None = None, q: float = 1.1, lut: VoigtLUT | None = None) -> StaticCloud:
    """Initialisation of SetupStaticCloudTest (src/StaticCloudTest.jl:92-322).

    nu: optional ν interfaces (GenerateLyαArrays_FromSaved); default is the reference rule
    with q_fine = q_wing = q and Tmin = 30 K."""
    if tau_cone is None:
        tau_cone = 1e-6 * tau_cl
    L_lya = 1e5 * K.Lsun
    Rcl = K.pc
    Rmax = 1.1 * Rcl
    Zmax = Rmax
    Zc = 0.0 if bc_z_bot == REFLECT else 0.5 * Zmax
    c_tilde = K.c
    geom = setup_cylindrical_grid(Rmax, Zmax, n_r, n_z)
    T = np.full((n_z, n_r), float(T0))
    if nu is None:
        nu = reference_freq_grid(30.0, q, q)
    L = generate_lya_arrays(c_tilde, geom.r, geom.Z, T, nu, recoil=recoil, lut=lut)
    Rs = rs * Rcl
    dL_HI = Rcl - Rs if ionized_source else Rcl
    if ionized_source and not dL_HI > 0:
        raise ValueError(f"H I path length is ΔL = {dL_HI}, which is unphysical.")
    nHI0 = tau_cl / (compute_sigma0(T0) * dL_HI)
    ac0 = tau_abs / Rcl
    rc = geom.rCoM[None, :]
    zc = geom.ZCoM[:, None]
    RCoM = np.sqrt(rc * rc + (zc - Zc) * (zc - Zc))
    mu = np.minimum(np.abs(zc - Zc) / RCoM, 1.0)
    mucone = math.cos(theta_cone)
    ftau = cone_factor(mu, mucone, tau_cone, tau_cl, smooth_cone, dmu_cone)
    inside_cloud = RCoM < Rcl
    inside_source = RCoM < Rs
    xHII = np.where(inside_source & ionized_source, 1.0 - 1e-4, 0.0)
    nHI = np.where(inside_cloud, nHI0 * (1 - xHII) * ftau, 0.0)
    L["alpha"] = np.where(inside_cloud, nHI * compute_sigma0(T), 0.0)
    L["alpha_c"] = np.where(inside_cloud, ac0 * ftau, 0.0)
    L["pd"] = np.where(inside_cloud, pd0, 0.0)
    js = np.where(inside_cloud & inside_source, 3 * L_lya / ((K.Four_pi * K.Four_pi) * (Rs * Rs * Rs)), 0.0)
    if np.max(js) <= 0.0:
        i0 = int(np.argmin(np.abs(geom.rCoM - 0.0)))
        j0 = int(np.argmin(np.abs(geom.ZCoM - Zc)))
        js[j0, i0] = L_lya / (K.Four_pi * geom.dV[j0, i0])
    L["js"] = js
    L["alpha_F"] = L["alpha"][:, :, None] * L["Hvoigt"]
    umax_cgs = umax * K.kms
    u_r, u_Z, durdr, durdz, duzdr, duzdz, urr = hubble_flow(rc, zc, umax_cgs, Rcl, Zc)
    ur_star = np.ascontiguousarray(hubble_flow(geom.r[None, :], zc, umax_cgs, Rcl, Zc)[0])
    uz_star = np.ascontiguousarray(hubble_flow(rc, geom.Z[:, None], umax_cgs, Rcl, Zc)[1])
    vel = dict(ur_star=ur_star, uz_star=uz_star, dur_dr=durdr, dur_dz=durdz, duz_dr=duzdr, duz_dz=duzdz,
               ur_over_r=urr, u_r=u_r, u_z=u_Z)
    for key in ("dur_dr", "dur_dz", "duz_dr", "duz_dz", "ur_over_r", "ur_star", "uz_star"):
        L[key] = vel[key]
    L["T"] = T
    dt = cfl_time_step(geom, c_tilde, u_r, u_Z, 0.4)
    a_tau = compute_a(T0) * tau_cl
    tdiff = max(a_tau ** (1 / 3), 3.0) * Rcl / c_tilde
    params = dict(tau_cl=tau_cl, T0=T0, tau_abs=tau_abs, pd0=pd0, rs=rs, umax=umax, recoil=recoil,
                  theta_cone=theta_cone, tau_cone=tau_cone, smooth_cone=smooth_cone, dmu_cone=dmu_cone,
                  n_r=n_r, n_z=n_z, n_nu=L.n_nu, L_lya=L_lya, Rcl=Rcl)
    return StaticCloud(geometry=geom, lya=L, T=T, vel=vel, c_tilde=c_tilde, dt=dt, tdiff=tdiff,
                       tmax=tmax_unitless * tdiff, Zc=Zc, bc_z_bot=bc_z_bot, params=params)


# ---------------------------------------------------------------------------
# Seeded workloads
# ---------------------------------------------------------------------------
DEFAULT_SEED = 20261202


def config_grid(n_nu: int, tau0: float, T_ref: float = 1e4) -> np.ndarray:
    """n_nu bins spanning ±4 (a τ0)^{1/3} Doppler widths at T_ref."""
    x_max = 4.0 * (compute_a(T_ref) * tau0) ** (1.0 / 3.0)
    return bins_freq_grid(n_nu, x_max, T_ref)


def point_source_cloud(n_r: int, n_z: int, n_nu: int, tau0: float, T0: float = 1e4, tau_abs: float = 0.0,
                       theta_cone: float = 0.0, tau_cone: float | None = None, umax: float = 0.0,
                       lut: VoigtLUT | None = None, bc_z_bot: int = REFLECT) -> StaticCloud:
    """Static cloud with a central point source of radius 0.5 cells (rs = 0.5 Δr / R_cl, which
    triggers the reference's single-cell emission branch, src/StaticCloudTest.jl:300-306)."""
    rs = 0.5 * (1.1 / n_r)
    return setup_static_cloud(tau0, T0, tau_abs=tau_abs, rs=rs, umax=umax, recoil=True, theta_cone=theta_cone,
                              tau_cone=tau_cone, smooth_cone=True, dmu_cone=0.025, tmax_unitless=4.0,
                              bc_z_bot=bc_z_bot, n_r=n_r, n_z=n_z, nu=config_grid(n_nu, tau0), lut=lut)


def static_cloud_32x32x64(lut: VoigtLUT | None = None) -> StaticCloud:
    """32×32 r-Z, 64 bins, τ0 = 1e6, T = 1e4 K, point source, no dust, no flow."""
    return point_source_cloud(32, 32, 64, 1e6, lut=lut)


def cone_dust_cloud_256x256x128(lut: VoigtLUT | None = None, n_r: int = 256, n_z: int = 256, n_nu: int = 128) -> StaticCloud:
    """256×256 r-Z, 128 bins, τ0 = 1e8, T = 1e4 K, polar cone of half-angle π/4 with
    τcone = 1e-6 τ0 (smooth, Δμ = 0.025), τ_abs = 1, point source of radius 0.5 cells."""
    return point_source_cloud(n_r, n_z, n_nu, 1e8, tau_abs=1.0, theta_cone=math.pi / 4, lut=lut)


def random_state(shape: tuple, seed: int = DEFAULT_SEED, scale: float = 1.0, jeps_fraction: float = 0.05,
                 free_fraction: float = 0.05):
    """Seeded random (J, H_r, H_Z): J lognormal(σ=2)·scale, H = f J n̂ with f ~ U(0, 0.999); a fraction
    of cells near free streaming (f in (0.999, 1)) and a fraction below Jeps, to exercise the
    closure and floor branches."""
    rng = np.random.Generator(np.random.PCG64(seed))
    J = scale * np.exp(rng.normal(0.0, 2.0, shape))
    f = rng.uniform(0.0, 0.999, shape)
    u = rng.uniform(0.0, 1.0, shape)
    f = np.where(u < free_fraction, rng.uniform(0.999, 1.0, shape), f)
    theta = rng.uniform(0.0, 2 * np.pi, shape)
    J = np.where((u >= free_fraction) & (u < free_fraction + jeps_fraction), K.Jeps * rng.uniform(0, 2, shape), J)
    Hr = f * J * np.cos(theta)
    Hz = f * J * np.sin(theta)
    return J, Hr, Hz


def developed_state(cloud: StaticCloud, seed: int = DEFAULT_SEED, floor: float = 1e-6, noise: float = 0.05,
                    f_esc0: float = 0.5):
    """A seeded, smooth 'developed' radiation field for benchmarking (not a solution of the
    equations): the static-sphere emergent-spectrum shape (x²/aτ) sech²(√(π³/54)|x|³/aτ)
    (src/ConstantsAndFunctions.jl:1018-1025) at every cell, scaled by 1/(1 + (R/Δr)²), with a
    relative floor so that every frequency bin is active, and a radial flux H = f J R̂ with
    f growing from 0.05 at the centre to 0.6 at the box edge.  The amplitude is tied to the
    source: the flux leaving the box carries f_esc0 of the luminosity L = 4π Σ ΔV ∫js
    (Compute_f_esc_Lyα_Cylindrical, src/ConstantsAndFunctions.jl:914-993)."""
    lya, g = cloud.lya, cloud.geometry
    rng = np.random.Generator(np.random.PCG64(seed))
    T0 = cloud.params["T0"]
    a_tau = compute_a(T0) * cloud.params["tau_cl"]
    x = (lya["nuC"] - K.nu_Lya) / float(np.asarray(lya["dnuD"]).reshape(-1)[0])
    arg = np.minimum(np.sqrt(np.pi ** 3 / 54.0) * np.abs(x) ** 3 / a_tau, 300.0)
    prof = (x * x / a_tau) / np.cosh(arg) ** 2
    prof = prof / prof.max() + floor
    rc = g.rCoM[None, :]
    zc = g.ZCoM[:, None] - cloud.Zc
    R = np.sqrt(rc * rc + zc * zc)
    dr = g.r[1] - g.r[0]
    amp = 1.0 / (1.0 + (R / dr) ** 2)
    J = amp[:, :, None] * prof[None, None, :] * np.exp(noise * rng.normal(size=(lya.n_z, lya.n_r, lya.n_nu)))
    f = 0.05 + 0.55 * (R / R.max())
    Hr = (f * rc / R)[:, :, None] * J
    Hz = (f * zc / R)[:, :, None] * J
    dnu = lya["dnu"]
    esc = K.Four_pi * (np.sum(g.A_r[:, -1, None] * np.maximum(Hr[:, -1, :], 0.0) * dnu)
                       + np.sum(g.A_Z[-1, :, None] * np.maximum(Hz[-1, :, :], 0.0) * dnu))
    if cloud.bc_z_bot == OUTFLOW:
        esc += K.Four_pi * np.sum(g.A_Z[0, :, None] * np.maximum(-Hz[0, :, :], 0.0) * dnu)
    L_sim = K.Four_pi * float(np.sum(g.dV * np.asarray(lya["js"])))
    s = f_esc0 * L_sim / esc if esc > 0.0 and L_sim > 0.0 else 1e-3
    J, Hr, Hz = J * s, Hr * s, Hz * s
    for k, v in (("J", J), ("H_r", Hr), ("H_z", Hz), ("J_old", J), ("H_r_old", Hr), ("H_z_old", Hz)):
        lya[k] = np.ascontiguousarray(v)
    return cloud
