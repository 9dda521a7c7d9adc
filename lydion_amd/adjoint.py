"""Reverse-mode gradients of the FP64 fused step (m1lya_adjoint_run).

The reference has no adjoint.  `adjoint_run` is the vector-Jacobian product a
ChainRules `rrule` around `m1lya_run_steps` needs: given seeds on the per-step
force multipliers (M_F, M_F_radial_net, M_F_radial_out) and on the escaping
spectrum of the final state, it returns the gradient of their weighted sum
with respect to the initial state and to every coefficient field the step
reads (α_Lyα, α_c, ∫js, ΔνD, T per cell; α_F, Hvoigt, ϕ_ν per cell and bin).

`physical_gradients` chains those through the static-cloud setup
(SetupStaticCloudTest, src/StaticCloudTest.jl:92-322) to the per-cell H I
density n_H, the per-cell temperature T and the source luminosity L:

    α = n_H σ0(T),  α_F = α Hvoigt(T),  ΔνD ∝ T^½,  σ0 ∝ T^-½,  js ∝ L,

with the temperature derivative of the Voigt tables (Hvoigt and the core
emission profile ϕ_ν) taken exactly for the tabulated profile
(update_hvoigt_dT / update_phi_dT: the piecewise-linear tables differentiated
segment by segment, then the quotient rule of the core normalisation).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import constants as K
from ._lib import OUTFLOW, REFLECT, ptr
from .lya_arrays import (VoigtLUT, compute_a, compute_dnuD, compute_sigma0, update_hvoigt, update_hvoigt_dT,
                         update_phi, update_phi_dT)

CELL_FIELDS = ("alpha", "alpha_c", "js", "dnuD", "T")
BIN_FIELDS = ("alpha_F", "Hvoigt", "phi")


@dataclass
class AdjointResult:
    objective: float
    history: np.ndarray        # (n_steps, 8) step scalars of the forward run
    grad_state: dict           # "J", "H_r", "H_z" -> (N_Z, N_r, N_ν)
    grad_cell: dict            # CELL_FIELDS -> (N_Z, N_r); "T" is the direct (recoil) term only
    grad_bin: dict             # BIN_FIELDS -> (N_Z, N_r, N_ν)


def adjoint_run(ctx, n_steps: int, c_tilde: float, dt: float, z_center: float, mf_seed=None, spec_seed=None,
                bc_r_right=OUTFLOW, bc_z_bot=REFLECT, bc_z_top=OUTFLOW, checkpoint_every: int = 0) -> AdjointResult:
    """Gradient of Σ_s mf_seed[s]·(M_F, net, out)(s) + spec_seed·S(final) over n_steps fused steps
    from the context's state.  mf_seed: (n_steps, 3) or None; spec_seed: (N_ν,) or None.
    The context ends in the state after n_steps, as run_steps leaves it."""
    nn, nr, nz = ctx.n_nu, ctx.n_r, ctx.n_z
    mf = None if mf_seed is None else np.ascontiguousarray(np.broadcast_to(mf_seed, (n_steps, 3)), dtype=np.float64)
    sp = None if spec_seed is None else np.ascontiguousarray(np.broadcast_to(spec_seed, (nn,)), dtype=np.float64)
    obj = C.c_double()
    hist = np.empty((n_steps, 8))
    gs = np.empty((3, nz, nr, nn))
    gc = np.empty((5, nz, nr))
    gb = np.empty((3, nz, nr, nn))
    ctx._check(ctx.lib.m1lya_adjoint_run(ctx.handle, int(n_steps), float(c_tilde), float(dt), float(z_center),
                                         int(bc_r_right), int(bc_z_bot), int(bc_z_top),
                                         ptr(mf) if mf is not None else None, ptr(sp) if sp is not None else None,
                                         int(checkpoint_every), C.byref(obj), ptr(hist), ptr(gs), ptr(gc), ptr(gb)))
    return AdjointResult(obj.value, hist, dict(zip(("J", "H_r", "H_z"), gs)), dict(zip(CELL_FIELDS, gc)),
                         dict(zip(BIN_FIELDS, gb)))


def escaping_spectrum(geometry, H_r, H_z, L_sum: float, bc_z_bot=REFLECT) -> np.ndarray:
    """S_k = 4π Σ_outer faces A max(H_out, 0) / L_sim (the integrand of f_esc,
    src/ConstantsAndFunctions.jl:914-993; L_sim = L_sum = 4π Σ ΔV js)."""
    s = (geometry.A_r[:, -1, None] * K.Four_pi * np.maximum(H_r[:, -1, :], 0.0)).sum(axis=0)
    s = s + (geometry.A_Z[-1, :, None] * K.Four_pi * np.maximum(H_z[-1, :, :], 0.0)).sum(axis=0)
    if bc_z_bot == OUTFLOW:
        s = s + (geometry.A_Z[0, :, None] * K.Four_pi * np.maximum(-H_z[0, :, :], 0.0)).sum(axis=0)
    return s / L_sum if L_sum > 0 else np.zeros_like(s)


def objective(history: np.ndarray, spectrum: np.ndarray | None, mf_seed=None, spec_seed=None) -> float:
    """The objective adjoint_run differentiates, from a forward history and final spectrum."""
    out = 0.0
    if mf_seed is not None:
        out += float(np.sum(np.broadcast_to(mf_seed, (history.shape[0], 3)) * history[:, :3]))
    if spec_seed is not None:
        out += float(np.dot(spec_seed, spectrum))
    return out


def setup_tables(cloud, T: np.ndarray, lut: VoigtLUT | None = None):
    """ΔνD, Hvoigt and ϕ_ν of the cloud's frequency grid at temperatures T (N_Z, N_r), as
    GenerateLyαArrays_FromSaved builds them (src/SetupLymanAlphaArrays.jl:244-393)."""
    L = cloud.lya
    lut = lut or VoigtLUT()
    dnuD = compute_dnuD(T)
    Hv = update_hvoigt(lut, L["dnu_rel"], L["dnuC_rel"], compute_a(T), 1.0 / dnuD)
    phi = update_phi(L["dnu"], Hv, *L.kcore)
    return dnuD, Hv, phi


def density(cloud) -> np.ndarray:
    """Per-cell H I density n_H = α / σ0(T) of a set-up cloud."""
    return cloud.lya["alpha"] / compute_sigma0(cloud.T)


def _dust_follows_density(cloud) -> bool:
    """True for clouds whose dust absorption α_c is proportional to n_H and part of α_F
    (the RHD-like state); the static clouds carry a fixed α_c outside α_F."""
    return "kappa_abs" in cloud.params


def setup_tables_dT(cloud, T: np.ndarray, lut: VoigtLUT | None = None):
    """d Hvoigt / dT and d ϕ_ν / dT of setup_tables at temperatures T (N_Z, N_r)."""
    L = cloud.lya
    lut = lut or VoigtLUT()
    Hv = update_hvoigt(lut, L["dnu_rel"], L["dnuC_rel"], compute_a(T), 1.0 / compute_dnuD(T))
    dHv = update_hvoigt_dT(lut, L["dnu_rel"], L["dnuC_rel"], T)
    return dHv, update_phi_dT(L["dnu"], Hv, dHv, *L.kcore)


def physical_gradients(cloud, res: AdjointResult, lut: VoigtLUT | None = None, with_T: bool = True) -> dict:
    """d Obj / d n_H (per cell), d Obj / d T (per cell; skipped when with_T is False) and
    d Obj / d L (scalar)."""
    L = cloud.lya
    T = np.asarray(cloud.T, dtype=np.float64)
    gc, gb = res.grad_cell, res.grad_bin
    sig = compute_sigma0(T)
    nH = density(cloud)
    alpha = L["alpha"]
    Hv = L["Hvoigt"]
    g_nH = sig * (gc["alpha"] + np.sum(Hv * gb["alpha_F"], axis=-1))
    if _dust_follows_density(cloud):
        # α_c ∝ n_H enters α_c and every bin of α_F
        dust = np.divide(L["alpha_c"], nH, out=np.zeros_like(nH), where=nH > 0)
        g_nH = g_nH + dust * (gc["alpha_c"] + np.sum(gb["alpha_F"], axis=-1))
    L_lya = cloud.params["L_lya"]
    g_L = float(np.sum(gc["js"] * L["js"])) / L_lya      # js ∝ L
    if not with_T:
        return {"n_H": g_nH, "T": None, "L": g_L}
    # temperature: σ0 ∝ T^-1/2, ΔνD ∝ T^1/2, Voigt tables differentiated exactly
    dHv, dphi = setup_tables_dT(cloud, T, lut)
    dalpha = nH * (-0.5 * sig / T)
    ddnuD = 0.5 * L["dnuD"] / T
    dalphaF = dalpha[..., None] * Hv + alpha[..., None] * dHv
    g_T = (gc["T"] + gc["dnuD"] * ddnuD + gc["alpha"] * dalpha
           + np.sum(gb["alpha_F"] * dalphaF + gb["Hvoigt"] * dHv + gb["phi"] * dphi, axis=-1))
    return {"n_H": g_nH, "T": g_T, "L": g_L}


def cloud_with(cloud, n_H: np.ndarray | None = None, T: np.ndarray | None = None, L_scale: float = 1.0,
               lut: VoigtLUT | None = None) -> dict:
    """Coefficient fields of the cloud with n_H, T (per cell) and L·L_scale replaced, as the
    setup computes them; keys are the context's upload names."""
    L = cloud.lya
    T0 = np.asarray(cloud.T, dtype=np.float64)
    T1 = T0 if T is None else np.asarray(T, dtype=np.float64)
    nH = density(cloud) if n_H is None else n_H
    out = {}
    if T is not None:
        dnuD, Hv, phi = setup_tables(cloud, T1, lut)
        out.update(dnuD=dnuD, Hvoigt=Hv, phi=phi, T=T1)
    else:
        Hv = L["Hvoigt"]
    alpha = nH * compute_sigma0(T1)
    out["alpha"] = alpha
    out["alpha_F"] = alpha[..., None] * Hv
    if _dust_follows_density(cloud):
        nH0 = density(cloud)
        dust = np.divide(L["alpha_c"], nH0, out=np.zeros_like(nH0), where=nH0 > 0)
        out["alpha_c"] = dust * nH
        out["alpha_F"] = out["alpha_F"] + out["alpha_c"][..., None]
    out["js"] = L["js"] * L_scale
    return out
