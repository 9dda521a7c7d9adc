"""Static-sphere setups and the reference's physics-level tables, shared by the CPU and the
GPU physics tests.

The cases follow RunLyaEscapeFractionStaticCloudTest.jl (T = 1e4 K, no recoil, p_d = 0,
reflecting bottom, 4 diffusion times): a uniform (rs = 1) or point (rs = 0.5 cells) source
in a dusty or dust-free H I sphere.  The escaping spectrum is the per-bin term of
Compute_f_esc_Lyα_Cylindrical (src/ConstantsAndFunctions.jl:914-993); f_esc is its
Δν-weighted sum, read as the final finite value of the history
(EscapeFractionTest_StaticCloud/PlotEscapeFractions.py:67-78).
"""
import math
import os

import numpy as np

from lydion_amd import constants as K
from lydion_amd.lya_arrays import compute_a
from lydion_amd.static_cloud import config_grid, setup_static_cloud

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS1_PER_X = 0.4707  # ϵ̃₁ = 0.4707 (aτ)^{1/3} τ_abs (PlotEscapeFractions.py:101-104)


def sphere(tau0, n, n_nu, x_dust, source, T0=1e4):
    """x_dust = (aτ0)^{1/3} τ_abs; source 'uniform' (rs = 1) or 'point' (rs = 0.5 cells)."""
    atau = compute_a(T0) * tau0
    rs = 1.0 if source == "uniform" else 0.5 * 1.1 / n
    return setup_static_cloud(tau0, T0, tau_abs=x_dust / atau ** (1.0 / 3.0), rs=rs, recoil=False,
                              tmax_unitless=4.0, n_r=n, n_z=n, nu=config_grid(n_nu, tau0, T0))


def n_steps(cloud):
    return int(math.ceil(cloud.tmax / cloud.dt))


def escaping_spectrum(cloud, H_r, H_z):
    """S_k ∝ Σ_outer faces A max(H_out, 0) (top and outer-r faces; the bottom reflects),
    normalised to ∫ S dx = 1 over x = (ν − ν_Lyα)/ΔνD.  Returns (x, dx, S)."""
    g, lya = cloud.geometry, cloud.lya
    S = (np.sum(g.A_r[:, -1, None] * np.maximum(H_r[:, -1, :], 0.0), axis=0)
         + np.sum(g.A_Z[-1, :, None] * np.maximum(H_z[-1, :, :], 0.0), axis=0))
    dnuD = float(np.asarray(lya["dnuD"]).reshape(-1)[0])
    x = (lya["nuC"] - K.nu_Lya) / dnuD
    dx = lya["dnu"] / dnuD
    return x, dx, S / np.sum(S * dx)


def j_sphere_point(x, atau):
    """Emergent spectrum of a static dust-free sphere with a central point source
    (J_sphere_point, src/ConstantsAndFunctions.jl:1018-1025)."""
    arg = np.minimum(math.sqrt(math.pi ** 3 / 54.0) * np.abs(x) ** 3 / atau, 300.0)
    return math.sqrt(math.pi ** 3 / 24.0) * (x * x / atau) / np.cosh(arg) ** 2


def analytic_f_esc(source, x_dust):
    eps, neg_fesc = np.loadtxt(os.path.join(GOLDEN, f"AnalyticalEscapeFraction_{source.capitalize()}Source.txt"),
                               unpack=True)
    return float(np.interp(EPS1_PER_X * x_dust, eps, -neg_fesc))


def mcrt_f_esc(source, x_dust):
    x, pct = np.loadtxt(os.path.join(GOLDEN, f"f_esc_tdMAX1.0_{source}_MCRT.txt"), unpack=True)
    return float(np.interp(x_dust, x, pct)) / 100.0


def analytic_mf_uniform(x_dust):
    """M_F / (aτ)^{1/3} of the uniform source."""
    eps, mf = np.loadtxt(os.path.join(GOLDEN, "AnalyticalMF_UniformSource.txt"), unpack=True)
    return float(np.interp(EPS1_PER_X * x_dust, eps, mf))
