"""Shared helpers of the parity tests."""
import numpy as np

import lydion_amd as L
from lydion_amd.static_cloud import random_state

STATE = ("J", "H_r", "H_z")


def small_cloud(n_r=12, n_z=10, n_nu=32, tau0=1e6, tau_abs=0.5, theta_cone=0.0, umax=0.0, bc_z_bot=L.REFLECT,
                lut=None):
    return L.point_source_cloud(n_r, n_z, n_nu, tau0, tau_abs=tau_abs, theta_cone=theta_cone, umax=umax,
                                lut=lut, bc_z_bot=bc_z_bot)


def randomize(cloud, seed=7, scale=1e-6):
    """Puts a seeded random state into the cloud (J, H and the *_old copies)."""
    lya = cloud.lya
    J, Hr, Hz = random_state((lya.n_z, lya.n_r, lya.n_nu), seed=seed, scale=scale)
    for k, v in (("J", J), ("H_r", Hr), ("H_z", Hz), ("J_old", J), ("H_r_old", Hr), ("H_z_old", Hz)):
        lya[k] = v.copy()
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    lya["pd"] = np.where(lya["alpha"] > 0, rng.uniform(0, 1e-3, lya["alpha"].shape), 0.0)
    return cloud


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def bitdiff(a, b):
    """Number of elements whose bits differ (NaN-safe)."""
    return int(np.count_nonzero(np.asarray(a).view(np.int64) != np.asarray(b).view(np.int64)))
