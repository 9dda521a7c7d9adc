"""Cylindrical r-Z grid of the Lya M1 step: host-side mirror of
SetupCylindricalGrid / Get_*_Cylindrical! / Compute_Grid_Differences!
(src/SpatialGrid.jl:26-169, 516-913).

Array layout follows the reference's Julia column-major arrays, expressed as
C-order numpy arrays with reversed axes: a Julia (N_r, N_Z) array is a numpy
(N_Z, N_r) array with identical memory.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib


def _linrange(start: float, stop: float, n: int) -> np.ndarray:
    """Julia LinRange(start, stop, n): element i = fma(t, stop, fma(-t, start, start)), t = i/(n-1).
    For start == 0 this is (i/(n-1)) * stop."""
    if start != 0.0:
        raise ValueError("only start = 0 grids are used by the Lya M1 setups")
    t = np.arange(n, dtype=np.float64) / float(n - 1)
    return t * stop


def grid_differences(x: np.ndarray, xc: np.ndarray):
    """Compute_Grid_Differences! (src/SpatialGrid.jl:516-586).
    Returns (dL, dR, inv_i_im1, inv_ip1_i, inv_ip2_ip1, inner[5], outer[5]) with inner/outer in the
    order {ΔL, ΔR, inv(i,i-1), inv(i+1,i), inv(i+2,i+1)}."""
    n = len(xc)
    c0 = x[0] - (xc[0] - x[0])
    cm1 = x[0] - (xc[1] - x[0])
    cNp1 = x[n] + (x[n] - xc[n - 1])
    cNp2 = x[n] + (x[n] - xc[n - 2])
    inner = np.array([x[0] - c0, xc[0] - x[0], 1.0 / (c0 - cm1), 1.0 / (xc[0] - c0), 1.0 / (xc[1] - xc[0])])
    cim1 = np.concatenate([[c0], xc[:-1]])
    cip1 = np.concatenate([xc[1:], [cNp1]])
    cip2 = np.concatenate([xc[2:], [cNp1, cNp2]])
    inv_i_im1 = 1.0 / (xc - cim1)
    inv_ip1_i = 1.0 / (cip1 - xc)
    inv_ip2_ip1 = 1.0 / (cip2 - cip1)
    dL = x[1:] - xc
    dR = cip1 - x[1:]
    outer = np.array([x[n] - xc[n - 1], cNp1 - x[n], 1.0 / (xc[n - 1] - xc[n - 2]), 1.0 / (cNp1 - xc[n - 1]),
                      1.0 / (cNp2 - cNp1)])
    return dL, dR, inv_i_im1, inv_ip1_i, inv_ip2_ip1, inner, outer


@dataclass
class GeometryCylindrical:
    """GeometryCylindrical{Float64} (src/Structs.jl:650-740), M1-relevant fields."""
    r: np.ndarray          # (N_r+1,)
    Z: np.ndarray          # (N_Z+1,)
    rCoM: np.ndarray       # (N_r,)
    ZCoM: np.ndarray       # (N_Z,)
    dV: np.ndarray         # (N_Z, N_r)   == Julia ΔV[i,j]
    inv_dV: np.ndarray     # (N_Z, N_r)
    A_r: np.ndarray        # (N_Z, N_r+1)
    A_Z: np.ndarray        # (N_Z+1, N_r)
    inv_r_cell: np.ndarray  # (N_r,)
    dr: tuple              # grid_differences(r, rCoM)
    dZ: tuple              # grid_differences(Z, ZCoM)

    @property
    def n_r(self) -> int:
        return len(self.rCoM)

    @property
    def n_z(self) -> int:
        return len(self.ZCoM)

    def to_c(self):
        """m1lya_geometry struct; keep the returned object alive during the call."""
        g = _lib.Geometry()
        keep = []

        def p(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _lib.ptr(a)

        g.r, g.Z, g.rCoM, g.ZCoM = p(self.r), p(self.Z), p(self.rCoM), p(self.ZCoM)
        g.dV, g.inv_dV, g.A_r, g.A_Z, g.inv_r_cell = p(self.dV), p(self.inv_dV), p(self.A_r), p(self.A_Z), p(
            self.inv_r_cell)
        g.dr_L, g.dr_R, g.inv_dr_i_im1, g.inv_dr_ip1_i, g.inv_dr_ip2_ip1 = (p(a) for a in self.dr[:5])
        g.dZ_L, g.dZ_R, g.inv_dZ_j_jm1, g.inv_dZ_jp1_j, g.inv_dZ_jp2_jp1 = (p(a) for a in self.dZ[:5])
        for m in range(5):
            g.dr_inner[m] = self.dr[5][m]
            g.dr_outer[m] = self.dr[6][m]
            g.dZ_inner[m] = self.dZ[5][m]
            g.dZ_outer[m] = self.dZ[6][m]
        return g, keep


def setup_cylindrical_grid(rmax: float, zmax: float, n_r: int, n_z: int) -> GeometryCylindrical:
    """SetupCylindricalGrid(rmax, rmax, Zmax, Zmax, N_r, 0, N_Z, 0, BC) for the uniform grids the
    static-cloud tests use (src/SpatialGrid.jl:606-913; N_r_log = N_Z_log_side = 0)."""
    if n_r < 3 or n_z < 3:
        raise ValueError("We need N_r ≥ 3 and N_Z ≥ 3 (src/SpatialGrid.jl:629-630)")
    r = _linrange(0.0, rmax, n_r + 1)
    Z = _linrange(0.0, zmax, n_z + 1)
    a, b = r[1:], r[:-1]
    rCoM = (2.0 / 3.0) * (a * a * a - b * b * b) / (a * a - b * b)
    ZCoM = 0.5 * (Z[:-1] + Z[1:])
    dz = (Z[1:] - Z[:-1])[:, None]
    ring = (r[1:] * r[1:] - r[:-1] * r[:-1])[None, :]
    dV = np.pi * ring * dz
    A_r = 2.0 * np.pi * r[None, :] * dz
    A_Z = np.broadcast_to(np.pi * ring, (n_z + 1, n_r)).copy()
    inv_r_cell = 2.0 / (r[1:] + r[:-1])
    return GeometryCylindrical(r=r, Z=Z, rCoM=rCoM, ZCoM=ZCoM, dV=np.ascontiguousarray(dV), inv_dV=1.0 / dV,
                               A_r=np.ascontiguousarray(A_r), A_Z=A_Z, inv_r_cell=inv_r_cell,
                               dr=grid_differences(r, rCoM), dZ=grid_differences(Z, ZCoM))
