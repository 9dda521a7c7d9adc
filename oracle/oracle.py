"""ctypes wrapper of the CPU FP64 oracle (oracle/liboracle_m1lya.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference the GPU results are compared
with.  The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_m1lya.so")
_dp = C.POINTER(C.c_double)
REFLECT, OUTFLOW = 0, 1


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class OrcGeom(C.Structure):
    _fields_ = [("n_r", C.c_int), ("n_z", C.c_int), ("r", _dp), ("z", _dp), ("rcom", _dp), ("zcom", _dp),
                ("dV", _dp), ("inv_dV", _dp), ("A_r", _dp), ("A_z", _dp), ("inv_r_cell", _dp),
                ("dr_L", _dp), ("dr_R", _dp), ("inv_dr_i_im1", _dp), ("inv_dr_ip1_i", _dp),
                ("inv_dr_ip2_ip1", _dp), ("dr_inner", C.c_double * 5), ("dr_outer", C.c_double * 5),
                ("dz_L", _dp), ("dz_R", _dp), ("inv_dz_j_jm1", _dp), ("inv_dz_jp1_j", _dp),
                ("inv_dz_jp2_jp1", _dp), ("dz_inner", C.c_double * 5), ("dz_outer", C.c_double * 5)]


_LYA_PTRS = ["nu", "nuC", "dnu", "inv_dnu", "inv_nuC", "J", "J_old", "H_r", "H_z", "H_r_old", "H_z_old",
             "Kphi_ip", "Kphi_im", "F_r", "F_r_Hr", "F_r_HZ", "F_z", "F_z_Hr", "F_z_HZ", "alpha_F", "Hvoigt",
             "phi", "Fkp", "Fk", "Fkm", "alpha", "alpha_c", "pd", "js", "dnuD", "drhou_r", "drhou_z", "f_r",
             "f_z", "e_lya"]


class OrcLya(C.Structure):
    _fields_ = [("n_nu", C.c_int), ("kcore_min", C.c_int), ("kcore_max", C.c_int), ("recoil", C.c_int),
                ("flux_hll", C.c_int), ("alpha_floor", C.c_double)] + [(n, _dp) for n in _LYA_PTRS]


_VEL = ["u_r_star", "u_z_star", "dur_dr", "dur_dz", "duz_dr", "duz_dz", "ur_over_r"]


class OrcVel(C.Structure):
    _fields_ = [(n, _dp) for n in _VEL]


def _declare(lib):
    d, i, p = C.c_double, C.c_int, C.c_void_p
    G, L, V = C.POINTER(OrcGeom), C.POINTER(OrcLya), C.POINTER(OrcVel)
    sigs = {
        "orc_exp": (d, [d]), "orc_minmod": (d, [d, d]), "orc_chi_L84": (d, [d]),
        "orc_D_edd": (None, [d, d, d, _dp]), "orc_D_edd_from_f_n": (None, [d, d, d, _dp]),
        "orc_hll_eigenvalues": (None, [d, d, _dp]), "orc_flux_glf": (None, [_dp, _dp]),
        "orc_flux_hll": (None, [_dp, _dp, _dp]),
        "orc_reconstruct_interface": (None, [_dp, _dp]), "orc_Kphiphi_inner_axis": (d, [d, d]),
        "orc_gtsv": (i, [i, _dp, _dp, _dp, _dp]), "orc_thomas": (None, [i, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_find_active_bins": (None, [L, G, C.POINTER(i), C.POINTER(i)]),
        "orc_advect_jandh_cyl": (None, [d, L, G, V, i, i, i, i, i, i]),
        "orc_advect_lya": (None, [d, L, G, V, i, i, i, i, C.POINTER(i), C.POINTER(i)]),
        "orc_emission_absorption": (None, [L, i, i, d, d]),
        "orc_update_freq_factors": (None, [L, i, i, _dp, V]),
        "orc_frequency_update_all_cells": (None, [d, L, i, i, _dp, V, i]),
        "orc_update_f_mf_l": (None, [L, G, d, d, i, _dp]),
        "orc_update_f_mf_l_grav": (None, [L, G, d, _dp, _dp, d, i, _dp]),
        "orc_f_esc_lya": (d, [L, G, d, i]), "orc_reset_drhou": (None, [L, i, i]),
        "orc_update_e_lya": (None, [L, i, i]), "orc_edge_count": (i, [i, i]),
        "orc_edge_spectra": (None, [L, i, i, _dp, _dp]),
        "orc_static_cloud_step": (None, [d, d, L, G, V, _dp, d, i, _dp]),
        "orc_cyl_grid_uniform": (None, [d, d, G]),
        "orc_lya_freq_grid": (i, [d, d, d, _dp, i]),
        "orc_compute_b": (d, [d]), "orc_compute_dnuD": (d, [d]), "orc_compute_a": (d, [d]),
        "orc_compute_sigma0": (d, [d]), "orc_lut_eval": (d, [_dp, i, d, d, d]),
        "orc_update_hvoigt": (None, [_dp, i, i, _dp, _dp, _dp, _dp, _dp, i, d]),
        "orc_update_phi": (None, [_dp, i, i, _dp, _dp, i, i]),
        "orc_cfl_dt": (d, [G, d, _dp, _dp, d]), "orc_hubble_flow": (None, [d, d, d, d, d, _dp]),
        "orc_cone_factor": (d, [d, d, d, d, i, d]),
        "orc_find_rsl": (d, [d, d, G, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_absorption_substep": (None, [L, i, i, _dp, d, _dp, _dp, C.POINTER(i), _dp]),
        "orc_prob_lya_sh95": (d, [d, d, i]), "orc_krec_hii": (d, [d, i]),
        "orc_js_rec_lya": (d, [d, d, d, i]), "orc_js_coll_lya": (d, [d, d, d]),
        "orc_pd_lya": (d, [d, d, d, d, d, d, d]),
        "orc_lya_sources": (None, [i, _dp, _dp, _dp, d, i, i, _dp]),
        "orc_set_threads": (None, [i]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    del p


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def f64(a):
    return np.ascontiguousarray(np.array(a, dtype=np.float64, copy=True))


class OracleGeom:
    """orc_geom built from a host GeometryCylindrical (arrays copied and kept alive)."""

    def __init__(self, geom):
        self.keep = {}
        g = OrcGeom()
        g.n_r, g.n_z = geom.n_r, geom.n_z
        src = dict(r=geom.r, z=geom.Z, rcom=geom.rCoM, zcom=geom.ZCoM, dV=geom.dV, inv_dV=geom.inv_dV,
                   A_r=geom.A_r, A_z=geom.A_Z, inv_r_cell=geom.inv_r_cell)
        names_r = ["dr_L", "dr_R", "inv_dr_i_im1", "inv_dr_ip1_i", "inv_dr_ip2_ip1"]
        names_z = ["dz_L", "dz_R", "inv_dz_j_jm1", "inv_dz_jp1_j", "inv_dz_jp2_jp1"]
        for n, a in zip(names_r, geom.dr[:5]):
            src[n] = a
        for n, a in zip(names_z, geom.dZ[:5]):
            src[n] = a
        for n, a in src.items():
            self.keep[n] = f64(a)
            setattr(g, n, ptr(self.keep[n]))
        for m in range(5):
            g.dr_inner[m], g.dr_outer[m] = geom.dr[5][m], geom.dr[6][m]
            g.dz_inner[m], g.dz_outer[m] = geom.dZ[5][m], geom.dZ[6][m]
        self.s = g


class OracleState:
    """A copy of a StaticCloud's state and coefficients in orc_lya/orc_vel form."""

    def __init__(self, cloud, hll: bool | None = None):
        load()
        L = cloud.lya
        nn, nr, nz = L.n_nu, L.n_r, L.n_z
        self.n_nu, self.n_r, self.n_z = nn, nr, nz
        self.geom = OracleGeom(cloud.geometry)
        self.T = f64(cloud.T)
        self.a = {}
        s3, s2 = (nz, nr, nn), (nz, nr)
        fr, fz = (nz, nr + 1, nn), (nz + 1, nr, nn)
        init = {"nu": L.nu, "nuC": L["nuC"], "dnu": L["dnu"], "inv_dnu": L["inv_dnu"], "inv_nuC": L["inv_nuC"]}
        for n in _LYA_PTRS:
            if n in init:
                self.a[n] = f64(init[n])
            elif n in ("F_r", "F_r_Hr", "F_r_HZ"):
                self.a[n] = np.zeros(fr)
            elif n in ("F_z", "F_z_Hr", "F_z_HZ"):
                self.a[n] = np.zeros(fz)
            elif n in ("alpha", "alpha_c", "pd", "js", "dnuD", "drhou_r", "drhou_z", "f_r", "f_z", "e_lya"):
                self.a[n] = f64(L[n]) if n in L.arrays else np.zeros(s2)
            else:
                self.a[n] = f64(L[n]) if n in L.arrays else np.zeros(s3)
        lya = OrcLya()
        lya.n_nu = nn
        lya.kcore_min, lya.kcore_max = L.kcore
        lya.recoil = int(bool(L.recoil))
        lya.flux_hll = int(bool(L.hll if hll is None else hll))
        lya.alpha_floor = L.alpha_floor
        for n in _LYA_PTRS:
            setattr(lya, n, ptr(self.a[n]))
        self.lya = lya
        self.v = {n: f64(cloud.vel[k]) for n, k in zip(_VEL, ["ur_star", "uz_star", "dur_dr", "dur_dz", "duz_dr",
                                                                 "duz_dz", "ur_over_r"])}
        vel = OrcVel()
        for n in _VEL:
            setattr(vel, n, ptr(self.v[n]))
        self.vel = vel
        self.c_tilde = cloud.c_tilde
        self.dt = cloud.dt
        self.Zc = cloud.Zc
        self.bc_z_bot = cloud.bc_z_bot

    def __getitem__(self, n):
        return self.a[n]

    def set(self, n, value):
        self.a[n][...] = value

    # ---- operators ----
    def find_active_bins(self):
        kmin, kmax = C.c_int(), C.c_int()
        load().orc_find_active_bins(C.byref(self.lya), C.byref(self.geom.s), C.byref(kmin), C.byref(kmax))
        return kmin.value, kmax.value

    def advect_lya(self, cdt, bc_r_right=OUTFLOW, bc_z_bot=None, bc_z_top=OUTFLOW):
        bc_z_bot = self.bc_z_bot if bc_z_bot is None else bc_z_bot
        kmin, kmax = C.c_int(), C.c_int()
        load().orc_advect_lya(cdt, C.byref(self.lya), C.byref(self.geom.s), C.byref(self.vel), REFLECT,
                              bc_r_right, bc_z_bot, bc_z_top, C.byref(kmin), C.byref(kmax))
        return kmin.value, kmax.value

    def emission_absorption(self, c_tilde, dt):
        load().orc_emission_absorption(C.byref(self.lya), self.n_r, self.n_z, c_tilde, dt)

    def update_freq_factors(self):
        load().orc_update_freq_factors(C.byref(self.lya), self.n_r, self.n_z, ptr(self.T), C.byref(self.vel))

    def frequency_update_all_cells(self, cdt, update_ff=True):
        load().orc_frequency_update_all_cells(cdt, C.byref(self.lya), self.n_r, self.n_z, ptr(self.T),
                                              C.byref(self.vel), int(update_ff))

    def update_f_mf_l(self, dt, z_center=None, bc_z_bot=None):
        out = np.zeros(4)
        load().orc_update_f_mf_l(C.byref(self.lya), C.byref(self.geom.s), dt,
                                 self.Zc if z_center is None else z_center,
                                 self.bc_z_bot if bc_z_bot is None else bc_z_bot, ptr(out))
        return tuple(out)

    def update_f_mf_l_grav(self, dt, fgr, fgz, z_center=None, bc_z_bot=None):
        out = np.zeros(5)
        fgr, fgz = f64(fgr), f64(fgz)
        load().orc_update_f_mf_l_grav(C.byref(self.lya), C.byref(self.geom.s), dt, ptr(fgr), ptr(fgz),
                                      self.Zc if z_center is None else z_center,
                                      self.bc_z_bot if bc_z_bot is None else bc_z_bot, ptr(out))
        return tuple(out)

    def f_esc(self, L_allspace, bc_z_bot=None):
        return load().orc_f_esc_lya(C.byref(self.lya), C.byref(self.geom.s), L_allspace,
                                    self.bc_z_bot if bc_z_bot is None else bc_z_bot)

    def reset_drhou(self):
        load().orc_reset_drhou(C.byref(self.lya), self.n_r, self.n_z)

    def update_e_lya(self):
        load().orc_update_e_lya(C.byref(self.lya), self.n_r, self.n_z)
        return self.a["e_lya"].copy()

    def edge_spectra(self):
        ne = load().orc_edge_count(self.n_r, self.n_z)
        raw = np.zeros((ne, self.n_nu))
        norm = np.zeros((ne, self.n_nu))
        load().orc_edge_spectra(C.byref(self.lya), self.n_r, self.n_z, ptr(raw), ptr(norm))
        return raw, norm

    def absorption_substep(self, dt, c_tilde, rho_d=None, kappa_F=(0.0, 0.0, 0.0)):
        """The Lyα block of the RHD absorption sub-step for every cell; dt scalar or (N_Z, N_r).
        Returns (redo (N_Z, N_r) ints, out (12, N_Z, N_r))."""
        dtc = f64(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n_z, self.n_r)))
        redo = np.zeros((self.n_z, self.n_r), dtype=np.int32)
        out = np.zeros((12, self.n_z, self.n_r))
        kF = f64(kappa_F)
        load().orc_absorption_substep(C.byref(self.lya), self.n_r, self.n_z, ptr(dtc), c_tilde,
                                      ptr(f64(rho_d)) if rho_d is not None else None, ptr(kF),
                                      redo.ctypes.data_as(C.POINTER(C.c_int)), ptr(out))
        return redo, out

    def step(self, c_tilde=None, dt=None, z_center=None, bc_z_bot=None):
        out = np.zeros(8)
        load().orc_static_cloud_step(self.c_tilde if c_tilde is None else c_tilde, self.dt if dt is None else dt,
                                     C.byref(self.lya), C.byref(self.geom.s), C.byref(self.vel), ptr(self.T),
                                     self.Zc if z_center is None else z_center,
                                     self.bc_z_bot if bc_z_bot is None else bc_z_bot, ptr(out))
        return out


def find_rsl(c_old, c_min, geom, u_r, u_z, gamma_ad, P, rho, rho_d, xHII, xH2, T, pd, kappa_abs):
    """FindRSL_LyαRT on the oracle (cell arrays (N_Z, N_r), rho_d (N_Z, N_r, 3))."""
    g = OracleGeom(geom)
    arrs = [f64(a) for a in (u_r, u_z, gamma_ad, P, rho, rho_d, xHII, xH2, T, pd, kappa_abs)]
    return load().orc_find_rsl(c_old, c_min, C.byref(g.s), *[ptr(a) for a in arrs])


def cfl_dt(geom, c_tilde, u_r, u_z, cfl):
    """GetRHDCFLTimeStep on the oracle."""
    g = OracleGeom(geom)
    return load().orc_cfl_dt(C.byref(g.s), c_tilde, ptr(f64(u_r)), ptr(f64(u_z)), cfl)


def lya_sources(chem, rho_d, kappa, z=0.0, ots=True, destruction=True):
    """The per-cell source terms of PrepareLyαSubCycling! on the oracle: chem (8, N_Z, N_r) =
    T, n_HI, n_H2, n_HII, n_HeI, n_e, n_OHx, n_CHx; returns (4, N_Z, N_r) = js, α_c, α_c,F, p_d."""
    chem = f64(chem)
    n = chem[0].size
    out = np.zeros((4,) + chem.shape[1:])
    rd = f64(rho_d) if rho_d is not None else np.zeros(chem.shape[1:] + (3,))
    load().orc_lya_sources(n, ptr(chem), ptr(rd), ptr(f64(kappa)), z, int(ots), int(destruction), ptr(out))
    return out
