"""Host-side mirror of the reference's Lya operator API, backed by libm1lya.

Each function keeps the name (in snake_case), argument meaning and return
value of the Julia function it replaces; the state lives in device memory in
an M1LyaContext (the Julia shim's `m1lya_ctx*`), created once per LyαArrays
and geometry.  Velocity and temperature arguments may be None: the field
already resident on the device is used.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FIELD, OUTFLOW, REFLECT, check, ptr
from .grid import GeometryCylindrical
from .lya_arrays import LyaArrays


def line_constants() -> "_lib.LineConstants":
    """m1lya_line_constants from the physical constants of the host setup."""
    from . import constants as K
    return _lib.LineConstants(K.nu_Lya, K.b_PREFACT, K.nu_Lya_over_c, K.a_PREFACT, K.sigma0_PREFACT)

# LyaArrays key -> C field id
_KEYS = {
    "J": "J", "J_old": "J_OLD", "H_r": "H_R", "H_z": "H_Z", "H_r_old": "H_R_OLD", "H_z_old": "H_Z_OLD",
    "alpha_F": "ALPHA_F", "Hvoigt": "HVOIGT", "phi": "PHI_NU", "Fkp": "FIJ_KP", "Fk": "FIJ_K", "Fkm": "FIJ_KM",
    "alpha": "ALPHA_LYA", "alpha_c": "ALPHA_C", "pd": "PD", "js": "JS", "T": "T", "dnuD": "DNU_D",
    "dur_dr": "DUR_DR", "dur_dz": "DUR_DZ", "duz_dr": "DUZ_DR", "duz_dz": "DUZ_DZ", "ur_over_r": "UR_OVER_R",
    "drhou_r": "DRHOU_R", "drhou_z": "DRHOU_Z", "f_r": "F_LYA_R", "f_z": "F_LYA_Z", "e_lya": "E_LYA",
    "ur_star": "UR_STAR", "uz_star": "UZ_STAR",
}
STATE_KEYS = ("J", "J_old", "H_r", "H_z", "H_r_old", "H_z_old")


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class M1LyaContext:
    """Device-resident Lya M1 state for one cloud (owns an m1lya_ctx*)."""

    def __init__(self, lya: LyaArrays, geometry: GeometryCylindrical, upload_state: bool = True):
        self.lib = _lib.load()
        self.lya = lya
        self.geometry = geometry
        self.n_nu, self.n_r, self.n_z = lya.n_nu, lya.n_r, lya.n_z
        h = C.c_void_p()
        check(self.lib.m1lya_create(self.n_nu, self.n_r, self.n_z, _lib.FP64, C.byref(h)), None)
        self.handle = h
        g, keep = geometry.to_c()
        self._check(self.lib.m1lya_set_geometry(self.handle, C.byref(g)))
        del keep
        fg = _lib.FreqGrid()
        arrs = [_f64(lya.nu), _f64(lya["nuC"]), _f64(lya["dnu"]), _f64(lya["inv_dnu"]), _f64(lya["inv_nuC"])]
        fg.nu, fg.nuC, fg.dnu, fg.inv_dnu, fg.inv_nuC = (ptr(a) for a in arrs)
        fg.kcore_min, fg.kcore_max = lya.kcore
        fg.recoil = int(bool(lya.recoil))
        fg.flux_method = _lib.HLL if lya.hll else _lib.GLF
        fg.alpha_floor = lya.alpha_floor
        self._check(self.lib.m1lya_set_freq_grid(self.handle, C.byref(fg)))
        del arrs
        for key in ("Hvoigt", "phi", "alpha_F", "alpha", "alpha_c", "pd", "js", "dnuD", "T",
                    "dur_dr", "dur_dz", "duz_dr", "duz_dz", "ur_over_r", "ur_star", "uz_star",
                    "drhou_r", "drhou_z", "e_lya"):
            if key in lya.arrays:
                self.upload(key, lya[key])
        if upload_state:
            for key in STATE_KEYS:
                if key in lya.arrays:
                    self.upload(key, lya[key])

    # ---- plumbing ----
    def _check(self, rc):
        check(rc, self.handle)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.m1lya_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, key):
        f = _KEYS[key]
        if _lib.FIELDS.index(f) < 12:
            return (self.n_z, self.n_r, self.n_nu)
        if f == "UR_STAR":
            return (self.n_z, self.n_r + 1)
        if f == "UZ_STAR":
            return (self.n_z + 1, self.n_r)
        return (self.n_z, self.n_r)

    def upload(self, key: str, arr) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), self._shape(key)))
        self._check(self.lib.m1lya_upload(self.handle, FIELD[_KEYS[key]], ptr(a), a.size))

    def download(self, key: str) -> np.ndarray:
        a = np.empty(self._shape(key))
        self._check(self.lib.m1lya_download(self.handle, FIELD[_KEYS[key]], ptr(a), a.size))
        return a

    def fill(self, key: str, value: float) -> None:
        self._check(self.lib.m1lya_fill(self.handle, FIELD[_KEYS[key]], float(value)))

    def synchronize(self) -> None:
        self._check(self.lib.m1lya_synchronize(self.handle))

    # ---- profiling ----
    def set_profiling(self, on) -> None:
        """0/False off; 1/True every launch bracketed by HIP events; 2 only the fused step's
        column kernel."""
        self._check(self.lib.m1lya_set_profiling(self.handle, int(on)))

    def set_adjoint_mode(self, mode: int) -> None:
        """0: interfaces differentiated once each in reverse mode (default); 1: dual-number
        passes per cell, the cross-check of mode 0."""
        self._check(self.lib.m1lya_set_adjoint_mode(self.handle, int(mode)))

    def kernel_stats(self) -> dict:
        n = self.lib.m1lya_kernel_stats_count(self.handle)
        if n < 0:
            self._check(n)
        out = {}
        for idx in range(n):
            name = C.create_string_buffer(128)
            launches = C.c_long()
            total = C.c_double()
            self._check(self.lib.m1lya_kernel_stats(self.handle, idx, name, 128, C.byref(launches), C.byref(total)))
            out[name.value.decode()] = (launches.value, total.value)
        return out

    def reset_kernel_stats(self) -> None:
        self._check(self.lib.m1lya_reset_kernel_stats(self.handle))

    # ---- device-side coefficient refresh ----
    def set_voigt_lut(self, lut) -> None:
        """Copies the Voigt tables of a VoigtLUT (CreateVoigtLUTs) and the line constants."""
        t = np.ascontiguousarray(lut.t, dtype=np.float64)
        self._check(self.lib.m1lya_set_voigt_lut(self.handle, ptr(t), int(lut.nx), float(lut.xmax),
                                                 C.byref(line_constants())))

    def prepare_subcycling(self, T, n_HI, alpha_c=None, alpha_c_F=None, js=None, pd=None,
                           phi_all_bins: bool = True) -> None:
        """PrepareLyαSubCycling! with Update_Hvoigt_And_ϕν_cell! on the device for every cell."""
        cell = (self.n_z, self.n_r)
        arrs = [None if v is None else np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), cell))
                for v in (T, n_HI, alpha_c, alpha_c_F, js, pd)]
        self._check(self.lib.m1lya_prepare_subcycling(self.handle, *[ptr(a) if a is not None else None for a in arrs],
                                                      int(bool(phi_all_bins))))


# ---------------------------------------------------------------------------
# Operator API (reference names)
# ---------------------------------------------------------------------------
def find_active_bins(ctx: M1LyaContext) -> tuple[int, int]:
    """FindActiveBins (src/UpdateLyaJandH.jl:47-97); 0-based (kmin, kmax)."""
    kmin, kmax = C.c_int(), C.c_int()
    ctx._check(ctx.lib.m1lya_find_active_bins(ctx.handle, C.byref(kmin), C.byref(kmax)))
    return kmin.value, kmax.value


def advect_lya(c_dt: float, ctx: M1LyaContext, u_r_star=None, u_z_star=None, dur_dr=None, dur_dz=None,
               duz_dr=None, duz_dz=None, ur_over_r=None, bc_r_left=REFLECT, bc_r_right=OUTFLOW,
               bc_z_bot=REFLECT, bc_z_top=OUTFLOW) -> None:
    """AdvectLyα! (src/UpdateLyaJandH.jl:107-171)."""
    arrs = [_f64(a) for a in (u_r_star, u_z_star, dur_dr, dur_dz, duz_dr, duz_dz, ur_over_r)]
    ctx._check(ctx.lib.m1lya_advect_lya(ctx.handle, float(c_dt), *(ptr(a) for a in arrs),
                                        int(bc_r_left), int(bc_r_right), int(bc_z_bot), int(bc_z_top)))


def lya_emission_absorption_update(ctx: M1LyaContext, c_tilde: float, dt: float) -> None:
    """LyαEmissionAbsorptionUpdate! (src/ConstantsAndFunctions.jl:483-573) on the resident fields."""
    ctx._check(ctx.lib.m1lya_emission_absorption_update(ctx.handle, float(c_tilde), float(dt)))


def update_freq_factors(ctx: M1LyaContext, T=None, dur_dr=None, dur_dz=None, duz_dr=None, duz_dz=None,
                        ur_over_r=None) -> None:
    """UpdateFreqFactors! (src/FreqFactors.jl:142-278)."""
    arrs = [_f64(a) for a in (T, dur_dr, dur_dz, duz_dr, duz_dz, ur_over_r)]
    ctx._check(ctx.lib.m1lya_update_freq_factors(ctx.handle, *(ptr(a) for a in arrs)))


def frequency_update_j_lya_all_cells(c_dt: float, ctx: M1LyaContext, T=None, dur_dr=None, dur_dz=None,
                                     duz_dr=None, duz_dz=None, ur_over_r=None,
                                     update_freq_factors: bool = True) -> None:
    """FrequencyUpdateJLyα_AllCells! (src/UpdateLyaJandH.jl:231-391)."""
    arrs = [_f64(a) for a in (T, dur_dr, dur_dz, duz_dr, duz_dz, ur_over_r)]
    ctx._check(ctx.lib.m1lya_frequency_update_all_cells(ctx.handle, float(c_dt), *(ptr(a) for a in arrs),
                                                        int(bool(update_freq_factors))))


def update_f_lya_mf_l_lya(ctx: M1LyaContext, dt: float, z_center: float, bc_z_bot=REFLECT, f_grav_r=None,
                          f_grav_z=None) -> tuple:
    """Update_f_Lyα_MF_L_Lyα! (src/ConstantsAndFunctions.jl:588-674); with f_grav_* the gravity
    variant (:691-804).  Returns (M_F, M_F_radial_net, M_F_radial_out, L_Lyα_tot) or
    (M_F, M_F_antigravity, M_F_radial_net, M_F_radial_out, L_Lyα_tot)."""
    if f_grav_r is None:
        out = (C.c_double * 4)()
        ctx._check(ctx.lib.m1lya_update_f_mf_l(ctx.handle, float(dt), float(z_center), int(bc_z_bot), out))
        return tuple(out)
    out = (C.c_double * 5)()
    gr, gz = _f64(f_grav_r), _f64(f_grav_z)
    ctx._check(ctx.lib.m1lya_update_f_mf_l_gravity(ctx.handle, float(dt), ptr(gr), ptr(gz), float(z_center),
                                                   int(bc_z_bot), out))
    return tuple(out)


def compute_f_esc_lya_cylindrical(ctx: M1LyaContext, L_allspace: float, bc_z_bot=REFLECT) -> float:
    """Compute_f_esc_Lyα_Cylindrical (src/ConstantsAndFunctions.jl:969-993)."""
    out = C.c_double()
    ctx._check(ctx.lib.m1lya_compute_f_esc(ctx.handle, float(L_allspace), int(bc_z_bot), C.byref(out)))
    return out.value


def reset_drhou(ctx: M1LyaContext) -> None:
    """Reset_Δρu! (src/ConstantsAndFunctions.jl:809-816)."""
    ctx._check(ctx.lib.m1lya_reset_drhou(ctx.handle))


def update_e_lya(ctx: M1LyaContext) -> np.ndarray:
    """Update_e_Lyα! (src/ConstantsAndFunctions.jl:416-446); returns e_Lyα (N_Z, N_r)."""
    ctx._check(ctx.lib.m1lya_update_e_lya(ctx.handle))
    return ctx.download("e_lya")


def edge_spectra(ctx: M1LyaContext) -> tuple[np.ndarray, np.ndarray]:
    """Edge-cell spectra of SaveTimeStepHDF5 (src/SaveHDF5Data_StaticCloudTests.jl:392-420):
    (raw, normalised), each (N_edge, N_ν)."""
    ne = ctx.lib.m1lya_edge_count(ctx.handle)
    raw = np.empty((ne, ctx.n_nu))
    norm = np.empty((ne, ctx.n_nu))
    ctx._check(ctx.lib.m1lya_edge_spectra(ctx.handle, ptr(raw), ptr(norm), raw.size))
    return raw, norm


STEP_FIELDS = ("M_F", "M_F_radial_net", "M_F_radial_out", "L_Lya", "f_esc", "kmin", "kmax")


def step(ctx: M1LyaContext, c_tilde: float, dt: float, z_center: float, bc_r_right=OUTFLOW,
         bc_z_bot=REFLECT, bc_z_top=OUTFLOW) -> dict:
    """One fused pass of the static-cloud loop body (src/StaticCloudTest.jl:357-392)."""
    out = (C.c_double * 8)()
    ctx._check(ctx.lib.m1lya_step(ctx.handle, float(c_tilde), float(dt), float(z_center), int(bc_r_right),
                                  int(bc_z_bot), int(bc_z_top), out))
    return dict(zip(STEP_FIELDS, list(out)[:7]))


def run_steps(ctx: M1LyaContext, n_steps: int, c_tilde: float, dt: float, z_center: float,
              bc_r_right=OUTFLOW, bc_z_bot=REFLECT, bc_z_top=OUTFLOW, history: bool = True):
    """n_steps fused steps with no host synchronisation in between.  Returns the
    (n_steps, 8) history of step scalars, or None."""
    hist = np.empty((n_steps, 8)) if history else None
    ctx._check(ctx.lib.m1lya_run_steps(ctx.handle, int(n_steps), float(c_tilde), float(dt), float(z_center),
                                       int(bc_r_right), int(bc_z_bot), int(bc_z_top),
                                       ptr(hist) if hist is not None else None))
    return hist


def run_until_steady(ctx: M1LyaContext, max_steps: int, tol: float, c_tilde: float, dt: float, z_center: float,
                     bc_r_right=OUTFLOW, bc_z_bot=REFLECT, bc_z_top=OUTFLOW) -> np.ndarray:
    """Fused steps until ΔJ/J = Σ Δν ΔV |J_new - J_start| / Σ Δν ΔV max(J_new - Jeps, 0) < tol, at
    most max_steps.  The reference runs a fixed horizon (src/StaticCloudTest.jl:317-322); this is
    the steady-state option.  Returns the (steps run, 8) history; column 7 holds each step's ΔJ/J."""
    hist = np.empty((max(int(max_steps), 1), 8))
    done = C.c_int(0)
    ctx._check(ctx.lib.m1lya_run_until_steady(ctx.handle, int(max_steps), float(tol), float(c_tilde), float(dt),
                                              float(z_center), int(bc_r_right), int(bc_z_bot), int(bc_z_top),
                                              ptr(hist), C.byref(done)))
    return hist[:done.value].copy()


def selftest_math(op: int, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """Runs a/b (op 0), sqrt(a) (op 1) or the library's exp (op 2) on the device."""
    lib = _lib.load()
    a = _f64(a)
    b = _f64(b) if b is not None else None
    out = np.empty_like(a)
    check(lib.m1lya_selftest_math(int(op), ptr(a), ptr(b), ptr(out), a.size), None)
    return out
