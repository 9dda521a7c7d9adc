"""Mixed-precision batched M1 Lyα step (the m1lya_batch_* entry points).

A batch advances many clouds -- of any mix of grid sizes, optical depths and
frequency resolutions -- by the fused time step of the static-cloud loop body
(src/StaticCloudTest.jl:357-392), one device launch per step for the whole
batch.  State is held in FP32 (J and H divided by a per-cloud scale J0); the
flux sweep and closure run in FP32, emission/absorption, the frequency factors
and the tridiagonal frequency solve in FP64, with compensated reductions for
M_F, L and f_esc.  A batch of one cloud is the single-cloud FP32 step.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import FIELD, OUTFLOW, REFLECT, check_batch, ptr

STEP_FIELDS = ("M_F", "M_F_radial_net", "M_F_radial_out", "L_Lya", "f_esc", "kmin", "kmax")

# LyaArrays key -> C field id, for the fields a batch holds
_KEYS = {
    "J": "J", "H_r": "H_R", "H_z": "H_Z", "alpha_F": "ALPHA_F", "Hvoigt": "HVOIGT", "phi": "PHI_NU",
    "alpha": "ALPHA_LYA", "alpha_c": "ALPHA_C", "pd": "PD", "js": "JS", "T": "T", "dnuD": "DNU_D",
    "dur_dr": "DUR_DR", "dur_dz": "DUR_DZ", "duz_dr": "DUZ_DR", "duz_dz": "DUZ_DZ", "ur_over_r": "UR_OVER_R",
    "ur_star": "UR_STAR", "uz_star": "UZ_STAR", "f_r": "F_LYA_R", "f_z": "F_LYA_Z",
}
_COEFF_KEYS = ("Hvoigt", "phi", "alpha_F", "alpha", "alpha_c", "pd", "js", "dnuD", "T",
               "dur_dr", "dur_dz", "duz_dr", "duz_dz", "ur_over_r", "ur_star", "uz_star")
_VEL_KEYS = ("dur_dr", "dur_dz", "duz_dr", "duz_dz", "ur_over_r")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_scale(lya, c_tilde: float, dt: float) -> float:
    """J0 for a cloud: the larger of max J and the largest one-step emission
    js·c̃·Δt·max ϕ, so that J/J0 stays O(1) at the top of the range."""
    jmax = float(np.max(lya["J"])) if "J" in lya.arrays else 0.0
    src = float(np.max(lya["js"])) * c_tilde * dt * float(np.max(lya["phi"])) if "js" in lya.arrays else 0.0
    s = max(jmax, src)
    return s if (s > 0.0 and np.isfinite(s)) else 1.0


class M1LyaBatch:
    """Device-resident batch of clouds (owns an m1lya_batch*).

    clouds: StaticCloud objects (or anything with .lya, .geometry, .c_tilde,
    .dt, .Zc and .bc_z_bot).  scales: per-cloud J0 (None: default_scale).
    """

    def __init__(self, clouds: Sequence, scales: Sequence[float] | None = None, upload_state: bool = True):
        self.lib = _lib.load()
        self.clouds = list(clouds)
        n = len(self.clouds)
        if n == 0:
            raise ValueError("a batch needs at least one cloud")
        dims = [(c.lya.n_nu, c.lya.n_r, c.lya.n_z) for c in self.clouds]
        nn = (C.c_int * n)(*[d[0] for d in dims])
        nr = (C.c_int * n)(*[d[1] for d in dims])
        nz = (C.c_int * n)(*[d[2] for d in dims])
        h = C.c_void_p()
        check_batch(self.lib.m1lya_batch_create(n, nn, nr, nz, C.byref(h)), None)
        self.handle = h
        self.dims = dims
        for idx, cl in enumerate(self.clouds):
            lya = cl.lya
            g, keep = cl.geometry.to_c()
            self._check(self.lib.m1lya_batch_set_geometry(self.handle, idx, C.byref(g)))
            del keep
            fg = _lib.FreqGrid()
            arrs = [_f64(lya.nu), _f64(lya["nuC"]), _f64(lya["dnu"]), _f64(lya["inv_dnu"]), _f64(lya["inv_nuC"])]
            fg.nu, fg.nuC, fg.dnu, fg.inv_dnu, fg.inv_nuC = (ptr(a) for a in arrs)
            fg.kcore_min, fg.kcore_max = lya.kcore
            fg.recoil = int(bool(lya.recoil))
            fg.flux_method = _lib.HLL if lya.hll else _lib.GLF
            fg.alpha_floor = lya.alpha_floor
            self._check(self.lib.m1lya_batch_set_freq_grid(self.handle, idx, C.byref(fg)))
            del arrs
            J0 = scales[idx] if scales is not None else default_scale(lya, cl.c_tilde, cl.dt)
            self.set_scale(idx, J0)
            for key in _COEFF_KEYS:
                if key in lya.arrays:
                    if key in _VEL_KEYS and not np.any(np.asarray(lya[key])):
                        continue
                    self.upload(idx, key, lya[key])
            if upload_state:
                for key in ("J", "H_r", "H_z"):
                    if key in lya.arrays:
                        self.upload(idx, key, lya[key])
            self.set_step(idx, cl.c_tilde, cl.dt, cl.Zc, bc_z_bot=cl.bc_z_bot)

    # ---- plumbing ----
    def _check(self, rc):
        check_batch(rc, self.handle)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.m1lya_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.dims)

    def _shape(self, idx, key):
        nn, nr, nz = self.dims[idx]
        f = _KEYS[key]
        if f in _lib.SPECTRAL_FIELDS:
            return (nz, nr, nn)
        if f == "UR_STAR":
            return (nz, nr + 1)
        if f == "UZ_STAR":
            return (nz + 1, nr)
        return (nz, nr)

    def upload(self, idx: int, key: str, arr) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), self._shape(idx, key)))
        self._check(self.lib.m1lya_batch_upload(self.handle, idx, FIELD[_KEYS[key]], ptr(a), a.size))

    def download(self, idx: int, key: str) -> np.ndarray:
        a = np.empty(self._shape(idx, key))
        self._check(self.lib.m1lya_batch_download(self.handle, idx, FIELD[_KEYS[key]], ptr(a), a.size))
        return a

    def set_scale(self, idx: int, J0: float) -> None:
        self._check(self.lib.m1lya_batch_set_scale(self.handle, idx, float(J0)))

    def scale(self, idx: int) -> float:
        v = C.c_double()
        self._check(self.lib.m1lya_batch_get_scale(self.handle, idx, C.byref(v)))
        return v.value

    def set_step(self, idx: int, c_tilde: float, dt: float, z_center: float, bc_r_right: int = OUTFLOW,
                 bc_z_bot: int = REFLECT, bc_z_top: int = OUTFLOW) -> None:
        self._check(self.lib.m1lya_batch_set_step(self.handle, idx, float(c_tilde), float(dt), float(z_center),
                                                  int(bc_r_right), int(bc_z_bot), int(bc_z_top)))

    # ---- stepping ----
    def run_steps(self, n_steps: int, history: bool = True):
        """n_steps steps of every cloud; returns the (n_steps, n_clouds, 8) history or None."""
        hist = np.empty((n_steps, len(self), 8)) if history else None
        self._check(self.lib.m1lya_batch_run_steps(self.handle, int(n_steps), ptr(hist) if history else None))
        return hist

    def run_until_steady(self, max_steps: int, tol: float):
        """Steps until every cloud's ΔJ/J < tol (or max_steps).  Returns (history of the steps run,
        (steps, n_clouds, 8) with ΔJ/J in column 7; per-cloud step count at which each became steady)."""
        hist = np.empty((max(int(max_steps), 1), len(self), 8))
        steady = (C.c_int * len(self))()
        done = C.c_int(0)
        self._check(self.lib.m1lya_batch_run_until_steady(self.handle, int(max_steps), float(tol), ptr(hist),
                                                          steady, C.byref(done)))
        return hist[:done.value].copy(), np.array(steady[:], dtype=int)

    def step(self) -> list[dict]:
        h = self.run_steps(1)
        return [dict(zip(STEP_FIELDS, row[:7])) for row in h[0]]

    def scalars(self) -> np.ndarray:
        out = np.empty((len(self), 8))
        self._check(self.lib.m1lya_batch_scalars(self.handle, ptr(out)))
        return out

    def active_bins(self) -> np.ndarray:
        out = (C.c_int * (2 * len(self)))()
        self._check(self.lib.m1lya_batch_active_bins(self.handle, out))
        return np.array(list(out), dtype=np.int64).reshape(len(self), 2)

    def synchronize(self) -> None:
        self._check(self.lib.m1lya_batch_synchronize(self.handle))

    # ---- tiles, autotuner, graph ----
    def set_tiles(self, seg_len: int, waves_per_block: int) -> None:
        self._check(self.lib.m1lya_batch_set_tiles(self.handle, int(seg_len), int(waves_per_block)))

    def tiles(self) -> tuple[int, int]:
        s, w = C.c_int(), C.c_int()
        self._check(self.lib.m1lya_batch_get_tiles(self.handle, C.byref(s), C.byref(w)))
        return s.value, w.value

    def autotune(self, trial_steps: int = 3) -> tuple[int, int, float]:
        s, w, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.m1lya_batch_autotune(self.handle, int(trial_steps), C.byref(s), C.byref(w),
                                                  C.byref(ms)))
        return s.value, w.value, ms.value

    def set_graph(self, on: bool) -> None:
        self._check(self.lib.m1lya_batch_set_graph(self.handle, int(bool(on))))

    # ---- device-side coefficient refresh ----
    def set_voigt_lut(self, lut) -> None:
        """Copies the Voigt tables of a VoigtLUT and the line constants to the device."""
        from .m1lya import line_constants
        t = np.ascontiguousarray(lut.t, dtype=np.float64)
        self._check(self.lib.m1lya_batch_set_voigt_lut(self.handle, ptr(t), int(lut.nx), float(lut.xmax),
                                                       C.byref(line_constants())))

    def prepare_subcycling(self, idx: int, T, n_HI, alpha_c=None, alpha_c_F=None, js=None, pd=None,
                           phi_all_bins: bool = True) -> None:
        """PrepareLyαSubCycling! with Update_Hvoigt_And_ϕν_cell! on the device for one cloud."""
        nn, nr, nz = self.dims[idx]
        arrs = [None if v is None else np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (nz, nr)))
                for v in (T, n_HI, alpha_c, alpha_c_F, js, pd)]
        self._check(self.lib.m1lya_batch_prepare_subcycling(self.handle, int(idx),
                                                            *[ptr(a) if a is not None else None for a in arrs],
                                                            int(bool(phi_all_bins))))

    # ---- profiling ----
    def set_profiling(self, on) -> None:
        """True / 1: time every launch; 2: time k_step only; False / 0: off."""
        self._check(self.lib.m1lya_batch_set_profiling(self.handle, int(on)))

    def kernel_stats(self) -> dict:
        out = {}
        for idx in range(self.lib.m1lya_batch_kernel_stats_count(self.handle)):
            name = C.create_string_buffer(128)
            launches, total = C.c_long(), C.c_double()
            self._check(self.lib.m1lya_batch_kernel_stats(self.handle, idx, name, 128, C.byref(launches),
                                                          C.byref(total)))
            out[name.value.decode()] = (launches.value, total.value)
        return out

    def reset_kernel_stats(self) -> None:
        self._check(self.lib.m1lya_batch_reset_kernel_stats(self.handle))
