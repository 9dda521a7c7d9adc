"""ctypes binding of libm1lya.so (include/m1lya.h).

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libm1lya.so")

# enum m1lya_status
OK, ERR_ARG, ERR_DIMS, ERR_HIP, ERR_NONFINITE, ERR_STATE, ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
REFLECT, OUTFLOW = 0, 1
GLF, HLL = 0, 1
FP64 = 0

# enum m1lya_field
FIELDS = [
    "J", "J_OLD", "H_R", "H_Z", "H_R_OLD", "H_Z_OLD", "ALPHA_F", "HVOIGT", "PHI_NU",
    "FIJ_KP", "FIJ_K", "FIJ_KM",
    "ALPHA_LYA", "ALPHA_C", "PD", "JS", "T", "DNU_D", "DUR_DR", "DUR_DZ", "DUZ_DR", "DUZ_DZ",
    "UR_OVER_R", "DRHOU_R", "DRHOU_Z", "F_LYA_R", "F_LYA_Z", "E_LYA",
    "UR_STAR", "UZ_STAR",
]
FIELD = {name: i for i, name in enumerate(FIELDS)}
SPECTRAL_FIELDS = set(FIELDS[:12])

_dp = C.POINTER(C.c_double)


class Geometry(C.Structure):
    _fields_ = [
        ("r", _dp), ("Z", _dp), ("rCoM", _dp), ("ZCoM", _dp),
        ("dV", _dp), ("inv_dV", _dp), ("A_r", _dp), ("A_Z", _dp), ("inv_r_cell", _dp),
        ("dr_L", _dp), ("dr_R", _dp), ("inv_dr_i_im1", _dp), ("inv_dr_ip1_i", _dp), ("inv_dr_ip2_ip1", _dp),
        ("dr_inner", C.c_double * 5), ("dr_outer", C.c_double * 5),
        ("dZ_L", _dp), ("dZ_R", _dp), ("inv_dZ_j_jm1", _dp), ("inv_dZ_jp1_j", _dp), ("inv_dZ_jp2_jp1", _dp),
        ("dZ_inner", C.c_double * 5), ("dZ_outer", C.c_double * 5),
    ]


class FreqGrid(C.Structure):
    _fields_ = [
        ("nu", _dp), ("nuC", _dp), ("dnu", _dp), ("inv_dnu", _dp), ("inv_nuC", _dp),
        ("kcore_min", C.c_int), ("kcore_max", C.c_int), ("recoil", C.c_int), ("flux_method", C.c_int),
        ("alpha_floor", C.c_double),
    ]


class LineConstants(C.Structure):
    """m1lya_line_constants"""
    _fields_ = [("nu_lya", C.c_double), ("b_prefact", C.c_double), ("nu_over_c", C.c_double),
                ("a_prefact", C.c_double), ("sigma0_prefact", C.c_double)]


class RslInputs(C.Structure):
    """m1lya_rsl_inputs"""
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("u_r", "u_z", "gamma_ad", "P", "rho", "rho_d", "xHII", "xH2",
                                                     "T")] + [("kappa_abs", C.c_double * 3), ("lc", LineConstants)]


EXPORTED = [
    "m1lya_abi_version", "m1lya_create", "m1lya_destroy", "m1lya_last_error", "m1lya_set_geometry",
    "m1lya_set_freq_grid", "m1lya_upload", "m1lya_download", "m1lya_fill", "m1lya_synchronize",
    "m1lya_find_active_bins", "m1lya_advect_lya", "m1lya_emission_absorption_update",
    "m1lya_update_freq_factors", "m1lya_frequency_update_all_cells", "m1lya_update_f_mf_l",
    "m1lya_update_f_mf_l_gravity", "m1lya_compute_f_esc", "m1lya_reset_drhou", "m1lya_update_e_lya",
    "m1lya_edge_count", "m1lya_edge_spectra", "m1lya_step", "m1lya_run_steps", "m1lya_run_until_steady",
    "m1lya_adjoint_run",
    "m1lya_set_voigt_lut", "m1lya_prepare_subcycling",
    "m1lya_set_profiling", "m1lya_set_adjoint_mode",
    "m1lya_kernel_stats_count", "m1lya_kernel_stats", "m1lya_reset_kernel_stats", "m1lya_selftest_math",
    "m1lya_batch_create", "m1lya_batch_destroy", "m1lya_batch_last_error", "m1lya_batch_count",
    "m1lya_batch_set_geometry", "m1lya_batch_set_freq_grid", "m1lya_batch_set_scale", "m1lya_batch_get_scale",
    "m1lya_batch_upload", "m1lya_batch_download", "m1lya_batch_set_step", "m1lya_batch_run_steps",
    "m1lya_batch_run_until_steady",
    "m1lya_batch_scalars", "m1lya_batch_active_bins", "m1lya_batch_set_tiles", "m1lya_batch_get_tiles",
    "m1lya_batch_autotune", "m1lya_batch_set_graph", "m1lya_batch_synchronize", "m1lya_batch_set_profiling",
    "m1lya_batch_kernel_stats_count", "m1lya_batch_kernel_stats", "m1lya_batch_reset_kernel_stats",
    "m1lya_batch_set_voigt_lut", "m1lya_batch_prepare_subcycling",
    "m1lya_find_active_bins_noalloc", "m1lya_noalloc_last_error",
    "m1lya_cfl_time_step", "m1lya_find_rsl", "m1lya_absorption_substep", "m1lya_prepare_subcycling_chem",
]

_lib = None


def load() -> C.CDLL:
    """Loads libm1lya.so and declares its signatures.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libm1lya.so not found at {LIB_PATH}; build it with `python -m lydion_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i, d, sz = C.c_int, C.c_double, C.c_size_t
    ip = C.POINTER(C.c_int)
    sig = {
        "m1lya_abi_version": (i, []),
        "m1lya_create": (i, [i, i, i, i, C.POINTER(vp)]),
        "m1lya_destroy": (None, [vp]),
        "m1lya_last_error": (i, [vp, C.c_char_p, sz]),
        "m1lya_set_geometry": (i, [vp, C.POINTER(Geometry)]),
        "m1lya_set_freq_grid": (i, [vp, C.POINTER(FreqGrid)]),
        "m1lya_upload": (i, [vp, i, _dp, sz]),
        "m1lya_download": (i, [vp, i, _dp, sz]),
        "m1lya_fill": (i, [vp, i, d]),
        "m1lya_synchronize": (i, [vp]),
        "m1lya_find_active_bins": (i, [vp, C.POINTER(i), C.POINTER(i)]),
        "m1lya_advect_lya": (i, [vp, d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i, i, i, i]),
        "m1lya_emission_absorption_update": (i, [vp, d, d]),
        "m1lya_update_freq_factors": (i, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "m1lya_frequency_update_all_cells": (i, [vp, d, _dp, _dp, _dp, _dp, _dp, _dp, i]),
        "m1lya_update_f_mf_l": (i, [vp, d, d, i, _dp]),
        "m1lya_update_f_mf_l_gravity": (i, [vp, d, _dp, _dp, d, i, _dp]),
        "m1lya_compute_f_esc": (i, [vp, d, i, _dp]),
        "m1lya_reset_drhou": (i, [vp]),
        "m1lya_update_e_lya": (i, [vp]),
        "m1lya_edge_count": (i, [vp]),
        "m1lya_edge_spectra": (i, [vp, _dp, _dp, sz]),
        "m1lya_step": (i, [vp, d, d, d, i, i, i, _dp]),
        "m1lya_run_steps": (i, [vp, i, d, d, d, i, i, i, _dp]),
        "m1lya_run_until_steady": (i, [vp, i, d, d, d, d, i, i, i, _dp, C.POINTER(C.c_int)]),
        "m1lya_find_active_bins_noalloc": (i, [vp, vp, vp, i, i, i, i, i, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "m1lya_noalloc_last_error": (i, [C.c_char_p, sz]),
        "m1lya_cfl_time_step": (i, [vp, d, _dp, _dp, d, _dp]),
        "m1lya_find_rsl": (i, [vp, d, d, C.POINTER(RslInputs), _dp]),
        "m1lya_absorption_substep": (i, [vp, _dp, d, d, _dp, _dp, C.POINTER(C.c_int), _dp]),
        "m1lya_prepare_subcycling_chem": (i, [vp, _dp, _dp, _dp, d, i, i, i]),
        "m1lya_adjoint_run": (i, [vp, i, d, d, d, i, i, i, _dp, _dp, i, _dp, _dp, _dp, _dp, _dp]),
        "m1lya_set_voigt_lut": (i, [vp, _dp, i, d, C.POINTER(LineConstants)]),
        "m1lya_prepare_subcycling": (i, [vp, _dp, _dp, _dp, _dp, _dp, _dp, i]),
        "m1lya_set_profiling": (i, [vp, i]),
        "m1lya_set_adjoint_mode": (i, [vp, i]),
        "m1lya_kernel_stats_count": (i, [vp]),
        "m1lya_kernel_stats": (i, [vp, i, C.c_char_p, sz, C.POINTER(C.c_long), _dp]),
        "m1lya_reset_kernel_stats": (i, [vp]),
        "m1lya_selftest_math": (i, [i, _dp, _dp, _dp, sz]),
        "m1lya_batch_create": (i, [i, ip, ip, ip, C.POINTER(vp)]),
        "m1lya_batch_destroy": (None, [vp]),
        "m1lya_batch_set_voigt_lut": (i, [vp, _dp, i, d, C.POINTER(LineConstants)]),
        "m1lya_batch_prepare_subcycling": (i, [vp, i, _dp, _dp, _dp, _dp, _dp, _dp, i]),
        "m1lya_batch_last_error": (i, [vp, C.c_char_p, sz]),
        "m1lya_batch_count": (i, [vp]),
        "m1lya_batch_set_geometry": (i, [vp, i, C.POINTER(Geometry)]),
        "m1lya_batch_set_freq_grid": (i, [vp, i, C.POINTER(FreqGrid)]),
        "m1lya_batch_set_scale": (i, [vp, i, d]),
        "m1lya_batch_get_scale": (i, [vp, i, _dp]),
        "m1lya_batch_upload": (i, [vp, i, i, _dp, sz]),
        "m1lya_batch_download": (i, [vp, i, i, _dp, sz]),
        "m1lya_batch_set_step": (i, [vp, i, d, d, d, i, i, i]),
        "m1lya_batch_run_steps": (i, [vp, i, _dp]),
        "m1lya_batch_run_until_steady": (i, [vp, i, d, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "m1lya_batch_scalars": (i, [vp, _dp]),
        "m1lya_batch_active_bins": (i, [vp, ip]),
        "m1lya_batch_set_tiles": (i, [vp, i, i]),
        "m1lya_batch_get_tiles": (i, [vp, ip, ip]),
        "m1lya_batch_autotune": (i, [vp, i, ip, ip, _dp]),
        "m1lya_batch_set_graph": (i, [vp, i]),
        "m1lya_batch_synchronize": (i, [vp]),
        "m1lya_batch_set_profiling": (i, [vp, i]),
        "m1lya_batch_kernel_stats_count": (i, [vp]),
        "m1lya_batch_kernel_stats": (i, [vp, i, C.c_char_p, sz, C.POINTER(C.c_long), _dp]),
        "m1lya_batch_reset_kernel_stats": (i, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


class M1LyaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libm1lya error {code}: {msg}")
        self.code = code


def check(rc: int, handle=None) -> None:
    if rc != OK:
        buf = C.create_string_buffer(1024)
        load().m1lya_last_error(handle, buf, 1024)
        raise M1LyaError(rc, buf.value.decode(errors="replace"))


def check_batch(rc: int, handle=None) -> None:
    if rc != OK:
        buf = C.create_string_buffer(1024)
        load().m1lya_batch_last_error(handle, buf, 1024)
        raise M1LyaError(rc, buf.value.decode(errors="replace"))
