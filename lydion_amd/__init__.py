"""lydion_amd -- MI355X-native multi-frequency M1 Lyα radiative-transfer step.

The hot path (advection, closure, emission/absorption, frequency solve and the
step reductions) runs in libm1lya.so, hand-written HIP for gfx950 behind the C
ABI of include/m1lya.h.  This package holds the host side: the ctypes binding,
the host setup mirrored from the reference (grid, ν grid, Voigt profile, static
cloud initialisation), the operator API with the reference's names, and the
mixed-precision batched step (M1LyaBatch).
"""
from . import constants
from ._lib import GLF, HLL, OUTFLOW, REFLECT, M1LyaError
from .grid import GeometryCylindrical, setup_cylindrical_grid
from .lya_arrays import LyaArrays, VoigtLUT, bins_freq_grid, generate_lya_arrays, reference_freq_grid
from .m1lya import (M1LyaContext, advect_lya, compute_f_esc_lya_cylindrical, edge_spectra, find_active_bins,
                    frequency_update_j_lya_all_cells, lya_emission_absorption_update, reset_drhou, run_steps,
                    step, update_e_lya, update_f_lya_mf_l_lya, update_freq_factors)
from .batch import M1LyaBatch
from .static_cloud import StaticCloud, static_cloud_32x32x64, cone_dust_cloud_256x256x128, point_source_cloud, setup_static_cloud

__all__ = [
    "constants", "GLF", "HLL", "OUTFLOW", "REFLECT", "M1LyaError", "GeometryCylindrical", "setup_cylindrical_grid",
    "LyaArrays", "VoigtLUT", "bins_freq_grid", "generate_lya_arrays", "reference_freq_grid", "M1LyaContext",
    "advect_lya", "compute_f_esc_lya_cylindrical", "edge_spectra", "find_active_bins",
    "frequency_update_j_lya_all_cells", "lya_emission_absorption_update", "reset_drhou", "run_steps", "step",
    "update_e_lya", "update_f_lya_mf_l_lya", "update_freq_factors", "StaticCloud", "static_cloud_32x32x64", "cone_dust_cloud_256x256x128",
    "point_source_cloud", "setup_static_cloud", "M1LyaBatch",
]
