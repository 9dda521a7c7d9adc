"""CPU checks of the drop-in boundary: libm1lya.so loads, exports every entry
point include/m1lya.h declares, and the host-side setup matches the reference's
rules.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "m1lya.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(m1lya_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_operator_api():
    names = _declared_functions()
    for required in ("m1lya_advect_lya", "m1lya_emission_absorption_update", "m1lya_frequency_update_all_cells",
                     "m1lya_update_freq_factors", "m1lya_update_f_mf_l", "m1lya_compute_f_esc", "m1lya_step"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) <= set(_declared_functions())
    assert _lib.load().m1lya_abi_version() == 1


def test_bad_arguments_fail_loudly_without_a_device():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.m1lya_create(2, 8, 8, 0, C.byref(h)) == _lib.ERR_DIMS
    assert lib.m1lya_create(64, 2, 8, 0, C.byref(h)) == _lib.ERR_DIMS
    assert lib.m1lya_create(577, 8, 8, 0, C.byref(h)) == _lib.ERR_DIMS
    buf = C.create_string_buffer(256)
    lib.m1lya_last_error(None, buf, 256)
    assert b"N_" in buf.value
    assert lib.m1lya_step(None, 1.0, 1.0, 0.0, 1, 0, 1, None) == _lib.ERR_ARG
    with pytest.raises(_lib.M1LyaError):
        _lib.check(lib.m1lya_create(1, 1, 1, 0, C.byref(h)))


def test_batch_bad_arguments_fail_loudly_without_a_device():
    lib = _lib.load()
    h = C.c_void_p()
    ints = lambda *v: (C.c_int * len(v))(*v)
    # dimension checks run before any HIP call
    assert lib.m1lya_batch_create(2, ints(64, 257), ints(8, 8), ints(8, 8), C.byref(h)) == _lib.ERR_DIMS
    assert lib.m1lya_batch_create(1, ints(64), ints(2), ints(8), C.byref(h)) == _lib.ERR_DIMS
    assert lib.m1lya_batch_create(0, ints(64), ints(8), ints(8), C.byref(h)) == _lib.ERR_ARG
    buf = C.create_string_buffer(256)
    lib.m1lya_batch_last_error(None, buf, 256)
    assert buf.value
    assert lib.m1lya_batch_run_steps(None, 1, None) == _lib.ERR_ARG
    assert lib.m1lya_batch_set_tiles(None, 16, 4) == _lib.ERR_ARG
    with pytest.raises(_lib.M1LyaError):
        _lib.check_batch(lib.m1lya_batch_create(1, ints(2), ints(8), ints(8), C.byref(h)))


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_bins_freq_grid_shape():
    for n_nu, tau in ((32, 1e5), (64, 1e6), (128, 1e8), (128, 1e9)):
        nu = L.static_cloud.config_grid(n_nu, tau)
        assert len(nu) == n_nu + 1
        assert np.all(np.diff(nu) > 0)
        assert np.allclose(nu + nu[::-1], 2 * L.constants.nu_Lya, rtol=1e-15)
        x_max = 4 * (L.lya_arrays.compute_a(1e4) * tau) ** (1 / 3)
        assert (nu[-1] - L.constants.nu_Lya) / L.lya_arrays.compute_dnuD(1e4) == pytest.approx(x_max, rel=1e-12)


def test_static_cloud_setup_point_source_and_luminosity():
    c = L.static_cloud_32x32x64()
    js = c.lya["js"]
    assert np.count_nonzero(js) == 1 and js[0, 0] > 0          # point source in the axis cell at Z = 0
    Lsrc = 4 * np.pi * (js * c.geometry.dV).sum()
    assert Lsrc == pytest.approx(1e5 * L.constants.Lsun, rel=1e-14)
    assert c.lya.n_nu == 64 and c.lya.kcore[0] < 32 < c.lya.kcore[1]
    # opacity: τ0 = α R_cl in the cloud
    assert c.lya["alpha"][0, 0] * L.constants.pc == pytest.approx(1e6, rel=1e-12)
    assert c.dt == pytest.approx(0.4 * 0.5 * (1.1 * L.constants.pc / 32) / L.constants.c, rel=1e-12)
