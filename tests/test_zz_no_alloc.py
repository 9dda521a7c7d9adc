"""m1lya_find_active_bins_noalloc: FindActiveBins on caller-owned device arrays with no device
allocation (partial sums in the library's fixed device array).  Checked against the context
entry point m1lya_find_active_bins on the same state, bit for bit (integer outputs), on two
calls in a row, the second with a smaller input, and its size limit."""
import ctypes as C

import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import _lib
from lydion_amd import m1lya as M
from tests.helpers import small_cloud

pytestmark = pytest.mark.gpu


def _windowed(cloud, k0, k1, seed):
    """A random state whose energy sits in bins k0..k1 (zero elsewhere), so the active window is
    interior and the threshold decides its ends."""
    lya = cloud.lya
    rng = np.random.Generator(np.random.PCG64(seed))
    J = np.zeros((lya.n_z, lya.n_r, lya.n_nu))
    J[:, :, k0:k1 + 1] = np.exp(rng.normal(0.0, 2.0, (lya.n_z, lya.n_r, k1 - k0 + 1))) * 1e-6
    J[:, :, k0] *= 1e-12  # the lowest bin sits near the 1e-10 threshold
    lya["J"] = J
    lya["J_old"] = J.copy()
    return cloud


class _Hip:
    """The HIP runtime the library is linked against (the test allocates the caller's device
    arrays with it; torch ships a runtime of its own)."""

    def __init__(self):
        _lib.load()   # the runtime libm1lya.so already holds, matched by its soname
        self.rt = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # hipMemcpyHostToDevice
        return p

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(cloud, hip):
    lya = cloud.lya
    J, dnu, dV = hip.upload(lya["J"]), hip.upload(lya["dnu"]), hip.upload(cloud.geometry.dV)
    kmin, kmax = C.c_int(-1), C.c_int(-1)
    kc0, kc1 = lya.kcore
    rc = _lib.load().m1lya_find_active_bins_noalloc(J, dnu, dV, lya.n_nu, lya.n_r, lya.n_z, kc0, kc1,
                                                      C.byref(kmin), C.byref(kmax))
    if rc != 0:
        buf = C.create_string_buffer(512)
        _lib.load().m1lya_noalloc_last_error(buf, 512)
        raise _lib.M1LyaError(rc, buf.value.decode())
    return kmin.value, kmax.value


def _reference(cloud):
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    r = M.find_active_bins(ctx)
    ctx.close()
    return r


def test_noalloc_matches_context_entry_point_twice():
    lut = L.VoigtLUT()
    big = _windowed(small_cloud(n_r=40, n_z=36, n_nu=128, tau0=1e8, lut=lut), 30, 100, seed=5)
    small = _windowed(small_cloud(n_r=12, n_z=10, n_nu=32, tau0=1e6, lut=lut), 6, 21, seed=6)
    r_big, r_small = _reference(big), _reference(small)
    hip = _Hip()
    got_big = _noalloc(big, hip)
    got_small = _noalloc(small, hip)   # the second call, with a smaller input, on the same array
    hip.free()
    print(f"noalloc: {got_big} (context {r_big}), then {got_small} (context {r_small})")
    assert got_big == r_big
    assert got_small == r_small
    assert r_big != r_small


def test_noalloc_rejects_more_bins_than_its_scratch():
    from lydion_amd.static_cloud import setup_static_cloud
    cloud = setup_static_cloud(1e6, 1e4, n_r=4, n_z=4, nu=L.reference_freq_grid(30.0, 1.03, 1.03))
    assert cloud.lya.n_nu > 256
    hip = _Hip()
    with pytest.raises(_lib.M1LyaError) as ei:
        _noalloc(cloud, hip)
    hip.free()
    assert ei.value.code == _lib.ERR_DIMS and "scratch maximum" in str(ei.value)
