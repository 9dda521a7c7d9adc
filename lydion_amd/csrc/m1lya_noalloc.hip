// m1lya_noalloc.hip -- FindActiveBins on caller-owned device arrays with no device
// allocation: the partial sums live in a file-scope device array (m1lya_find_active_bins_noalloc
// in include/m1lya.h).
//
// The array is the same on every call on a device, and it keeps the last call's contents.
// Every row the partial kernel writes is written in full before it is read, so nothing needs
// clearing.  Calls that overlap on one device (two host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/m1lya.h"
#include "m1lya_internal.h"

namespace {

// Partial rows of the active-bin energies: at most kRows blocks of 2 N_ν doubles, N_ν <= kMaxNu,
// then the 2 N_ν column sums and the 2-int result.  4 MiB per device.
constexpr int kMaxNu = 256;
constexpr int kRows = 1024;
constexpr size_t kPartDoubles = (size_t)kRows * 2 * kMaxNu;
constexpr size_t kScratchDoubles = kPartDoubles + 2 * kMaxNu + 8;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// The device address of this device's copy of `scratch`, looked up once per device with that
// device current.  The array's name is a host-side stand-in and is never passed to a kernel.
int scratch_ptr(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(M1LYA_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(M1LYA_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(M1LYA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = (double*)p;
    }
    *out = g_ptr[dev];
    return M1LYA_OK;
}

}  // namespace

extern "C" {

int m1lya_noalloc_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return M1LYA_ERR_ARG;
    size_t n = g_err.size() < len - 1 ? g_err.size() : len - 1;
    g_err.copy(buf, n);
    buf[n] = '\0';
    return M1LYA_OK;
}

int m1lya_find_active_bins_noalloc(const double* J, const double* dnu, const double* dV, int n_nu, int n_r,
                                   int n_z, int kcore_min, int kcore_max, int* kmin, int* kmax)
{
    if (!J || !dnu || !dV || !kmin || !kmax) return fail(M1LYA_ERR_ARG, "NULL argument");
    if (n_nu < 3 || n_r < 1 || n_z < 1) return fail(M1LYA_ERR_DIMS, "needs N_nu >= 3, N_r >= 1, N_Z >= 1");
    if (n_nu > kMaxNu)
        return fail(M1LYA_ERR_DIMS, "N_nu = " + std::to_string(n_nu) + " exceeds the scratch maximum of " +
                                        std::to_string(kMaxNu) + " bins");
    if (kcore_min < 0 || kcore_max >= n_nu || kcore_min > kcore_max)
        return fail(M1LYA_ERR_ARG, "core bins outside [0, N_nu)");
    const long long ncells = (long long)n_r * n_z;
    if (ncells > 0x7fffffffLL) return fail(M1LYA_ERR_DIMS, "too many cells");
    double* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc != M1LYA_OK) return rc;
    // the block partition of m1lya_find_active_bins, so both sum in the same order
    const int nblocks = (int)(ncells < kRows ? ncells : kRows);
    const int pstride = 2 * n_nu;
    double* part = base;
    double* red = base + kPartDoubles;
    int* krange = (int*)(red + 2 * kMaxNu);
    hipStream_t s = nullptr;
    m1::launch_active_partial(s, J, dnu, dV, n_nu, (int)ncells, nblocks, part, pstride);
    m1::launch_reduce_cols(s, part, nblocks, pstride, 2 * n_nu, red);
    m1::launch_active_final(s, red, n_nu, kcore_min, kcore_max, krange);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(M1LYA_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    int h[2] = {0, 0};
    e = hipMemcpy(h, krange, sizeof(h), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(M1LYA_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    *kmin = h[0];
    *kmax = h[1];
    return M1LYA_OK;
}

}  // extern "C"
