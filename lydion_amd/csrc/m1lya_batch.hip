// m1lya_batch.hip -- the mixed-precision, batched M1 Lya step (m1lya_batch_* in
// include/m1lya.h): kernels and C ABI.
//
// A batch holds clouds of different sizes (N_r, N_Z >= 3, 3 <= N_ν <= 256).  Their
// state J, H_r, H_Z is stored in FP32, normalised per cloud by a scale J0.  One step
// of every cloud is one launch of k_step plus two small reduction launches:
//
//   * one wavefront marches along r over a segment of one cell row; lanes hold the
//     frequency bins (lane l owns bins lR .. lR+R-1, R = 1 for N_ν <= 64, 2 up to 128, 4 up to 256);
//   * FP32 flux sweep + M1 closure (src/AdvectJandH.jl:526-1390, src/M1Closure.jl):
//     the r interface between consecutive cells of the segment is evaluated once
//     and carried in registers, the two Z interfaces of each cell in place;
//   * FP64 emission / absorption (src/ConstantsAndFunctions.jl:483-573), frequency
//     factors (src/FreqFactors.jl:142-278; each bin's exponentials are evaluated
//     once and passed to the neighbouring rows by lane shuffles) and the frequency
//     solve of src/UpdateLyaJandH.jl:178-391 as a wavefront-parallel cyclic
//     reduction: local levels over the lane's bins (one for two, two for four), then parallel cyclic
//     reduction across the 64 lanes;
//   * H rescale and causality, FP32 store into the other state buffer;
//   * Neumaier-compensated FP64 accumulation of M_F, L, f_esc and the active-bin
//     energies: per block (k_step), per slice of a cloud's blocks (k_reduce), then per
//     cloud (k_final).
//
// Work items (one per wavefront) and partial-sum rows (one per block) cover every
// cloud of the batch, so uneven clouds share one launch.  The three launches of a
// step can be captured into a hipGraph per buffer parity; the segment length and
// the number of wavefronts per block are chosen by a timing autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/m1lya.h"
#include "m1lya_fast.h"
#include "m1lya_internal.h"

// FP contraction is on only inside the FP32 advection (m1lya_fast.h and the cell update
// below).  The FP64 emission, frequency factors, solve and reductions round as written:
// with contraction the backend fuses a multiply into an add or not depending on the code
// around it, so kernel variants that share this code (the velocity-shear variant, for
// one) would round it differently and a cloud's result would depend on its batch.

namespace mb {

constexpr double kC = 2.99792458e10;
constexpr double kInvC = 1.0 / kC;
constexpr double kFourPi = 4.0 * 3.141592653589793;
constexpr double kHoverKB = 6.626069573e-27 / 1.380648813e-16;
constexpr double kJeps = 1e-30;          // J floor in units of J0 (m1f::kJeps32)
constexpr double kHeps2 = 1e-200;        // |H|^2 floor of the cell-centred closure
constexpr double kEpsActive = 1e-10;     // FindActiveBins threshold
constexpr double kPmin = 1e-3, kPmax = 40.0;
constexpr int kMaxBins = 256;                    // four bins per lane
constexpr int kNScal = 7;                        // MF, L, net, out, esc, ΔJ sum, non-finite count
constexpr int kPartStride = 2 * kMaxBins + 8;    // e1[N_ν], e10[N_ν], scalars
constexpr int kMaxWaves = 8;
constexpr int kSlices = 16;                      // row slices of the first reduction stage

// Per-cell inputs of the step, one 160-byte record per cell (read with scalar loads).
struct CellRec {
    double alpha, alpha_c, pd, js, T, dnuD, dV, rcom, zcom;
    double durdr, durdz, duzdr, duzdz, urr;
    double A_rout, A_ztop, A_zbot;   // areas of the cell's faces on the outer r / top / bottom boundary, else 0
    float gRl, gRr, gZl, gZr;        // A/ΔV of the four faces
    float inv_r, pad;
};
static_assert(sizeof(CellRec) == 160, "CellRec layout");

// Per-bin constants of the frequency factors.
struct BinRec {
    double nuC, dnu, inv_dnu, inv_nuC;
    double half_R, half_L;   // ν_{k+1} - νC_k and νC_k - ν_k
};

struct Cloud {
    int nn, nr, nz, R;
    int pair;                      // N_ν <= 32: two cells per wavefront (k_step<..., 1, true>)
    long long o3, oc, ofr, ofz, obin, oivr, oivz;
    int blk0, blk1;                // this cloud's partial rows (= its blocks)
    int kcore_min, kcore_max, recoil, hll;
    int bc_r_right, bc_z_bot, bc_z_top, has_vel;
    float cdt, cdtc;               // c̃Δt and c̃Δt/c for the FP32 flux sweep
    double c_tilde, dt, z_center, J0, alpha_floor;
    // derived on the host when the descriptors are uploaded (sync_clouds)
    double cdt64;                  // c̃Δt
    double dw;                     // 1/(c̃Δt), the diagonal of the backward-Euler rows
    double src_scale;              // c̃Δt/J0: emission ∫js·ϕ_ν in units of J0
    double mf_scale;               // (4π/c)(1/c̃)·J0/Δt: Σ Δν ΔH in units of J0 -> f_Lyα
    // this cloud's slices of the batch arrays (set in sync_clouds)
    float* sJ[2];
    float* sHr[2];
    float* sHz[2];
    const float *saF, *sHv, *sphi, *sivr, *sivz, *sucr, *sucz;
    const CellRec* srec;
    double *sfr, *sfz;
};

struct Seg {
    int cloud, j, i0, i1;          // cells [i0, i1) of row j (j >= N_Z: a ghost row of the block)
};

struct Ctrl {
    double* hist;                  // [cap][n_clouds][8]
    int cap, step, next;           // history cursor of the running step / of the next step
    int halt;                      // set once every cloud is steady: later launches return at once
    int live;                      // k_reduce's view of halt, read by k_final (halt may change while
                                   // k_final runs, since one of its blocks sets it)
    int nsteady;                   // clouds whose ΔJ/J has fallen below tol
    int bad, bad_step;             // first cloud (+1) with a non-finite J, H_r or H_Z, and its step
    double tol;                    // steady-state tolerance of ΔJ/J (0: no stop)
    int* steady;                   // [n_clouds]: 1 + the step at which the cloud became steady, else 0
};

struct Arrays {
    float* J[2];
    float* Hr[2];
    float* Hz[2];
    float *aF, *Hv, *phi;
    const CellRec* rec;
    const BinRec* bin;
    float *uc_r, *uc_z;
    float *ivr, *ivz;
    double *f_r, *f_z;
};

// ------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------
struct KSum {   // Neumaier compensated sum
    double s, c;
    __device__ void add(double x)
    {
        const double t = s + x;
        c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

// Cross-lane moves through DPP (register to register).  CTRL: 0x138 wave_shr:1 (lane l
// receives lane l-1), 0x130 wave_shl:1 (lane l receives lane l+1), 0x110+n row_shr:n
// (within rows of 16 lanes); lanes without a source receive 0.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double from_prev_lane(double v) { return dpp_f64<0x138>(v); }   // 0 in lane 0
__device__ __forceinline__ double from_next_lane(double v) { return dpp_f64<0x130>(v); }   // 0 in lane 63

__device__ __forceinline__ float bperm_f32(int addr, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// The value of lane addr/4 (ds_bpermute, any lane of the wavefront).
__device__ __forceinline__ double bperm_f64(int addr, double v)
{
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sum over the 64 lanes, in every lane: scans within rows of 16, then the four row totals
// (a fixed order, so the result does not depend on anything but the inputs).
__device__ __forceinline__ double wave_sum64(double v)
{
    v += dpp_f64<0x111>(v);
    v += dpp_f64<0x112>(v);
    v += dpp_f64<0x114>(v);
    v += dpp_f64<0x118>(v);
    return (lane_value(v, 15) + lane_value(v, 31)) + (lane_value(v, 47) + lane_value(v, 63));
}

// Sum over each half-wavefront (lanes 0-31, 32-63), in every lane of the half; the same
// fixed order as wave_sum64 within a half.
__device__ __forceinline__ double half_sum32(double v, int half)
{
    v += dpp_f64<0x111>(v);
    v += dpp_f64<0x112>(v);
    v += dpp_f64<0x114>(v);
    v += dpp_f64<0x118>(v);
    const double s0 = lane_value(v, 15) + lane_value(v, 31), s1 = lane_value(v, 47) + lane_value(v, 63);
    return half ? s1 : s0;
}

// 1/x from the hardware estimate (relative error ~2^-24): drcp0 as is, for values that
// only set a relaxation rate; drcp1 after one Newton step (~2e-15), for pivots and
// quantities that set an equilibrium; drcp after two (a few ulp).
__device__ __forceinline__ double drcp0(double x) { return __builtin_amdgcn_rcp(x); }

__device__ __forceinline__ double drcp1(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

__device__ __forceinline__ double drcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// exp(x) for |x| <= kPmax = 40, the range of the clamped exponents of the frequency factors:
// Cody-Waite reduction x = n ln2 + r with ln2 in two parts (n·ln2_hi is exact for |n| < 2^11),
// |r| <= ln2/2, the degree-12 Taylor polynomial of exp(r) in Horner form (truncation below
// 2e-16 relative) and the scale 2^n; no overflow or underflow handling is needed in this range.
__device__ __forceinline__ double exp_clamped(double x)
{
    constexpr double kLog2e = 1.44269504088896338700e+00;
    constexpr double kLn2Hi = 6.93147180369123816490e-01;
    constexpr double kLn2Lo = 1.90821492927058770002e-10;
    const double n = __builtin_rint(x * kLog2e);
    const double r = fma(-n, kLn2Lo, fma(-n, kLn2Hi, x));
    double p = 1.0 / 479001600.0;
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)n);
}

// exp(x) for |x| <= kPmax to about 1e-7 relative: the power of two 2^n in FP64 and the fraction
// 2^r, |r| <= 1/2, from the hardware's FP32 exp2 (v_exp_f32, ~1 ulp).  The frequency factors
// are formed in conservative interface form, so an error in the exponentials moves the
// equilibrium ratio of neighbouring bins by that error and does not break photon conservation.
__device__ __forceinline__ double exp_fast(double x)
{
    constexpr double kLog2e = 1.44269504088896338700e+00;
    const double y = x * kLog2e;
    const double n = __builtin_rint(y);
    return __builtin_ldexp((double)__builtin_amdgcn_exp2f((float)(y - n)), (int)n);
}

// OneMinusϕ_L_and_χ_L / OneMinusϕ_R_and_χ_R (src/FreqFactors.jl:54-133) in FP64
__device__ __forceinline__ void omphi_L64(double P, double& o, double& chi)
{
    if (P >= kPmax) { o = exp(-kPmax); chi = (1.0 - o) * (1.0 / kPmax); return; }
    if (P <= -kPmax) { o = exp(kPmax); chi = (1.0 - o) * (-1.0 / kPmax); return; }
    if (fabs(P) < kPmin) { o = 1.0 - P + 0.5 * (P * P); chi = 1.0 - 0.5 * P; return; }
    o = exp_clamped(-P);
    chi = (1.0 - o) * drcp1(P);
}
__device__ __forceinline__ void omphi_R64(double P, double& o, double& chi)
{
    if (P >= kPmax) { o = exp(kPmax); chi = (1.0 - o) * (1.0 / kPmax); return; }
    if (P <= -kPmax) { o = exp(-kPmax); chi = (1.0 - o) * (-1.0 / kPmax); return; }
    if (fabs(P) < kPmin) { o = 1.0 + P + 0.5 * (P * P); chi = -1.0 - 0.5 * P; return; }
    o = exp_clamped(P);
    chi = (1.0 - o) * drcp1(P);
}

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// range of logical blocks so that neighbouring rows share its L2.
__device__ __forceinline__ int xcd_block(int b, int grid)
{
    const int per = grid >> 3;   // grid is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

// Ghost-aware FP32 loads (the ghost values the reference builds at each boundary):
//   r axis (i < 0): mirror, H_r negated                  src/AdvectJandH.jl:666-673
//   r outer (i >= N_r): :Reflect mirror / :Outflow copy with H_r = max(H_r,0)  :853-871
//   Z bottom (j < 0): :Reflect mirror / :Outflow copy with H_Z = min(H_Z,0)    :1000-1018
//   Z top (j >= N_Z): :Reflect mirror / :Outflow copy with H_Z = max(H_Z,0)    :1185-1203
__device__ __forceinline__ void ld_r(const float* J, const float* Hr, const float* Hz, int nn, int nr, int j,
                                     int k, int c, int bc_right, float& vj, float& vr, float& vz)
{
    int m;
    float sr = 1.0f;
    bool outflow = false;
    if (c < 0) {
        m = -c - 1; sr = -1.0f;
    } else if (c < nr) {
        m = c;
    } else if (bc_right == M1LYA_REFLECT) {
        m = 2 * nr - 1 - c; sr = -1.0f;
    } else {
        m = nr - 1; outflow = true;
    }
    const size_t idx = ((size_t)j * nr + m) * nn + k;
    vj = J[idx];
    const float h = Hr[idx];
    vr = outflow ? m1f::fmaxf_(h, 0.0f) : sr * h;
    vz = Hz[idx];
}

__device__ __forceinline__ void ld_z(const float* J, const float* Hr, const float* Hz, int nn, int nr, int nz,
                                     int i, int k, int c, int bc_bot, int bc_top, float& vj, float& vr,
                                     float& vz)
{
    int m, mode = 0;   // 0 plain, 1 negate, 2 min(H_Z,0), 3 max(H_Z,0)
    if (c < 0) {
        if (bc_bot == M1LYA_REFLECT) { m = -c - 1; mode = 1; }
        else { m = 0; mode = 2; }
    } else if (c < nz) {
        m = c;
    } else if (bc_top == M1LYA_REFLECT) {
        m = 2 * nz - 1 - c; mode = 1;
    } else {
        m = nz - 1; mode = 3;
    }
    const size_t idx = ((size_t)m * nr + i) * nn + k;
    vj = J[idx];
    vr = Hr[idx];
    const float h = Hz[idx];
    vz = (mode == 0) ? h : (mode == 1) ? -h : (mode == 2) ? m1f::fminf_(h, 0.0f) : m1f::fmaxf_(h, 0.0f);
}

// Compute_D_Edd (src/M1Closure.jl:46-83) in FP64 with the cancellation-free a, b.
__device__ inline void D_edd64(double Hr, double Hz, double J, double& Drr, double& Dzz, double& Drz)
{
    const double Hm2 = Hr * Hr + Hz * Hz;
    if (J <= kJeps || Hm2 <= kHeps2 || Hm2 <= 1e-20 * J * J) {
        Drr = 1.0 / 3.0; Dzz = 1.0 / 3.0; Drz = 0.0;
        return;
    }
    const double Hm = sqrt(Hm2);
    const double iH = drcp(Hm);
    double f = Hm * drcp(J);
    f = f < 1e-10 ? 1e-10 : (f > 1.0 - 1e-10 ? 1.0 - 1e-10 : f);
    double a, b;
    m1f::levermore_ab64(f, a, b);
    const double nr = Hr * iH, nz = Hz * iH;
    Drr = a + b * nr * nr;
    Dzz = a + b * nz * nz;
    Drz = b * nr * nz;
}

// Bins of one lane: R consecutive bins starting at lane*R.
template <int NQ>
struct LaneBins {
    int kb[NQ];
    bool val[NQ], act[NQ];
};

// ------------------------------------------------------------------------
// k_step: one full M1 step of every cloud of the batch.
// Block = wpb wavefronts, one row segment each; all segments of a block belong
// to one cloud, whose partial row the block writes.
// ------------------------------------------------------------------------
// One row of UpdateFreqFactors! (src/FreqFactors.jl:142-278) as the reference builds it, with
// bin k's own VEL = D:∇u / c in the exponents of bins k-1, k, k+1 (cells whose velocity
// gradient has shear or anisotropy).
__device__ __forceinline__ void aniso_row(const BinRec* __restrict__ bins, const float* __restrict__ Hv, int k, int nn,
                                       double hr, double hz, double jo, double durdr, double durdz, double duzdr,
                                       double duzdz, double urr, double PREF, double R0, int recoil, double dw,
                                       double& ra, double& rb, double& rc, double& rd)
{
    // contracted: only the shear variant of k_step holds this code, so every cell that takes
    // it rounds it the same way, and the fused form needs fewer registers
#pragma clang fp contract(fast)
    double Dr, Dz, Dx;
    D_edd64(hr, hz, jo, Dr, Dz, Dx);
    double Dpp = 1.0 - Dr - Dz;
    Dpp = Dpp > 0.0 ? Dpp : 0.0;
    const double v = (Dr * durdr + Dx * (durdz + duzdr) + Dz * duzdz + Dpp * urr) * kInvC;
    auto expo = [&](int x, double half, double& qx) {
        const double Kx = PREF * (double)Hv[x];
        qx = half * drcp1(Kx);
        double Cx = v * bins[x].nuC;
        if (recoil) Cx += Kx * (R0 - 2.0 * bins[x].inv_nuC);
        return fmin(fmax(Cx * qx, -kPmax), kPmax);
    };
    double qk, oLk, xLk, oRk, xRk;
    const double Pk = expo(k, bins[k].half_R, qk);
    omphi_L64(Pk, oLk, xLk);
    if (fabs(Pk) < kPmin || fabs(Pk) >= kPmax) {
        omphi_R64(Pk, oRk, xRk);
    } else {
        // exp(P) = 1 / exp(-P): one exponential for both sides of bin k
        oRk = drcp1(oLk);
        xRk = (1.0 - oRk) * drcp1(Pk);
    }
    const double idn = bins[k].inv_dnu;
    double qkp, oRkp, xRkp;
    omphi_R64(expo(k + 1, bins[k + 1].half_L, qkp), oRkp, xRkp);
    const double den_kp = idn * drcp1(xLk * qk - xRkp * qkp);
    double qkm, oLkm, xLkm;
    omphi_L64(expo(k - 1, bins[k - 1].half_R, qkm), oLkm, xLkm);
    const double den_km = idn * drcp1(xLkm * qkm - xRk * qk);
    ra = (k == 1) ? 0.0 : oLkm * den_km;
    rc = (k == nn - 2) ? 0.0 : oRkp * den_kp;
    rb = -((oLk * den_kp + oRk * den_km) + dw);
    rd = -dw * jo;
}

// LDS of a k_step block of wpb wavefronts (dynamic, so that 4-wavefront blocks keep four
// blocks per CU while 8-wavefront blocks take two):
//   sbin  BinRec[64 NQ]                 the cloud's frequency-bin constants, read every cell
//   sh    double[wpb][2 (64 NQ) + 8]    per-wavefront active-bin energies and step sums
//   sacc  double[wpb][2 kNScal]         per-wavefront compensated sums (s, c)
//   zrow  float[wpb + 4][3][64 NQ]      Z stencil rows j0-2 .. j0+wpb+1: J, H_r/J, H_Z/J
//   zflx  float[wpb + 1][3][64 NQ]      Z interfaces j0 .. j0+wpb: F_J, F_Hr, F_HZ
// (NQ: bins per lane of the instantiation, 2 for N_ν <= 128, 4 for N_ν <= 256)
// (bin k of lane l at [q*64 + l] in zrow and zflx).
constexpr size_t step_lds_bytes(int wpb, int nq)
{
    return sizeof(BinRec) * 64 * nq + sizeof(double) * (size_t)wpb * ((2 * 64 * nq + 8) + 2 * kNScal) +
           sizeof(float) * 3 * 64 * nq * (size_t)(2 * wpb + 5);
}

// kPair (with NQ = 1, clouds of N_ν <= 32): the two halves of a wavefront hold two neighbouring
// cells of its row, i (lanes 0-31) and i+1 (lanes 32-63), one bin per lane, so that such clouds
// fill every lane; the march advances two cells at a time.  The launch covers the blocks
// [blk0, blk1) of the batch.
template <bool kShear, int NQ, bool kPair>
__global__ __launch_bounds__(64 * kMaxWaves, NQ <= 2 ? 4 : 2) void k_step(const Cloud* __restrict__ clouds, Arrays A,
                                                            const Seg* __restrict__ segs, int blk0, int blk1,
                                                            int wpb, int parity, const int* __restrict__ krange,
                                                            double* __restrict__ part, const Ctrl* __restrict__ ctrl)
{
    static_assert(!kPair || NQ == 1, "pair mode holds one bin per lane");
    if (ctrl->halt) return;
    constexpr int kB = 64 * NQ;          // bins a wavefront holds
    constexpr int kSh = 2 * kB + 8;      // LDS row of the per-wavefront sums
    extern __shared__ __align__(16) unsigned char smem[];
    BinRec* sbin = reinterpret_cast<BinRec*>(smem);
    double (*sh)[kSh] = reinterpret_cast<double (*)[kSh]>(sbin + kB);
    double (*sacc)[2 * kNScal] = reinterpret_cast<double (*)[2 * kNScal]>(&sh[wpb][0]);
    float (*zrow)[3][kB] = reinterpret_cast<float (*)[3][kB]>(&sacc[wpb][0]);
    float (*zflx)[3][kB] = zrow + (wpb + 4);
    const int lb = blk0 + xcd_block(blockIdx.x, gridDim.x);
    if (lb >= blk1) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Seg sg = segs[(size_t)lb * wpb + wave];
    const int cidx = segs[(size_t)lb * wpb].cloud;
    const int j0 = segs[(size_t)lb * wpb].j;
    constexpr int kStep = kPair ? 2 : 1;            // cells per step of the march
    const int half = kPair ? (lane >> 5) : 0;       // the lane's cell is i + half

    // per-wavefront compensated sums of M_F, L, radial net/out and f_esc, kept by lane 0
    // in LDS (s, c pairs) rather than in registers of every lane
    enum { S_MF = 0, S_L, S_NET, S_OUT, S_ESC, S_CONV, S_BAD };
    if (lane < 2 * kNScal) sacc[wave][lane] = 0.0;
    auto kadd = [&](int m, double x) {
        if (lane == 0) {
            const double s0 = sacc[wave][2 * m], t = s0 + x;
            sacc[wave][2 * m + 1] += (fabs(s0) >= fabs(x)) ? (s0 - t) + x : (x - t) + s0;
            sacc[wave][2 * m] = t;
        }
    };
    const Cloud& cl = clouds[cidx];
    // the wavefront's row of sh accumulates the active-bin energies e1[k], e10[k] (in pair mode
    // each half its own: e1 at [32 half + k], e10 at [64 + 32 half + k], folded at the end)
    for (int t = lane; t < (kPair ? 128 : 2 * cl.nn); t += 64) sh[wave][t] = 0.0;
    const int nn = cl.nn, nr = cl.nr, nz = cl.nz, R = cl.R, j = sg.j;
    const bool row_valid = (j < nz);                    // rows above the cloud are ghost rows
    const int nvalid = min(wpb, nz - j0);
    const int hll = cl.hll, bcr = cl.bc_r_right, bcb = cl.bc_z_bot, bct = cl.bc_z_top;
    // The cloud's descriptor (step constants and array slices) is read where it is used,
    // through the constant address space: the loads are scalar and invariant, so a field
    // can be loaded again instead of being held across the cell loop in an SGPR (the kernel
    // is at the SGPR limit, and spilled SGPRs cost VALU lane reads and writes).
    using ConstCloud = __attribute__((address_space(4))) const Cloud;
    auto cloud_now = [&]() -> ConstCloud& { return *(ConstCloud*)(clouds + cidx); };
    for (int t = threadIdx.x; t < nn; t += blockDim.x) sbin[t] = A.bin[cl.obin + t];
    __syncthreads();
    const BinRec* bins = sbin;
    const int kmin = krange[2 * cidx], kmax = krange[2 * cidx + 1];

    // The lane's bins and their lane masks.  Inside the cell loop they are rebuilt from a
    // lane index made opaque in every iteration: a compare costs one instruction there,
    // while masks hoisted out of the loop occupy SGPR pairs that spill to VGPR lanes.
    auto lane_bins = [&](int ln) {
        LaneBins<NQ> B;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            B.kb[q] = kPair ? (ln & 31) : ln * R + q;
            B.val[q] = (q < R) && (B.kb[q] < nn);
            B.act[q] = B.val[q] && B.kb[q] >= kmin && B.kb[q] <= kmax;
        }
        return B;
    };
    LaneBins<NQ> L = lane_bins(lane);

    const int i0 = sg.i0;

    // Ghost-aware (J, H_r/J, H_Z/J) of cell (i, row) for the Z exchange; rows beyond
    // N_Z + 1 are never part of a needed stencil and are clamped.
    auto load_zrow = [&](int i, int row, int slot) {
        row = min(row, nz + 1);
        if (kPair) i = min(i + half, nr - 1);   // the lane's cell (a cell past the row is not used)
        auto& cp = cloud_now();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (!L.act[q]) continue;
            float vj, vr, vz, fr, fz;
            ld_z(cp.sJ[parity], cp.sHr[parity], cp.sHz[parity], nn, nr, nz, i, L.kb[q], row, bcb, bct, vj, vr, vz);
            m1f::flux_ratios(vj, vr, vz, fr, fz);
            const int p = q * 64 + lane;
            zrow[slot][0][p] = vj;
            zrow[slot][1][p] = fr;
            zrow[slot][2][p] = fz;
        }
    };

    // r window (J, H_r/J, H_Z/J) of cells (i-1, i, i+1, i+2); FL / kimR: flux and
    // K_ϕϕ of the left interface of cell i.
    float wJ[NQ][4], wR[NQ][4], wZ[NQ][4], FL[NQ][3], kimR[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            wJ[q][m] = 0.f; wR[q][m] = 0.f; wZ[q][m] = 0.f;
            // cells i0-2, i0-1, i0 in slots 1..3: the first shift and load of the
            // march (i = i0-1) make the window (i0-2, i0-1, i0, i0+1).  In pair mode the lane's
            // cells c-1, c (c = i0-2+half) in slots 2, 3: the first shift by two and two loads
            // make (c-1, c, c+1, c+2).
            if (row_valid && L.val[q] && m >= (kPair ? 2 : 1)) {
                auto& cp = cloud_now();
                float vr, vz;
                ld_r(cp.sJ[parity], cp.sHr[parity], cp.sHz[parity], nn, nr, j, L.kb[q],
                     kPair ? i0 - 5 + m + half : i0 - 3 + m, bcr, wJ[q][m], vr, vz);
                m1f::flux_ratios(wJ[q][m], vr, vz, wR[q][m], wZ[q][m]);
            }
        }
        FL[q][0] = FL[q][1] = FL[q][2] = 0.f;
        kimR[q] = 0.f;
    }

    // per lane: Σ Δν ΔV |J_new - J_start| (the cloud's ΔJ/J) and the non-finite outputs
    double conv = 0.0;
    int bad = 0;
    // lane t: Σ_k Δν (ΔH_r, ΔH_Z) of the emission/absorption update of cell i0 + t
    double seg_sr = 0.0, seg_sz = 0.0;
    // pair mode: flux and K_ϕϕ of the right interface of the other half's cell, one step back
    float Xp[3] = {0.f, 0.f, 0.f}, Xkp = 0.f;

    // All wavefronts of the block march over the same cells [i0, i1) in lockstep;
    // i = i0-1 only evaluates the left r interface of cell i0 (pair mode: i = i0-2, whose
    // upper half evaluates it).
    for (int i = i0 - kStep; i < sg.i1; i += kStep) {
        int ln = lane;
        asm volatile("" : "+v"(ln));
        L = lane_bins(ln);
        const bool first = (i < i0);
        const int ci = i + half;                          // the lane's cell
        const bool cell_ok = !kPair || ci < sg.i1;       // pair mode: i+1 may lie past the segment
        const int cc = kPair ? min(max(ci, 0), nr - 1) : i;   // its index for loads (clamped in pair mode)
        const size_t cell = (size_t)j * nr + cc;   // used only for valid rows, i >= i0
        const size_t base = cell * nn;

        float af_c[NQ] = {};   // α_F of cell i (the left side of its right r interface)
        // ---- right r interface of cell i (FP32) ----
        float FR[NQ][3], kip[NQ], kimN[NQ];
        {
        auto& cp = cloud_now();
        const float* aF = cp.saF;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            FR[q][0] = FR[q][1] = FR[q][2] = 0.f;
            kip[q] = 0.f; kimN[q] = 0.f;
            if (!row_valid || !L.val[q]) continue;
            const int k = L.kb[q];
            if (kPair) {
                wJ[q][0] = wJ[q][2]; wJ[q][1] = wJ[q][3];
                wR[q][0] = wR[q][2]; wR[q][1] = wR[q][3];
                wZ[q][0] = wZ[q][2]; wZ[q][1] = wZ[q][3];
#pragma unroll
                for (int m = 2; m < 4; ++m) {
                    float vr, vz;
                    ld_r(cp.sJ[parity], cp.sHr[parity], cp.sHz[parity], nn, nr, j, k, ci + m - 1, bcr, wJ[q][m], vr,
                         vz);
                    m1f::flux_ratios(wJ[q][m], vr, vz, wR[q][m], wZ[q][m]);
                }
            } else {
                wJ[q][0] = wJ[q][1]; wJ[q][1] = wJ[q][2]; wJ[q][2] = wJ[q][3];
                wR[q][0] = wR[q][1]; wR[q][1] = wR[q][2]; wR[q][2] = wR[q][3];
                wZ[q][0] = wZ[q][1]; wZ[q][1] = wZ[q][2]; wZ[q][2] = wZ[q][3];
                float vr, vz;
                ld_r(cp.sJ[parity], cp.sHr[parity], cp.sHz[parity], nn, nr, j, k, i + 2, bcr, wJ[q][3], vr, vz);
                m1f::flux_ratios(wJ[q][3], vr, vz, wR[q][3], wZ[q][3]);
            }
            af_c[q] = aF[((size_t)j * nr + max(cc, 0)) * nn + k];
            if (!L.act[q]) continue;
            m1f::Stencil32 st;
            st.J[0] = wJ[q][1]; st.J[1] = wJ[q][2]; st.J[2] = wJ[q][3]; st.J[3] = wJ[q][0];
            st.fr[0] = wR[q][1]; st.fr[1] = wR[q][2]; st.fr[2] = wR[q][3]; st.fr[3] = wR[q][0];
            st.fz[0] = wZ[q][1]; st.fz[1] = wZ[q][2]; st.fz[2] = wZ[q][3]; st.fz[3] = wZ[q][0];
            const float aL = af_c[q];
            // the right face of the lane's cell (pair mode: clamped to the row's faces for the
            // lanes whose cell is not used)
            const int fr_ = kPair ? min(max(ci + 1, 0), nr) : i + 1;
            const float aR = aF[((size_t)j * nr + min(fr_, nr - 1)) * nn + k];
            const float u = cp.sucr[(size_t)j * (nr + 1) + fr_];
            const m1f::Face32 f =
                m1f::interface_flux(0, hll, st, cp.sivr + (size_t)fr_ * m1f::kIvStride, aL, aR, u);
            FR[q][0] = f.FJ; FR[q][1] = f.FHr; FR[q][2] = f.FHz;
            kip[q] = f.kphi_L;
            kimN[q] = (fr_ == 0) ? m1f::kphi_inner_axis(f.J_R, f.Hz_R) : f.kphi_R;
        }
        }
        if constexpr (kPair) {
            // the left interface of cell i+1 is the right one of cell i (the other half, this
            // step); that of cell i is the right one of cell i-1 (the other half, last step)
            const int sw = (lane ^ 32) << 2;
            float X[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) X[m] = bperm_f32(sw, FR[0][m]);
            const float Xk = bperm_f32(sw, kimN[0]);
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                FL[0][m] = half ? X[m] : Xp[m];
                Xp[m] = X[m];
            }
            kimR[0] = half ? Xk : Xkp;
            Xkp = Xk;
            if (first) continue;
        } else if (first) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                FL[q][0] = FR[q][0]; FL[q][1] = FR[q][1]; FL[q][2] = FR[q][2];
                kimR[q] = kimN[q];
            }
            continue;
        }

        // ---- Z interfaces through LDS ----
        if (row_valid) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (!L.act[q]) continue;
                const int p = q * 64 + lane;
                zrow[wave + 2][0][p] = wJ[q][1];
                zrow[wave + 2][1][p] = wR[q][1];
                zrow[wave + 2][2][p] = wZ[q][1];
            }
        } else {
            load_zrow(i, j, wave + 2);
        }
        for (int h = wave; h < 4; h += wpb) {
            const int slot = (h < 2) ? h : wpb + h;
            load_zrow(i, j0 - 2 + slot, slot);
        }
        __syncthreads();
        // Wavefront w evaluates interface w.  The block's top interface (f = wpb) is split
        // between wavefronts 0 and 1 by bin parity when every lane holds two bins, so that no
        // wavefront evaluates two full interfaces before the barrier.
        // (with four bins per lane: the lower two to wavefront 0, the upper two to wavefront 1)
        const bool split_top = (R >= 2) && (wpb >= 2);
        constexpr int full = (1 << NQ) - 1;
        const int lo = (R == 4) ? 0x3 : 0x1, hi = (R == 4) ? 0xC : 0x2;
#pragma unroll 1
        for (int it = 0; it < 2; ++it) {
            const int f = it ? wpb : wave;
            const int qm = !it ? full : (wave == 0) ? (split_top ? lo : full) : (wave == 1 && split_top) ? hi : 0;
            if (f > nvalid || qm == 0) continue;
            const int g = j0 + f;
            auto& cp = cloud_now();
            const float* aF = cp.saF;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (!L.act[q] || !((qm >> q) & 1)) continue;
                const int k = L.kb[q];
                const int p = q * 64 + lane;
                m1f::Stencil32 st;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int slot = f + ((m + 1) & 3);   // cells (g-1, g, g+1, g-2)
                    st.J[m] = zrow[slot][0][p];
                    st.fr[m] = zrow[slot][1][p];
                    st.fz[m] = zrow[slot][2][p];
                }
                const float aL = aF[((size_t)max(g - 1, 0) * nr + cc) * nn + k];
                const float aR = aF[((size_t)min(g, nz - 1) * nr + cc) * nn + k];
                const m1f::Face32 r = m1f::interface_flux(1, hll, st, cp.sivz + (size_t)g * m1f::kIvStride, aL, aR,
                                                          cp.sucz[(size_t)g * nr + cc]);
                zflx[f][0][p] = r.FJ;
                zflx[f][1][p] = r.FHr;
                zflx[f][2][p] = r.FHz;
            }
        }
        __syncthreads();
        if (!row_valid) continue;

        auto& cv = cloud_now();
        const CellRec cr = cv.srec[cell];
        double jo[NQ], hr[NQ], hz[NQ];
        // ---- cell update (FP32): src/AdvectJandH.jl:1319-1379 ----
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
#pragma clang fp contract(fast)
            jo[q] = 0.0; hr[q] = 0.0; hz[q] = 0.0;
            if (!L.val[q]) continue;
            const int k = L.kb[q];
            const float Js = wJ[q][1], Hrs = cv.sHr[parity][base + k], Hzs = cv.sHz[parity][base + k];
            float Jn = Js, Hrn = Hrs, Hzn = Hzs;
            if (L.act[q]) {
                const int p = q * 64 + lane;
                const float gRl = cr.gRl, gRr = cr.gRr, gZl = cr.gZl, gZr = cr.gZr;
                const float dJ = gZr * zflx[wave + 1][0][p] - gZl * zflx[wave][0][p];
                const float dHr = gZr * zflx[wave + 1][1][p] - gZl * zflx[wave][1][p];
                const float dHz = gZr * zflx[wave + 1][2][p] - gZl * zflx[wave][2][p];
                Jn = Js - cv.cdt * ((gRr * FR[q][0] - gRl * FL[q][0]) + dJ);
                Jn = m1f::fmaxf_(Jn, m1f::kJeps32);
                Hzn = Hzs - cv.cdt * ((gRr * FR[q][2] - gRl * FL[q][2]) + dHz);
                Hrn = Hrs - cv.cdt * ((gRr * FR[q][1] - gRl * FL[q][1]) + dHr);
                Hrn += cv.cdt * (0.5f * (kimR[q] + kip[q])) * cr.inv_r;
                if (cv.has_vel) {
                    const float durdr = (float)cr.durdr, durdz = (float)cr.durdz;
                    const float duzdr = (float)cr.duzdr, duzdz = (float)cr.duzdz, urr = (float)cr.urr;
                    const m1f::Edd32 D = m1f::D_edd(Hrs, Hzs, Js);
                    const float Dpp = m1f::fmaxf_(1.0f - D.rr - D.zz, 0.0f);
                    const float Ddu = D.rr * durdr + D.rz * (durdz + duzdr) + D.zz * duzdz + Dpp * urr;
                    Hrn += -cv.cdtc * (Hrs * durdr + Hzs * durdz);
                    Hzn += -cv.cdtc * (Hrs * duzdr + Hzs * duzdz);
                    Jn = m1f::fmaxf_(Jn - cv.cdtc * Ddu * Js, m1f::kJeps32);
                }
                // EnforceCausalReducedFlux: src/AdvectJandH.jl:478-493
                m1f::causal_limit(Jn, Hrn, Hzn);
            }
            if (!kPair) {
                FL[q][0] = FR[q][0]; FL[q][1] = FR[q][1]; FL[q][2] = FR[q][2];
                kimR[q] = kimN[q];
            }
            jo[q] = Jn; hr[q] = Hrn; hz[q] = Hzn;
        }

        // ================= emission / absorption (FP64) =================
        // AbsorptionUpdateJorH (src/ConstantsAndFunctions.jl:451-462), J in units of J0;
        // the H deposit hr - hr/(1+x) is formed as hr·x/(1+x) (no cancellation)
        const double al_pd = cr.alpha * cr.pd, src = cr.js * cv.src_scale;
        const float* Hvp = cv.sHv;
        const float* php = cv.sphi;
        double hv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) hv[q] = 1.0;
        double sr = 0.0, sz = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (!L.val[q]) continue;
            const size_t idx = base + L.kb[q];
            const double h = Hvp[idx];
            hv[q] = h;
            const double aJ = cr.alpha_c + al_pd * h;
            const double jn = (jo[q] + src * (double)php[idx]) * drcp1(1.0 + aJ * cv.cdt64);
            const double xH = (double)af_c[q] * cv.cdt64;
            const double gH = xH * drcp1(1.0 + xH);          // 1 - 1/(1+x)
            const double dk = bins[L.kb[q]].dnu;
            sr += dk * (hr[q] * gH);
            sz += dk * (hz[q] * gH);
            jo[q] = jn;
            hr[q] -= hr[q] * gH;
            hz[q] -= hz[q] * gH;
        }
        // the cell's Σ Δν ΔH goes to lane i - i0: the per-cell terms of Update_f_Lyα_MF_L_Lyα!
        // are formed after the march, one cell per lane, instead of by the whole wavefront
        if constexpr (kPair) {
            sr = half_sum32(sr, half);
            sz = half_sum32(sz, half);
            const double sr0 = lane_value(sr, 0), sz0 = lane_value(sz, 0);
            const double sr1 = lane_value(sr, 32), sz1 = lane_value(sz, 32);
            if (ln == i - i0) {
                seg_sr = sr0;
                seg_sz = sz0;
            }
            if (ln == i + 1 - i0 && i + 1 < sg.i1) {
                seg_sr = sr1;
                seg_sz = sz1;
            }
        } else {
            sr = wave_sum64(sr);
            sz = wave_sum64(sz);
            if (ln == i - i0) {
                seg_sr = sr;
                seg_sz = sz;
            }
        }

        // ================= frequency factors (FP64) =================
        // Exponential fitting of src/FreqFactors.jl:142-278 in conservative interface
        // form.  Per bin m: K_m, C_m, P_m over the right half-width (the "dk" of row
        // m) and P'_m over the left half-width (the "dkp" of row m-1), oL_m = exp(-P_m),
        // oR'_m = exp(P'_m), A_m = χL(P_m)·q_m, B'_m = χR(P'_m)·q'_m.  Per interface
        // m+1/2: G = 1/(A_m - B'_{m+1}), shared by rows m and m+1, so the frequency
        // flux leaving one bin is exactly the flux entering the next whatever the
        // rounding of G.  Row k: Fkp = oR'_{k+1}·G_{k+1/2}/Δν_k, Fkm = oL_{k-1}·
        // G_{k-1/2}/Δν_k, Fk = (oL_k·G_{k+1/2} + oR'_k·G_{k-1/2})/Δν_k.  (The reference
        // row k+1 uses P_{k+1} where this uses P'_{k+1}; the two differ by the rounding
        // of the bin centre, ~1e-12 relative.)
        const bool vel = cv.has_vel && (cr.durdr != 0.0 || cr.durdz != 0.0 || cr.duzdr != 0.0 ||
                                        cr.duzdz != 0.0 || cr.urr != 0.0);
        // With an isotropic velocity gradient (∂u_r/∂r = ∂u_Z/∂Z = u_r/r, no shear) D:∇u is the
        // trace times the rate for every bin, so VEL is the same in all bins and the shared-G
        // interface form below equals the reference rows.  Otherwise VEL differs between bins and
        // each row is built as the reference builds it, with its own bin's VEL in the exponents of
        // its three neighbours (src/FreqFactors.jl:226-254).
        const bool aniso =
            kShear && vel && !((cr.durdz + cr.duzdr) == 0.0 && cr.durdr == cr.duzdz && cr.duzdz == cr.urr);
        const double PREF = 0.5 * (cr.alpha > cv.alpha_floor ? cr.alpha : cv.alpha_floor) * (cr.dnuD * cr.dnuD);
        const double R0 = cv.recoil ? kHoverKB / cr.T : 0.0;
        const int R1 = R - 1;
        double ra[NQ], rb[NQ], rc[NQ], rd[NQ];
        if (!aniso) {
            double Aq[NQ], oLq[NQ], Bpq[NQ], oRpq[NQ];
    #pragma unroll
            for (int q = 0; q < NQ; ++q) {
                Aq[q] = 1.0; oLq[q] = 1.0; Bpq[q] = 0.0; oRpq[q] = 1.0;
                if (!L.val[q]) continue;
                const BinRec bk = bins[L.kb[q]];
                double VEL = 0.0;
                if (vel) {
                    double Dr, Dz, Dx;
                    D_edd64(hr[q], hz[q], jo[q], Dr, Dz, Dx);
                    double Dpp = 1.0 - Dr - Dz;
                    Dpp = Dpp > 0.0 ? Dpp : 0.0;
                    VEL = (Dr * cr.durdr + Dx * (cr.durdz + cr.duzdr) + Dz * cr.duzdz + Dpp * cr.urr) * kInvC;
                }
                const double K = PREF * hv[q];
                const double iK = drcp1(K);
                double C = VEL * bk.nuC;
                if (cv.recoil) C += K * (R0 - 2.0 * bk.inv_nuC);
                const double qR = bk.half_R * iK, qL = bk.half_L * iK;
                const double P = fmin(fmax(C * qR, -kPmax), kPmax);
                const double Pp = fmin(fmax(C * qL, -kPmax), kPmax);
                double oL, xL, oRp, xRp;
                if (fabs(P) < kPmin) {
                    oL = 1.0 - P + 0.5 * (P * P);
                    xL = 1.0 - 0.5 * P;
                } else {
                    oL = exp_fast(-P);
                    xL = (1.0 - oL) * drcp0(P);
                }
                if (fabs(Pp) < kPmin) {
                    oRp = 1.0 + Pp + 0.5 * (Pp * Pp);
                    xRp = -1.0 - 0.5 * Pp;
                } else {
                    // exp(P') = exp(P)·exp(P' - P); P' - P is the rounding of the bin centre
                    const double d = Pp - P;
                    oRp = (fabs(d) < 1e-4) ? drcp1(oL) * (1.0 + d * (1.0 + d * (0.5 + d * (1.0 / 6.0)))) : exp_fast(Pp);
                    xRp = (1.0 - oRp) * drcp0(Pp);
                }
                Aq[q] = xL * qR;
                oLq[q] = oL;
                Bpq[q] = xRp * qL;
                oRpq[q] = oRp;
            }
            // neighbours: bin k-1 -> oL (lane above for q = 0), bin k+1 -> B', oR' (lane below for q = R-1)
            auto lane_last = [&](const double(&v)[NQ]) {   // the value of the lane's last bin, q = R-1
                double r = v[0];
#pragma unroll
                for (int q = 1; q < NQ; ++q) r = (q == R1) ? v[q] : r;
                return r;
            };
            const double oL_prev = from_prev_lane(lane_last(oLq));
            const double Bp_next = from_next_lane(Bpq[0]), oRp_next = from_next_lane(oRpq[0]);
            // interfaces owned by this lane: Gq[q] = G at the top of bin lR+q
            double Gq[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                Gq[q] = drcp0(Aq[q] - ((q == R1) ? Bp_next : Bpq[q + 1 < NQ ? q + 1 : NQ - 1]));
            const double G_prev = from_prev_lane(lane_last(Gq));

#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                ra[q] = 0.0; rb[q] = 1.0; rc[q] = 0.0; rd[q] = L.val[q] ? jo[q] : 0.0;
                const int k = L.kb[q];
                if (!L.val[q] || k < 1 || k > nn - 2) continue;
                const double idn = bins[k].inv_dnu;
                const double oLm = (q == 0) ? oL_prev : oLq[q > 0 ? q - 1 : 0];
                const double oRpp = (q == R1) ? oRp_next : oRpq[q + 1 < NQ ? q + 1 : NQ - 1];
                const double Gm = (q == 0) ? G_prev : Gq[q > 0 ? q - 1 : 0];
                const double Gp = Gq[q];
                const double Fkp = oRpp * Gp * idn;
                const double Fkm = oLm * Gm * idn;
                const double Fk = (oLq[q] * Gp + oRpq[q] * Gm) * idn;
                // rows of FrequencyUpdateJLyα! (src/UpdateLyaJandH.jl:200-215); gtsv drops
                // the couplings to bins 0 and N_ν-1, which keep their value
                ra[q] = (k == 1) ? 0.0 : Fkm;
                rc[q] = (k == nn - 2) ? 0.0 : Fkp;
                rb[q] = -(Fk + cv.dw);
                rd[q] = -cv.dw * jo[q];
            }
        } else {
            // one bin at a time (a scheduling barrier between them keeps the two rows' temporaries
            // from being live together)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                ra[q] = 0.0; rb[q] = 1.0; rc[q] = 0.0; rd[q] = L.val[q] ? jo[q] : 0.0;
                const int k = L.kb[q];
                if (L.val[q] && k >= 1 && k <= nn - 2)
                    aniso_row(bins, Hvp + base, k, nn, hr[q], hz[q], jo[q], cr.durdr, cr.durdz, cr.duzdr,
                              cr.duzdz, cr.urr, PREF, R0, cv.recoil, cv.dw, ra[q], rb[q], rc[q], rd[q]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }

        // ================= tridiagonal solve: CR levels + PCR over the lanes =================
        double a, b, c, d, ua = 0.0, uc = 0.0, ud = 0.0;
        // four bins per lane: a first cyclic-reduction level eliminates bins 1 and 3 of the lane into
        // bins 0 and 2 (bin 3 of the lane above couples to bin 0); rows 0' and 2' then form the
        // two-bins-per-lane system the level below reduces
        double u1a = 0.0, u1c = 0.0, u1d = 0.0, u3a = 0.0, u3c = 0.0, u3d = 0.0;
        if (NQ == 4 && R == 4) {
            const double ib1 = drcp1(rb[1]), ib3 = drcp1(rb[NQ - 1]);
            u1a = ra[1] * ib1; u1c = rc[1] * ib1; u1d = rd[1] * ib1;
            u3a = ra[NQ - 1] * ib3; u3c = rc[NQ - 1] * ib3; u3d = rd[NQ - 1] * ib3;
            const double pa = from_prev_lane(u3a), pc = from_prev_lane(u3c), pd3 = from_prev_lane(u3d);
            const double b0 = rb[0] - ra[0] * pc - rc[0] * u1a;
            const double d0 = rd[0] - ra[0] * pd3 - rc[0] * u1d;
            const double a0 = -ra[0] * pa;
            const double c0 = -rc[0] * u1c;
            const double b2 = rb[2 % NQ] - ra[2 % NQ] * u1c - rc[2 % NQ] * u3a;
            const double d2 = rd[2 % NQ] - ra[2 % NQ] * u1d - rc[2 % NQ] * u3d;
            const double a2 = -ra[2 % NQ] * u1a;
            const double c2 = -rc[2 % NQ] * u3c;
            ra[0] = a0; rb[0] = b0; rc[0] = c0; rd[0] = d0;
            ra[1] = a2; rb[1] = b2; rc[1] = c2; rd[1] = d2;
        }
        if (NQ >= 2 && R >= 2) {
            // eliminate the odd row of each lane's pair into the even rows
            const double ib1 = drcp1(rb[1]);
            ua = ra[1] * ib1; uc = rc[1] * ib1; ud = rd[1] * ib1;
            const double pua = from_prev_lane(ua), puc = from_prev_lane(uc), pud = from_prev_lane(ud);
            b = rb[0] - ra[0] * puc - rc[0] * ua;
            d = rd[0] - ra[0] * pud - rc[0] * ud;
            a = -ra[0] * pua;
            c = -rc[0] * uc;
        } else {
            a = ra[0]; b = rb[0]; c = rc[0]; d = rd[0];
        }
        {
            // level 1: neighbours through DPP
            const double ib = drcp1(b);
            const double u = a * ib, v = c * ib, w = d * ib;
            const double ub = from_prev_lane(u), vb = from_prev_lane(v), wb = from_prev_lane(w);
            const double uf = from_next_lane(u), vf = from_next_lane(v), wf = from_next_lane(w);
            const double nb = b - a * vb - c * uf;
            const double nd = d - a * wb - c * wf;
            a = -a * ub;
            c = -c * vf;
            b = nb;
            d = nd;
        }
        // Levels 2..32 read lanes l -/+ s through ds_bpermute, wrapping around the wavefront.
        // The wrapped values meet zero coefficients: the lanes below s have a = 0 and those
        // above 63 - s have c = 0 at level s (bin 0, bin N_ν-1 and the bins past N_ν are
        // identity rows, and level s/2 zeroed the next s/2 lanes by multiplying with them),
        // and the values are finite, so no select is needed.
        // (pair mode: each half is its own system of 32 rows; the lanes wrap within the half, and
        // level 1's DPP moves across the halves meet bin 0's a = 0 and bin 31's c = 0)
#pragma unroll
        for (int s = 2; s < (kPair ? 32 : 64); s <<= 1) {
            const double ib = drcp1(b);
            const double u = a * ib, v = c * ib, w = d * ib;
            const int au = (kPair ? (((ln - s) & 31) | (ln & 32)) : ((ln - s) & 63)) << 2;
            const int ad = (kPair ? (((ln + s) & 31) | (ln & 32)) : ((ln + s) & 63)) << 2;
            const double ub = bperm_f64(au, u), vb = bperm_f64(au, v), wb = bperm_f64(au, w);
            const double uf = bperm_f64(ad, u), vf = bperm_f64(ad, v), wf = bperm_f64(ad, w);
            const double nb = b - a * vb - c * uf;
            const double nd = d - a * wb - c * wf;
            a = -a * ub;
            c = -c * vf;
            b = nb;
            d = nd;
        }
        double xs[NQ];
        xs[0] = d * drcp1(b);
#pragma unroll
        for (int q = 1; q < NQ; ++q) xs[q] = 0.0;
        if (NQ >= 2 && R >= 2) {
            const double x0n = from_next_lane(xs[0]);
            xs[1] = ud - ua * xs[0] - uc * x0n;
            if (NQ == 4 && R == 4) {
                // back through the first level: bins 2, 1, 3 of the lane from bins 0, 2 and the next lane's 0
                const double x2 = xs[1];
                xs[2 % NQ] = x2;
                xs[1] = u1d - u1a * xs[0] - u1c * x2;
                xs[NQ - 1] = u3d - u3a * x2 - u3c * x0n;
            }
        }

        // ================= H correction, store, reductions =================
        const bool edge = (cr.A_rout != 0.0) || (cr.A_ztop != 0.0) || (cr.A_zbot != 0.0 && bcb == M1LYA_OUTFLOW);
        const double Azb = (bcb == M1LYA_OUTFLOW) ? cr.A_zbot : 0.0;
        double escl = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (!L.val[q] || !cell_ok) continue;
            const int k = L.kb[q];
            // J floor, H rescale and causality: src/UpdateLyaJandH.jl:335-379
            const double jn = (k >= 1 && k <= nn - 2) ? xs[q] : jo[q];
            const double Jk = jn > kJeps ? jn : kJeps;
            const double ratio = Jk * drcp0(jo[q] > kJeps ? jo[q] : kJeps);
            double h_r = hr[q] * ratio, h_z = hz[q] * ratio;
            const double Hm2 = h_r * h_r + h_z * h_z;
            if (Hm2 > Jk * Jk) {
                const double s = Jk * drcp1(sqrt(Hm2));
                h_r *= s;
                h_z *= s;
            }
            const float Jf = (float)Jk, hrf = (float)h_r, hzf = (float)h_z;
            // the floor would hide a NaN in J (NaN > Jeps is false): test the solve's output
            // (a NaN or an infinity in any of the three makes the sum non-finite; the state is
            // O(1) in units of J0, far from FLT_MAX, so a finite sum cannot overflow)
            bad |= !__builtin_isfinite(((float)jn + hrf) + hzf);
            const size_t idx = base + k;
            auto& cw = cloud_now();
            cw.sJ[parity ^ 1][idx] = Jf;
            cw.sHr[parity ^ 1][idx] = hrf;
            cw.sHz[parity ^ 1][idx] = hzf;
            const double Jd = Jf, dk = bins[k].dnu;
            conv += dk * (double)fabsf(Jf - wJ[q][1]) * cr.dV;
            const int e1k = kPair ? 32 * half + k : k, e10k = kPair ? 64 + 32 * half + k : nn + k;
            sh[wave][e1k] += dk * ((Jd - kJeps) > 0.0 ? (Jd - kJeps) : 0.0) * cr.dV;
            sh[wave][e10k] += dk * ((Jd - 10.0 * kJeps) > 0.0 ? (Jd - 10.0 * kJeps) : 0.0) * cr.dV;
            // Compute_f_esc_Cylindrical terms (src/ConstantsAndFunctions.jl:937-953)
            if (edge)
                escl += dk * kFourPi *
                        (cr.A_rout * fmax((double)hrf, 0.0) + cr.A_ztop * fmax((double)hzf, 0.0) +
                         Azb * fmax(-(double)hzf, 0.0));
        }
        if (kPair ? (__ballot(edge) != 0) : edge) kadd(S_ESC, wave_sum64(escl) * cv.J0);
    }

    // ---- Update_f_Lyα_MF_L_Lyα! per-cell terms (src/ConstantsAndFunctions.jl:612-660) of the
    //      segment's cells, one per lane, summed over the lanes in a fixed order ----
    if (row_valid) {
        auto& cv = cloud_now();
        double mf = 0.0, ls = 0.0, net = 0.0, out = 0.0;
        const int i = i0 + lane;
        if (i < sg.i1) {
            const size_t cell = (size_t)j * nr + i;
            const CellRec* cr = cv.srec + cell;
            const double fr = seg_sr * cv.mf_scale, fz = seg_sz * cv.mf_scale;
            cv.sfr[cell] = fr;
            cv.sfz[cell] = fz;
            const double dV = cr->dV;
            mf = dV * sqrt(fr * fr + fz * fz);
            ls = kFourPi * dV * cr->js;
            const double ddr = cr->rcom, ddz = cr->zcom - cv.z_center;
            const double d2 = ddr * ddr + ddz * ddz;
            double frad = 0.0;
            if (d2 > 0.0) frad = (fr * ddr + fz * ddz) * drcp1(sqrt(d2));
            net = dV * frad;
            out = dV * (frad > 0.0 ? frad : 0.0);
        }
        kadd(S_MF, wave_sum64(mf));
        kadd(S_L, wave_sum64(ls));
        kadd(S_NET, wave_sum64(net));
        kadd(S_OUT, wave_sum64(out));
    }

    // ---- block partial row: per-wavefront values in LDS, summed in wavefront order ----
    const Cloud& cb = clouds[cidx];
    const int nnb = cb.nn;
    {
        const double cs = wave_sum64(conv);
        const double nb = (double)__popcll(__ballot(bad != 0));
        kadd(S_CONV, cs * cb.J0);
        kadd(S_BAD, nb);
    }
    if constexpr (kPair) {
        // fold the two halves' active-bin energies into the row layout e1[k], e10[k]
        double e1 = 0.0, e10 = 0.0;
        if (lane < 32) {
            e1 = sh[wave][lane] + sh[wave][32 + lane];
            e10 = sh[wave][64 + lane] + sh[wave][96 + lane];
        }
        __syncthreads();
        if (lane < nnb) {
            sh[wave][lane] = e1;
            sh[wave][nnb + lane] = e10;
        }
    }
    for (int t = lane; t < 2 * nnb; t += 64) sh[wave][t] *= cb.J0;
    if (lane == 0) {
        for (int m = 0; m < kNScal; ++m) sh[wave][2 * nnb + m] = sacc[wave][2 * m] + sacc[wave][2 * m + 1];
    }
    __syncthreads();
    const int ncol = 2 * nnb + kNScal;
    for (int t = threadIdx.x; t < ncol; t += blockDim.x) {
        KSum s{0, 0};
        for (int w = 0; w < wpb; ++w) s.add(sh[w][t]);
        part[(size_t)lb * kPartStride + t] = s.value();
    }
}

// Active-bin energies straight from the state (before the first step after an
// upload): block rows in k_step's format, scalars zero.
__global__ __launch_bounds__(64 * kMaxWaves) void k_active_rows(const Cloud* __restrict__ clouds, Arrays A,
                                                     const Seg* __restrict__ segs, int nblk, int wpb, int parity,
                                                     double* __restrict__ part)
{
    __shared__ double sh[kMaxWaves][kPartStride];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    if (lb >= nblk) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Seg sg = segs[(size_t)lb * wpb + wave];
    const Cloud& cb = clouds[segs[(size_t)lb * wpb].cloud];
    const int nnb = cb.nn, R = cb.R;
    double e1[4] = {0.0, 0.0, 0.0, 0.0}, e10[4] = {0.0, 0.0, 0.0, 0.0};
    if (sg.j < cb.nz) {
        const float* J = A.J[parity] + cb.o3;
        for (int i = sg.i0; i < sg.i1; ++i) {
            const size_t cell = (size_t)sg.j * cb.nr + i;
            const double dV = A.rec[cb.oc + cell].dV;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = lane * R + q;
                if (q >= R || k >= nnb) continue;
                const double Jd = J[cell * nnb + k];
                const double dk = A.bin[cb.obin + k].dnu;
                e1[q] += dk * ((Jd - kJeps) > 0.0 ? (Jd - kJeps) : 0.0) * dV;
                e10[q] += dk * ((Jd - 10.0 * kJeps) > 0.0 ? (Jd - 10.0 * kJeps) : 0.0) * dV;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = lane * R + q;
        if (q < R && k < nnb) {
            sh[wave][k] = e1[q] * cb.J0;
            sh[wave][nnb + k] = e10[q] * cb.J0;
        }
    }
    if (lane < kNScal) sh[wave][2 * nnb + lane] = 0.0;
    __syncthreads();
    const int ncol = 2 * nnb + kNScal;
    for (int t = threadIdx.x; t < ncol; t += blockDim.x) {
        KSum s{0, 0};
        for (int w = 0; w < wpb; ++w) s.add(sh[w][t]);
        part[(size_t)lb * kPartStride + t] = s.value();
    }
}

// Per-cloud sums of the block rows, stage 1: block (column chunk, cloud, slice) sums
// one slice of the cloud's rows; lanes over columns, wavefronts over rows, Neumaier
// sums combined in a fixed order (deterministic).  Block (0,0,0) also advances the
// history cursor to the value the previous step's k_final left.
__global__ __launch_bounds__(256) void k_reduce(const Cloud* __restrict__ clouds, const double* __restrict__ part,
                                                double* __restrict__ red, Ctrl* __restrict__ ctrl)
{
    __shared__ double ss[4][64], sc[4][64];
    // the halt flag and the cloud's extents are loaded together (one memory latency)
    const int halt = ctrl->halt;
    const Cloud& cl = clouds[blockIdx.y];
    const int nn = cl.nn, blk0 = cl.blk0, blk1 = cl.blk1;
    const bool halted = halt != 0;
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
        ctrl->live = !halted;
        if (!halted) ctrl->step = ctrl->next;
    }
    if (halted) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + lane;
    const int ncol = 2 * nn + kNScal;
    if ((int)blockIdx.x * 64 >= ncol) return;
    const int nrow = blk1 - blk0;
    const int per = (nrow + kSlices - 1) / kSlices;
    const int r0 = blk0 + (int)blockIdx.z * per, r1 = min(blk1, r0 + per);
    KSum acc{0, 0};
    if (col < ncol) {
        // kRows rows of the wavefront's share are loaded before any is added, so their loads
        // overlap; the compensated sum takes them in row order, as one row at a time would
        constexpr int kRows = 8;
        for (int r = r0 + wave; r < r1; r += 4 * kRows) {
            double v[kRows];
#pragma unroll
            for (int q = 0; q < kRows; ++q) {
                const int rq = min(r + 4 * q, r1 - 1);
                v[q] = part[(size_t)rq * kPartStride + col];
            }
#pragma unroll
            for (int q = 0; q < kRows; ++q)
                if (r + 4 * q < r1) acc.add(v[q]);
        }
    }
    ss[wave][lane] = acc.s;
    sc[wave][lane] = acc.c;
    __syncthreads();
    if (wave == 0 && col < ncol) {
        KSum t{ss[0][lane], sc[0][lane] + sc[1][lane] + sc[2][lane] + sc[3][lane]};
        t.add(ss[1][lane]);
        t.add(ss[2][lane]);
        t.add(ss[3][lane]);
        red[((size_t)blockIdx.y * kSlices + blockIdx.z) * kPartStride + col] = t.value();
    }
}

// Step scalars, one block of 256 threads per cloud: the kSlices rows of k_reduce are summed
// per column in slice order (compensated), then Update_f_Lyα_MF_L_Lyα! normalisation
// (src/ConstantsAndFunctions.jl:663-673), Compute_f_esc_Lyα_Cylindrical (:969-993) and
// FindActiveBins (src/UpdateLyaJandH.jl:47-97) for the next step.
__global__ __launch_bounds__(256) void k_final(const Cloud* __restrict__ clouds, int n, const double* __restrict__ red,
                                               int* __restrict__ krange, double* __restrict__ scalars,
                                               Ctrl* __restrict__ ctrl, int active_only)
{
    __shared__ double col[kPartStride];
    __shared__ double tree[256];
    __shared__ double vout[8];
    __shared__ int kmn_s, kmx_s, sstep, skr[2], ssteady, scap;
    __shared__ double stol;
    __shared__ double* shist;
    if (!ctrl->live) return;
    const int c = blockIdx.x;
    const Cloud& cl = clouds[c];
    const int nn = cl.nn, ncol = 2 * nn + kNScal, t = threadIdx.x;
    // the serial tail's inputs from global memory, loaded by a thread each while the columns
    // are summed, so that none of them is a dependent load on the tail
    if (t == 0) sstep = ctrl->step;
    else if (t == 1) skr[0] = krange[2 * c];
    else if (t == 2) skr[1] = krange[2 * c + 1];
    else if (t == 3) stol = ctrl->tol;
    else if (t == 4) ssteady = (ctrl->tol > 0.0) ? ctrl->steady[c] : 0;   // read only with a stop tolerance
    else if (t == 5) shist = ctrl->hist;
    else if (t == 6) scap = ctrl->cap;
    for (int m = t; m < ncol; m += blockDim.x) {
        KSum acc{0, 0};
#pragma unroll
        for (int z = 0; z < kSlices; ++z) acc.add(red[((size_t)c * kSlices + z) * kPartStride + m]);
        col[m] = acc.value();
    }
    if (t == 0) { kmn_s = nn - 1; kmx_s = 0; }
    __syncthreads();
    // Etot = Σ e1[k]: fixed-shape tree
    tree[t] = (t < nn) ? col[t] : 0.0;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (t < w) tree[t] += tree[t + w];
        __syncthreads();
    }
    const double Etot = tree[0];
    if (t < nn && col[nn + t] > kEpsActive * Etot) {
        atomicMin(&kmn_s, t);
        atomicMax(&kmx_s, t);
    }
    const int step = sstep;
    if (!active_only && t == 0) {
        double MF = col[2 * nn], L = col[2 * nn + 1], nt = col[2 * nn + 2], ow = col[2 * nn + 3];
        const double es = col[2 * nn + 4];
        if (L > 0.0) {
            const double sc = 1.0 / (L / kC);
            MF *= sc; nt *= sc; ow *= sc;
        } else {
            MF = 0.0; nt = 0.0; ow = 0.0;
        }
        const double Ltot = (cl.bc_z_bot == M1LYA_REFLECT) ? 2.0 * L : L;
        double fesc = 0.0;
        if (Ltot > 0.0) {
            const double Lsim = (cl.bc_z_bot == M1LYA_REFLECT) ? 0.5 * Ltot : Ltot;
            fesc = (Lsim > 0.0) ? es / Lsim : 0.0;
        }
        // ΔJ/J = Σ Δν ΔV |J_new - J_start| / E_tot (E_tot: the energy FindActiveBins measures)
        const double rel = (Etot > 0.0) ? col[2 * nn + 5] / Etot : 0.0;
        vout[0] = MF; vout[1] = nt; vout[2] = ow; vout[3] = Ltot; vout[4] = fesc;
        vout[5] = (double)skr[0]; vout[6] = (double)skr[1]; vout[7] = rel;
        if (col[2 * nn + 6] != 0.0 && atomicCAS(&ctrl->bad, 0, c + 1) == 0) ctrl->bad_step = step;
        if (stol > 0.0 && rel < stol && ssteady == 0) {
            ctrl->steady[c] = step + 1;
            if (atomicAdd(&ctrl->nsteady, 1) == n - 1) ctrl->halt = 1;
        }
        if (c == 0) ctrl->next = step + 1;
    }
    __syncthreads();
    if (!active_only && t < 8) {
        // the scalars and the history row, one value per thread
        scalars[(size_t)c * 8 + t] = vout[t];
        if (shist && step < scap) shist[((size_t)step * n + c) * 8 + t] = vout[t];
    }
    if (t != 0) return;
    krange[2 * c] = min(kmn_s, cl.kcore_min);
    krange[2 * c + 1] = max(kmx_s, cl.kcore_max);
}

// FP64 -> FP32 copy (the batch stores the refreshed coefficient tables in FP32)
__global__ void k_to_f32(float* __restrict__ dst, const double* __restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

__global__ void k_scale(float* a, float* b, float* c, size_t n, float s)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) { a[t] *= s; b[t] *= s; c[t] *= s; }
}

}  // namespace mb

// =========================================================================
// Host side
// =========================================================================
using namespace mb;

namespace {

thread_local std::string g_batch_error;

struct BFail {
    int code;
};

struct KStatB {
    long launches = 0;
    double total_ms = 0.0;
};

}  // namespace

struct m1lya_batch {
    int n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Cloud> cl;                 // host copy of the descriptors
    std::vector<char> geom_set, grid_set;
    Cloud* d_cl = nullptr;
    bool cl_dirty = true;

    size_t n3 = 0, nc = 0, nfr = 0, nfz = 0, nbin = 0, nivr = 0, nivz = 0;
    std::vector<CellRec> h_rec;            // host mirror of the cell records
    std::vector<std::vector<double>> h_nu; // ν interfaces of each cloud (coefficient refresh)
    // device-side coefficient refresh: Voigt tables, line constants and FP64 scratch
    double* lut = nullptr;
    int lut_nx = 0;
    double lut_xmax = 0.0;
    m1::LineConst lc{};
    double* rscratch = nullptr;            // sized for the largest cloud
    double nu_lya = 0.0;
    bool shear = false;                    // a cell's velocity gradient is sheared or anisotropic
    bool wide = false;                     // a cloud has more than 128 bins (four per lane)
    std::vector<BinRec> h_bin;
    CellRec* d_rec = nullptr;
    BinRec* d_bin = nullptr;
    Arrays A{};
    int parity = 0;
    bool krange_valid = false;

    // tile shape and the work / partial-row tables it implies
    int seg_len = 16, wpb = 8;   // the autotuner's choice on the 256x256x128 cone cloud
    std::vector<Seg> segs;
    Seg* d_segs = nullptr;
    size_t segs_cap = 0;
    int nblk = 0, grid = 0;
    int nblk_pair = 0;                     // blocks [0, nblk_pair) belong to clouds in pair mode
    double* part = nullptr;
    size_t part_rows_cap = 0;
    double* red = nullptr;
    int* krange = nullptr;
    double* scalars = nullptr;
    Ctrl* ctrl = nullptr;
    Ctrl h_ctrl{nullptr, 0, 0, 0, 0, 1, 0, 0, 0, 0.0, nullptr};
    int* steady = nullptr;                 // [n]: Ctrl::steady
    double* hist = nullptr;
    int hist_cap = 0;

    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};

    std::vector<void*> allocs;
    std::string err;
    int profiling = 0;                     // 0 off, 1 every launch, 2 k_step only
    std::map<std::string, KStatB> stats;
    struct Pending {
        const char* name;
        hipEvent_t e0, e1;
    };
    std::vector<Pending> pending;         // event pairs not yet read
    std::vector<hipEvent_t> event_pool;
};

namespace {

void bfail(m1lya_batch* b, int code, const std::string& msg)
{
    if (b) b->err = msg;
    g_batch_error = msg;
    throw BFail{code};
}

void bchk(m1lya_batch* b, hipError_t e, const char* what)
{
    if (e != hipSuccess) bfail(b, M1LYA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* balloc(m1lya_batch* b, size_t n)
{
    void* p = nullptr;
    bchk(b, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    b->allocs.push_back(p);
    bchk(b, hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)), "hipMemset");
    return (T*)p;
}

void bfree(m1lya_batch* b, void* p)
{
    if (!p) return;
    (void)hipFree(p);
    b->allocs.erase(std::remove(b->allocs.begin(), b->allocs.end(), p), b->allocs.end());
}

template <class F>
int bguard(m1lya_batch* b, F&& f)
{
    try {
        f();
        return M1LYA_OK;
    } catch (const BFail& e) {
        return e.code;
    } catch (const std::exception& e) {
        if (b) b->err = e.what();
        g_batch_error = e.what();
        return M1LYA_ERR_HIP;
    }
}

void drop_graphs(m1lya_batch* b)
{
    for (int p = 0; p < 2; ++p)
        if (b->gexec[p]) {
            (void)hipGraphExecDestroy(b->gexec[p]);
            b->gexec[p] = nullptr;
        }
}

void check_cloud(m1lya_batch* b, int c)
{
    if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
    if (c < 0 || c >= b->n) bfail(b, M1LYA_ERR_ARG, "cloud index out of range");
}

// Work table of the current tile shape: per cloud, row segments of seg_len cells,
// wpb of them (same i range, consecutive rows) per block; blocks never mix clouds.  The
// blocks of clouds in pair mode come first: they are launched on their own.
void build_tiles(m1lya_batch* b)
{
    std::vector<Seg> segs;
    int blk = 0;
    std::vector<int> order;
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < b->n; ++c)
            if ((b->cl[c].pair != 0) == (pass == 0)) order.push_back(c);
    int nblk_pair = 0;
    for (int c : order) {
        Cloud& cl = b->cl[c];
        cl.blk0 = blk;
        const int nseg_r = (cl.nr + b->seg_len - 1) / b->seg_len;
        for (int j0 = 0; j0 < cl.nz; j0 += b->wpb)
            for (int s = 0; s < nseg_r; ++s) {
                // rows j0+w >= N_Z are ghost rows: their wavefronts only take part in
                // the Z exchange of the block
                for (int w = 0; w < b->wpb; ++w)
                    segs.push_back(Seg{c, j0 + w, s * b->seg_len, std::min(cl.nr, (s + 1) * b->seg_len)});
                ++blk;
            }
        cl.blk1 = blk;
        if (cl.pair) nblk_pair = blk;
    }
    b->nblk = blk;
    b->nblk_pair = nblk_pair;
    b->grid = (blk + 7) / 8 * 8;
    if (segs.size() > b->segs_cap) {
        bfree(b, b->d_segs);
        b->d_segs = balloc<Seg>(b, segs.size());
        b->segs_cap = segs.size();
    }
    if ((size_t)blk > b->part_rows_cap) {
        bfree(b, b->part);
        b->part = balloc<double>(b, (size_t)blk * kPartStride);
        b->part_rows_cap = blk;
    }
    bchk(b, hipMemcpy(b->d_segs, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice), "segments");
    b->segs = std::move(segs);
    b->cl_dirty = true;
    drop_graphs(b);
}

void sync_clouds(m1lya_batch* b)
{
    if (!b->cl_dirty) return;
    for (Cloud& cl : b->cl) {
        cl.cdt64 = cl.c_tilde * cl.dt;
        cl.dw = 1.0 / cl.cdt64;
        cl.src_scale = cl.cdt64 / cl.J0;
        cl.mf_scale = (kFourPi / kC) / cl.c_tilde * cl.J0 / cl.dt;
        for (int p = 0; p < 2; ++p) {
            cl.sJ[p] = b->A.J[p] + cl.o3;
            cl.sHr[p] = b->A.Hr[p] + cl.o3;
            cl.sHz[p] = b->A.Hz[p] + cl.o3;
        }
        cl.saF = b->A.aF + cl.o3;
        cl.sHv = b->A.Hv + cl.o3;
        cl.sphi = b->A.phi + cl.o3;
        cl.sivr = b->A.ivr + cl.oivr;
        cl.sivz = b->A.ivz + cl.oivz;
        cl.sucr = b->A.uc_r + cl.ofr;
        cl.sucz = b->A.uc_z + cl.ofz;
        cl.srec = b->A.rec + cl.oc;
        cl.sfr = b->A.f_r + cl.oc;
        cl.sfz = b->A.f_z + cl.oc;
    }
    bchk(b, hipMemcpy(b->d_cl, b->cl.data(), b->n * sizeof(Cloud), hipMemcpyHostToDevice), "cloud descriptors");
    b->cl_dirty = false;
}

void push_records(m1lya_batch* b, int cloud)
{
    const Cloud& cl = b->cl[cloud];
    const size_t ncell = (size_t)cl.nr * cl.nz;
    bchk(b, hipMemcpy(b->d_rec + cl.oc, b->h_rec.data() + cl.oc, ncell * sizeof(CellRec), hipMemcpyHostToDevice),
         "cell records");
}

void set_ctrl(m1lya_batch* b, double* hist, int cap, double tol = 0.0)
{
    b->h_ctrl = Ctrl{hist, cap, 0, 0, 0, 1, 0, 0, 0, tol, b->steady};
    bchk(b, hipMemcpy(b->ctrl, &b->h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice), "ctrl");
    bchk(b, hipMemset(b->steady, 0, sizeof(int) * b->n), "steady");
}

hipEvent_t pool_event(m1lya_batch* b)
{
    if (!b->event_pool.empty()) {
        hipEvent_t e = b->event_pool.back();
        b->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    bchk(b, hipEventCreate(&e), "hipEventCreate");
    return e;
}

// Brackets one launch with events on the batch stream when profiling; the pair is
// read later (collect_stats), so profiling adds no host synchronisation.
struct BProf {
    m1lya_batch* b;
    const char* name;
    bool on;
    hipEvent_t e0{}, e1{};
    BProf(m1lya_batch* b_, const char* n, bool step_kernel = false)
        : b(b_), name(n), on(b_->profiling == 1 || (b_->profiling == 2 && step_kernel))
    {
        if (on) {
            e0 = pool_event(b);
            e1 = pool_event(b);
            bchk(b, hipEventRecord(e0, b->stream), "hipEventRecord");
        }
    }
    void done()
    {
        bchk(b, hipGetLastError(), name);
        if (on) {
            bchk(b, hipEventRecord(e1, b->stream), "hipEventRecord");
            b->pending.push_back(m1lya_batch::Pending{name, e0, e1});
        }
    }
};

void collect_stats(m1lya_batch* b)
{
    if (b->pending.empty()) return;
    bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    for (auto& p : b->pending) {
        float ms = 0.f;
        bchk(b, hipEventElapsedTime(&ms, p.e0, p.e1), "hipEventElapsedTime");
        KStatB& st = b->stats[p.name];
        st.launches += 1;
        st.total_ms += ms;
        b->event_pool.push_back(p.e0);
        b->event_pool.push_back(p.e1);
    }
    b->pending.clear();
}

void launch_step_kernels(m1lya_batch* b, int parity)
{
    hipStream_t s = b->stream;
    const dim3 block(64 * b->wpb);
    if (b->nblk_pair > 0) {
        // clouds of N_ν <= 32, two cells per wavefront
        BProf p(b, "k_step_pair", true);
        const size_t lds = step_lds_bytes(b->wpb, 1);
        const dim3 grid((b->nblk_pair + 7) / 8 * 8);
        if (b->shear)
            hipLaunchKernelGGL((k_step<true, 1, true>), grid, block, lds, s, b->d_cl, b->A, b->d_segs, 0, b->nblk_pair,
                               b->wpb, parity, b->krange, b->part, b->ctrl);
        else
            hipLaunchKernelGGL((k_step<false, 1, true>), grid, block, lds, s, b->d_cl, b->A, b->d_segs, 0,
                               b->nblk_pair, b->wpb, parity, b->krange, b->part, b->ctrl);
        p.done();
    }
    if (b->nblk > b->nblk_pair) {
        BProf p(b, "k_step", true);
        // the variant with the reference-row path for sheared / anisotropic velocity gradients only
        // when a cell of the batch needs it: it costs registers the common path does not pay for
        // and the four-bins-per-lane variant only when a cloud has more than 128 bins
        const size_t lds = step_lds_bytes(b->wpb, b->wide ? 4 : 2);
        const int b0 = b->nblk_pair, b1 = b->nblk;
        const dim3 grid((b1 - b0 + 7) / 8 * 8);
        if (b->wide) {
            if (b->shear)
                hipLaunchKernelGGL((k_step<true, 4, false>), grid, block, lds, s, b->d_cl, b->A, b->d_segs, b0, b1,
                                   b->wpb, parity, b->krange, b->part, b->ctrl);
            else
                hipLaunchKernelGGL((k_step<false, 4, false>), grid, block, lds, s, b->d_cl, b->A, b->d_segs, b0, b1,
                                   b->wpb, parity, b->krange, b->part, b->ctrl);
        } else if (b->shear) {
            hipLaunchKernelGGL((k_step<true, 2, false>), grid, block, lds, s, b->d_cl, b->A, b->d_segs, b0, b1,
                               b->wpb, parity, b->krange, b->part, b->ctrl);
        } else {
            hipLaunchKernelGGL((k_step<false, 2, false>), grid, block, lds, s, b->d_cl, b->A, b->d_segs, b0, b1,
                               b->wpb, parity, b->krange, b->part, b->ctrl);
        }
        p.done();
    }
    {
        BProf p(b, "k_reduce");
        hipLaunchKernelGGL(k_reduce, dim3((2 * kMaxBins + kNScal + 63) / 64, b->n, kSlices), dim3(256), 0, s, b->d_cl,
                           b->part, b->red, b->ctrl);
        p.done();
    }
    {
        BProf p(b, "k_final");
        hipLaunchKernelGGL(k_final, dim3(b->n), dim3(256), 0, s, b->d_cl, b->n, b->red, b->krange, b->scalars, b->ctrl, 0);
        p.done();
    }
}

void compute_krange(m1lya_batch* b)
{
    sync_clouds(b);
    hipStream_t s = b->stream;
    hipLaunchKernelGGL(k_active_rows, dim3(b->grid), dim3(64 * b->wpb), 0, s, b->d_cl, b->A, b->d_segs, b->nblk,
                       b->wpb, b->parity, b->part);
    bchk(b, hipGetLastError(), "k_active_rows");
    hipLaunchKernelGGL(k_reduce, dim3((2 * kMaxBins + kNScal + 63) / 64, b->n, kSlices), dim3(256), 0, s, b->d_cl,
                       b->part, b->red, b->ctrl);
    bchk(b, hipGetLastError(), "k_reduce");
    hipLaunchKernelGGL(k_final, dim3(b->n), dim3(256), 0, s, b->d_cl, b->n, b->red, b->krange, b->scalars, b->ctrl, 1);
    bchk(b, hipGetLastError(), "k_final");
    b->krange_valid = true;
}

void require_ready(m1lya_batch* b)
{
    for (int c = 0; c < b->n; ++c) {
        const std::string who = "cloud " + std::to_string(c);
        if (!b->geom_set[c]) bfail(b, M1LYA_ERR_STATE, who + ": geometry not set");
        if (!b->grid_set[c]) bfail(b, M1LYA_ERR_STATE, who + ": frequency grid not set");
        if (!(b->cl[c].dt > 0.0) || !(b->cl[c].c_tilde > 0.0))
            bfail(b, M1LYA_ERR_STATE, who + ": step parameters not set");
    }
}

void capture_graphs(m1lya_batch* b)
{
    const int prof = b->profiling;
    b->profiling = 0;
    for (int p = 0; p < 2; ++p) {
        if (b->gexec[p]) continue;
        hipGraph_t g;
        bchk(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        launch_step_kernels(b, p);
        bchk(b, hipStreamEndCapture(b->stream, &g), "hipStreamEndCapture");
        bchk(b, hipGraphInstantiate(&b->gexec[p], g, nullptr, nullptr, 0), "hipGraphInstantiate");
        (void)hipGraphDestroy(g);
    }
    b->profiling = prof;
}

// Reports a non-finite J, H_r or H_Z that a step of the last run wrote (k_final records the
// first cloud and step), then clears the device controls.
void check_nonfinite(m1lya_batch* b, const char* what)
{
    Ctrl h{};
    bchk(b, hipMemcpy(&h, b->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost), "ctrl");
    if (h.bad) {
        set_ctrl(b, nullptr, 0);
        bfail(b, M1LYA_ERR_NONFINITE, std::string(what) + ": cloud " + std::to_string(h.bad - 1) +
                                          ": non-finite J, H_r or H_Z after step " + std::to_string(h.bad_step) +
                                          " of the call");
    }
}

void ensure_hist(m1lya_batch* b, int n_steps)
{
    if (n_steps > b->hist_cap) {
        bfree(b, b->hist);
        b->hist_cap = std::max(n_steps, 64);
        b->hist = balloc<double>(b, (size_t)b->hist_cap * b->n * 8);
    }
}

void run_steps(m1lya_batch* b, int n_steps, double* history)
{
    require_ready(b);
    sync_clouds(b);
    if (history) ensure_hist(b, n_steps);
    bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    set_ctrl(b, history ? b->hist : nullptr, history ? b->hist_cap : 0);
    if (!b->krange_valid) compute_krange(b);
    const bool graph = b->use_graph && !b->profiling;
    if (graph) capture_graphs(b);
    for (int s = 0; s < n_steps; ++s) {
        if (graph)
            bchk(b, hipGraphLaunch(b->gexec[b->parity], b->stream), "hipGraphLaunch");
        else
            launch_step_kernels(b, b->parity);
        b->parity ^= 1;
    }
    if (history) {
        bchk(b, hipMemcpyAsync(history, b->hist, (size_t)n_steps * b->n * 8 * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream),
             "history download");
    }
    bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    check_nonfinite(b, "run_steps");
}

// Steps until every cloud's ΔJ/J has fallen below tol (k_final then sets the halt flag and
// the rest of the queued launches return at once), or max_steps.  steady[c] receives the
// number of steps after which cloud c first was below tol (0: never).
int run_until_steady(m1lya_batch* b, int max_steps, double tol, double* history, int* steady)
{
    require_ready(b);
    sync_clouds(b);
    if (history) ensure_hist(b, max_steps);
    bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    set_ctrl(b, history ? b->hist : nullptr, history ? b->hist_cap : 0, tol);
    if (!b->krange_valid) compute_krange(b);
    const bool graph = b->use_graph && !b->profiling;
    if (graph) capture_graphs(b);
    const int chunk = 64;
    int launched = 0;
    Ctrl h{};
    while (launched < max_steps) {
        const int n = std::min(chunk, max_steps - launched);
        for (int s = 0; s < n; ++s) {
            if (graph)
                bchk(b, hipGraphLaunch(b->gexec[b->parity], b->stream), "hipGraphLaunch");
            else
                launch_step_kernels(b, b->parity);
            b->parity ^= 1;
        }
        launched += n;
        bchk(b, hipMemcpyAsync(&h, b->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, b->stream), "ctrl");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        if (h.halt || h.bad) break;
    }
    // every launched step flipped the buffer parity on the host; the halted ones did not run
    const int done = h.next;
    if ((launched - done) & 1) b->parity ^= 1;
    if (history && done > 0)
        bchk(b, hipMemcpy(history, b->hist, (size_t)done * b->n * 8 * sizeof(double), hipMemcpyDeviceToHost),
             "history download");
    if (steady) bchk(b, hipMemcpy(steady, b->steady, sizeof(int) * b->n, hipMemcpyDeviceToHost), "steady");
    check_nonfinite(b, "run_until_steady");
    set_ctrl(b, nullptr, 0);
    return done;
}

void up_f32(m1lya_batch* b, float* dst, const double* src, size_t n, double scale)
{
    std::vector<float> h(n);
    for (size_t t = 0; t < n; ++t) h[t] = (float)(src[t] * scale);
    bchk(b, hipMemcpy(dst, h.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload");
}

}  // namespace

extern "C" {

int m1lya_batch_create(int n_clouds, const int* n_nu, const int* n_r, const int* n_z, m1lya_batch** out)
{
    if (!out || !n_nu || !n_r || !n_z || n_clouds < 1) {
        g_batch_error = "m1lya_batch_create: bad arguments";
        return M1LYA_ERR_ARG;
    }
    *out = nullptr;
    for (int c = 0; c < n_clouds; ++c)
        if (n_nu[c] < 3 || n_nu[c] > kMaxBins || n_r[c] < 3 || n_z[c] < 3) {
            g_batch_error = "m1lya_batch_create: cloud " + std::to_string(c) +
                            " needs 3 <= N_nu <= " + std::to_string(kMaxBins) + " and N_r, N_Z >= 3";
            return M1LYA_ERR_DIMS;
        }
    m1lya_batch* b = new m1lya_batch();
    const int rc = bguard(b, [&] {
        b->n = n_clouds;
        bchk(b, hipGetDevice(&b->device), "hipGetDevice");
        bchk(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        b->cl.resize(n_clouds);
        b->geom_set.assign(n_clouds, 0);
        b->grid_set.assign(n_clouds, 0);
        for (int c = 0; c < n_clouds; ++c) {
            Cloud& cl = b->cl[c];
            std::memset(&cl, 0, sizeof(Cloud));
            cl.nn = n_nu[c]; cl.nr = n_r[c]; cl.nz = n_z[c];
            cl.R = cl.nn <= 64 ? 1 : cl.nn <= 128 ? 2 : 4;
            cl.pair = cl.nn <= 32;
            if (cl.R == 4) b->wide = true;
            const size_t ncell = (size_t)cl.nr * cl.nz;
            cl.o3 = (long long)b->n3; b->n3 += ncell * cl.nn;
            cl.oc = (long long)b->nc; b->nc += ncell;
            cl.ofr = (long long)b->nfr; b->nfr += (size_t)(cl.nr + 1) * cl.nz;
            cl.ofz = (long long)b->nfz; b->nfz += (size_t)cl.nr * (cl.nz + 1);
            cl.obin = (long long)b->nbin; b->nbin += cl.nn;
            cl.oivr = (long long)b->nivr; b->nivr += (size_t)(cl.nr + 1) * m1f::kIvStride;
            cl.oivz = (long long)b->nivz; b->nivz += (size_t)(cl.nz + 1) * m1f::kIvStride;
            cl.J0 = 1.0;
            cl.bc_r_right = M1LYA_OUTFLOW; cl.bc_z_bot = M1LYA_REFLECT; cl.bc_z_top = M1LYA_OUTFLOW;
        }
        b->h_rec.assign(b->nc, CellRec{});
        b->h_bin.assign(b->nbin, BinRec{});
        Arrays& A = b->A;
        for (int p = 0; p < 2; ++p) {
            A.J[p] = balloc<float>(b, b->n3);
            A.Hr[p] = balloc<float>(b, b->n3);
            A.Hz[p] = balloc<float>(b, b->n3);
        }
        A.aF = balloc<float>(b, b->n3);
        A.Hv = balloc<float>(b, b->n3);
        A.phi = balloc<float>(b, b->n3);
        b->d_rec = balloc<CellRec>(b, b->nc);
        b->d_bin = balloc<BinRec>(b, b->nbin);
        A.rec = b->d_rec;
        A.bin = b->d_bin;
        A.uc_r = balloc<float>(b, b->nfr);
        A.uc_z = balloc<float>(b, b->nfz);
        A.ivr = balloc<float>(b, b->nivr);
        A.ivz = balloc<float>(b, b->nivz);
        A.f_r = balloc<double>(b, b->nc);
        A.f_z = balloc<double>(b, b->nc);
        b->d_cl = balloc<Cloud>(b, n_clouds);
        b->red = balloc<double>(b, (size_t)n_clouds * kSlices * kPartStride);
        b->krange = balloc<int>(b, (size_t)n_clouds * 2);
        b->scalars = balloc<double>(b, (size_t)n_clouds * 8);
        b->ctrl = balloc<Ctrl>(b, 1);
        b->steady = balloc<int>(b, (size_t)n_clouds);
        build_tiles(b);
    });
    if (rc != M1LYA_OK) {
        m1lya_batch_destroy(b);
        return rc;
    }
    *out = b;
    return M1LYA_OK;
}

void m1lya_batch_destroy(m1lya_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (auto& p : b->pending) {
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    for (hipEvent_t e : b->event_pool) (void)hipEventDestroy(e);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int m1lya_batch_last_error(const m1lya_batch* b, char* buf, size_t len)
{
    if (!buf || len == 0) return M1LYA_ERR_ARG;
    const std::string& s = b ? b->err : g_batch_error;
    std::snprintf(buf, len, "%s", s.c_str());
    return M1LYA_OK;
}

int m1lya_batch_count(const m1lya_batch* b) { return b ? b->n : 0; }

int m1lya_batch_set_geometry(m1lya_batch* b, int cloud, const m1lya_geometry* g)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        if (!g || !g->dV || !g->inv_dV || !g->A_r || !g->A_Z || !g->inv_r_cell || !g->rCoM || !g->ZCoM || !g->dr_L ||
            !g->dr_R || !g->inv_dr_i_im1 || !g->inv_dr_ip1_i || !g->inv_dr_ip2_ip1 || !g->dZ_L || !g->dZ_R ||
            !g->inv_dZ_j_jm1 || !g->inv_dZ_jp1_j || !g->inv_dZ_jp2_jp1)
            bfail(b, M1LYA_ERR_ARG, "geometry has a NULL array");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        const Cloud& cl = b->cl[cloud];
        const int nr = cl.nr, nz = cl.nz;
        CellRec* rec = b->h_rec.data() + cl.oc;
        for (int j = 0; j < nz; ++j)
            for (int i = 0; i < nr; ++i) {
                const size_t c = (size_t)j * nr + i;
                CellRec& r = rec[c];
                // A/ΔV of each cell face: the flux-divergence factors of src/AdvectJandH.jl:1319-1340
                r.gRl = (float)(g->A_r[(size_t)j * (nr + 1) + i] * g->inv_dV[c]);
                r.gRr = (float)(g->A_r[(size_t)j * (nr + 1) + i + 1] * g->inv_dV[c]);
                r.gZl = (float)(g->A_Z[(size_t)j * nr + i] * g->inv_dV[c]);
                r.gZr = (float)(g->A_Z[(size_t)(j + 1) * nr + i] * g->inv_dV[c]);
                r.inv_r = (float)g->inv_r_cell[i];
                r.dV = g->dV[c];
                r.rcom = g->rCoM[i];
                r.zcom = g->ZCoM[j];
                // boundary faces of Compute_f_esc_Cylindrical (src/ConstantsAndFunctions.jl:937-953)
                r.A_rout = (i == nr - 1) ? g->A_r[(size_t)j * (nr + 1) + nr] : 0.0;
                r.A_ztop = (j == nz - 1) ? g->A_Z[(size_t)nz * nr + i] : 0.0;
                r.A_zbot = (j == 0) ? g->A_Z[i] : 0.0;
            }
        push_records(b, cloud);
        // one interval set per interface: face 0 inner boundary, face N outer boundary,
        // face f in 1..N-1 the interior arrays at f-1 (src/AdvectJandH.jl:632-639, 803-812);
        // the three inverse spacings are stored times ΔL or ΔR (half-cell units)
        auto pack = [&](int n, const double* dL, const double* dR, const double* a, const double* bb,
                        const double* cc, const double* inner, const double* outer, float* dst) {
            const int S = m1f::kIvStride;
            std::vector<double> h((size_t)(n + 1) * S);
            auto put = [&](int f, double vL, double vR, double ia, double ib, double ic) {
                double* o = h.data() + (size_t)f * S;
                o[0] = vL; o[1] = vR;
                o[2] = vL * ia; o[3] = vL * ib;
                o[4] = vR * ib; o[5] = vR * ic;
            };
            put(0, inner[0], inner[1], inner[2], inner[3], inner[4]);
            put(n, outer[0], outer[1], outer[2], outer[3], outer[4]);
            for (int f = 1; f < n; ++f) put(f, dL[f - 1], dR[f - 1], a[f - 1], bb[f - 1], cc[f - 1]);
            up_f32(b, dst, h.data(), h.size(), 1.0);
        };
        pack(nr, g->dr_L, g->dr_R, g->inv_dr_i_im1, g->inv_dr_ip1_i, g->inv_dr_ip2_ip1, g->dr_inner, g->dr_outer,
             b->A.ivr + cl.oivr);
        pack(nz, g->dZ_L, g->dZ_R, g->inv_dZ_j_jm1, g->inv_dZ_jp1_j, g->inv_dZ_jp2_jp1, g->dZ_inner, g->dZ_outer,
             b->A.ivz + cl.oivz);
        b->geom_set[cloud] = 1;
        b->krange_valid = false;
    });
}

int m1lya_batch_set_freq_grid(m1lya_batch* b, int cloud, const m1lya_freq_grid* g)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        Cloud& cl = b->cl[cloud];
        if (!g || !g->nu || !g->nuC || !g->dnu || !g->inv_dnu || !g->inv_nuC)
            bfail(b, M1LYA_ERR_ARG, "frequency grid has a NULL array");
        if (g->kcore_min < 0 || g->kcore_max >= cl.nn || g->kcore_min > g->kcore_max)
            bfail(b, M1LYA_ERR_ARG, "kcore range out of bounds");
        if (g->flux_method != M1LYA_GLF && g->flux_method != M1LYA_HLL) bfail(b, M1LYA_ERR_ARG, "unknown flux method");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        BinRec* br = b->h_bin.data() + cl.obin;
        for (int k = 0; k < cl.nn; ++k) {
            br[k].nuC = g->nuC[k];
            br[k].dnu = g->dnu[k];
            br[k].inv_dnu = g->inv_dnu[k];
            br[k].inv_nuC = g->inv_nuC[k];
            br[k].half_R = g->nu[k + 1] - g->nuC[k];
            br[k].half_L = g->nuC[k] - g->nu[k];
        }
        bchk(b, hipMemcpy(b->d_bin + cl.obin, br, cl.nn * sizeof(BinRec), hipMemcpyHostToDevice), "bins");
        cl.kcore_min = g->kcore_min;
        cl.kcore_max = g->kcore_max;
        cl.recoil = g->recoil ? 1 : 0;
        cl.hll = g->flux_method == M1LYA_HLL ? 1 : 0;
        cl.alpha_floor = g->alpha_floor;
        if ((int)b->h_nu.size() < b->n) b->h_nu.resize(b->n);
        b->h_nu[cloud].assign(g->nu, g->nu + cl.nn + 1);
        b->grid_set[cloud] = 1;
        b->cl_dirty = true;
        b->krange_valid = false;
    });
}

int m1lya_batch_set_voigt_lut(m1lya_batch* b, const double* luts, int nx, double xmax, const m1lya_line_constants* lc)
{
    return bguard(b, [&] {
        if (!luts || !lc || nx < 2 || !(xmax > 0.0)) bfail(b, M1LYA_ERR_ARG, "set_voigt_lut: bad arguments");
        if (b->lut && b->lut_nx != nx) bfail(b, M1LYA_ERR_ARG, "set_voigt_lut: table size changed");
        if (!b->lut) {
            size_t n3max = 0, ncmax = 0;
            int nnmax = 0;
            for (const Cloud& cl : b->cl) {
                n3max = std::max(n3max, (size_t)cl.nn * cl.nr * cl.nz);
                ncmax = std::max(ncmax, (size_t)cl.nr * cl.nz);
                nnmax = std::max(nnmax, cl.nn);
            }
            b->lut = balloc<double>(b, (size_t)5 * nx);
            // Hv, phi, aF [n3]; T, nHI, acF, alpha, dnuD [ncell]; dnu_rel, dnuC_rel, dnu [N_ν + 1]
            b->rscratch = balloc<double>(b, 3 * n3max + 5 * ncmax + 3 * ((size_t)nnmax + 1));
        }
        bchk(b, hipMemcpy(b->lut, luts, (size_t)5 * nx * sizeof(double), hipMemcpyHostToDevice), "lut");
        b->lut_nx = nx;
        b->lut_xmax = xmax;
        b->lc = m1::LineConst{lc->b_prefact, lc->nu_over_c, lc->a_prefact, lc->sigma0_prefact};
        b->nu_lya = lc->nu_lya;
    });
}

int m1lya_batch_prepare_subcycling(m1lya_batch* b, int cloud, const double* T, const double* n_HI,
                                   const double* alpha_c, const double* alpha_c_F, const double* js,
                                   const double* pd, int phi_all_bins)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        if (!b->lut) bfail(b, M1LYA_ERR_STATE, "prepare_subcycling: Voigt tables not set (m1lya_batch_set_voigt_lut)");
        if (!b->grid_set[cloud] || (int)b->h_nu.size() <= cloud || b->h_nu[cloud].empty())
            bfail(b, M1LYA_ERR_STATE, "prepare_subcycling: frequency grid not set");
        if (!T || !n_HI) bfail(b, M1LYA_ERR_ARG, "prepare_subcycling: T and n_HI are required");
        const Cloud& cl = b->cl[cloud];
        const int nn = cl.nn;
        const size_t ncell = (size_t)cl.nr * cl.nz, n3 = ncell * nn;
        size_t n3max = 0, ncmax = 0;
        for (const Cloud& c2 : b->cl) {
            n3max = std::max(n3max, (size_t)c2.nn * c2.nr * c2.nz);
            ncmax = std::max(ncmax, (size_t)c2.nr * c2.nz);
        }
        double* Hv = b->rscratch;
        double* phi = Hv + n3max;
        double* aF = phi + n3max;
        double* dT = aF + n3max;
        double* dn = dT + ncmax;
        double* dac = dn + ncmax;
        double* dal = dac + ncmax;
        double* ddD = dal + ncmax;
        double* drel = ddD + ncmax;
        double* dcrel = drel + (nn + 1);
        double* ddnu = dcrel + (nn + 1);
        // ν - ν_Lyα and Δν formed as the host setup forms them
        const std::vector<double>& nu = b->h_nu[cloud];
        std::vector<double> rel(nn + 1), crel(nn), dnu(nn);
        for (int k = 0; k <= nn; ++k) rel[k] = nu[k] - b->nu_lya;
        for (int k = 0; k < nn; ++k) {
            dnu[k] = nu[k + 1] - nu[k];
            crel[k] = (nu[k] + 0.5 * dnu[k]) - b->nu_lya;
        }
        hipStream_t st = b->stream;
        bchk(b, hipMemcpyAsync(drel, rel.data(), rel.size() * sizeof(double), hipMemcpyHostToDevice, st), "dnu_rel");
        bchk(b, hipMemcpyAsync(dcrel, crel.data(), crel.size() * sizeof(double), hipMemcpyHostToDevice, st), "dnuC_rel");
        bchk(b, hipMemcpyAsync(ddnu, dnu.data(), dnu.size() * sizeof(double), hipMemcpyHostToDevice, st), "dnu");
        bchk(b, hipMemcpyAsync(dT, T, ncell * sizeof(double), hipMemcpyHostToDevice, st), "T");
        bchk(b, hipMemcpyAsync(dn, n_HI, ncell * sizeof(double), hipMemcpyHostToDevice, st), "n_HI");
        if (alpha_c_F)
            bchk(b, hipMemcpyAsync(dac, alpha_c_F, ncell * sizeof(double), hipMemcpyHostToDevice, st), "alpha_c_F");
        m1::RefreshArgs r;
        r.nn = nn; r.ncell = (int)ncell; r.kmin = cl.kcore_min; r.kmax = cl.kcore_max;
        r.phi_all_bins = phi_all_bins ? 1 : 0;
        r.lut = b->lut; r.nx = b->lut_nx; r.xmax = b->lut_xmax; r.lc = b->lc;
        r.dnu_rel = drel; r.dnuC_rel = dcrel; r.dnu = ddnu;
        r.T = dT; r.nHI = dn; r.acF = alpha_c_F ? dac : nullptr;
        r.Hv = Hv; r.phi = phi; r.aF = aF; r.alpha = dal; r.dnuD = ddD;
        m1::launch_refresh(st, r);
        const Arrays& A = b->A;
        const unsigned nb = (unsigned)std::min<size_t>((n3 + 255) / 256, 4096);
        hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, st, A.Hv + cl.o3, Hv, n3);
        hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, st, A.phi + cl.o3, phi, n3);
        hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, st, A.aF + cl.o3, aF, n3);
        bchk(b, hipGetLastError(), "refresh");
        std::vector<double> al(ncell), dD(ncell);
        bchk(b, hipMemcpyAsync(al.data(), dal, ncell * sizeof(double), hipMemcpyDeviceToHost, st), "alpha");
        bchk(b, hipMemcpyAsync(dD.data(), ddD, ncell * sizeof(double), hipMemcpyDeviceToHost, st), "dnuD");
        bchk(b, hipStreamSynchronize(st), "hipStreamSynchronize");
        CellRec* rec = b->h_rec.data() + cl.oc;
        for (size_t c = 0; c < ncell; ++c) {
            rec[c].alpha = al[c];
            rec[c].dnuD = dD[c];
            rec[c].T = T[c];
            if (alpha_c) rec[c].alpha_c = alpha_c[c];
            if (js) rec[c].js = js[c];
            if (pd) rec[c].pd = pd[c];
        }
        push_records(b, cloud);
    });
}

int m1lya_batch_set_scale(m1lya_batch* b, int cloud, double J0)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        if (!(J0 > 0.0) || !std::isfinite(J0)) bfail(b, M1LYA_ERR_ARG, "J0 must be finite and positive");
        Cloud& cl = b->cl[cloud];
        const double f = cl.J0 / J0;
        if (f != 1.0) {
            // keep the resident state's physical values
            const size_t n = (size_t)cl.nn * cl.nr * cl.nz;
            const int p = b->parity;
            hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, b->A.J[p] + cl.o3,
                               b->A.Hr[p] + cl.o3, b->A.Hz[p] + cl.o3, n, (float)f);
            bchk(b, hipGetLastError(), "k_scale");
            bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        }
        cl.J0 = J0;
        b->cl_dirty = true;
        b->krange_valid = false;
    });
}

int m1lya_batch_get_scale(const m1lya_batch* b, int cloud, double* J0)
{
    if (!b || !J0 || cloud < 0 || cloud >= b->n) return M1LYA_ERR_ARG;
    *J0 = b->cl[cloud].J0;
    return M1LYA_OK;
}

int m1lya_batch_upload(m1lya_batch* b, int cloud, int field, const double* host, size_t count)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        if (!host) bfail(b, M1LYA_ERR_ARG, "host pointer is NULL");
        Cloud& cl = b->cl[cloud];
        const size_t ncell = (size_t)cl.nr * cl.nz, n3 = ncell * cl.nn;
        const size_t nfr = (size_t)(cl.nr + 1) * cl.nz, nfz = (size_t)cl.nr * (cl.nz + 1);
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        Arrays& A = b->A;
        const int p = b->parity;
        auto need = [&](size_t n) {
            if (count != n)
                bfail(b, M1LYA_ERR_ARG,
                      "field " + std::to_string(field) + ": count " + std::to_string(count) + " != " +
                          std::to_string(n));
        };
        CellRec* rec = b->h_rec.data() + cl.oc;
        auto cellfield = [&](double CellRec::*m) {
            need(ncell);
            for (size_t c = 0; c < ncell; ++c) rec[c].*m = host[c];
            push_records(b, cloud);
        };
        const double inv_J0 = 1.0 / cl.J0;
        switch (field) {
        case M1LYA_J: need(n3); up_f32(b, A.J[p] + cl.o3, host, n3, inv_J0); b->krange_valid = false; break;
        case M1LYA_H_R: need(n3); up_f32(b, A.Hr[p] + cl.o3, host, n3, inv_J0); break;
        case M1LYA_H_Z: need(n3); up_f32(b, A.Hz[p] + cl.o3, host, n3, inv_J0); break;
        case M1LYA_ALPHA_F: need(n3); up_f32(b, A.aF + cl.o3, host, n3, 1.0); break;
        case M1LYA_HVOIGT: need(n3); up_f32(b, A.Hv + cl.o3, host, n3, 1.0); break;
        case M1LYA_PHI_NU: need(n3); up_f32(b, A.phi + cl.o3, host, n3, 1.0); break;
        case M1LYA_ALPHA_LYA: cellfield(&CellRec::alpha); break;
        case M1LYA_ALPHA_C: cellfield(&CellRec::alpha_c); break;
        case M1LYA_PD: cellfield(&CellRec::pd); break;
        case M1LYA_JS: cellfield(&CellRec::js); break;
        case M1LYA_T: cellfield(&CellRec::T); break;
        case M1LYA_DNU_D: cellfield(&CellRec::dnuD); break;
        case M1LYA_DUR_DR: cellfield(&CellRec::durdr); break;
        case M1LYA_DUR_DZ: cellfield(&CellRec::durdz); break;
        case M1LYA_DUZ_DR: cellfield(&CellRec::duzdr); break;
        case M1LYA_DUZ_DZ: cellfield(&CellRec::duzdz); break;
        case M1LYA_UR_OVER_R: cellfield(&CellRec::urr); break;
        case M1LYA_UR_STAR: need(nfr); up_f32(b, A.uc_r + cl.ofr, host, nfr, kInvC); break;
        case M1LYA_UZ_STAR: need(nfz); up_f32(b, A.uc_z + cl.ofz, host, nfz, kInvC); break;
        default: bfail(b, M1LYA_ERR_ARG, "field " + std::to_string(field) + " is not held by a batch");
        }
        if (field >= M1LYA_DUR_DR && field <= M1LYA_UR_OVER_R) {
            cl.has_vel = 1;
            b->cl_dirty = true;
            // recompute whether any cell needs the per-row VEL path (see k_step)
            bool sh = false;
            for (const CellRec& r : b->h_rec)
                if ((r.durdz + r.duzdr) != 0.0 || r.durdr != r.duzdz || r.duzdz != r.urr) { sh = true; break; }
            if (sh != b->shear) {
                b->shear = sh;
                drop_graphs(b);
            }
        }
    });
}

int m1lya_batch_download(m1lya_batch* b, int cloud, int field, double* host, size_t count)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        if (!host) bfail(b, M1LYA_ERR_ARG, "host pointer is NULL");
        const Cloud& cl = b->cl[cloud];
        const size_t ncell = (size_t)cl.nr * cl.nz, n3 = ncell * cl.nn;
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        const int p = b->parity;
        auto need = [&](size_t n) {
            if (count != n) bfail(b, M1LYA_ERR_ARG, "field " + std::to_string(field) + ": wrong count");
        };
        auto down32 = [&](const float* src, size_t n, double scale) {
            std::vector<float> h(n);
            bchk(b, hipMemcpy(h.data(), src, n * sizeof(float), hipMemcpyDeviceToHost), "download");
            for (size_t t = 0; t < n; ++t) host[t] = (double)h[t] * scale;
        };
        auto down64 = [&](const double* src, size_t n) {
            bchk(b, hipMemcpy(host, src, n * sizeof(double), hipMemcpyDeviceToHost), "download");
        };
        switch (field) {
        case M1LYA_J: need(n3); down32(b->A.J[p] + cl.o3, n3, cl.J0); break;
        case M1LYA_H_R: need(n3); down32(b->A.Hr[p] + cl.o3, n3, cl.J0); break;
        case M1LYA_H_Z: need(n3); down32(b->A.Hz[p] + cl.o3, n3, cl.J0); break;
        case M1LYA_F_LYA_R: need(ncell); down64(b->A.f_r + cl.oc, ncell); break;
        case M1LYA_F_LYA_Z: need(ncell); down64(b->A.f_z + cl.oc, ncell); break;
        default: bfail(b, M1LYA_ERR_ARG, "field " + std::to_string(field) + " cannot be downloaded from a batch");
        }
    });
}

int m1lya_batch_set_step(m1lya_batch* b, int cloud, double c_tilde, double dt, double z_center, int bc_r_right,
                         int bc_z_bot, int bc_z_top)
{
    return bguard(b, [&] {
        check_cloud(b, cloud);
        if (!(c_tilde > 0.0) || !(dt > 0.0)) bfail(b, M1LYA_ERR_ARG, "c_tilde and dt must be positive");
        for (int bc : {bc_r_right, bc_z_bot, bc_z_top})
            if (bc != M1LYA_REFLECT && bc != M1LYA_OUTFLOW) bfail(b, M1LYA_ERR_ARG, "invalid boundary condition");
        Cloud& cl = b->cl[cloud];
        cl.c_tilde = c_tilde;
        cl.dt = dt;
        cl.z_center = z_center;
        cl.cdt = (float)(c_tilde * dt);
        cl.cdtc = (float)(c_tilde * dt * kInvC);
        cl.bc_r_right = bc_r_right;
        cl.bc_z_bot = bc_z_bot;
        cl.bc_z_top = bc_z_top;
        b->cl_dirty = true;
    });
}

int m1lya_batch_run_steps(m1lya_batch* b, int n_steps, double* history)
{
    return bguard(b, [&] {
        if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
        if (n_steps < 0) bfail(b, M1LYA_ERR_ARG, "n_steps < 0");
        run_steps(b, n_steps, history);
    });
}

int m1lya_batch_run_until_steady(m1lya_batch* b, int max_steps, double tol, double* history, int* steady,
                                 int* steps_done)
{
    return bguard(b, [&] {
        if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
        if (max_steps < 0) bfail(b, M1LYA_ERR_ARG, "max_steps < 0");
        if (!(tol > 0.0)) bfail(b, M1LYA_ERR_ARG, "tol must be > 0");
        const int done = run_until_steady(b, max_steps, tol, history, steady);
        if (steps_done) *steps_done = done;
    });
}

int m1lya_batch_scalars(m1lya_batch* b, double* out)
{
    return bguard(b, [&] {
        if (!b || !out) bfail(b, M1LYA_ERR_ARG, "NULL argument");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        bchk(b, hipMemcpy(out, b->scalars, (size_t)b->n * 8 * sizeof(double), hipMemcpyDeviceToHost), "scalars");
    });
}

int m1lya_batch_active_bins(m1lya_batch* b, int* out)
{
    return bguard(b, [&] {
        if (!b || !out) bfail(b, M1LYA_ERR_ARG, "NULL argument");
        require_ready(b);
        if (!b->krange_valid) compute_krange(b);
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        bchk(b, hipMemcpy(out, b->krange, (size_t)b->n * 2 * sizeof(int), hipMemcpyDeviceToHost), "krange");
    });
}

int m1lya_batch_set_tiles(m1lya_batch* b, int seg_len, int waves_per_block)
{
    return bguard(b, [&] {
        if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
        if (seg_len < 1 || seg_len > 64) bfail(b, M1LYA_ERR_ARG, "seg_len out of range [1, 64]");
        if (waves_per_block < 1 || waves_per_block > kMaxWaves) bfail(b, M1LYA_ERR_ARG, "waves_per_block must be 1..8");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        b->seg_len = seg_len;
        b->wpb = waves_per_block;
        build_tiles(b);
    });
}

int m1lya_batch_get_tiles(const m1lya_batch* b, int* seg_len, int* waves_per_block)
{
    if (!b || !seg_len || !waves_per_block) return M1LYA_ERR_ARG;
    *seg_len = b->seg_len;
    *waves_per_block = b->wpb;
    return M1LYA_OK;
}

int m1lya_batch_set_graph(m1lya_batch* b, int on)
{
    return bguard(b, [&] {
        if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        b->use_graph = on != 0;
        if (!b->use_graph) drop_graphs(b);
    });
}

// Times every tile shape of the candidate set on the live batch and keeps the
// fastest.  State, active-bin ranges and buffer parity are restored, so the call
// does not advance the simulation (the step scalars are those of the last trial).
// Everything m1lya_batch_autotune changes, put back on every exit: the trial buffers and
// events are released, profiling is restored, and unless the sweep finished the tile shape,
// the saved state and the buffer parity return to what they were before the call.
struct AutotuneScope {
    m1lya_batch* b;
    size_t bytes = 0;
    float* sJ = nullptr;
    float* sR = nullptr;
    float* sZ = nullptr;
    int* sK = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int prof, seg, wpb, parity;
    bool saved = false, finished = false;

    explicit AutotuneScope(m1lya_batch* bb)
        : b(bb), prof(bb->profiling), seg(bb->seg_len), wpb(bb->wpb), parity(bb->parity) {}

    void copy(void* dst, const void* src, size_t n)
    {
        bchk(b, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, b->stream), "autotune copy");
    }
    void save()
    {
        bytes = b->n3 * sizeof(float);
        sJ = balloc<float>(b, b->n3);
        sR = balloc<float>(b, b->n3);
        sZ = balloc<float>(b, b->n3);
        sK = balloc<int>(b, (size_t)b->n * 2);
        copy(sJ, b->A.J[parity], bytes);
        copy(sR, b->A.Hr[parity], bytes);
        copy(sZ, b->A.Hz[parity], bytes);
        copy(sK, b->krange, b->n * 2 * sizeof(int));
        saved = true;
    }
    void restore()
    {
        copy(b->A.J[parity], sJ, bytes);
        copy(b->A.Hr[parity], sR, bytes);
        copy(b->A.Hz[parity], sZ, bytes);
        copy(b->krange, sK, b->n * 2 * sizeof(int));
        b->parity = parity;
    }
    ~AutotuneScope()
    {
        if (!finished) {
            // an error left a trial candidate in place: best effort back to the entry state,
            // the error itself is what the call reports
            try {
                if (b->seg_len != seg || b->wpb != wpb) {
                    b->seg_len = seg;
                    b->wpb = wpb;
                    build_tiles(b);
                    sync_clouds(b);
                }
                if (saved) restore();
            } catch (...) {
            }
            b->parity = parity;
        }
        (void)hipStreamSynchronize(b->stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        bfree(b, sJ);
        bfree(b, sR);
        bfree(b, sZ);
        bfree(b, sK);
        b->profiling = prof;
    }
};

int m1lya_batch_autotune(m1lya_batch* b, int trial_steps, int* seg_len, int* waves_per_block, double* best_ms)
{
    return bguard(b, [&] {
        if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
        if (trial_steps < 1) bfail(b, M1LYA_ERR_ARG, "trial_steps must be >= 1");
        require_ready(b);
        sync_clouds(b);
        if (!b->krange_valid) compute_krange(b);
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        AutotuneScope scope(b);
        scope.save();
        b->profiling = 0;
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        set_ctrl(b, nullptr, 0);
        // segment lengths: powers of two (16 gives 4 wavefronts per SIMD, the kernel's
        // occupancy, on 256 rows of 256 cells) and the lengths that give 3 or 6 wavefronts
        // per SIMD on 256-cell rows; 2, 4 or 8 wavefronts per block (taller blocks share
        // more Z interfaces and halo rows, shorter ones fit more blocks per CU)
        static const int seg_opts[] = {6, 8, 11, 16, 22, 32, 43, 64};
        static const int wpb_opts[] = {2, 4, 8};
        int best_s = b->seg_len, best_w = b->wpb;
        double best = 1e300;
        bchk(b, hipEventCreate(&scope.e0), "hipEventCreate");
        bchk(b, hipEventCreate(&scope.e1), "hipEventCreate");
        // two passes over the candidates, each candidate keeping its faster pass: the first
        // candidates of a single pass ran while the clocks were still rising
        for (int pass = 0; pass < 2; ++pass)
        for (int s : seg_opts)
            for (int w : wpb_opts) {
                b->seg_len = s;
                b->wpb = w;
                build_tiles(b);
                sync_clouds(b);
                scope.restore();
                launch_step_kernels(b, b->parity);   // warm-up
                b->parity ^= 1;
                bchk(b, hipEventRecord(scope.e0, b->stream), "hipEventRecord");
                for (int t = 0; t < trial_steps; ++t) {
                    launch_step_kernels(b, b->parity);
                    b->parity ^= 1;
                }
                bchk(b, hipEventRecord(scope.e1, b->stream), "hipEventRecord");
                bchk(b, hipEventSynchronize(scope.e1), "hipEventSynchronize");
                float ms = 0.f;
                bchk(b, hipEventElapsedTime(&ms, scope.e0, scope.e1), "hipEventElapsedTime");
                const double per = ms / trial_steps;
                if (per < best) {
                    best = per;
                    best_s = s;
                    best_w = w;
                }
            }
        b->seg_len = best_s;
        b->wpb = best_w;
        build_tiles(b);
        sync_clouds(b);
        scope.restore();
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        scope.finished = true;
        if (seg_len) *seg_len = best_s;
        if (waves_per_block) *waves_per_block = best_w;
        if (best_ms) *best_ms = best;
    });
}

int m1lya_batch_synchronize(m1lya_batch* b)
{
    return bguard(b, [&] {
        if (!b) bfail(nullptr, M1LYA_ERR_ARG, "batch is NULL");
        bchk(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    });
}

int m1lya_batch_set_profiling(m1lya_batch* b, int on)
{
    if (!b) return M1LYA_ERR_ARG;
    b->profiling = (on == 2) ? 2 : (on != 0 ? 1 : 0);
    return M1LYA_OK;
}

int m1lya_batch_kernel_stats_count(m1lya_batch* b)
{
    if (!b) return 0;
    const int rc = bguard(b, [&] { collect_stats(b); });
    return rc == M1LYA_OK ? (int)b->stats.size() : rc;
}

int m1lya_batch_kernel_stats(m1lya_batch* b, int index, char* name, size_t name_len, long* launches, double* total_ms)
{
    if (!b) return M1LYA_ERR_ARG;
    const int rc = bguard(b, [&] { collect_stats(b); });
    if (rc != M1LYA_OK) return rc;
    if (index < 0 || index >= (int)b->stats.size()) return M1LYA_ERR_ARG;
    auto it = b->stats.begin();
    std::advance(it, index);
    if (name && name_len) std::snprintf(name, name_len, "%s", it->first.c_str());
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.total_ms;
    return M1LYA_OK;
}

int m1lya_batch_reset_kernel_stats(m1lya_batch* b)
{
    if (!b) return M1LYA_ERR_ARG;
    return bguard(b, [&] {
        collect_stats(b);
        b->stats.clear();
    });
}

}  // extern "C"
