// m1lya_fast.h -- FP32 device math of the mixed-precision M1 Lya step.
//
// Same formulas as the FP64 path (src/AdvectJandH.jl, src/M1Closure.jl), with
// native reciprocal / sqrt and FMA contraction, rearranged where FP32's range or
// rounding would change the result:
//  - the state is normalised by a per-cloud scale J0 (J' = J/J0, H' = H/J0) and the
//    floor Jeps becomes kJeps32 in those units;
//  - limited slopes are formed in units of the half-cell width (minmod(s·x, s·y) =
//    s·minmod(x, y) for s > 0), so they never sit near the bottom of FP32's range
//    the way per-cm slopes (~1e-22) and their products do;
//  - the limiter is the median of (A, B, 0), not a test of the sign of a product that can
//    underflow;
//  - flux magnitudes are formed from H/J ratios, so |H|² never underflows;
//  - the thick/thin blend uses 1 - ϵ = τ²/(num + τ²) and (1 - ϵ)/τ = τ/(num + τ²)
//    (no cancellation, no division by τ);
//  - the Levermore closure uses χ = (3 + 4f²)/(5 + 2s) = (5 - 2s)/3, s = √(4 - 3f²)
//    (exact), so D = aI + b n⊗n with a = (s - 1)/3, b = 2 - s: one square root and no
//    division; a carries an absolute error of a few 1e-8, which is what enters the
//    fluxes D·J; the cell-centred closure switches to FP64 for f > 0.9.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(fast)

namespace m1f {

constexpr float kJeps32 = 1e-30f;
constexpr float kFmin32 = 1e-10f;
constexpr float kFmax32 = 1.0f - 1e-7f;
constexpr float kThird = 1.0f / 3.0f;
constexpr float kFreeStreamF = 0.9f;

// Interval set of one interface in half-cell units: {ΔL, ΔR, ΔL/(x_i - x_{i-1}),
// ΔL/(x_{i+1} - x_i), ΔR/(x_{i+1} - x_i), ΔR/(x_{i+2} - x_{i+1})}.
constexpr int kIvStride = 6;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
// max / min / clamp as one v_med3_f32 each (the median of (a, b, +inf) is max(a, b));
// the state holds no NaN, and med3 needs no canonicalising of its inputs
__device__ __forceinline__ float fmaxf_(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, __builtin_inff()); }
__device__ __forceinline__ float fminf_(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, -__builtin_inff()); }
__device__ __forceinline__ float fclamp(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }

// MinMod (src/ConstantsAndFunctions.jl:1571-1582): 0 unless A and B share a sign, else the
// one of smaller magnitude, which is the median of (A, B, 0): one v_med3_f32, no product
// that could underflow, and 0 when either argument is 0.
__device__ __forceinline__ float minmod(float A, float B) { return __builtin_amdgcn_fmed3f(A, B, 0.0f); }

// a = (1-χ)/2 and b = (3χ-1)/2 of the Levermore closure (src/M1Closure.jl:28-33) from
// f²: χ = (5 - 2s)/3 with s = √(4 - 3f²).
__device__ __forceinline__ void levermore_ab(float f2, float& a, float& b)
{
    const float s = fsqrt(4.0f - 3.0f * f2);
    a = (s - 1.0f) * kThird;
    b = 2.0f - s;
}

__device__ __forceinline__ void levermore_ab64(double f, double& a, double& b)
{
    const double s = sqrt(4.0 - 3.0 * f * f);
    a = (s - 1.0) * (1.0 / 3.0);
    b = 2.0 - s;
}

struct Edd32 { float rr, zz, rz; };

// Compute_D_Edd (src/M1Closure.jl:46-83), cell-centred, from the ratios H/J;
// FP64 near free streaming.
__device__ inline Edd32 D_edd(float Hr, float Hz, float J)
{
    Edd32 D;
    D.rr = kThird; D.zz = kThird; D.rz = 0.0f;
    if (J <= kJeps32) return D;
    const float iJ = rcp(J);
    const float xr = Hr * iJ, xz = Hz * iJ;
    const float f2 = xr * xr + xz * xz;
    if (!(f2 > kFmin32 * kFmin32)) return D;
    const float f = fsqrt(f2);
    if (f > kFreeStreamF) {
        const double hr = Hr, hz = Hz, j = J;
        const double hm = sqrt(hr * hr + hz * hz);
        double f64 = hm / j;
        f64 = f64 < 1e-10 ? 1e-10 : (f64 > 1.0 - 1e-10 ? 1.0 - 1e-10 : f64);
        double a, b;
        levermore_ab64(f64, a, b);
        const double nr = hr / hm, nz = hz / hm;
        D.rr = (float)(a + b * nr * nr);
        D.zz = (float)(a + b * nz * nz);
        D.rz = (float)(b * nr * nz);
        return D;
    }
    float a, b;
    const float fc = fclamp(f, kFmin32, kFmax32);
    levermore_ab(fc * fc, a, b);
    const float inf_ = rcp(f);
    const float nr = xr * inf_, nz = xz * inf_;
    const float bnr = b * nr;
    D.rr = a + bnr * nr;
    D.zz = a + b * nz * nz;
    D.rz = bnr * nz;
    return D;
}

// Compute_D_Edd_from_f_n (src/M1Closure.jl:91-119)
__device__ inline Edd32 D_edd_f_n(float f, float nr, float nz)
{
    Edd32 D;
    if (f > kFmin32) {
        f = fminf_(f, kFmax32);
        float a, b;
        levermore_ab(f * f, a, b);
        float bnr = b * nr;
        D.rr = a + bnr * nr;
        D.zz = a + b * nz * nz;
        D.rz = bnr * nz;
    } else {
        D.rr = kThird; D.zz = kThird; D.rz = 0.0f;
    }
    return D;
}

// One reconstructed side of an interface: limits |f| to 1 (src/AdvectJandH.jl:430-452),
// then the direction n = f/|f| and the closure D = aI + b n⊗n (D_ϕϕ = a).
struct Side32 {
    float fr, fz, f, nr, nz, a, b;
};

__device__ __forceinline__ Side32 side_state(float fr, float fz)
{
    Side32 s;
    const float f2 = fr * fr + fz * fz;
    // No branches: below |f| = 1e-10 the closure is isotropic without a special case, since
    // 4 - 3 f² then rounds to 4, so b = 0 and a = 1/3 exactly and the (finite) direction
    // drops out of D; above |f| = 1 the limited f is the direction itself.
    const float i = frsq(fmaxf_(f2, kFmin32 * kFmin32));
    s.nr = fr * i;
    s.nz = fz * i;
    const bool lim = f2 > 1.0f;
    s.fr = lim ? s.nr : fr;
    s.fz = lim ? s.nz : fz;
    s.f = fminf_(f2 * i, 1.0f);
    levermore_ab(fminf_(f2, kFmax32 * kFmax32), s.a, s.b);
    return s;
}

__device__ inline void hll_eigen(float f, float mu, float& l1, float& l3)
{
    f = fclamp(f, kFmin32, kFmax32);
    mu = fclamp(mu, -1.0f, 1.0f);
    float f2 = f * f, mu2 = mu * mu, muf = mu * f;
    float SQ = fsqrt(fmaxf_(4.0f - 3.0f * f2, 0.0f));
    float iSQ = rcp(SQ);
    float FAC = fsqrt(fmaxf_((2.0f / 3.0f) * (4.0f - 3.0f * f2 - SQ) + 2.0f * mu2 * (2.0f - f2 - SQ), 0.0f));
    l1 = fclamp((muf - FAC) * iSQ, -1.0f, 1.0f);
    l3 = fclamp((muf + FAC) * iSQ, -1.0f, 1.0f);
}

// Cells (i, i+1, i+2, i-1) of one interface: J and the flux ratios H/J.
struct Stencil32 {
    float J[4], fr[4], fz[4];
};

// H/J of one cell (J floored as in the reference's reconstruction).
__device__ __forceinline__ void flux_ratios(float J, float Hr, float Hz, float& fr, float& fz)
{
    const float iJ = rcp(fmaxf_(J, kJeps32));
    fr = Hr * iJ;
    fz = Hz * iJ;
}

struct Face32 {
    float FJ, FHr, FHz;
    float J_L, J_R, Hz_R;
    float kphi_L, kphi_R;
};

// ReconstructInterface -> closure -> GLF/HLL flux -> upwind advection flux
// (src/AdvectJandH.jl:81-118, 155-204, 276-282, 359-473).  iv: kIvStride values;
// the stencil carries J and H/J of its four cells.
__device__ inline Face32 interface_flux(int dir, int hll, const Stencil32& s, const float* iv, float aF_L,
                                        float aF_R, float u_over_c)
{
    const float dL = iv[0], dR = iv[1], cL1 = iv[2], cL2 = iv[3], cR1 = iv[4], cR2 = iv[5];
    const float tau_L = aF_L * dL, tau_R = aF_R * dR;
    const float fr0 = s.fr[0], fr1 = s.fr[1], fr2 = s.fr[2], fr3 = s.fr[3];
    const float fz0 = s.fz[0], fz1 = s.fz[1], fz2 = s.fz[2], fz3 = s.fz[3];
    // left state: x_i + minmod(slope_{i-1/2}, slope_{i+1/2})·ΔL, slopes times ΔL
    const float dJc = s.J[1] - s.J[0], dfrc = fr1 - fr0, dfzc = fz1 - fz0;
    float J_L = fmaxf_(s.J[0] + minmod(cL1 * (s.J[0] - s.J[3]), cL2 * dJc), kJeps32);
    float J_R = fmaxf_(s.J[1] - minmod(cR2 * (s.J[2] - s.J[1]), cR1 * dJc), kJeps32);
    float fr_L = fr0 + minmod(cL1 * (fr0 - fr3), cL2 * dfrc);
    float fr_R = fr1 - minmod(cR2 * (fr2 - fr1), cR1 * dfrc);
    float fz_L = fz0 + minmod(cL1 * (fz0 - fz3), cL2 * dfzc);
    float fz_R = fz1 - minmod(cR2 * (fz2 - fz1), cR1 * dfzc);
    const Side32 SL = side_state(fr_L, fz_L), SR = side_state(fr_R, fz_R);
    const float Hr_L = SL.fr * J_L, Hr_R = SR.fr * J_R, Hz_L = SL.fz * J_L, Hz_R = SR.fz * J_R;
    Edd32 DL, DR;
    DL.rr = SL.a + SL.b * SL.nr * SL.nr; DL.zz = SL.a + SL.b * SL.nz * SL.nz; DL.rz = SL.b * SL.nr * SL.nz;
    DR.rr = SR.a + SR.b * SR.nr * SR.nr; DR.zz = SR.a + SR.b * SR.nz * SR.nz; DR.rz = SR.b * SR.nr * SR.nz;
    const float f_L = SL.f, f_R = SR.f, nr_L = SL.nr, nz_L = SL.nz, nr_R = SR.nr, nz_R = SR.nz;
    float FJ_L, FJ_R, FHr_L, FHr_R, FHz_L, FHz_R;
    if (dir == 0) {
        FJ_L = Hr_L; FJ_R = Hr_R;
        FHr_L = DL.rr * J_L; FHr_R = DR.rr * J_R;
        FHz_L = DL.rz * J_L; FHz_R = DR.rz * J_R;
    } else {
        FJ_L = Hz_L; FJ_R = Hz_R;
        FHr_L = DL.rz * J_L; FHr_R = DR.rz * J_R;
        FHz_L = DL.zz * J_L; FHz_R = DR.zz * J_R;
    }
    // ϵ = (0.2 + τ)/(0.2 + τ + τ²) and the Fick flux (1 - ϵ)·(-ΔJ/(3τ)), both from one
    // reciprocal: 1 - ϵ = τ²·den and (1 - ϵ)/τ = τ·den, so nothing is divided by τ and
    // 1 - ϵ does not cancel (in vacuum, τ = 0, both vanish exactly).
    const float tau = tau_L + tau_R;
    const float num = 0.2f + tau;
    const float den = rcp(num + tau * tau);
    const float eps = num * den;
    const float ome = tau * tau * den;
    const float F_J_diff_ome = -kThird * dJc * (tau * den);
    float FJ, FHr, FHz;
    if (hll) {
        float l1L, l3L, l1R, l3R;
        hll_eigen(f_L, dir == 0 ? nr_L : nz_L, l1L, l3L);
        hll_eigen(f_R, dir == 0 ? nr_R : nz_R, l1R, l3R);
        const float S_R = fmaxf_(fmaxf_(l3L, l3R), 0.0f), S_L = fminf_(fminf_(l1L, l1R), 0.0f);
        const float idS = rcp(fmaxf_(S_R - S_L, 1e-30f));
        const float SRSL = S_R * S_L;
        FJ = eps * ((S_R * FJ_L - S_L * FJ_R + SRSL * (J_R - J_L)) * idS) + F_J_diff_ome;
        FHr = eps * ((S_R * FHr_L - S_L * FHr_R + SRSL * (Hr_R - Hr_L)) * idS) + 0.5f * ome * (FHr_L + FHr_R);
        FHz = eps * ((S_R * FHz_L - S_L * FHz_R + SRSL * (Hz_R - Hz_L)) * idS) + 0.5f * ome * (FHz_L + FHz_R);
    } else {
        FJ = eps * (0.5f * (FJ_L + FJ_R) - 0.5f * (J_R - J_L)) + F_J_diff_ome;
        const float he = 0.5f * eps;
        FHr = 0.5f * (FHr_L + FHr_R) - he * (Hr_R - Hr_L);
        FHz = 0.5f * (FHz_L + FHz_R) - he * (Hz_R - Hz_L);
    }
    Face32 r;
    r.FJ = FJ;
    r.FHr = FHr;
    r.FHz = FHz;
    if (u_over_c != 0.0f) {   // the face velocity is one value for all bins: a uniform branch
        const bool up = (u_over_c < 0.0f);
        r.FJ += u_over_c * (up ? J_R : J_L);
        r.FHr += u_over_c * (up ? Hr_R : Hr_L);
        r.FHz += u_over_c * (up ? Hz_R : Hz_L);
    }
    r.J_L = J_L; r.J_R = J_R; r.Hz_R = Hz_R;
    // clamp(1 - D_rr - D_ZZ, 0, 1)·J with 1 - D_rr - D_ZZ = D_ϕϕ = a
    r.kphi_L = fclamp(SL.a, 0.0f, 1.0f) * J_L;
    r.kphi_R = fclamp(SR.a, 0.0f, 1.0f) * J_R;
    return r;
}

// Compute_K_ϕϕ_InnerAxis (src/AdvectJandH.jl:499-522)
__device__ inline float kphi_inner_axis(float J_axis, float Hz_axis)
{
    J_axis = fmaxf_(J_axis, kJeps32);
    const float fz = fclamp(Hz_axis * rcp(J_axis), -1.0f, 1.0f);
    const float f = fabsf(fz);
    const float nz = (f > 0.0f) ? (fz > 0.0f ? 1.0f : -1.0f) : 0.0f;
    const Edd32 D = D_edd_f_n(f, 0.0f, nz);
    return fclamp(1.0f - D.rr - D.zz, 0.0f, 1.0f) * J_axis;
}

// EnforceCausalReducedFlux (src/AdvectJandH.jl:478-493) in ratio form: scales
// (Hr, Hz) to |H| = J when |H| > J, without forming |H|² (which can underflow).
__device__ __forceinline__ void causal_limit(float J, float& Hr, float& Hz)
{
    const float iJ = rcp(J);
    const float xr = Hr * iJ, xz = Hz * iJ;
    const float m2 = xr * xr + xz * xz;
    if (m2 > 1.0f) {
        const float fac = frsq(m2);
        Hr *= fac;
        Hz *= fac;
    }
}

}  // namespace m1f

// contraction ends with this header: code that includes it rounds as written again (the
// library is built with -ffp-contract=off; this target has no float_control push / pop)
#pragma clang fp contract(off)
