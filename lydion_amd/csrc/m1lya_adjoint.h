// m1lya_adjoint.h -- the FP64 step's formulas over a generic scalar, for the
// hand-written reverse sweep (m1lya_adjoint.hip).
//
// Every routine restates the forward routine of m1lya_device.h with the same
// branches, taken on the value part.  Instantiated with double it is the
// forward formula; with Dl (value + one tangent) it gives the directional
// derivative along the seeded input, which the reverse kernels contract with
// the output cotangents (vector-Jacobian products of small local maps: one
// interface flux, one closure, one row of the frequency-coupling matrix).
// Kinks (limiters, max/min, clamps) take the derivative of the branch the
// value selects; ties go the way the forward breaks them.
#pragma once
#include <hip/hip_runtime.h>

#include "m1lya_device.h"

namespace m1a {

using namespace m1;

struct Dl {
    double v, d;
    __device__ Dl() : v(0.0), d(0.0) {}
    __device__ Dl(double x) : v(x), d(0.0) {}
    __device__ Dl(double x, double dx) : v(x), d(dx) {}
};

// 1/x for tangents: the hardware estimate and two Newton steps (a few ulp).  Values are
// formed exactly as the forward forms them; only their derivatives use this.
__device__ __forceinline__ double trcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ Dl operator+(Dl a, Dl b) { return Dl(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dl operator-(Dl a, Dl b) { return Dl(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dl operator-(Dl a) { return Dl(-a.v, -a.d); }
__device__ __forceinline__ Dl operator*(Dl a, Dl b) { return Dl(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dl operator/(Dl a, Dl b)
{
    double q = a.v / b.v;
    return Dl(q, (a.d - q * b.d) * trcp(b.v));
}
__device__ __forceinline__ Dl operator+(Dl a, double b) { return Dl(a.v + b, a.d); }
__device__ __forceinline__ Dl operator+(double a, Dl b) { return Dl(a + b.v, b.d); }
__device__ __forceinline__ Dl operator-(Dl a, double b) { return Dl(a.v - b, a.d); }
__device__ __forceinline__ Dl operator-(double a, Dl b) { return Dl(a - b.v, -b.d); }
__device__ __forceinline__ Dl operator*(Dl a, double b) { return Dl(a.v * b, a.d * b); }
__device__ __forceinline__ Dl operator*(double a, Dl b) { return Dl(a * b.v, a * b.d); }
__device__ __forceinline__ Dl operator/(Dl a, double b) { return Dl(a.v / b, a.d * trcp(b)); }
__device__ __forceinline__ Dl operator/(double a, Dl b)
{
    double q = a / b.v;
    return Dl(q, -q * b.d * trcp(b.v));
}
__device__ __forceinline__ Dl& operator+=(Dl& a, Dl b) { a = a + b; return a; }
__device__ __forceinline__ Dl& operator*=(Dl& a, Dl b) { a = a * b; return a; }

// Value and two tangents: two directional derivatives in one pass.
struct Dl2 {
    double v, d0, d1;
    __device__ Dl2() : v(0.0), d0(0.0), d1(0.0) {}
    __device__ Dl2(double x) : v(x), d0(0.0), d1(0.0) {}
    __device__ Dl2(double x, double a, double b) : v(x), d0(a), d1(b) {}
};

__device__ __forceinline__ Dl2 operator+(Dl2 a, Dl2 b) { return Dl2(a.v + b.v, a.d0 + b.d0, a.d1 + b.d1); }
__device__ __forceinline__ Dl2 operator-(Dl2 a, Dl2 b) { return Dl2(a.v - b.v, a.d0 - b.d0, a.d1 - b.d1); }
__device__ __forceinline__ Dl2 operator-(Dl2 a) { return Dl2(-a.v, -a.d0, -a.d1); }
__device__ __forceinline__ Dl2 operator*(Dl2 a, Dl2 b)
{
    return Dl2(a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1);
}
__device__ __forceinline__ Dl2 operator/(Dl2 a, Dl2 b)
{
    const double q = a.v / b.v, r = trcp(b.v);
    return Dl2(q, (a.d0 - q * b.d0) * r, (a.d1 - q * b.d1) * r);
}
__device__ __forceinline__ Dl2 operator+(Dl2 a, double b) { return Dl2(a.v + b, a.d0, a.d1); }
__device__ __forceinline__ Dl2 operator+(double a, Dl2 b) { return Dl2(a + b.v, b.d0, b.d1); }
__device__ __forceinline__ Dl2 operator-(Dl2 a, double b) { return Dl2(a.v - b, a.d0, a.d1); }
__device__ __forceinline__ Dl2 operator-(double a, Dl2 b) { return Dl2(a - b.v, -b.d0, -b.d1); }
__device__ __forceinline__ Dl2 operator*(Dl2 a, double b) { return Dl2(a.v * b, a.d0 * b, a.d1 * b); }
__device__ __forceinline__ Dl2 operator*(double a, Dl2 b) { return Dl2(a * b.v, a * b.d0, a * b.d1); }
__device__ __forceinline__ Dl2 operator/(Dl2 a, double b)
{
    const double r = trcp(b);
    return Dl2(a.v / b, a.d0 * r, a.d1 * r);
}
__device__ __forceinline__ Dl2 operator/(double a, Dl2 b)
{
    const double q = a / b.v, r = -q * trcp(b.v);
    return Dl2(q, r * b.d0, r * b.d1);
}
__device__ __forceinline__ Dl2& operator+=(Dl2& a, Dl2 b) { a = a + b; return a; }
__device__ __forceinline__ Dl2& operator*=(Dl2& a, Dl2 b) { a = a * b; return a; }

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(const Dl& x) { return x.v; }
__device__ __forceinline__ double val(const Dl2& x) { return x.v; }
__device__ __forceinline__ double der(double) { return 0.0; }
__device__ __forceinline__ double der(const Dl& x) { return x.d; }

__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ Dl tsqrt(Dl x)
{
    double s = sqrt(x.v);
    return Dl(s, s > 0.0 ? 0.5 * x.d * trcp(s) : 0.0);
}
__device__ __forceinline__ Dl2 tsqrt(Dl2 x)
{
    const double s = sqrt(x.v);
    const double h = s > 0.0 ? 0.5 * trcp(s) : 0.0;
    return Dl2(s, h * x.d0, h * x.d1);
}
__device__ __forceinline__ double texp(double x) { return exp_det(x); }
__device__ __forceinline__ Dl texp(Dl x)
{
    double e = exp_det(x.v);
    return Dl(e, e * x.d);
}
__device__ __forceinline__ Dl2 texp(Dl2 x)
{
    const double e = exp_det(x.v);
    return Dl2(e, e * x.d0, e * x.d1);
}
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ Dl tabs(Dl x) { return x.v < 0.0 ? -x : x; }
__device__ __forceinline__ Dl2 tabs(Dl2 x) { return x.v < 0.0 ? -x : x; }

// max/min/clamp with the forward's tie rules (m1lya_device.h dmax/dmin/dclamp)
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return (val(b) > val(a)) ? b : a; }
template <class T> __device__ __forceinline__ T tmin(T a, T b) { return (val(b) < val(a)) ? b : a; }
template <class T> __device__ __forceinline__ T tclamp(T x, double lo, double hi)
{
    return (val(x) > hi) ? T(hi) : ((val(x) < lo) ? T(lo) : x);
}

// MinMod: src/ConstantsAndFunctions.jl:1571-1582
template <class T> __device__ __forceinline__ T tminmod(T A, T B)
{
    if (val(A) * val(B) <= 0.0) return T(0.0);
    return (fabs(val(A)) < fabs(val(B))) ? A : B;
}

// Levermore closure factor: src/M1Closure.jl:28-33
template <class T> __device__ __forceinline__ T tchi(T f)
{
    T f2 = f * f;
    return (3.0 + 4.0 * f2) / (5.0 + 2.0 * tsqrt(4.0 - 3.0 * f2));
}

template <class T> struct EddT { T rr, zz, rz; };

// Compute_D_Edd: src/M1Closure.jl:46-83
template <class T> __device__ EddT<T> tD_edd(T H_r, T H_z, T J)
{
    const double Heps2 = kHeps * kHeps;
    const double fmin2 = kFmin * kFmin;
    T Hmag2 = H_r * H_r + H_z * H_z;
    EddT<T> D;
    if (val(J) <= kJeps || val(Hmag2) <= Heps2 || val(Hmag2) <= fmin2 * val(J) * val(J)) {
        D.rr = T(kOneThird); D.zz = T(kOneThird); D.rz = T(0.0);
        return D;
    }
    T Hmag = tsqrt(Hmag2);
    T f = tclamp(Hmag / J, kFmin, kFmax);
    T chi = tchi(f);
    T inv_Hmag = 1.0 / Hmag;
    T a = (1.0 - chi) * 0.5;
    T b = (3.0 * chi - 1.0) * 0.5;
    T nr = H_r * inv_Hmag;
    T nz = H_z * inv_Hmag;
    T b_nr = b * nr;
    D.rr = a + b_nr * nr;
    D.zz = a + b * nz * nz;
    D.rz = b_nr * nz;
    return D;
}

// Compute_D_Edd_from_f_n: src/M1Closure.jl:91-119
template <class T> __device__ EddT<T> tD_edd_f_n(T f, T nr, T nz)
{
    EddT<T> D;
    if (val(f) > kFmin) {
        f = tmin(f, T(kFmax));
        T chi = tchi(f);
        T a = (1.0 - chi) * 0.5;
        T b = (3.0 * chi - 1.0) * 0.5;
        T b_nr = b * nr;
        D.rr = a + b_nr * nr;
        D.zz = a + b * nz * nz;
        D.rz = b_nr * nz;
    } else {
        D.rr = T(kOneThird); D.zz = T(kOneThird); D.rz = T(0.0);
    }
    return D;
}

// HLL_Eigenvalues: src/M1Closure.jl:273-296
template <class T> __device__ void thll_eigen(T f, T mu, T& l1, T& l3)
{
    f = tclamp(f, kFmin, kFmax);
    mu = tclamp(mu, -1.0, 1.0);
    T f2 = f * f, mu2 = mu * mu, muf = mu * f;
    T SQ = tsqrt(tmax(4.0 - 3.0 * f2, T(0.0)));
    T inv_SQ = 1.0 / SQ;
    T FAC = tsqrt(tmax(kTwoThirds * (4.0 - 3.0 * f2 - SQ) + 2.0 * mu2 * (2.0 - f2 - SQ), T(0.0)));
    l1 = tclamp((muf - FAC) * inv_SQ, -1.0, 1.0);
    l3 = tclamp((muf + FAC) * inv_SQ, -1.0, 1.0);
}

template <class T> struct StencilT { T J[4], Hr[4], Hz[4]; };

template <class T> struct ReconT {
    T J_L, J_R, Hr_L, Hr_R, Hz_L, Hz_R, tau_L, tau_R, f_L, f_R, nr_L, nr_R, nz_L, nz_R;
};

// ReconstructInterface: src/AdvectJandH.jl:359-473
template <class T>
__device__ ReconT<T> treconstruct(const StencilT<T>& s, const double* iv, T aF_i, T aF_ip1)
{
    const double dL = iv[0], dR = iv[1], inv_i_im1 = iv[2], inv_ip1_i = iv[3], inv_ip2_ip1 = iv[4];
    ReconT<T> o;
    o.tau_L = aF_i * dL;
    o.tau_R = aF_ip1 * dR;
    T inv_J_i = 1.0 / tmax(s.J[0], T(kJeps));
    T inv_J_ip1 = 1.0 / tmax(s.J[1], T(kJeps));
    T inv_J_ip2 = 1.0 / tmax(s.J[2], T(kJeps));
    T inv_J_im1 = 1.0 / tmax(s.J[3], T(kJeps));
    T fr_i = s.Hr[0] * inv_J_i, fr_ip1 = s.Hr[1] * inv_J_ip1;
    T fr_ip2 = s.Hr[2] * inv_J_ip2, fr_im1 = s.Hr[3] * inv_J_im1;
    T fz_i = s.Hz[0] * inv_J_i, fz_ip1 = s.Hz[1] * inv_J_ip1;
    T fz_ip2 = s.Hz[2] * inv_J_ip2, fz_im1 = s.Hz[3] * inv_J_im1;

    T dJc = (s.J[1] - s.J[0]) * inv_ip1_i;
    T dJL = tminmod((s.J[0] - s.J[3]) * inv_i_im1, dJc);
    T dJR = tminmod((s.J[2] - s.J[1]) * inv_ip2_ip1, dJc);
    T dfrc = (fr_ip1 - fr_i) * inv_ip1_i;
    T dfrL = tminmod((fr_i - fr_im1) * inv_i_im1, dfrc);
    T dfrR = tminmod((fr_ip2 - fr_ip1) * inv_ip2_ip1, dfrc);
    T dfzc = (fz_ip1 - fz_i) * inv_ip1_i;
    T dfzL = tminmod((fz_i - fz_im1) * inv_i_im1, dfzc);
    T dfzR = tminmod((fz_ip2 - fz_ip1) * inv_ip2_ip1, dfzc);

    T J_L = tmax(s.J[0] + dJL * dL, T(kJeps));
    T J_R = tmax(s.J[1] - dJR * dR, T(kJeps));
    T fr_L = fr_i + dfrL * dL;
    T fr_R = fr_ip1 - dfrR * dR;
    T fz_L = fz_i + dfzL * dL;
    T fz_R = fz_ip1 - dfzR * dR;

    T f_L = tsqrt(fr_L * fr_L + fz_L * fz_L);
    if (val(f_L) > 1.0) {
        T inv_f = 1.0 / f_L;
        fr_L = fr_L * inv_f; fz_L = fz_L * inv_f; f_L = T(1.0);
    }
    T f_R = tsqrt(fr_R * fr_R + fz_R * fz_R);
    if (val(f_R) > 1.0) {
        T inv_f = 1.0 / f_R;
        fr_R = fr_R * inv_f; fz_R = fz_R * inv_f; f_R = T(1.0);
    }
    o.nr_L = T(0.0); o.nz_L = T(0.0); o.nr_R = T(0.0); o.nz_R = T(0.0);
    if (val(f_L) > 0.0) { T inv = 1.0 / f_L; o.nr_L = fr_L * inv; o.nz_L = fz_L * inv; }
    if (val(f_R) > 0.0) { T inv = 1.0 / f_R; o.nr_R = fr_R * inv; o.nz_R = fz_R * inv; }
    o.J_L = J_L; o.J_R = J_R;
    o.Hr_L = fr_L * J_L; o.Hr_R = fr_R * J_R;
    o.Hz_L = fz_L * J_L; o.Hz_R = fz_R * J_R;
    o.f_L = f_L; o.f_R = f_R;
    return o;
}

// Compute_K_ϕϕ_InnerAxis: src/AdvectJandH.jl:499-522
template <class T> __device__ T tkphi_inner_axis(T J_axis, T Hz_axis)
{
    J_axis = tmax(J_axis, T(kJeps));
    T fz = Hz_axis / J_axis;
    fz = tclamp(fz, -1.0, 1.0);
    T f = tabs(fz);
    double nz = 0.0;
    if (val(f) > 0.0) nz = (val(fz) > 0.0) ? 1.0 : ((val(fz) < 0.0) ? -1.0 : val(fz));
    EddT<T> D = tD_edd_f_n(f, T(0.0), T(nz));
    return tclamp(1.0 - D.rr - D.zz, 0.0, 1.0) * J_axis;
}

template <class T> struct FaceT {
    T FJ, FHr, FHz;
    T kphi_L;   // clamp(1 - D_rr - D_ZZ, 0, 1) J on the left side  (K_ϕϕ of the left cell, r faces)
    T kphi_R;   // ... on the right side; at the axis face the inner-axis K_ϕϕ of cell 0
};

// One interface: ReconstructInterface -> closure -> GetFluxM1 -> + GetAdvectionFlux
// (src/AdvectJandH.jl:682-716 and its copies; :207-282 for HLL and the upwind term).
template <class T>
__device__ FaceT<T> tface(int dir, int hll, bool axis, const StencilT<T>& s, const double* iv, T aF_L, T aF_R,
                          double u_over_c)
{
    ReconT<T> q = treconstruct(s, iv, aF_L, aF_R);
    EddT<T> DL = tD_edd_f_n(q.f_L, q.nr_L, q.nz_L);
    EddT<T> DR = tD_edd_f_n(q.f_R, q.nr_R, q.nz_R);
    T FJ_L, FJ_R, FHr_L, FHr_R, FHz_L, FHz_R;
    if (dir == 0) {
        FJ_L = q.Hr_L; FJ_R = q.Hr_R;
        FHr_L = DL.rr * q.J_L; FHr_R = DR.rr * q.J_R;
        FHz_L = DL.rz * q.J_L; FHz_R = DR.rz * q.J_R;
    } else {
        FJ_L = q.Hz_L; FJ_R = q.Hz_R;
        FHr_L = DL.rz * q.J_L; FHr_R = DR.rz * q.J_R;
        FHz_L = DL.zz * q.J_L; FHz_R = DR.zz * q.J_R;
    }
    T tau_cell = q.tau_L + q.tau_R;
    T num = 0.2 + tau_cell;
    T eps = num / (num + tau_cell * tau_cell);
    T ome = 1.0 - eps;
    T tau_eff = tmax(tau_cell, T(1e-300));
    T F_J_diff = -kOneThird * (s.J[1] - s.J[0]) / tau_eff;
    FaceT<T> r;
    if (hll) {
        // GetFluxHLL: src/AdvectJandH.jl:155-204
        T l1L, l3L, l1R, l3R;
        thll_eigen(q.f_L, dir == 0 ? q.nr_L : q.nz_L, l1L, l3L);
        thll_eigen(q.f_R, dir == 0 ? q.nr_R : q.nz_R, l1R, l3R);
        T half_ome = 0.5 * ome;
        T S_R = tmax(tmax(l3L, l3R), T(0.0));
        T S_L = tmin(tmin(l1L, l1R), T(0.0));
        T inv_dS = 1.0 / tmax(S_R - S_L, T(1e-300));
        T SRSL = S_R * S_L;
        T FJ = (S_R * FJ_L - S_L * FJ_R + SRSL * (q.J_R - q.J_L)) * inv_dS;
        T FHr = (S_R * FHr_L - S_L * FHr_R + SRSL * (q.Hr_R - q.Hr_L)) * inv_dS;
        T FHz = (S_R * FHz_L - S_L * FHz_R + SRSL * (q.Hz_R - q.Hz_L)) * inv_dS;
        r.FJ = eps * FJ + ome * F_J_diff;
        r.FHr = eps * FHr + half_ome * (FHr_L + FHr_R);
        r.FHz = eps * FHz + half_ome * (FHz_L + FHz_R);
    } else {
        // GetFluxGLF: src/AdvectJandH.jl:81-118
        T F_J_GLF = 0.5 * (FJ_L + FJ_R) - 0.5 * (q.J_R - q.J_L);
        r.FJ = eps * F_J_GLF + ome * F_J_diff;
        T half_eps = 0.5 * eps;
        r.FHr = 0.5 * (FHr_L + FHr_R) - half_eps * (q.Hr_R - q.Hr_L);
        r.FHz = 0.5 * (FHz_L + FHz_R) - half_eps * (q.Hz_R - q.Hz_L);
    }
    // GetAdvectionFlux: src/AdvectJandH.jl:276-282
    const bool up = (u_over_c < 0.0);
    r.FJ = r.FJ + u_over_c * (up ? q.J_R : q.J_L);
    r.FHr = r.FHr + u_over_c * (up ? q.Hr_R : q.Hr_L);
    r.FHz = r.FHz + u_over_c * (up ? q.Hz_R : q.Hz_L);
    r.kphi_L = tclamp(1.0 - DL.rr - DL.zz, 0.0, 1.0) * q.J_L;
    r.kphi_R = axis ? tkphi_inner_axis(q.J_R, q.Hz_R) : tclamp(1.0 - DR.rr - DR.zz, 0.0, 1.0) * q.J_R;
    return r;
}

// D_colon_∇u_Cylindrical: src/ConstantsAndFunctions.jl:883-895
template <class T>
__device__ __forceinline__ T tD_colon_grad_u(const EddT<T>& D, double durdr, double durdz, double duzdr,
                                            double duzdz, double urr)
{
    T Dpp = tmax(1.0 - D.rr - D.zz, T(0.0));
    return D.rr * durdr + D.rz * (durdz + duzdr) + D.zz * duzdz + Dpp * urr;
}

// OneMinusϕ_L_and_χ_L / OneMinusϕ_R_and_χ_R: src/FreqFactors.jl:54-133
template <class T> __device__ void tomphi_L(T P, T& o, T& chi)
{
    if (val(P) >= kPmax) { o = T(exp_det(-kPmax)); chi = (1.0 - o) / kPmax; return; }
    if (val(P) <= -kPmax) { o = T(exp_det(kPmax)); chi = (1.0 - o) / (-kPmax); return; }
    if (fabs(val(P)) < kPmin) { o = 1.0 - P + 0.5 * (P * P); chi = 1.0 - 0.5 * P; return; }
    o = texp(-P);
    chi = (1.0 - o) / P;
}
template <class T> __device__ void tomphi_R(T P, T& o, T& chi)
{
    if (val(P) >= kPmax) { o = T(exp_det(kPmax)); chi = (1.0 - o) / kPmax; return; }
    if (val(P) <= -kPmax) { o = T(exp_det(-kPmax)); chi = (1.0 - o) / (-kPmax); return; }
    if (fabs(val(P)) < kPmin) { o = 1.0 + P + 0.5 * (P * P); chi = -1.0 - 0.5 * P; return; }
    o = texp(P);
    chi = (1.0 - o) / P;
}

// Constants of one interior row k of the frequency-coupling matrix.
struct RowConst {
    double nuCkp, nuCk, nuCkm;       // bin centres k+1, k, k-1
    double dkp, dk, dkm;             // νC_{k+1}-ν_{k+1}, ν_{k+1}-νC_k, ν_k-νC_{k-1}
    double inv_dnu_k;
    double two_inv_kp, two_inv_k, two_inv_km;
    int recoil;
};

template <class T> struct RowT { T Fkp, Fk, Fkm; };

// One row of UpdateFreqFactors! (src/FreqFactors.jl:142-278): the coefficients of
// J_{k+1}, J_k, J_{k-1} from Hvoigt at k-1, k, k+1, PREFACT = ½ max(α, α_floor) ΔνD²,
// RECOIL_0 = h/(k_B T) and VEL = D:∇u / c of bin k.
template <class T>
__device__ RowT<T> tfreq_row(const RowConst& c, T Hvm, T Hv0, T Hvp, T PREFACT, T RECOIL_0, T VEL)
{
    T Kkp = PREFACT * Hvp, Kk = PREFACT * Hv0, Kkm = PREFACT * Hvm;
    T Ckp = VEL * c.nuCkp, Ck = VEL * c.nuCk, Ckm = VEL * c.nuCkm;
    if (c.recoil) {
        Ckp = Ckp + Kkp * (RECOIL_0 - c.two_inv_kp);
        Ck = Ck + Kk * (RECOIL_0 - c.two_inv_k);
        Ckm = Ckm + Kkm * (RECOIL_0 - c.two_inv_km);
    }
    T qkp = c.dkp * (1.0 / Kkp), qk = c.dk * (1.0 / Kk), qkm = c.dkm * (1.0 / Kkm);
    T Pkp = tclamp(Ckp * qkp, -kPmax, kPmax);
    T Pk = tclamp(Ck * qk, -kPmax, kPmax);
    T Pkm = tclamp(Ckm * qkm, -kPmax, kPmax);
    T oRkp, xRkp, oLk, xLk, oRk, xRk, oLkm, xLkm;
    tomphi_R(Pkp, oRkp, xRkp);
    tomphi_L(Pk, oLk, xLk);
    tomphi_R(Pk, oRk, xRk);
    tomphi_L(Pkm, oLkm, xLkm);
    T den_kp = c.inv_dnu_k * (1.0 / (xLk * qk - xRkp * qkp));
    T den_km = c.inv_dnu_k * (1.0 / (xLkm * qkm - xRk * qk));
    RowT<T> r;
    r.Fkp = oRkp * den_kp;
    r.Fkm = oLkm * den_km;
    r.Fk = oLk * den_kp + oRk * den_km;
    return r;
}

// 1 - ϕ and χ of tomphi_L / tomphi_R with their derivatives along P (0 where P is clamped
// to ±kPmax, where the forward's values are constants)
__device__ __forceinline__ void omphi_d(bool left, double P, double& o, double& chi, double& do_, double& dchi)
{
    if (P >= kPmax || P <= -kPmax) {
        const double e = left ? (P >= kPmax ? -kPmax : kPmax) : (P >= kPmax ? kPmax : -kPmax);
        o = exp_det(e);
        chi = (1.0 - o) / (P >= kPmax ? kPmax : -kPmax);
        do_ = 0.0; dchi = 0.0;
        return;
    }
    if (fabs(P) < kPmin) {
        if (left) { o = 1.0 - P + 0.5 * (P * P); chi = 1.0 - 0.5 * P; do_ = -1.0 + P; }
        else { o = 1.0 + P + 0.5 * (P * P); chi = -1.0 - 0.5 * P; do_ = 1.0 + P; }
        dchi = -0.5;
        return;
    }
    o = exp_det(left ? -P : P);
    chi = (1.0 - o) / P;
    do_ = left ? -o : o;
    dchi = (-do_ - chi) / P;
}

// Reverse mode of tfreq_row: cotangents of (F_{k+1}, F_k, F_{k-1}) -> of its six inputs
// g = (Hvoigt_{k-1}, Hvoigt_k, Hvoigt_{k+1}, PREFACT, RECOIL_0, VEL); the values follow the
// forward's operations and branches, so the result equals the dual-number passes to rounding.
__device__ inline void freq_row_vjp(const RowConst& c, double Hvm, double Hv0, double Hvp, double PREFACT,
                                    double RECOIL_0, double VEL, double lFkp, double lFk, double lFkm, double g[6])
{
    const double Kkp = PREFACT * Hvp, Kk = PREFACT * Hv0, Kkm = PREFACT * Hvm;
    double Ckp = VEL * c.nuCkp, Ck = VEL * c.nuCk, Ckm = VEL * c.nuCkm;
    if (c.recoil) {
        Ckp = Ckp + Kkp * (RECOIL_0 - c.two_inv_kp);
        Ck = Ck + Kk * (RECOIL_0 - c.two_inv_k);
        Ckm = Ckm + Kkm * (RECOIL_0 - c.two_inv_km);
    }
    const double qkp = c.dkp * (1.0 / Kkp), qk = c.dk * (1.0 / Kk), qkm = c.dkm * (1.0 / Kkm);
    const double xkp = Ckp * qkp, xk = Ck * qk, xkm = Ckm * qkm;
    const bool pkp = !(xkp > kPmax) && !(xkp < -kPmax), pk = !(xk > kPmax) && !(xk < -kPmax);
    const bool pkm = !(xkm > kPmax) && !(xkm < -kPmax);
    const double Pkp = (xkp > kPmax) ? kPmax : ((xkp < -kPmax) ? -kPmax : xkp);
    const double Pk = (xk > kPmax) ? kPmax : ((xk < -kPmax) ? -kPmax : xk);
    const double Pkm = (xkm > kPmax) ? kPmax : ((xkm < -kPmax) ? -kPmax : xkm);
    double oRkp, xRkp, oLk, xLk, oRk, xRk, oLkm, xLkm;
    double doRkp, dxRkp, doLk, dxLk, doRk, dxRk, doLkm, dxLkm;
    omphi_d(false, Pkp, oRkp, xRkp, doRkp, dxRkp);
    omphi_d(true, Pk, oLk, xLk, doLk, dxLk);
    omphi_d(false, Pk, oRk, xRk, doRk, dxRk);
    omphi_d(true, Pkm, oLkm, xLkm, doLkm, dxLkm);
    const double D1 = xLk * qk - xRkp * qkp, D2 = xLkm * qkm - xRk * qk;
    const double den_kp = c.inv_dnu_k * (1.0 / D1), den_km = c.inv_dnu_k * (1.0 / D2);
    // F_{k+1} = oR'_{k+1} den_kp, F_{k-1} = oL_{k-1} den_km, F_k = oL_k den_kp + oR_k den_km
    const double loRkp = lFkp * den_kp, loLkm = lFkm * den_km;
    const double loLk = lFk * den_kp, loRk = lFk * den_km;
    const double lden_kp = lFkp * oRkp + lFk * oLk, lden_km = lFkm * oLkm + lFk * oRk;
    const double lD1 = -lden_kp * den_kp / D1, lD2 = -lden_km * den_km / D2;
    const double lxLk = lD1 * qk, lxRkp = -lD1 * qkp, lxLkm = lD2 * qkm, lxRk = -lD2 * qk;
    double lqk = lD1 * xLk - lD2 * xRk, lqkp = -lD1 * xRkp, lqkm = lD2 * xLkm;
    const double lPkp = loRkp * doRkp + lxRkp * dxRkp;
    const double lPk = loLk * doLk + lxLk * dxLk + loRk * doRk + lxRk * dxRk;
    const double lPkm = loLkm * doLkm + lxLkm * dxLkm;
    // P = clamp(C q)
    const double lCkp = pkp ? lPkp * qkp : 0.0, lCk = pk ? lPk * qk : 0.0, lCkm = pkm ? lPkm * qkm : 0.0;
    lqkp += pkp ? lPkp * Ckp : 0.0;
    lqk += pk ? lPk * Ck : 0.0;
    lqkm += pkm ? lPkm * Ckm : 0.0;
    // q = dk / K, C = VEL νC (+ K (RECOIL_0 - 2/νC))
    double lKkp = -lqkp * qkp / Kkp, lKk = -lqk * qk / Kk, lKkm = -lqkm * qkm / Kkm;
    double lR0 = 0.0;
    if (c.recoil) {
        lKkp += lCkp * (RECOIL_0 - c.two_inv_kp);
        lKk += lCk * (RECOIL_0 - c.two_inv_k);
        lKkm += lCkm * (RECOIL_0 - c.two_inv_km);
        lR0 = lCkp * Kkp + lCk * Kk + lCkm * Kkm;
    }
    g[0] = lKkm * PREFACT;
    g[1] = lKk * PREFACT;
    g[2] = lKkp * PREFACT;
    g[3] = lKkp * Hvp + lKk * Hv0 + lKkm * Hvm;
    g[4] = lR0;
    g[5] = lCkp * c.nuCkp + lCk * c.nuCk + lCkm * c.nuCkm;
}

}  // namespace m1a
