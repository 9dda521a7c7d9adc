// m1lya_device.h -- device math of the FP64 M1 Lya step (gfx950).
//
// Every routine evaluates its formula in the operation order the reference
// writes it, and the library is compiled with -ffp-contract=off, so the FP64
// step reproduces the reference arithmetic rounding-for-rounding (division
// and sqrt are the IEEE correctly rounded sequences on CDNA).  Reference
// citations are file:line into olofnebrin/LYDION.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace m1 {

// ---- constants: src/ConstantsAndFunctions.jl:45-170, src/M1Closure.jl:36-41 ----
constexpr double kC = 2.99792458e10;
constexpr double kInvC = 1.0 / kC;
constexpr double kHoverKB = 6.626069573e-27 / 1.380648813e-16;
constexpr double kPi = 3.141592653589793;
constexpr double kFourPi = 4.0 * kPi;
constexpr double kOneThird = 1.0 / 3.0;
constexpr double kTwoThirds = 2.0 / 3.0;
constexpr double kJeps = 1e-60;
constexpr double kHeps = 1e-100;
constexpr double kFmin = 1e-10;
constexpr double kFmax = 1.0 - 1e-10;
constexpr double kPmin = 1e-3;   // src/FreqFactors.jl:16
constexpr double kPmax = 40.0;   // src/FreqFactors.jl:17
constexpr double kEpsActive = 1e-10;  // src/UpdateLyaJandH.jl:43

enum { BC_REFLECT = 0, BC_OUTFLOW = 1 };

__device__ __forceinline__ double dmax(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dclamp(double x, double lo, double hi)
{
    return (x > hi) ? hi : ((x < lo) ? lo : x);
}

// exp: Cody-Waite reduction by ln2 plus a degree-13 Taylor polynomial in
// Horner form, then an exact ldexp.  Only +, *, floor and ldexp are used,
// so the result is the same on every IEEE double machine that performs the
// operations as written; it is within 2 ulp of the correctly rounded exp.
__device__ inline double exp_det(double x)
{
    if (x > 709.0) return __longlong_as_double(0x7ff0000000000000LL);
    if (x < -745.0) return 0.0;
    const double inv_ln2 = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double n = floor(x * inv_ln2 + 0.5);
    double r = (x - n * ln2_hi) - n * ln2_lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)n);
}

// MinMod limiter: src/ConstantsAndFunctions.jl:1571-1582
__device__ __forceinline__ double minmod(double A, double B)
{
    if (A * B <= 0.0) return 0.0;
    return (fabs(A) < fabs(B)) ? A : B;
}

// Levermore (1984) closure factor: src/M1Closure.jl:28-33
__device__ __forceinline__ double chi_L84(double f)
{
    double f2 = f * f;
    return (3.0 + 4.0 * f2) / (5.0 + 2.0 * sqrt(4.0 - 3.0 * f2));
}

struct Edd { double rr, zz, rz; };

// Compute_D_Edd: src/M1Closure.jl:46-83
__device__ inline Edd D_edd(double H_r, double H_z, double J)
{
    const double Heps2 = kHeps * kHeps;
    const double fmin2 = kFmin * kFmin;
    double Hmag2 = H_r * H_r + H_z * H_z;
    Edd D;
    if (J <= kJeps || Hmag2 <= Heps2 || Hmag2 <= fmin2 * J * J) {
        D.rr = kOneThird; D.zz = kOneThird; D.rz = 0.0;
        return D;
    }
    double Hmag = sqrt(Hmag2);
    double f = dclamp(Hmag / J, kFmin, kFmax);
    double chi = chi_L84(f);
    double inv_Hmag = 1.0 / Hmag;
    double a = (1.0 - chi) * 0.5;
    double b = (3.0 * chi - 1.0) * 0.5;
    double nr = H_r * inv_Hmag;
    double nz = H_z * inv_Hmag;
    double b_nr = b * nr;
    D.rr = a + b_nr * nr;
    D.zz = a + b * nz * nz;
    D.rz = b_nr * nz;
    return D;
}

// Compute_D_Edd_from_f_n: src/M1Closure.jl:91-119
__device__ inline Edd D_edd_f_n(double f, double nr, double nz)
{
    Edd D;
    if (f > kFmin) {
        f = dmin(f, kFmax);
        double chi = chi_L84(f);
        double a = (1.0 - chi) * 0.5;
        double b = (3.0 * chi - 1.0) * 0.5;
        double b_nr = b * nr;
        D.rr = a + b_nr * nr;
        D.zz = a + b * nz * nz;
        D.rz = b_nr * nz;
    } else {
        D.rr = kOneThird; D.zz = kOneThird; D.rz = 0.0;
    }
    return D;
}

// HLL_Eigenvalues: src/M1Closure.jl:273-296 -> {λ1, λ3}
__device__ inline void hll_eigen(double f, double mu, double& l1, double& l3)
{
    f = dclamp(f, kFmin, kFmax);
    mu = dclamp(mu, -1.0, 1.0);
    double f2 = f * f, mu2 = mu * mu, muf = mu * f;
    double SQ = sqrt(dmax(4.0 - 3.0 * f2, 0.0));
    double inv_SQ = 1.0 / SQ;
    double FAC = sqrt(dmax(kTwoThirds * (4.0 - 3.0 * f2 - SQ) + 2.0 * mu2 * (2.0 - f2 - SQ), 0.0));
    l1 = dclamp((muf - FAC) * inv_SQ, -1.0, 1.0);
    l3 = dclamp((muf + FAC) * inv_SQ, -1.0, 1.0);
}

// State of the four cells of one interface stencil, in the reference's
// argument order: (i, i+1, i+2, i-1) where i is the cell left of the face.
struct Stencil {
    double J[4], Hr[4], Hz[4];
};

struct Recon {
    double J_L, J_R, Hr_L, Hr_R, Hz_L, Hz_R, tau_L, tau_R, f_L, f_R, nr_L, nr_R, nz_L, nz_R;
};

// ReconstructInterface: src/AdvectJandH.jl:359-473.
// iv = {ΔL, ΔR, 1/(x_i-x_{i-1}), 1/(x_{i+1}-x_i), 1/(x_{i+2}-x_{i+1})}
__device__ inline Recon reconstruct(const Stencil& s, const double* iv, double aF_i, double aF_ip1)
{
    const double dL = iv[0], dR = iv[1], inv_i_im1 = iv[2], inv_ip1_i = iv[3], inv_ip2_ip1 = iv[4];
    Recon o;
    o.tau_L = aF_i * dL;
    o.tau_R = aF_ip1 * dR;
    double inv_J_i = 1.0 / dmax(s.J[0], kJeps);
    double inv_J_ip1 = 1.0 / dmax(s.J[1], kJeps);
    double inv_J_ip2 = 1.0 / dmax(s.J[2], kJeps);
    double inv_J_im1 = 1.0 / dmax(s.J[3], kJeps);
    double fr_i = s.Hr[0] * inv_J_i, fr_ip1 = s.Hr[1] * inv_J_ip1;
    double fr_ip2 = s.Hr[2] * inv_J_ip2, fr_im1 = s.Hr[3] * inv_J_im1;
    double fz_i = s.Hz[0] * inv_J_i, fz_ip1 = s.Hz[1] * inv_J_ip1;
    double fz_ip2 = s.Hz[2] * inv_J_ip2, fz_im1 = s.Hz[3] * inv_J_im1;

    double dJc = (s.J[1] - s.J[0]) * inv_ip1_i;
    double dJL = minmod((s.J[0] - s.J[3]) * inv_i_im1, dJc);
    double dJR = minmod((s.J[2] - s.J[1]) * inv_ip2_ip1, dJc);
    double dfrc = (fr_ip1 - fr_i) * inv_ip1_i;
    double dfrL = minmod((fr_i - fr_im1) * inv_i_im1, dfrc);
    double dfrR = minmod((fr_ip2 - fr_ip1) * inv_ip2_ip1, dfrc);
    double dfzc = (fz_ip1 - fz_i) * inv_ip1_i;
    double dfzL = minmod((fz_i - fz_im1) * inv_i_im1, dfzc);
    double dfzR = minmod((fz_ip2 - fz_ip1) * inv_ip2_ip1, dfzc);

    double J_L = dmax(s.J[0] + dJL * dL, kJeps);
    double J_R = dmax(s.J[1] - dJR * dR, kJeps);
    double fr_L = fr_i + dfrL * dL;
    double fr_R = fr_ip1 - dfrR * dR;
    double fz_L = fz_i + dfzL * dL;
    double fz_R = fz_ip1 - dfzR * dR;

    double f_L = sqrt(fr_L * fr_L + fz_L * fz_L);
    if (f_L > 1.0) {
        double inv_f = 1.0 / f_L;
        fr_L *= inv_f; fz_L *= inv_f; f_L = 1.0;
    }
    double f_R = sqrt(fr_R * fr_R + fz_R * fz_R);
    if (f_R > 1.0) {
        double inv_f = 1.0 / f_R;
        fr_R *= inv_f; fz_R *= inv_f; f_R = 1.0;
    }
    o.nr_L = 0.0; o.nz_L = 0.0; o.nr_R = 0.0; o.nz_R = 0.0;
    if (f_L > 0.0) { double inv = 1.0 / f_L; o.nr_L = fr_L * inv; o.nz_L = fz_L * inv; }
    if (f_R > 0.0) { double inv = 1.0 / f_R; o.nr_R = fr_R * inv; o.nz_R = fz_R * inv; }
    o.J_L = J_L; o.J_R = J_R;
    o.Hr_L = fr_L * J_L; o.Hr_R = fr_R * J_R;
    o.Hz_L = fz_L * J_L; o.Hz_R = fz_R * J_R;
    o.f_L = f_L; o.f_R = f_R;
    return o;
}

struct Flux3 { double J, Hr, Hz; };

// GetFluxGLF: src/AdvectJandH.jl:81-118 (F* are the physical fluxes at the two sides)
__device__ inline Flux3 flux_glf(double FJ_L, double FJ_R, double FHr_L, double FHr_R,
                                 double FHz_L, double FHz_R, const Recon& q, double J_i, double J_ip1)
{
    double tau_cell = q.tau_L + q.tau_R;
    double num = 0.2 + tau_cell;
    double eps = num / (num + tau_cell * tau_cell);
    double ome = 1.0 - eps;
    double tau_eff = dmax(tau_cell, 1e-300);
    double F_J_diff = -kOneThird * (J_ip1 - J_i) / tau_eff;
    double F_J_GLF = 0.5 * (FJ_L + FJ_R) - 0.5 * (q.J_R - q.J_L);
    Flux3 F;
    F.J = eps * F_J_GLF + ome * F_J_diff;
    double half_eps = 0.5 * eps;
    F.Hr = 0.5 * (FHr_L + FHr_R) - half_eps * (q.Hr_R - q.Hr_L);
    F.Hz = 0.5 * (FHz_L + FHz_R) - half_eps * (q.Hz_R - q.Hz_L);
    return F;
}

// GetFluxHLL: src/AdvectJandH.jl:155-204
__device__ inline Flux3 flux_hll(double FJ_L, double FJ_R, double FHr_L, double FHr_R,
                                 double FHz_L, double FHz_R, const Recon& q, double J_i, double J_ip1,
                                 double l1L, double l1R, double l3L, double l3R)
{
    double tau_cell = q.tau_L + q.tau_R;
    double num = 0.2 + tau_cell;
    double eps = num / (num + tau_cell * tau_cell);
    double ome = 1.0 - eps;
    double half_ome = 0.5 * ome;
    double tau_eff = dmax(tau_cell, 1e-300);
    double F_J_diff = -kOneThird * (J_ip1 - J_i) / tau_eff;
    double S_R = dmax(dmax(l3L, l3R), 0.0);
    double S_L = dmin(dmin(l1L, l1R), 0.0);
    double inv_dS = 1.0 / dmax(S_R - S_L, 1e-300);
    double SRSL = S_R * S_L;
    double FJ = (S_R * FJ_L - S_L * FJ_R + SRSL * (q.J_R - q.J_L)) * inv_dS;
    double FHr = (S_R * FHr_L - S_L * FHr_R + SRSL * (q.Hr_R - q.Hr_L)) * inv_dS;
    double FHz = (S_R * FHz_L - S_L * FHz_R + SRSL * (q.Hz_R - q.Hz_L)) * inv_dS;
    Flux3 F;
    F.J = eps * FJ + ome * F_J_diff;
    F.Hr = eps * FHr + half_ome * (FHr_L + FHr_R);
    F.Hz = eps * FHz + half_ome * (FHz_L + FHz_R);
    return F;
}

// Compute_K_ϕϕ_InnerAxis: src/AdvectJandH.jl:499-522
__device__ inline double kphi_inner_axis(double J_axis, double Hz_axis)
{
    J_axis = dmax(J_axis, kJeps);
    double fz = Hz_axis / J_axis;
    fz = dclamp(fz, -1.0, 1.0);
    double f = fabs(fz);
    double nz = 0.0;
    if (f > 0.0) nz = (fz > 0.0) ? 1.0 : ((fz < 0.0) ? -1.0 : fz);
    Edd D = D_edd_f_n(f, 0.0, nz);
    return dclamp(1.0 - D.rr - D.zz, 0.0, 1.0) * J_axis;
}

// One interface of the M1 sweep: reconstruction -> closure -> K -> numerical
// flux -> upwind advection flux (src/AdvectJandH.jl:682-716 and its copies for
// the other interfaces and the Z direction, :1021-1280).  dir 0 = r, 1 = Z.
struct FaceResult {
    Flux3 F;
    double J_L, J_R, Hz_R;
    double kphi_L, kphi_R;   // clamp(1 - D_rr - D_ZZ, 0, 1) * J at the two sides (r only)
};

__device__ inline FaceResult interface_flux(int dir, int hll, const Stencil& s, const double* iv,
                                            double aF_L, double aF_R, double u_over_c)
{
    Recon q = reconstruct(s, iv, aF_L, aF_R);
    Edd DL = D_edd_f_n(q.f_L, q.nr_L, q.nz_L);
    Edd DR = D_edd_f_n(q.f_R, q.nr_R, q.nz_R);
    double FJ_L, FJ_R, FHr_L, FHr_R, FHz_L, FHz_R;
    if (dir == 0) {
        FJ_L = q.Hr_L; FJ_R = q.Hr_R;
        FHr_L = DL.rr * q.J_L; FHr_R = DR.rr * q.J_R;
        FHz_L = DL.rz * q.J_L; FHz_R = DR.rz * q.J_R;
    } else {
        FJ_L = q.Hz_L; FJ_R = q.Hz_R;
        FHr_L = DL.rz * q.J_L; FHr_R = DR.rz * q.J_R;
        FHz_L = DL.zz * q.J_L; FHz_R = DR.zz * q.J_R;
    }
    Flux3 F;
    if (hll) {
        double l1L, l3L, l1R, l3R;
        hll_eigen(q.f_L, dir == 0 ? q.nr_L : q.nz_L, l1L, l3L);
        hll_eigen(q.f_R, dir == 0 ? q.nr_R : q.nz_R, l1R, l3R);
        F = flux_hll(FJ_L, FJ_R, FHr_L, FHr_R, FHz_L, FHz_R, q, s.J[0], s.J[1], l1L, l1R, l3L, l3R);
    } else {
        F = flux_glf(FJ_L, FJ_R, FHr_L, FHr_R, FHz_L, FHz_R, q, s.J[0], s.J[1]);
    }
    // GetAdvectionFlux: src/AdvectJandH.jl:276-282
    bool up = (u_over_c < 0.0);
    FaceResult r;
    r.F.J = F.J + u_over_c * (up ? q.J_R : q.J_L);
    r.F.Hr = F.Hr + u_over_c * (up ? q.Hr_R : q.Hr_L);
    r.F.Hz = F.Hz + u_over_c * (up ? q.Hz_R : q.Hz_L);
    r.J_L = q.J_L; r.J_R = q.J_R; r.Hz_R = q.Hz_R;
    r.kphi_L = dclamp(1.0 - DL.rr - DL.zz, 0.0, 1.0) * q.J_L;
    r.kphi_R = dclamp(1.0 - DR.rr - DR.zz, 0.0, 1.0) * q.J_R;
    return r;
}

// D_colon_∇u_Cylindrical: src/ConstantsAndFunctions.jl:883-895
__device__ __forceinline__ double D_colon_grad_u(const Edd& D, double durdr, double durdz,
                                                 double duzdr, double duzdz, double urr)
{
    double Dpp = dmax(1.0 - D.rr - D.zz, 0.0);
    return D.rr * durdr + D.rz * (durdz + duzdr) + D.zz * duzdz + Dpp * urr;
}

// OneMinusϕ_L_and_χ_L / OneMinusϕ_R_and_χ_R: src/FreqFactors.jl:54-133
__device__ inline void one_minus_phi_L(double P, double& o, double& chi)
{
    if (P >= kPmax) { o = exp_det(-kPmax); chi = (1.0 - o) / kPmax; return; }
    if (P <= -kPmax) { o = exp_det(kPmax); chi = (1.0 - o) / (-kPmax); return; }
    if (fabs(P) < kPmin) { o = 1.0 - P + 0.5 * (P * P); chi = 1.0 - 0.5 * P; return; }
    o = exp_det(-P);
    chi = (1.0 - o) / P;
}
__device__ inline void one_minus_phi_R(double P, double& o, double& chi)
{
    if (P >= kPmax) { o = exp_det(kPmax); chi = (1.0 - o) / kPmax; return; }
    if (P <= -kPmax) { o = exp_det(-kPmax); chi = (1.0 - o) / (-kPmax); return; }
    if (fabs(P) < kPmin) { o = 1.0 + P + 0.5 * (P * P); chi = -1.0 - 0.5 * P; return; }
    o = exp_det(P);
    chi = (1.0 - o) / P;
}

}  // namespace m1
