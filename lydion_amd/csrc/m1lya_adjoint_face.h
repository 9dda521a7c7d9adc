// m1lya_adjoint_face.h -- reverse mode of one interface of the FP64 advection.
//
// face_vjp evaluates the interface exactly as m1a::tface<double> does (same
// operations, same branches, same tie rules), then runs the chain rule backwards:
// given the cotangents of the interface outputs (F_J, F_Hr, F_HZ, K_ϕϕ on the two
// sides) it returns the cotangents of its inputs, the four stencil cells' J, H_r,
// H_Z and the two α_F values.  At a kink (limiter, max/min, clamp) it takes the
// derivative of the branch the value selects, as the dual-number evaluation does,
// so the two agree to rounding.  Forward formulas: src/AdvectJandH.jl:81-118,
// 155-204, 276-282, 359-473, 499-522; src/M1Closure.jl:28-33, 91-119, 273-296.
#pragma once
#include <hip/hip_runtime.h>

#include "m1lya_adjoint.h"

namespace m1a {

struct FaceGrad {
    double J[4], Hr[4], Hz[4];   // stencil slots (cells i, i+1, i+2, i-1 of the interface)
    double aF_L, aF_R;
};

namespace facevjp {

// max(x, lo) as the forward's tmax(x, T(lo)): x passes when !(lo > x)
__device__ __forceinline__ bool pass_max(double x, double lo) { return !(lo > x); }

// forward minmod (tminmod): returns A, B or 0; which = 0 (A), 1 (B), 2 (zero)
__device__ __forceinline__ double minmod_f(double A, double B, int& which)
{
    if (A * B <= 0.0) { which = 2; return 0.0; }
    if (fabs(A) < fabs(B)) { which = 0; return A; }
    which = 1;
    return B;
}

// Compute_D_Edd_from_f_n forward, with the intermediates its reverse needs
struct EddF {
    bool on, pass_f;
    double fc, SQ, den, chi, b, nr, nz;
    double rr, zz, rz;
};

__device__ __forceinline__ EddF edd_f_n(double f, double nr, double nz)
{
    EddF e;
    e.on = f > kFmin;
    e.nr = nr; e.nz = nz;
    if (!e.on) {
        e.pass_f = false; e.fc = 0.0; e.SQ = 0.0; e.den = 1.0; e.chi = 0.0; e.b = 0.0;
        e.rr = kOneThird; e.zz = kOneThird; e.rz = 0.0;
        return e;
    }
    e.pass_f = !(kFmax < f);                 // tmin(f, kFmax)
    e.fc = e.pass_f ? f : kFmax;
    const double f2 = e.fc * e.fc;
    e.SQ = sqrt(4.0 - 3.0 * f2);
    const double num = 3.0 + 4.0 * f2;
    e.den = 5.0 + 2.0 * e.SQ;
    e.chi = num / e.den;
    const double a = (1.0 - e.chi) * 0.5;
    e.b = (3.0 * e.chi - 1.0) * 0.5;
    const double b_nr = e.b * nr;
    e.rr = a + b_nr * nr;
    e.zz = a + e.b * nz * nz;
    e.rz = b_nr * nz;
    return e;
}

// reverse of edd_f_n: cotangents of (rr, zz, rz) -> of (f, nr, nz), accumulated
__device__ __forceinline__ void edd_f_n_rev(const EddF& e, double lrr, double lzz, double lrz, double& lf,
                                            double& lnr, double& lnz)
{
    if (!e.on) return;
    const double la = lrr + lzz;
    const double lb = lrr * e.nr * e.nr + lzz * e.nz * e.nz + lrz * e.nr * e.nz;
    lnr += lrr * 2.0 * e.b * e.nr + lrz * e.b * e.nz;
    lnz += lzz * 2.0 * e.b * e.nz + lrz * e.b * e.nr;
    const double lchi = -0.5 * la + 1.5 * lb;
    const double lnum = lchi / e.den;
    const double lden = -lchi * e.chi / e.den;
    const double lSQ = 2.0 * lden;
    const double lt = e.SQ > 0.0 ? 0.5 * lSQ / e.SQ : 0.0;   // t = 4 - 3 f²
    const double lf2 = 4.0 * lnum - 3.0 * lt;
    if (e.pass_f) lf += 2.0 * e.fc * lf2;
}

// tclamp(x, lo, hi) passes x when neither bound is taken
__device__ __forceinline__ bool pass_clamp(double x, double lo, double hi) { return !(x > hi) && !(x < lo); }
__device__ __forceinline__ double clamp_f(double x, double lo, double hi) { return (x > hi) ? hi : ((x < lo) ? lo : x); }

// HLL_Eigenvalues forward and reverse
struct EigF {
    bool pf, pmu, pq, pr, p1, p3;
    double f, mu, f2, mu2, SQ, inv_SQ, Q, FAC, l1, l3;
};

__device__ __forceinline__ EigF eig_f(double f0, double mu0)
{
    EigF e;
    e.pf = pass_clamp(f0, kFmin, kFmax);
    e.f = clamp_f(f0, kFmin, kFmax);
    e.pmu = pass_clamp(mu0, -1.0, 1.0);
    e.mu = clamp_f(mu0, -1.0, 1.0);
    e.f2 = e.f * e.f;
    e.mu2 = e.mu * e.mu;
    const double muf = e.mu * e.f;
    const double q = 4.0 - 3.0 * e.f2;
    e.pq = !(0.0 > q);
    e.SQ = sqrt(e.pq ? q : 0.0);
    e.inv_SQ = 1.0 / e.SQ;
    const double r = kTwoThirds * (4.0 - 3.0 * e.f2 - e.SQ) + 2.0 * e.mu2 * (2.0 - e.f2 - e.SQ);
    e.pr = !(0.0 > r);
    e.Q = e.pr ? r : 0.0;
    e.FAC = sqrt(e.Q);
    const double x1 = (muf - e.FAC) * e.inv_SQ, x3 = (muf + e.FAC) * e.inv_SQ;
    e.p1 = pass_clamp(x1, -1.0, 1.0);
    e.p3 = pass_clamp(x3, -1.0, 1.0);
    e.l1 = clamp_f(x1, -1.0, 1.0);
    e.l3 = clamp_f(x3, -1.0, 1.0);
    return e;
}

__device__ __forceinline__ void eig_rev(const EigF& e, double ll1, double ll3, double& lf0, double& lmu0)
{
    const double lx1 = e.p1 ? ll1 : 0.0, lx3 = e.p3 ? ll3 : 0.0;
    const double muf = e.mu * e.f;
    // x = (muf ∓ FAC) inv_SQ
    const double lmuf = (lx1 + lx3) * e.inv_SQ;
    const double lFAC = (-lx1 + lx3) * e.inv_SQ;
    const double linv = lx1 * (muf - e.FAC) + lx3 * (muf + e.FAC);
    double lSQ = -linv * e.inv_SQ * e.inv_SQ;
    const double lQ = e.FAC > 0.0 ? 0.5 * lFAC / e.FAC : 0.0;
    const double lr = e.pr ? lQ : 0.0;
    // r = (2/3)(4 - 3 f2 - SQ) + 2 mu2 (2 - f2 - SQ)
    double lf2 = lr * (-3.0 * kTwoThirds - 2.0 * e.mu2);
    const double lmu2 = lr * 2.0 * (2.0 - e.f2 - e.SQ);
    lSQ += lr * (-kTwoThirds - 2.0 * e.mu2);
    const double lqv = e.SQ > 0.0 ? 0.5 * lSQ / e.SQ : 0.0;
    const double lq = e.pq ? lqv : 0.0;
    lf2 += -3.0 * lq;
    const double lf = 2.0 * e.f * lf2 + lmuf * e.mu;
    const double lmu = 2.0 * e.mu * lmu2 + lmuf * e.f;
    if (e.pf) lf0 += lf;
    if (e.pmu) lmu0 += lmu;
}

}  // namespace facevjp

// One interface and its vector-Jacobian product.  lFJ, lFHr, lFHz: cotangents of the
// three fluxes; lkL, lkR: of the two K_ϕϕ outputs (tface's kphi_L, kphi_R).
__device__ inline FaceGrad face_vjp(int dir, int hll, bool axis, const double (&J)[4], const double (&Hr)[4],
                                    const double (&Hz)[4], const double* iv, double aF_L, double aF_R,
                                    double u_over_c, double lFJ, double lFHr, double lFHz, double lkL, double lkR)
{
    using namespace facevjp;
    FaceGrad g;
#pragma unroll
    for (int m = 0; m < 4; ++m) { g.J[m] = 0.0; g.Hr[m] = 0.0; g.Hz[m] = 0.0; }

    // ================= forward (treconstruct) =================
    const double dL = iv[0], dR = iv[1], c01 = iv[2], c12 = iv[3], c23 = iv[4];
    const double tau_L = aF_L * dL, tau_R = aF_R * dR;
    bool pJ[4];
    double iJ[4], fr[4], fz[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        pJ[m] = pass_max(J[m], kJeps);
        iJ[m] = 1.0 / (pJ[m] ? J[m] : kJeps);
        fr[m] = Hr[m] * iJ[m];
        fz[m] = Hz[m] * iJ[m];
    }
    const double dJc = (J[1] - J[0]) * c12;
    int wJL, wJR, wrL, wrR, wzL, wzR;
    const double dJL = minmod_f((J[0] - J[3]) * c01, dJc, wJL);
    const double dJR = minmod_f((J[2] - J[1]) * c23, dJc, wJR);
    const double dfrc = (fr[1] - fr[0]) * c12;
    const double dfrL = minmod_f((fr[0] - fr[3]) * c01, dfrc, wrL);
    const double dfrR = minmod_f((fr[2] - fr[1]) * c23, dfrc, wrR);
    const double dfzc = (fz[1] - fz[0]) * c12;
    const double dfzL = minmod_f((fz[0] - fz[3]) * c01, dfzc, wzL);
    const double dfzR = minmod_f((fz[2] - fz[1]) * c23, dfzc, wzR);

    const double JL0 = J[0] + dJL * dL, JR0 = J[1] - dJR * dR;
    const bool pJL = pass_max(JL0, kJeps), pJR = pass_max(JR0, kJeps);
    const double J_L = pJL ? JL0 : kJeps, J_R = pJR ? JR0 : kJeps;
    const double frL0 = fr[0] + dfrL * dL, frR0 = fr[1] - dfrR * dR;
    const double fzL0 = fz[0] + dfzL * dL, fzR0 = fz[1] - dfzR * dR;
    const double fL0 = sqrt(frL0 * frL0 + fzL0 * fzL0);
    const double fR0 = sqrt(frR0 * frR0 + fzR0 * fzR0);
    const bool nL = fL0 > 1.0, nR = fR0 > 1.0;   // limited to |f| = 1
    const double ifL0 = nL ? 1.0 / fL0 : 1.0, ifR0 = nR ? 1.0 / fR0 : 1.0;
    const double frL = nL ? frL0 * ifL0 : frL0, fzL = nL ? fzL0 * ifL0 : fzL0;
    const double frR = nR ? frR0 * ifR0 : frR0, fzR = nR ? fzR0 * ifR0 : fzR0;
    const double f_L = nL ? 1.0 : fL0, f_R = nR ? 1.0 : fR0;
    const bool dnL = f_L > 0.0, dnR = f_R > 0.0;
    const double invnL = dnL ? 1.0 / f_L : 0.0, invnR = dnR ? 1.0 / f_R : 0.0;
    const double nr_L = dnL ? frL * invnL : 0.0, nz_L = dnL ? fzL * invnL : 0.0;
    const double nr_R = dnR ? frR * invnR : 0.0, nz_R = dnR ? fzR * invnR : 0.0;
    const double HrL = frL * J_L, HrR = frR * J_R, HzL = fzL * J_L, HzR = fzR * J_R;

    // ================= forward (closure, flux) =================
    const EddF DL = edd_f_n(f_L, nr_L, nz_L), DR = edd_f_n(f_R, nr_R, nz_R);
    double FJ_L, FJ_R, FHr_L, FHr_R, FHz_L, FHz_R;
    if (dir == 0) {
        FJ_L = HrL; FJ_R = HrR;
        FHr_L = DL.rr * J_L; FHr_R = DR.rr * J_R;
        FHz_L = DL.rz * J_L; FHz_R = DR.rz * J_R;
    } else {
        FJ_L = HzL; FJ_R = HzR;
        FHr_L = DL.rz * J_L; FHr_R = DR.rz * J_R;
        FHz_L = DL.zz * J_L; FHz_R = DR.zz * J_R;
    }
    const double tau = tau_L + tau_R;
    const double num = 0.2 + tau;
    const double dnm = num + tau * tau;
    const double eps = num / dnm;
    const bool ptau = pass_max(tau, 1e-300);
    const double tau_eff = ptau ? tau : 1e-300;
    const double FJd = -kOneThird * (J[1] - J[0]) / tau_eff;

    // ================= reverse =================
    // upwind term: F += u (up ? right : left)
    const bool up = (u_over_c < 0.0);
    double lJ_L = 0.0, lJ_R = 0.0, lHrL = 0.0, lHrR = 0.0, lHzL = 0.0, lHzR = 0.0;
    if (up) { lJ_R += u_over_c * lFJ; lHrR += u_over_c * lFHr; lHzR += u_over_c * lFHz; }
    else { lJ_L += u_over_c * lFJ; lHrL += u_over_c * lFHr; lHzL += u_over_c * lFHz; }
    double lFJ_L = 0.0, lFJ_R = 0.0, lFHr_L = 0.0, lFHr_R = 0.0, lFHz_L = 0.0, lFHz_R = 0.0;
    double leps = 0.0, lFJd = 0.0;
    double lfL = 0.0, lfR = 0.0, lnrL = 0.0, lnrR = 0.0, lnzL = 0.0, lnzR = 0.0;
    const double ome = 1.0 - eps;
    if (hll) {
        const EigF EL = eig_f(f_L, dir == 0 ? nr_L : nz_L), ER = eig_f(f_R, dir == 0 ? nr_R : nz_R);
        // S_R = max(max(l3L, l3R), 0), S_L = min(min(l1L, l1R), 0) with the forward's tie rules
        const bool r3R = EL.l3 < ER.l3;                    // tmax(l3L, l3R) picks l3R
        const double m3 = r3R ? ER.l3 : EL.l3;
        const bool pSR = !(0.0 > m3);
        const double S_R = pSR ? m3 : 0.0;
        const bool r1R = ER.l1 < EL.l1;                    // tmin(l1L, l1R) picks l1R
        const double m1 = r1R ? ER.l1 : EL.l1;
        const bool pSL = !(0.0 < m1);
        const double S_L = pSL ? m1 : 0.0;
        const double dS0 = S_R - S_L;
        const bool pdS = pass_max(dS0, 1e-300);
        const double inv_dS = 1.0 / (pdS ? dS0 : 1e-300);
        const double SRSL = S_R * S_L;
        const double NJ = S_R * FJ_L - S_L * FJ_R + SRSL * (J_R - J_L);
        const double NHr = S_R * FHr_L - S_L * FHr_R + SRSL * (HrR - HrL);
        const double NHz = S_R * FHz_L - S_L * FHz_R + SRSL * (HzR - HzL);
        // FJ = eps NJ inv + ome FJd; FHr = eps NHr inv + (ome/2)(FHr_L + FHr_R); FHz likewise
        const double lNJ = lFJ * eps * inv_dS, lNHr = lFHr * eps * inv_dS, lNHz = lFHz * eps * inv_dS;
        leps += lFJ * NJ * inv_dS + lFHr * NHr * inv_dS + lFHz * NHz * inv_dS;
        double lome = lFJ * FJd + 0.5 * (lFHr * (FHr_L + FHr_R) + lFHz * (FHz_L + FHz_R));
        lFJd += lFJ * ome;
        lFHr_L += 0.5 * ome * lFHr; lFHr_R += 0.5 * ome * lFHr;
        lFHz_L += 0.5 * ome * lFHz; lFHz_R += 0.5 * ome * lFHz;
        leps -= lome;
        const double linv = lFJ * eps * NJ + lFHr * eps * NHr + lFHz * eps * NHz;
        const double ldS = pdS ? -linv * inv_dS * inv_dS : 0.0;
        double lS_R = ldS, lS_L = -ldS;
        // N = S_R F_L - S_L F_R + S_R S_L (U_R - U_L)
        lS_R += lNJ * (FJ_L + S_L * (J_R - J_L)) + lNHr * (FHr_L + S_L * (HrR - HrL)) +
                lNHz * (FHz_L + S_L * (HzR - HzL));
        lS_L += lNJ * (-FJ_R + S_R * (J_R - J_L)) + lNHr * (-FHr_R + S_R * (HrR - HrL)) +
                lNHz * (-FHz_R + S_R * (HzR - HzL));
        lFJ_L += lNJ * S_R; lFJ_R -= lNJ * S_L;
        lFHr_L += lNHr * S_R; lFHr_R -= lNHr * S_L;
        lFHz_L += lNHz * S_R; lFHz_R -= lNHz * S_L;
        lJ_R += lNJ * SRSL; lJ_L -= lNJ * SRSL;
        lHrR += lNHr * SRSL; lHrL -= lNHr * SRSL;
        lHzR += lNHz * SRSL; lHzL -= lNHz * SRSL;
        const double lm3 = pSR ? lS_R : 0.0, lm1 = pSL ? lS_L : 0.0;
        const double ll3L = r3R ? 0.0 : lm3, ll3R = r3R ? lm3 : 0.0;
        const double ll1L = r1R ? 0.0 : lm1, ll1R = r1R ? lm1 : 0.0;
        double lmuL = 0.0, lmuR = 0.0;
        eig_rev(EL, ll1L, ll3L, lfL, lmuL);
        eig_rev(ER, ll1R, ll3R, lfR, lmuR);
        if (dir == 0) { lnrL += lmuL; lnrR += lmuR; }
        else { lnzL += lmuL; lnzR += lmuR; }
    } else {
        // FJ = eps (0.5 (FJ_L + FJ_R) - 0.5 (J_R - J_L)) + ome FJd
        const double FJg = 0.5 * (FJ_L + FJ_R) - 0.5 * (J_R - J_L);
        leps += lFJ * FJg - lFJ * FJd;
        lFJd += lFJ * ome;
        lFJ_L += 0.5 * eps * lFJ; lFJ_R += 0.5 * eps * lFJ;
        lJ_R -= 0.5 * eps * lFJ; lJ_L += 0.5 * eps * lFJ;
        // FHr = 0.5 (FHr_L + FHr_R) - (eps/2)(HrR - HrL)
        const double he = 0.5 * eps;
        lFHr_L += 0.5 * lFHr; lFHr_R += 0.5 * lFHr;
        lHrR -= he * lFHr; lHrL += he * lFHr;
        lFHz_L += 0.5 * lFHz; lFHz_R += 0.5 * lFHz;
        lHzR -= he * lFHz; lHzL += he * lFHz;
        leps += -0.5 * (lFHr * (HrR - HrL) + lFHz * (HzR - HzL));
    }
    // eps = num / (num + tau²), FJd = -(1/3)(J1 - J0)/tau_eff
    const double lnum = leps / dnm - leps * eps / dnm;
    const double ltau_q = -leps * eps / dnm * 2.0 * tau;
    double ltau = lnum + ltau_q;
    const double lJdiff = -kOneThird * lFJd / tau_eff;
    g.J[1] += lJdiff;
    g.J[0] -= lJdiff;
    if (ptau) ltau += -lFJd * FJd / tau_eff;
    g.aF_L = ltau * dL;
    g.aF_R = ltau * dR;

    // K_ϕϕ outputs
    {
        const double xL = 1.0 - DL.rr - DL.zz;
        if (pass_clamp(xL, 0.0, 1.0)) {
            lJ_L += lkL * xL;
            const double lx = lkL * J_L;
            edd_f_n_rev(DL, -lx, -lx, 0.0, lfL, lnrL, lnzL);
        } else {
            lJ_L += lkL * clamp_f(xL, 0.0, 1.0);
        }
    }
    if (axis) {
        // Compute_K_ϕϕ_InnerAxis(J_R, HzR)
        const bool pa = pass_max(J_R, kJeps);
        const double Ja = pa ? J_R : kJeps;
        const double fz0 = HzR / Ja;
        const bool pfz = pass_clamp(fz0, -1.0, 1.0);
        const double fzc = clamp_f(fz0, -1.0, 1.0);
        const double f = fzc < 0.0 ? -fzc : fzc;
        double nzc = 0.0;
        if (f > 0.0) nzc = (fzc > 0.0) ? 1.0 : ((fzc < 0.0) ? -1.0 : fzc);
        const EddF DA = edd_f_n(f, 0.0, nzc);
        const double x = 1.0 - DA.rr - DA.zz;
        double lJa = 0.0, lf = 0.0, lnr_d = 0.0, lnz_d = 0.0;
        if (pass_clamp(x, 0.0, 1.0)) {
            lJa += lkR * x;
            const double lx = lkR * Ja;
            edd_f_n_rev(DA, -lx, -lx, 0.0, lf, lnr_d, lnz_d);
        } else {
            lJa += lkR * clamp_f(x, 0.0, 1.0);
        }
        const double lfzc = fzc < 0.0 ? -lf : lf;
        const double lfz0 = pfz ? lfzc : 0.0;
        lHzR += lfz0 / Ja;
        lJa += -lfz0 * fz0 / Ja;
        if (pa) lJ_R += lJa;
    } else {
        const double xR = 1.0 - DR.rr - DR.zz;
        if (pass_clamp(xR, 0.0, 1.0)) {
            lJ_R += lkR * xR;
            const double lx = lkR * J_R;
            edd_f_n_rev(DR, -lx, -lx, 0.0, lfR, lnrR, lnzR);
        } else {
            lJ_R += lkR * clamp_f(xR, 0.0, 1.0);
        }
    }

    // side fluxes
    double lDLrr = 0.0, lDLzz = 0.0, lDLrz = 0.0, lDRrr = 0.0, lDRzz = 0.0, lDRrz = 0.0;
    if (dir == 0) {
        lHrL += lFJ_L; lHrR += lFJ_R;
        lDLrr += lFHr_L * J_L; lJ_L += lFHr_L * DL.rr;
        lDRrr += lFHr_R * J_R; lJ_R += lFHr_R * DR.rr;
        lDLrz += lFHz_L * J_L; lJ_L += lFHz_L * DL.rz;
        lDRrz += lFHz_R * J_R; lJ_R += lFHz_R * DR.rz;
    } else {
        lHzL += lFJ_L; lHzR += lFJ_R;
        lDLrz += lFHr_L * J_L; lJ_L += lFHr_L * DL.rz;
        lDRrz += lFHr_R * J_R; lJ_R += lFHr_R * DR.rz;
        lDLzz += lFHz_L * J_L; lJ_L += lFHz_L * DL.zz;
        lDRzz += lFHz_R * J_R; lJ_R += lFHz_R * DR.zz;
    }
    edd_f_n_rev(DL, lDLrr, lDLzz, lDLrz, lfL, lnrL, lnzL);
    edd_f_n_rev(DR, lDRrr, lDRzz, lDRrz, lfR, lnrR, lnzR);

    // ================= reverse (treconstruct) =================
    // H = fr J on each side
    double lfrL = lHrL * J_L, lfzL = lHzL * J_L, lfrR = lHrR * J_R, lfzR = lHzR * J_R;
    lJ_L += lHrL * frL + lHzL * fzL;
    lJ_R += lHrR * frR + lHzR * fzR;
    // n = fr / f when f > 0
    if (dnL) {
        lfrL += lnrL * invnL; lfzL += lnzL * invnL;
        lfL += -(lnrL * nr_L + lnzL * nz_L) * invnL;
    }
    if (dnR) {
        lfrR += lnrR * invnR; lfzR += lnzR * invnR;
        lfR += -(lnrR * nr_R + lnzR * nz_R) * invnR;
    }
    // limiting to |f| = 1: fr = fr0 / f0, f = 1 (constant); else f = f0
    double lfrL0, lfzL0, lfL0, lfrR0, lfzR0, lfR0;
    if (nL) {
        const double s = lfrL * frL + lfzL * fzL;   // d(fr0/f0) along f0
        lfrL0 = lfrL * ifL0; lfzL0 = lfzL * ifL0;
        lfL0 = -s * ifL0;
    } else {
        lfrL0 = lfrL; lfzL0 = lfzL; lfL0 = lfL;
    }
    if (nR) {
        const double s = lfrR * frR + lfzR * fzR;
        lfrR0 = lfrR * ifR0; lfzR0 = lfzR * ifR0;
        lfR0 = -s * ifR0;
    } else {
        lfrR0 = lfrR; lfzR0 = lfzR; lfR0 = lfR;
    }
    // f0 = sqrt(fr0² + fz0²)
    if (fL0 > 0.0) { lfrL0 += lfL0 * frL0 / fL0; lfzL0 += lfL0 * fzL0 / fL0; }
    if (fR0 > 0.0) { lfrR0 += lfR0 * frR0 / fR0; lfzR0 += lfR0 * fzR0 / fR0; }
    // J_L = max(J0 + dJL dL, Jeps), J_R = max(J1 - dJR dR, Jeps)
    const double lJL0 = pJL ? lJ_L : 0.0, lJR0 = pJR ? lJ_R : 0.0;
    g.J[0] += lJL0;
    g.J[1] += lJR0;
    const double ldJL = lJL0 * dL, ldJR = -lJR0 * dR;
    double lfr[4] = {0.0, 0.0, 0.0, 0.0}, lfz[4] = {0.0, 0.0, 0.0, 0.0};
    lfr[0] += lfrL0; lfr[1] += lfrR0;
    lfz[0] += lfzL0; lfz[1] += lfzR0;
    const double ldfrL = lfrL0 * dL, ldfrR = -lfrR0 * dR;
    const double ldfzL = lfzL0 * dL, ldfzR = -lfzR0 * dR;
    // limited slopes: minmod(a, c) passes its chosen argument
    double ldJc = 0.0, ldfrc = 0.0, ldfzc = 0.0;
    auto mm_rev = [&](int which, double l, double c_slope, double& lc, int mL, int mR, double* x) {
        // a = (x[mL] - x[mR]) c_slope
        if (which == 0) { x[mL] += l * c_slope; x[mR] -= l * c_slope; }
        else if (which == 1) lc += l;
    };
    mm_rev(wJL, ldJL, c01, ldJc, 0, 3, g.J);
    mm_rev(wJR, ldJR, c23, ldJc, 2, 1, g.J);
    mm_rev(wrL, ldfrL, c01, ldfrc, 0, 3, lfr);
    mm_rev(wrR, ldfrR, c23, ldfrc, 2, 1, lfr);
    mm_rev(wzL, ldfzL, c01, ldfzc, 0, 3, lfz);
    mm_rev(wzR, ldfzR, c23, ldfzc, 2, 1, lfz);
    g.J[1] += ldJc * c12; g.J[0] -= ldJc * c12;
    lfr[1] += ldfrc * c12; lfr[0] -= ldfrc * c12;
    lfz[1] += ldfzc * c12; lfz[0] -= ldfzc * c12;
    // fr = Hr / max(J, Jeps)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        g.Hr[m] += lfr[m] * iJ[m];
        g.Hz[m] += lfz[m] * iJ[m];
        if (pJ[m]) g.J[m] += -(lfr[m] * fr[m] + lfz[m] * fz[m]) * iJ[m];
    }
    return g;
}

}  // namespace m1a
