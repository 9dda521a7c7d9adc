/*
 * m1lya_oracle.c -- CPU FP64 restatement of LYDION's multi-frequency M1 Lya step.
 *
 * TEST INFRASTRUCTURE ONLY (see m1lya_oracle.h).  Built with
 *   gcc -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -fopenmp
 * so that every floating-point operation is performed exactly as written,
 * in the order the reference writes it.
 *
 * Reference citations are file:line into olofnebrin/LYDION.
 * Known bit-level gaps to the reference (documented, "pinned to rounding"):
 *   - Julia's exp is replaced by orc_exp (range reduction + degree-13 Taylor,
 *     <= 2 ulp, checked against libm in the tests);
 *   - Julia @simd/@fastmath loops may reassociate reductions; sums here run
 *     in plain sequential order (j outer, i inner, k innermost).
 */
#include "m1lya_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* Constants: src/ConstantsAndFunctions.jl:45-170, src/M1Closure.jl:36-41,  */
/* src/FreqFactors.jl:16-18, src/UpdateLyaJandH.jl:43                      */
/* ---------------------------------------------------------------------- */
#define C_LIGHT   2.99792458e10
#define H_PLANCK  6.626069573e-27
#define K_B       1.380648813e-16
#define AMU       1.660539069e-24
#define M_H       (1.0080 * AMU)
#define M_E       9.109382917e-28
#define E_CHARGE  4.80320451e-10
#define NU_LYA    2.466068e15
#define NU_LYC    3.289842e15
#define A_LYA     6.265e8
#define F_LU_LYA  0.41641

static const double OneThird  = 1.0 / 3.0;
static const double TwoThirds = 2.0 / 3.0;
static const double PI_D      = 3.141592653589793;
#define FOUR_PI   (4.0 * PI_D)
#define INV_C     (1.0 / C_LIGHT)

static const double Jeps  = 1e-60;
static const double Heps  = 1e-100;
static const double fmin_ = 1e-10;
static const double fmax_ = 1.0 - 1e-10;

static const double Pmin = 1e-3;
static const double Pmax = 40.0;
static const double eps_active = 1e-10;

static inline double jmax(double a, double b) { return (b > a) ? b : a; }
static inline double jmin(double a, double b) { return (b < a) ? b : a; }
static inline double jclamp(double x, double lo, double hi)
{
    return (x > hi) ? hi : ((x < lo) ? lo : x);
}

/* ---------------------------------------------------------------------- */
/* exp: Cody-Waite reduction by ln2 and a degree-13 Taylor polynomial in   */
/* Horner form.  Written with plain +,*,floor,ldexp so that the identical  */
/* arithmetic is reproducible on any IEEE double machine.  Arguments on    */
/* the hot path are clamped to |x| <= 40 (src/FreqFactors.jl:252-254).     */
/* ---------------------------------------------------------------------- */
double orc_exp(double x)
{
    if (x > 709.0) return INFINITY;
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

/* MinMod limiter: src/ConstantsAndFunctions.jl:1571-1582 */
double orc_minmod(double A, double B)
{
    if (A * B <= 0.0) return 0.0;
    return (fabs(A) < fabs(B)) ? A : B;
}

/* ---------------------------------------------------------------------- */
/* M1 closure: src/M1Closure.jl:28-119                                     */
/* ---------------------------------------------------------------------- */
double orc_chi_L84(double f)
{
    double f2 = f * f;
    return (3.0 + 4.0 * f2) / (5.0 + 2.0 * sqrt(4.0 - 3.0 * f2));
}

/* Compute_D_Edd: src/M1Closure.jl:46-83.  out = {D_rr, D_ZZ, D_rZ} */
void orc_D_edd(double H_r, double H_z, double J, double out[3])
{
    const double Heps2 = Heps * Heps;
    const double fmin2 = fmin_ * fmin_;
    double Hmag2 = H_r * H_r + H_z * H_z;
    if (J <= Jeps || Hmag2 <= Heps2 || Hmag2 <= fmin2 * J * J) {
        out[0] = OneThird; out[1] = OneThird; out[2] = 0.0;
        return;
    }
    double Hmag = sqrt(Hmag2);
    double f = jclamp(Hmag / J, fmin_, fmax_);
    double chi = orc_chi_L84(f);
    double inv_Hmag = 1.0 / Hmag;
    double a = (1.0 - chi) * 0.5;
    double b = (3.0 * chi - 1.0) * 0.5;
    double nr = H_r * inv_Hmag;
    double nz = H_z * inv_Hmag;
    double b_nr = b * nr;
    out[0] = a + b_nr * nr;
    out[1] = a + b * nz * nz;
    out[2] = b_nr * nz;
}

/* Compute_D_Edd_from_f_n: src/M1Closure.jl:91-119 */
void orc_D_edd_from_f_n(double f, double nr, double nz, double out[3])
{
    if (f > fmin_) {
        f = jmin(f, fmax_);
        double chi = orc_chi_L84(f);
        double a = (1.0 - chi) * 0.5;
        double b = (3.0 * chi - 1.0) * 0.5;
        double b_nr = b * nr;
        out[0] = a + b_nr * nr;
        out[1] = a + b * nz * nz;
        out[2] = b_nr * nz;
    } else {
        out[0] = OneThird; out[1] = OneThird; out[2] = 0.0;
    }
}

/* HLL_Eigenvalues: src/M1Closure.jl:273-296.  out = {λ1, λ3} */
void orc_hll_eigenvalues(double f, double mu, double out[2])
{
    f = jclamp(f, fmin_, fmax_);
    mu = jclamp(mu, -1.0, 1.0);
    double f2 = f * f, mu2 = mu * mu, muf = mu * f;
    double SQRT = sqrt(jmax(4.0 - 3.0 * f2, 0.0));
    double inv_SQRT = 1.0 / SQRT;
    double FAC = sqrt(jmax(TwoThirds * (4.0 - 3.0 * f2 - SQRT) + 2.0 * mu2 * (2.0 - f2 - SQRT), 0.0));
    out[0] = jclamp((muf - FAC) * inv_SQRT, -1.0, 1.0);
    out[1] = jclamp((muf + FAC) * inv_SQRT, -1.0, 1.0);
}

/* GetFluxGLF: src/AdvectJandH.jl:81-118.
 * in = {FJ_L,FJ_R,FHr_L,FHr_R,FHZ_L,FHZ_R,J_L,J_R,Hr_L,Hr_R,HZ_L,HZ_R,τ_L,τ_R,J_i,J_ip1} */
void orc_flux_glf(const double in[16], double out[3])
{
    double FJ_L = in[0], FJ_R = in[1], FHr_L = in[2], FHr_R = in[3], FHZ_L = in[4], FHZ_R = in[5];
    double J_L = in[6], J_R = in[7], Hr_L = in[8], Hr_R = in[9], HZ_L = in[10], HZ_R = in[11];
    double tau_L = in[12], tau_R = in[13], J_i = in[14], J_ip1 = in[15];
    double tau_cell = tau_L + tau_R;
    double num = 0.2 + tau_cell;
    double eps = num / (num + tau_cell * tau_cell);
    double ome = 1.0 - eps;
    double tau_eff = jmax(tau_cell, 1e-300);
    double F_J_diff = -OneThird * (J_ip1 - J_i) / tau_eff;
    double F_J_GLF = 0.5 * (FJ_L + FJ_R) - 0.5 * (J_R - J_L);
    out[0] = eps * F_J_GLF + ome * F_J_diff;
    double half_eps = 0.5 * eps;
    out[1] = 0.5 * (FHr_L + FHr_R) - half_eps * (Hr_R - Hr_L);
    out[2] = 0.5 * (FHZ_L + FHZ_R) - half_eps * (HZ_R - HZ_L);
}

/* GetFluxHLL: src/AdvectJandH.jl:155-204 (extra inputs λ1_L, λ1_R, λ3_L, λ3_R) */
static void flux_hll(const double in[16], double l1L, double l1R, double l3L, double l3R,
                     double out[3])
{
    double FJ_L = in[0], FJ_R = in[1], FHr_L = in[2], FHr_R = in[3], FHZ_L = in[4], FHZ_R = in[5];
    double J_L = in[6], J_R = in[7], Hr_L = in[8], Hr_R = in[9], HZ_L = in[10], HZ_R = in[11];
    double tau_L = in[12], tau_R = in[13], J_i = in[14], J_ip1 = in[15];
    double tau_cell = tau_L + tau_R;
    double num = 0.2 + tau_cell;
    double eps = num / (num + tau_cell * tau_cell);
    double ome = 1.0 - eps;
    double half_ome = 0.5 * ome;
    double tau_eff = jmax(tau_cell, 1e-300);
    double F_J_diff = -OneThird * (J_ip1 - J_i) / tau_eff;
    double S_R = jmax(jmax(l3L, l3R), 0.0);
    double S_L = jmin(jmin(l1L, l1R), 0.0);
    double inv_dS = 1.0 / jmax(S_R - S_L, 1e-300);
    double SRSL = S_R * S_L;
    double FJ  = (S_R * FJ_L  - S_L * FJ_R  + SRSL * (J_R - J_L)) * inv_dS;
    double FHr = (S_R * FHr_L - S_L * FHr_R + SRSL * (Hr_R - Hr_L)) * inv_dS;
    double FHZ = (S_R * FHZ_L - S_L * FHZ_R + SRSL * (HZ_R - HZ_L)) * inv_dS;
    out[0] = eps * FJ + ome * F_J_diff;
    out[1] = eps * FHr + half_ome * (FHr_L + FHr_R);
    out[2] = eps * FHZ + half_ome * (FHZ_L + FHZ_R);
}

/* GetFluxHLL with the wave speeds as inputs: in as orc_flux_glf, l = {λ1_L, λ1_R, λ3_L, λ3_R}. */
void orc_flux_hll(const double in[16], const double l[4], double out[3])
{
    flux_hll(in, l[0], l[1], l[2], l[3], out);
}

/* ReconstructInterface: src/AdvectJandH.jl:359-473.
 * in  = {J_i,J_ip1,J_ip2,J_im1, Hr_i,Hr_ip1,Hr_ip2,Hr_im1, HZ_i,HZ_ip1,HZ_ip2,HZ_im1,
 *        Δr_L,Δr_R, inv_i_im1, inv_ip1_i, inv_ip2_ip1, αF_i, αF_ip1}
 * out = {J_L,J_R,Hr_L,Hr_R,HZ_L,HZ_R,τ_L,τ_R,f_L,f_R,nr_L,nr_R,nZ_L,nZ_R} */
void orc_reconstruct_interface(const double in[19], double out[14])
{
    double J_i = in[0], J_ip1 = in[1], J_ip2 = in[2], J_im1 = in[3];
    double Hr_i = in[4], Hr_ip1 = in[5], Hr_ip2 = in[6], Hr_im1 = in[7];
    double HZ_i = in[8], HZ_ip1 = in[9], HZ_ip2 = in[10], HZ_im1 = in[11];
    double dL = in[12], dR = in[13], inv_i_im1 = in[14], inv_ip1_i = in[15], inv_ip2_ip1 = in[16];
    double aF_i = in[17], aF_ip1 = in[18];

    double tau_L = aF_i * dL;
    double tau_R = aF_ip1 * dR;

    double inv_J_i   = 1.0 / jmax(J_i, Jeps);
    double inv_J_ip1 = 1.0 / jmax(J_ip1, Jeps);
    double inv_J_ip2 = 1.0 / jmax(J_ip2, Jeps);
    double inv_J_im1 = 1.0 / jmax(J_im1, Jeps);

    double fr_i = Hr_i * inv_J_i, fr_ip1 = Hr_ip1 * inv_J_ip1;
    double fr_ip2 = Hr_ip2 * inv_J_ip2, fr_im1 = Hr_im1 * inv_J_im1;
    double fZ_i = HZ_i * inv_J_i, fZ_ip1 = HZ_ip1 * inv_J_ip1;
    double fZ_ip2 = HZ_ip2 * inv_J_ip2, fZ_im1 = HZ_im1 * inv_J_im1;

    double dJ_c  = (J_ip1 - J_i) * inv_ip1_i;
    double dJ_L  = orc_minmod((J_i - J_im1) * inv_i_im1, dJ_c);
    double dJ_R  = orc_minmod((J_ip2 - J_ip1) * inv_ip2_ip1, dJ_c);
    double dfr_c = (fr_ip1 - fr_i) * inv_ip1_i;
    double dfr_L = orc_minmod((fr_i - fr_im1) * inv_i_im1, dfr_c);
    double dfr_R = orc_minmod((fr_ip2 - fr_ip1) * inv_ip2_ip1, dfr_c);
    double dfZ_c = (fZ_ip1 - fZ_i) * inv_ip1_i;
    double dfZ_L = orc_minmod((fZ_i - fZ_im1) * inv_i_im1, dfZ_c);
    double dfZ_R = orc_minmod((fZ_ip2 - fZ_ip1) * inv_ip2_ip1, dfZ_c);

    double J_L = jmax(J_i + dJ_L * dL, Jeps);
    double J_R = jmax(J_ip1 - dJ_R * dR, Jeps);
    double fr_L = fr_i + dfr_L * dL;
    double fr_R = fr_ip1 - dfr_R * dR;
    double fZ_L = fZ_i + dfZ_L * dL;
    double fZ_R = fZ_ip1 - dfZ_R * dR;

    double f_L = sqrt(fr_L * fr_L + fZ_L * fZ_L);
    if (f_L > 1.0) {
        double inv_f = 1.0 / f_L;
        fr_L *= inv_f; fZ_L *= inv_f; f_L = 1.0;
    }
    double f_R = sqrt(fr_R * fr_R + fZ_R * fZ_R);
    if (f_R > 1.0) {
        double inv_f = 1.0 / f_R;
        fr_R *= inv_f; fZ_R *= inv_f; f_R = 1.0;
    }
    double nr_L = 0.0, nZ_L = 0.0, nr_R = 0.0, nZ_R = 0.0;
    if (f_L > 0.0) { double inv = 1.0 / f_L; nr_L = fr_L * inv; nZ_L = fZ_L * inv; }
    if (f_R > 0.0) { double inv = 1.0 / f_R; nr_R = fr_R * inv; nZ_R = fZ_R * inv; }

    out[0] = J_L; out[1] = J_R;
    out[2] = fr_L * J_L; out[3] = fr_R * J_R;
    out[4] = fZ_L * J_L; out[5] = fZ_R * J_R;
    out[6] = tau_L; out[7] = tau_R;
    out[8] = f_L; out[9] = f_R;
    out[10] = nr_L; out[11] = nr_R; out[12] = nZ_L; out[13] = nZ_R;
}

/* Compute_K_ϕϕ_InnerAxis: src/AdvectJandH.jl:499-522 */
double orc_Kphiphi_inner_axis(double J_axis, double HZ_axis)
{
    J_axis = jmax(J_axis, Jeps);
    double fZ = HZ_axis / J_axis;
    fZ = jclamp(fZ, -1.0, 1.0);
    double f = fabs(fZ);
    double nz = 0.0;
    if (f > 0.0) nz = (fZ > 0.0) ? 1.0 : ((fZ < 0.0) ? -1.0 : fZ);
    double D[3];
    orc_D_edd_from_f_n(f, 0.0, nz, D);
    return jclamp(1.0 - D[0] - D[1], 0.0, 1.0) * J_axis;
}

/* One interface: ReconstructInterface -> closure -> K -> GetFluxM1 -> +advection flux.
 * This factors the block repeated at src/AdvectJandH.jl:682-716 (and its copies at
 * :745-780, :793-833, :882-914, :932-965 for r; :1021-1052, :1059-1087,
 * :1110-1150, :1206-1240, :1246-1280 for Z) without changing its arithmetic.
 * st = {J_i,J_ip1,J_ip2,J_im1, Hr_i..., HZ_i...}; iv = {dL,dR,inv_i_im1,inv_ip1_i,inv_ip2_ip1}.
 * dir 0 = r (FJ=H_r, FHr=K_rr, FHZ=K_rZ), dir 1 = Z (FJ=H_Z, FHr=K_rZ, FHZ=K_ZZ).
 * fl  = {F_J, F_Hr, F_HZ} including the upwind advection flux.
 * rec = reconstructed state (for K_ϕϕ), D_L, D_R = closures at the two sides. */
static void interface_flux(int dir, int hll, const double st[12], const double iv[5],
                           double aF_L, double aF_R, double u_over_c,
                           double fl[3], double rec[14], double D_L[3], double D_R[3])
{
    double in[19];
    memcpy(in, st, 12 * sizeof(double));
    memcpy(in + 12, iv, 5 * sizeof(double));
    in[17] = aF_L; in[18] = aF_R;
    orc_reconstruct_interface(in, rec);
    double J_L = rec[0], J_R = rec[1], Hr_L = rec[2], Hr_R = rec[3], HZ_L = rec[4], HZ_R = rec[5];
    orc_D_edd_from_f_n(rec[8], rec[10], rec[12], D_L);
    orc_D_edd_from_f_n(rec[9], rec[11], rec[13], D_R);
    double g[16];
    if (dir == 0) {
        double Krr_L = D_L[0] * J_L, Krr_R = D_R[0] * J_R;
        double KrZ_L = D_L[2] * J_L, KrZ_R = D_R[2] * J_R;
        g[0] = Hr_L; g[1] = Hr_R; g[2] = Krr_L; g[3] = Krr_R; g[4] = KrZ_L; g[5] = KrZ_R;
    } else {
        double KZZ_L = D_L[1] * J_L, KZZ_R = D_R[1] * J_R;
        double KrZ_L = D_L[2] * J_L, KrZ_R = D_R[2] * J_R;
        g[0] = HZ_L; g[1] = HZ_R; g[2] = KrZ_L; g[3] = KrZ_R; g[4] = KZZ_L; g[5] = KZZ_R;
    }
    g[6] = J_L; g[7] = J_R; g[8] = Hr_L; g[9] = Hr_R; g[10] = HZ_L; g[11] = HZ_R;
    g[12] = rec[6]; g[13] = rec[7]; g[14] = st[0]; g[15] = st[1];
    double F[3];
    if (hll) {
        /* GetFluxM1(Val(:HLL), Val(:r|:Z)): src/AdvectJandH.jl:207-270, M1WaveSpeeds :301-326 */
        double eL[2], eR[2];
        double muL = (dir == 0) ? rec[10] : rec[12];
        double muR = (dir == 0) ? rec[11] : rec[13];
        orc_hll_eigenvalues(rec[8], muL, eL);
        orc_hll_eigenvalues(rec[9], muR, eR);
        flux_hll(g, eL[0], eR[0], eL[1], eR[1], F);
    } else {
        orc_flux_glf(g, F);
    }
    /* GetAdvectionFlux: src/AdvectJandH.jl:276-282 */
    fl[0] = F[0] + u_over_c * ((u_over_c < 0.0) ? J_R : J_L);
    fl[1] = F[1] + u_over_c * ((u_over_c < 0.0) ? Hr_R : Hr_L);
    fl[2] = F[2] + u_over_c * ((u_over_c < 0.0) ? HZ_R : HZ_L);
}

/* ---------------------------------------------------------------------- */
/* FindActiveBins: src/UpdateLyaJandH.jl:47-97                             */
/* ---------------------------------------------------------------------- */
void orc_find_active_bins(const orc_lya *L, const orc_geom *G, int *kmin, int *kmax)
{
    const int nn = L->n_nu, nr = G->n_r, nz = G->n_z;
    double Etot = 0.0;
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nr; ++i) {
            double dV = G->dV[j * nr + i];
            const double *Jc = L->J + ((size_t)j * nr + i) * nn;
            for (int k = 0; k < nn; ++k)
                Etot += L->dnu[k] * jmax(Jc[k] - Jeps, 0.0) * dV;
        }
    int kmn = nn - 1, kmx = 0;
    for (int k = 0; k < nn; ++k) {
        double Ebin = 0.0;
        for (int j = 0; j < nz; ++j)
            for (int i = 0; i < nr; ++i)
                Ebin += L->dnu[k] * jmax(L->J[((size_t)j * nr + i) * nn + k] - 10 * Jeps, 0.0)
                        * G->dV[j * nr + i];
        if (Ebin > eps_active * Etot) {
            if (k < kmn) kmn = k;
            if (k > kmx) kmx = k;
        }
    }
    if (L->kcore_max > kmx) kmx = L->kcore_max;
    if (L->kcore_min < kmn) kmn = L->kcore_min;
    *kmin = kmn; *kmax = kmx;
}

/* ---------------------------------------------------------------------- */
/* AdvectJandH_Cylindrical!: src/AdvectJandH.jl:526-1390                   */
/* ---------------------------------------------------------------------- */
#define IDX(k, i, j) (((size_t)(j) * nr + (i)) * nn + (k))
#define IDXR(k, i, j) (((size_t)(j) * (nr + 1) + (i)) * nn + (k))
#define IDXZ(k, i, j) (((size_t)(j) * nr + (i)) * nn + (k))

static void load_iv(double iv[5], const double *dL, const double *dR, const double *a,
                    const double *b, const double *c, int idx)
{
    iv[0] = dL[idx]; iv[1] = dR[idx]; iv[2] = a[idx]; iv[3] = b[idx]; iv[4] = c[idx];
}

void orc_advect_jandh_cyl(double cdt, orc_lya *L, const orc_geom *G, const orc_vel *V,
                          int update_jold, int bc_r_right, int bc_z_bot, int bc_z_top,
                          int kmin, int kmax)
{
    const int nn = L->n_nu, nr = G->n_r, nz = G->n_z;
    const int hll = L->flux_hll;
    const int k0 = (kmin > 0) ? kmin : 0;
    const int k1 = (kmax < nn - 1) ? kmax : nn - 1;
    const double cdt_over_c = cdt * INV_C;
    double *J = L->J, *Hr = L->H_r, *HZ = L->H_z, *aF = L->alpha_F;
    double *Kip = L->Kphi_ip, *Kim = L->Kphi_im;
    double *Fr = L->F_r, *FrHr = L->F_r_Hr, *FrHZ = L->F_r_HZ;
    double *Fz = L->F_z, *FzHr = L->F_z_Hr, *FzHZ = L->F_z_HZ;

    /* ---- r-direction fluxes: src/AdvectJandH.jl:641-968 ---- */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nz; ++j) {
        double st[12], iv[5], fl[3], rec[14], DL[3], DR[3];
        const double u1 = V->u_r_star[j * (nr + 1) + 0] * INV_C;
        const double u2 = V->u_r_star[j * (nr + 1) + 1] * INV_C;
        for (int k = k0; k <= k1; ++k) {
            double J1 = J[IDX(k, 0, j)], J2 = J[IDX(k, 1, j)];
            double Hr1 = Hr[IDX(k, 0, j)], Hr2 = Hr[IDX(k, 1, j)];
            double HZ1 = HZ[IDX(k, 0, j)], HZ2 = HZ[IDX(k, 1, j)];
            double aF1 = aF[IDX(k, 0, j)];
            /* interface i = 1/2: inner axis always reflects (:660-673) */
            double J0 = J1, Hr0 = -Hr1, HZ0 = HZ1;
            double Jm1 = J2, Hrm1 = -Hr2, HZm1 = HZ2;
            double s0[12] = {J0, J1, J2, Jm1, Hr0, Hr1, Hr2, Hrm1, HZ0, HZ1, HZ2, HZm1};
            interface_flux(0, hll, s0, G->dr_inner, aF1, aF1, u1, fl, rec, DL, DR);
            Fr[IDXR(k, 0, j)] = fl[0]; FrHr[IDXR(k, 0, j)] = fl[1]; FrHZ[IDXR(k, 0, j)] = fl[2];
            Kim[IDX(k, 0, j)] = orc_Kphiphi_inner_axis(rec[1], rec[5]);
            /* interface i = 3/2 (:740-797) */
            double s1[12] = {J1, J2, J[IDX(k, 2, j)], J0, Hr1, Hr2, Hr[IDX(k, 2, j)], Hr0,
                             HZ1, HZ2, HZ[IDX(k, 2, j)], HZ0};
            load_iv(iv, G->dr_L, G->dr_R, G->inv_dr_i_im1, G->inv_dr_ip1_i, G->inv_dr_ip2_ip1, 0);
            interface_flux(0, hll, s1, iv, aF1, aF[IDX(k, 1, j)], u2, fl, rec, DL, DR);
            Fr[IDXR(k, 1, j)] = fl[0]; FrHr[IDXR(k, 1, j)] = fl[1]; FrHZ[IDXR(k, 1, j)] = fl[2];
            Kip[IDX(k, 0, j)] = jclamp(1.0 - DL[0] - DL[1], 0.0, 1.0) * rec[0];
            Kim[IDX(k, 1, j)] = jclamp(1.0 - DR[0] - DR[1], 0.0, 1.0) * rec[1];
        }
        (void)st;
        /* interior interfaces i+1/2, 2 <= i <= N_r-2 (Julia), :803-858 */
        for (int i = 1; i <= nr - 3; ++i) {
            const double u = V->u_r_star[j * (nr + 1) + i + 1] * INV_C;
            load_iv(iv, G->dr_L, G->dr_R, G->inv_dr_i_im1, G->inv_dr_ip1_i, G->inv_dr_ip2_ip1, i);
            for (int k = k0; k <= k1; ++k) {
                double s[12] = {J[IDX(k, i, j)], J[IDX(k, i + 1, j)], J[IDX(k, i + 2, j)], J[IDX(k, i - 1, j)],
                                Hr[IDX(k, i, j)], Hr[IDX(k, i + 1, j)], Hr[IDX(k, i + 2, j)], Hr[IDX(k, i - 1, j)],
                                HZ[IDX(k, i, j)], HZ[IDX(k, i + 1, j)], HZ[IDX(k, i + 2, j)], HZ[IDX(k, i - 1, j)]};
                interface_flux(0, hll, s, iv, aF[IDX(k, i, j)], aF[IDX(k, i + 1, j)], u, fl, rec, DL, DR);
                Fr[IDXR(k, i + 1, j)] = fl[0]; FrHr[IDXR(k, i + 1, j)] = fl[1]; FrHZ[IDXR(k, i + 1, j)] = fl[2];
                Kip[IDX(k, i, j)] = jclamp(1.0 - DL[0] - DL[1], 0.0, 1.0) * rec[0];
                Kim[IDX(k, i + 1, j)] = jclamp(1.0 - DR[0] - DR[1], 0.0, 1.0) * rec[1];
            }
        }
        /* outer r boundary: interfaces N_r-1/2 and N_r+1/2, :861-966 */
        const double uN = V->u_r_star[j * (nr + 1) + nr - 1] * INV_C;
        const double uNp1 = V->u_r_star[j * (nr + 1) + nr] * INV_C;
        for (int k = k0; k <= k1; ++k) {
            double JN = J[IDX(k, nr - 1, j)], JNm1 = J[IDX(k, nr - 2, j)];
            double HrN = Hr[IDX(k, nr - 1, j)], HrNm1 = Hr[IDX(k, nr - 2, j)];
            double HZN = HZ[IDX(k, nr - 1, j)], HZNm1 = HZ[IDX(k, nr - 2, j)];
            double aFN = aF[IDX(k, nr - 1, j)];
            double JNp1, HrNp1, HZNp1, JNp2, HrNp2, HZNp2;
            if (bc_r_right == ORC_REFLECT) {
                JNp1 = JN; HrNp1 = -HrN; HZNp1 = HZN;
                JNp2 = JNm1; HrNp2 = -HrNm1; HZNp2 = HZNm1;
            } else {
                JNp1 = JN; HrNp1 = jmax(HrN, 0.0); HZNp1 = HZN;
                JNp2 = JNp1; HrNp2 = HrNp1; HZNp2 = HZNp1;
            }
            double s[12] = {JNm1, JN, JNp1, J[IDX(k, nr - 3 >= 0 ? nr - 3 : 0, j)],
                            HrNm1, HrN, HrNp1, Hr[IDX(k, nr - 3 >= 0 ? nr - 3 : 0, j)],
                            HZNm1, HZN, HZNp1, HZ[IDX(k, nr - 3 >= 0 ? nr - 3 : 0, j)]};
            load_iv(iv, G->dr_L, G->dr_R, G->inv_dr_i_im1, G->inv_dr_ip1_i, G->inv_dr_ip2_ip1, nr - 2);
            interface_flux(0, hll, s, iv, aF[IDX(k, nr - 2, j)], aFN, uN, fl, rec, DL, DR);
            Fr[IDXR(k, nr - 1, j)] = fl[0]; FrHr[IDXR(k, nr - 1, j)] = fl[1]; FrHZ[IDXR(k, nr - 1, j)] = fl[2];
            Kip[IDX(k, nr - 2, j)] = jclamp(1.0 - DL[0] - DL[1], 0.0, 1.0) * rec[0];
            Kim[IDX(k, nr - 1, j)] = jclamp(1.0 - DR[0] - DR[1], 0.0, 1.0) * rec[1];
            double s2[12] = {JN, JNp1, JNp2, JNm1, HrN, HrNp1, HrNp2, HrNm1, HZN, HZNp1, HZNp2, HZNm1};
            interface_flux(0, hll, s2, G->dr_outer, aFN, aFN, uNp1, fl, rec, DL, DR);
            Fr[IDXR(k, nr, j)] = fl[0]; FrHr[IDXR(k, nr, j)] = fl[1]; FrHZ[IDXR(k, nr, j)] = fl[2];
            Kip[IDX(k, nr - 1, j)] = jclamp(1.0 - DL[0] - DL[1], 0.0, 1.0) * rec[0];
        }
    }

    /* ---- Z-direction fluxes, lower boundary: src/AdvectJandH.jl:978-1097 ---- */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) {
        double iv[5], fl[3], rec[14], DL[3], DR[3];
        const double u1 = V->u_z_star[0 * nr + i] * INV_C;
        const double u2 = V->u_z_star[1 * nr + i] * INV_C;
        for (int k = k0; k <= k1; ++k) {
            double J1 = J[IDX(k, i, 0)], J2 = J[IDX(k, i, 1)];
            double Hr1 = Hr[IDX(k, i, 0)], Hr2 = Hr[IDX(k, i, 1)];
            double HZ1 = HZ[IDX(k, i, 0)], HZ2 = HZ[IDX(k, i, 1)];
            double aF1 = aF[IDX(k, i, 0)];
            double J0, HZ0, Hr0, Jm1, HZm1, Hrm1;
            if (bc_z_bot == ORC_REFLECT) {
                J0 = J1; HZ0 = -HZ1; Hr0 = Hr1;
                Jm1 = J2; HZm1 = -HZ2; Hrm1 = Hr2;
            } else {
                J0 = J1; HZ0 = jmin(HZ1, 0.0); Hr0 = Hr1;
                Jm1 = J0; HZm1 = HZ0; Hrm1 = Hr0;
            }
            double s0[12] = {J0, J1, J2, Jm1, Hr0, Hr1, Hr2, Hrm1, HZ0, HZ1, HZ2, HZm1};
            interface_flux(1, hll, s0, G->dz_inner, aF1, aF1, u1, fl, rec, DL, DR);
            Fz[IDXZ(k, i, 0)] = fl[0]; FzHr[IDXZ(k, i, 0)] = fl[1]; FzHZ[IDXZ(k, i, 0)] = fl[2];
            double s1[12] = {J1, J2, J[IDX(k, i, 2)], J0, Hr1, Hr2, Hr[IDX(k, i, 2)], Hr0,
                             HZ1, HZ2, HZ[IDX(k, i, 2)], HZ0};
            load_iv(iv, G->dz_L, G->dz_R, G->inv_dz_j_jm1, G->inv_dz_jp1_j, G->inv_dz_jp2_jp1, 0);
            interface_flux(1, hll, s1, iv, aF1, aF[IDX(k, i, 1)], u2, fl, rec, DL, DR);
            Fz[IDXZ(k, i, 1)] = fl[0]; FzHr[IDXZ(k, i, 1)] = fl[1]; FzHZ[IDXZ(k, i, 1)] = fl[2];
        }
    }
    /* interior Z interfaces j+1/2, 2 <= j <= N_Z-2 (Julia): :1101-1160 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nz - 3; ++j) {
        double iv[5], fl[3], rec[14], DL[3], DR[3];
        load_iv(iv, G->dz_L, G->dz_R, G->inv_dz_j_jm1, G->inv_dz_jp1_j, G->inv_dz_jp2_jp1, j);
        for (int i = 0; i < nr; ++i) {
            const double u = V->u_z_star[(j + 1) * nr + i] * INV_C;
            for (int k = k0; k <= k1; ++k) {
                double s[12] = {J[IDX(k, i, j)], J[IDX(k, i, j + 1)], J[IDX(k, i, j + 2)], J[IDX(k, i, j - 1)],
                                Hr[IDX(k, i, j)], Hr[IDX(k, i, j + 1)], Hr[IDX(k, i, j + 2)], Hr[IDX(k, i, j - 1)],
                                HZ[IDX(k, i, j)], HZ[IDX(k, i, j + 1)], HZ[IDX(k, i, j + 2)], HZ[IDX(k, i, j - 1)]};
                interface_flux(1, hll, s, iv, aF[IDX(k, i, j)], aF[IDX(k, i, j + 1)], u, fl, rec, DL, DR);
                Fz[IDXZ(k, i, j + 1)] = fl[0]; FzHr[IDXZ(k, i, j + 1)] = fl[1]; FzHZ[IDXZ(k, i, j + 1)] = fl[2];
            }
        }
    }
    /* upper Z boundary: interfaces N_Z-1/2 and N_Z+1/2, :1167-1289 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) {
        double iv[5], fl[3], rec[14], DL[3], DR[3];
        const double uN = V->u_z_star[(nz - 1) * nr + i] * INV_C;
        const double uNp1 = V->u_z_star[nz * nr + i] * INV_C;
        for (int k = k0; k <= k1; ++k) {
            double JN = J[IDX(k, i, nz - 1)], JNm1 = J[IDX(k, i, nz - 2)];
            double HrN = Hr[IDX(k, i, nz - 1)], HrNm1 = Hr[IDX(k, i, nz - 2)];
            double HZN = HZ[IDX(k, i, nz - 1)], HZNm1 = HZ[IDX(k, i, nz - 2)];
            double aFN = aF[IDX(k, i, nz - 1)];
            double JNp1, HZNp1, HrNp1, JNp2, HZNp2, HrNp2;
            if (bc_z_top == ORC_REFLECT) {
                JNp1 = JN; HZNp1 = -HZN; HrNp1 = HrN;
                JNp2 = JNm1; HZNp2 = -HZNm1; HrNp2 = HrNm1;
            } else {
                JNp1 = JN; HZNp1 = jmax(HZN, 0.0); HrNp1 = HrN;
                JNp2 = JNp1; HZNp2 = HZNp1; HrNp2 = HrNp1;
            }
            int jm = nz - 3 >= 0 ? nz - 3 : 0;
            double s[12] = {JNm1, JN, JNp1, J[IDX(k, i, jm)], HrNm1, HrN, HrNp1, Hr[IDX(k, i, jm)],
                            HZNm1, HZN, HZNp1, HZ[IDX(k, i, jm)]};
            load_iv(iv, G->dz_L, G->dz_R, G->inv_dz_j_jm1, G->inv_dz_jp1_j, G->inv_dz_jp2_jp1, nz - 2);
            interface_flux(1, hll, s, iv, aF[IDX(k, i, nz - 2)], aFN, uN, fl, rec, DL, DR);
            Fz[IDXZ(k, i, nz - 1)] = fl[0]; FzHr[IDXZ(k, i, nz - 1)] = fl[1]; FzHZ[IDXZ(k, i, nz - 1)] = fl[2];
            double s2[12] = {JN, JNp1, JNp2, JNm1, HrN, HrNp1, HrNp2, HrNm1, HZN, HZNp1, HZNp2, HZNm1};
            interface_flux(1, hll, s2, G->dz_outer, aFN, aFN, uNp1, fl, rec, DL, DR);
            Fz[IDXZ(k, i, nz)] = fl[0]; FzHr[IDXZ(k, i, nz)] = fl[1]; FzHZ[IDXZ(k, i, nz)] = fl[2];
        }
    }

    /* ---- cell update: src/AdvectJandH.jl:1302-1382 ---- */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nz; ++j) {
        for (int i = 0; i < nr; ++i) {
            const double Ar_L = G->A_r[j * (nr + 1) + i], Ar_R = G->A_r[j * (nr + 1) + i + 1];
            const double Az_L = G->A_z[j * nr + i], Az_R = G->A_z[(j + 1) * nr + i];
            const double inv_dV = G->inv_dV[j * nr + i];
            const int c = j * nr + i;
            const double durdr = V->dur_dr[c], durdz = V->dur_dz[c], duzdr = V->duz_dr[c],
                         duzdz = V->duz_dz[c], urr = V->ur_over_r[c];
            const double invr = G->inv_r_cell[i];
            const double cdt_dV = cdt * inv_dV;
            for (int k = k0; k <= k1; ++k) {
                double Js = J[IDX(k, i, j)], Hrs = Hr[IDX(k, i, j)], HZs = HZ[IDX(k, i, j)];
                double D[3];
                orc_D_edd(Hrs, HZs, Js, D);
                double Drr = D[0], DZZ = D[1], DrZ = D[2];
                /* UpdateConservativeQuantity: src/AdvectJandH.jl:336-354 */
                double Jn = Js - cdt_dV * ((Ar_R * Fr[IDXR(k, i + 1, j)] - Ar_L * Fr[IDXR(k, i, j)])
                                           + (Az_R * Fz[IDXZ(k, i, j + 1)] - Az_L * Fz[IDXZ(k, i, j)]));
                Jn = jmax(Jn, Jeps);
                double HZn = HZs - cdt_dV * ((Ar_R * FrHZ[IDXR(k, i + 1, j)] - Ar_L * FrHZ[IDXR(k, i, j)])
                                             + (Az_R * FzHZ[IDXZ(k, i, j + 1)] - Az_L * FzHZ[IDXZ(k, i, j)]));
                double Hrn = Hrs - cdt_dV * ((Ar_R * FrHr[IDXR(k, i + 1, j)] - Ar_L * FrHr[IDXR(k, i, j)])
                                             + (Az_R * FzHr[IDXZ(k, i, j + 1)] - Az_L * FzHr[IDXZ(k, i, j)]));
                /* D_colon_∇u_Cylindrical: src/ConstantsAndFunctions.jl:883-895 */
                double Dpp = jmax(1.0 - Drr - DZZ, 0.0);
                double Ddu = Drr * durdr + DrZ * (durdz + duzdr) + DZZ * duzdz + Dpp * urr;
                double Kav = 0.5 * (Kim[IDX(k, i, j)] + Kip[IDX(k, i, j)]);
                Hrn += cdt * Kav * invr;
                /* ΔH_Aberration: src/AdvectJandH.jl:300-312 */
                double dHr = -cdt_over_c * (Hrs * durdr + HZs * durdz);
                double dHZ = -cdt_over_c * (Hrs * duzdr + HZs * duzdz);
                Hrn = Hrn + dHr;
                HZn = HZn + dHZ;
                Jn = jmax(Jn - cdt_over_c * Ddu * Js, Jeps);
                /* EnforceCausalReducedFlux: src/AdvectJandH.jl:478-493 */
                double Jc = jmax(Jn, Jeps);
                double Hm2 = Hrn * Hrn + HZn * HZn;
                if (Hm2 > Jc * Jc) {
                    double fac = Jc / sqrt(Hm2);
                    Hrn *= fac; HZn *= fac;
                }
                Hr[IDX(k, i, j)] = Hrn;
                HZ[IDX(k, i, j)] = HZn;
                J[IDX(k, i, j)] = Jn;
            }
        }
    }
    if (update_jold)
        memcpy(L->J_old, J, (size_t)nn * nr * nz * sizeof(double));
}

/* AdvectLyα!: src/UpdateLyaJandH.jl:107-171 (cylindrical geometry) */
void orc_advect_lya(double cdt, orc_lya *L, const orc_geom *G, const orc_vel *V,
                    int bc_r_left, int bc_r_right, int bc_z_bot, int bc_z_top,
                    int *kmin_out, int *kmax_out)
{
    (void)bc_r_left;   /* cylindrical inner axis always reflects (:660-662) */
    int kmin, kmax;
    orc_find_active_bins(L, G, &kmin, &kmax);
    orc_advect_jandh_cyl(cdt, L, G, V, 1, bc_r_right, bc_z_bot, bc_z_top, kmin, kmax);
    size_t n = (size_t)L->n_nu * G->n_r * G->n_z;
    memcpy(L->H_r_old, L->H_r, n * sizeof(double));
    memcpy(L->H_z_old, L->H_z, n * sizeof(double));
    if (kmin_out) *kmin_out = kmin;
    if (kmax_out) *kmax_out = kmax;
}

/* ---------------------------------------------------------------------- */
/* LyαEmissionAbsorptionUpdate!: src/ConstantsAndFunctions.jl:451-573      */
/* ---------------------------------------------------------------------- */
static inline double absorption_update(double Jold, double dt, double dJdt, double ac, double ct)
{
    return (Jold + dJdt * dt) / (1.0 + ac * ct * dt);
}

void orc_emission_absorption(orc_lya *L, int n_r, int n_z, double c_tilde, double dt)
{
    const int nn = L->n_nu, nr = n_r;
    const double inv_ct = 1.0 / c_tilde;
    const double fpc_ict = (FOUR_PI / C_LIGHT) * inv_ct;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n_r * n_z; ++c) {
        (void)nr;
        const double ac = L->alpha_c[c], al = L->alpha[c], pd = L->pd[c], js = L->js[c];
        const double al_pd = al * pd;
        const double js_ct = js * c_tilde;
        double *J = L->J + (size_t)c * nn, *Jo = L->J_old + (size_t)c * nn;
        double *Hr = L->H_r + (size_t)c * nn, *HZ = L->H_z + (size_t)c * nn;
        double *Hro = L->H_r_old + (size_t)c * nn, *HZo = L->H_z_old + (size_t)c * nn;
        const double *Hv = L->Hvoigt + (size_t)c * nn, *phi = L->phi + (size_t)c * nn;
        const double *aF = L->alpha_F + (size_t)c * nn;
        for (int k = 0; k < nn; ++k) {
            double aJ = ac + al_pd * Hv[k];
            double dJdt = js_ct * phi[k];
            J[k] = absorption_update(Jo[k], dt, dJdt, aJ, c_tilde);
            Hro[k] = Hr[k];
            HZo[k] = HZ[k];
            double aH = aF[k];
            Hr[k] = absorption_update(Hro[k], dt, 0.0, aH, c_tilde);
            HZ[k] = absorption_update(HZo[k], dt, 0.0, aH, c_tilde);
        }
        double sr = 0.0, sz = 0.0;
        for (int k = 0; k < nn; ++k) {
            double d = L->dnu[k];
            sr += d * (Hro[k] - Hr[k]);
            sz += d * (HZo[k] - HZ[k]);
        }
        L->drhou_r[c] += fpc_ict * sr;
        L->drhou_z[c] += fpc_ict * sz;
        for (int k = 0; k < nn; ++k) {
            Jo[k] = J[k];
            Hro[k] = Hr[k];
            HZo[k] = HZ[k];
        }
    }
}

/* ---------------------------------------------------------------------- */
/* UpdateFreqFactors!: src/FreqFactors.jl:54-278                           */
/* ---------------------------------------------------------------------- */
static void one_minus_phi_L(double P, double *o, double *chi)
{
    if (P >= Pmax) { *o = orc_exp(-Pmax); *chi = (1.0 - *o) / Pmax; return; }
    if (P <= -Pmax) { *o = orc_exp(Pmax); *chi = (1.0 - *o) / (-Pmax); return; }
    if (fabs(P) < Pmin) { *o = 1.0 - P + 0.5 * (P * P); *chi = 1.0 - 0.5 * P; return; }
    *o = orc_exp(-P);
    *chi = (1.0 - *o) / P;
}

static void one_minus_phi_R(double P, double *o, double *chi)
{
    if (P >= Pmax) { *o = orc_exp(Pmax); *chi = (1.0 - *o) / Pmax; return; }
    if (P <= -Pmax) { *o = orc_exp(-Pmax); *chi = (1.0 - *o) / (-Pmax); return; }
    if (fabs(P) < Pmin) { *o = 1.0 + P + 0.5 * (P * P); *chi = -1.0 - 0.5 * P; return; }
    *o = orc_exp(P);
    *chi = (1.0 - *o) / P;
}

static void freq_factors_cell(const orc_lya *L, int c, double T, double durdr, double durdz,
                              double duzdr, double duzdz, double urr,
                              double *Fkp, double *Fk, double *Fkm)
{
    const int nn = L->n_nu;
    const double h_over_kB = H_PLANCK / K_B;
    const double dnuD = L->dnuD[c];
    const double PREFACT = 0.5 * jmax(L->alpha[c], L->alpha_floor) * (dnuD * dnuD);
    const double RECOIL_0 = h_over_kB / T;
    const double *Jc = L->J + (size_t)c * nn, *Hrc = L->H_r + (size_t)c * nn;
    const double *HZc = L->H_z + (size_t)c * nn, *Hv = L->Hvoigt + (size_t)c * nn;
    const double *nu = L->nu, *nuC = L->nuC;
    for (int k = 1; k <= nn - 2; ++k) {
        double nuCk = nuC[k], nuCkp = nuC[k + 1], nuCkm = nuC[k - 1];
        double dkp = nuCkp - nu[k + 1];
        double dk = nu[k + 1] - nuCk;
        double dkm = nu[k] - nuCkm;
        double inv_dnu_k = L->inv_dnu[k];
        double two_inv_k = 2.0 * L->inv_nuC[k];
        double two_inv_kp = 2.0 * L->inv_nuC[k + 1];
        double two_inv_km = 2.0 * L->inv_nuC[k - 1];
        double D[3];
        orc_D_edd(Hrc[k], HZc[k], Jc[k], D);
        double Dpp = jmax(1.0 - D[0] - D[1], 0.0);
        double Ddu = D[0] * durdr + D[2] * (durdz + duzdr) + D[1] * duzdz + Dpp * urr;
        double VEL = Ddu * INV_C;
        double Kkp = PREFACT * Hv[k + 1], Kk = PREFACT * Hv[k], Kkm = PREFACT * Hv[k - 1];
        double Ckp = VEL * nuCkp, Ck = VEL * nuCk, Ckm = VEL * nuCkm;
        if (L->recoil) {
            Ckp += Kkp * (RECOIL_0 - two_inv_kp);
            Ck += Kk * (RECOIL_0 - two_inv_k);
            Ckm += Kkm * (RECOIL_0 - two_inv_km);
        }
        double qkp = dkp * (1.0 / Kkp), qk = dk * (1.0 / Kk), qkm = dkm * (1.0 / Kkm);
        double Pkp = jclamp(Ckp * qkp, -Pmax, Pmax);
        double Pk = jclamp(Ck * qk, -Pmax, Pmax);
        double Pkm = jclamp(Ckm * qkm, -Pmax, Pmax);
        double oRkp, xRkp, oLk, xLk, oRk, xRk, oLkm, xLkm;
        one_minus_phi_R(Pkp, &oRkp, &xRkp);
        one_minus_phi_L(Pk, &oLk, &xLk);
        one_minus_phi_R(Pk, &oRk, &xRk);
        one_minus_phi_L(Pkm, &oLkm, &xLkm);
        double den_kp = inv_dnu_k * (1.0 / (xLk * qk - xRkp * qkp));
        Fkp[k] = oRkp * den_kp;
        double den_km = inv_dnu_k * (1.0 / (xLkm * qkm - xRk * qk));
        Fkm[k] = oLkm * den_km;
        Fk[k] = oLk * den_kp + oRk * den_km;
    }
}

void orc_update_freq_factors(orc_lya *L, int n_r, int n_z, const double *T, const orc_vel *V)
{
    const int nn = L->n_nu;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n_r * n_z; ++c)
        freq_factors_cell(L, c, T[c], V->dur_dr[c], V->dur_dz[c], V->duz_dr[c], V->duz_dz[c],
                          V->ur_over_r[c], L->Fkp + (size_t)c * nn, L->Fk + (size_t)c * nn,
                          L->Fkm + (size_t)c * nn);
}

/* ---------------------------------------------------------------------- */
/* LAPACK dgtsv (NRHS = 1), restated from the published reference LAPACK   */
/* algorithm (Gaussian elimination with partial pivoting by row            */
/* interchange).  Called at src/TridiagonalSolver.jl:54.  Returns INFO.    */
/* dl, du have n-1 entries; d, b have n.  b is overwritten by x.           */
/* ---------------------------------------------------------------------- */
int orc_gtsv(int n, double *dl, double *d, double *du, double *b)
{
    if (n <= 0) return 0;
    for (int i = 0; i < n - 2; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                double fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return i + 1;
            }
            dl[i] = 0.0;
        } else {
            double fact = d[i] / dl[i];
            d[i] = dl[i];
            double temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            dl[i] = du[i + 1];
            du[i + 1] = -fact * dl[i];
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (n > 1) {
        int i = n - 2;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                double fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return i + 1;
            }
        } else {
            double fact = d[i] / dl[i];
            d[i] = dl[i];
            double temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (d[n - 1] == 0.0) return n;
    b[n - 1] = b[n - 1] / d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; --i)
        b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    return 0;
}

/* Plain Thomas algorithm (no pivoting) -- used by the tests as a second,
 * independent solver for the M-matrix systems of the frequency update. */
void orc_thomas(int n, const double *a, const double *b, const double *c, const double *d,
                double *cp, double *dp, double *x)
{
    cp[0] = c[0] / b[0];
    dp[0] = d[0] / b[0];
    for (int i = 1; i < n; ++i) {
        double m = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / m;
        dp[i] = (d[i] - a[i] * dp[i - 1]) / m;
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
}

/* ---------------------------------------------------------------------- */
/* FrequencyUpdateJLyα_AllCells!: src/UpdateLyaJandH.jl:178-391            */
/* ---------------------------------------------------------------------- */
void orc_frequency_update_all_cells(double cdt, orc_lya *L, int n_r, int n_z,
                                    const double *T, const orc_vel *V, int update_ff)
{
    const int nn = L->n_nu;
    const double dw = 1.0 / cdt;
    if (update_ff) orc_update_freq_factors(L, n_r, n_z, T, V);
    const int nint = nn - 2;
#pragma omp parallel
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)(nint > 0 ? nint : 1) * 4);
        double *B = A + nint, *Cc = B + nint, *Dd = Cc + nint;
#pragma omp for schedule(static)
        for (int c = 0; c < n_r * n_z; ++c) {
            double *J = L->J + (size_t)c * nn;
            const double *Jo = L->J_old + (size_t)c * nn;
            const double *Fkp = L->Fkp + (size_t)c * nn, *Fk = L->Fk + (size_t)c * nn;
            const double *Fkm = L->Fkm + (size_t)c * nn;
            for (int row = 0; row < nint; ++row) {
                int k = row + 1;
                A[row] = Fkm[k];
                B[row] = -(Fk[k] + dw);
                Cc[row] = Fkp[k];
                Dd[row] = -dw * Jo[k];
            }
            /* TridiagonalSolve: dl <- A[2:end], du <- C[1:end-1], x <- D (:45-54) */
            orc_gtsv(nint, A + 1, B, Cc, Dd);
            for (int row = 0; row < nint; ++row) J[row + 1] = Dd[row];
        }
        free(A);
    }
    /* flux correction: src/UpdateLyaJandH.jl:335-379 */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n_r * n_z; ++c) {
        double *J = L->J + (size_t)c * nn, *Jo = L->J_old + (size_t)c * nn;
        double *Hr = L->H_r + (size_t)c * nn, *HZ = L->H_z + (size_t)c * nn;
        double *Hro = L->H_r_old + (size_t)c * nn, *HZo = L->H_z_old + (size_t)c * nn;
        for (int k = 0; k < nn; ++k) {
            double Jk = jmax(J[k], Jeps);
            J[k] = Jk;
            double hr = Hr[k], hz = HZ[k];
            double inv_Jold = 1.0 / jmax(Jo[k], Jeps);
            double ratio = Jk * inv_Jold;
            hr = hr * ratio;
            hz = hz * ratio;
            double Hm2 = hr * hr + hz * hz;
            if (Hm2 > Jk * Jk) {
                double s = Jk / sqrt(Hm2);
                hr = hr * s;
                hz = hz * s;
            }
            Hr[k] = hr; HZ[k] = hz;
            Jo[k] = Jk;
            Hro[k] = hr; HZo[k] = hz;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Update_f_Lyα_MF_L_Lyα!: src/ConstantsAndFunctions.jl:588-674            */
/* out = {M_F, M_F_radial_net, M_F_radial_out, L_Lyα_tot}                  */
/* ---------------------------------------------------------------------- */
void orc_update_f_mf_l(orc_lya *L, const orc_geom *G, double dt, double z_center,
                       int bc_z_bot, double out[4])
{
    const int nr = G->n_r, nz = G->n_z;
    const double inv_dt = 1.0 / dt;
    for (int c = 0; c < nr * nz; ++c) {
        L->f_r[c] = L->drhou_r[c] * inv_dt;
        L->f_z[c] = L->drhou_z[c] * inv_dt;
    }
    double MF = 0.0, Lsum = 0.0;
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nr; ++i) {
            int c = j * nr + i;
            MF += G->dV[c] * sqrt(L->f_r[c] * L->f_r[c] + L->f_z[c] * L->f_z[c]);
            Lsum += FOUR_PI * G->dV[c] * L->js[c];
        }
    double net = 0.0, outw = 0.0;
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nr; ++i) {
            int c = j * nr + i;
            double dr = G->rcom[i], dz = G->zcom[j] - z_center;
            double dist = sqrt(dr * dr + dz * dz);
            double frad;
            if (dist > 0.0) {
                double inv_d = 1.0 / dist;
                frad = L->f_r[c] * (dr * inv_d) + L->f_z[c] * (dz * inv_d);
            } else {
                frad = 0.0;
            }
            net += G->dV[c] * frad;
            outw += G->dV[c] * jmax(frad, 0.0);
        }
    if (Lsum > 0.0) {
        double s = 1.0 / (Lsum / C_LIGHT);
        MF = MF * s; net = net * s; outw = outw * s;
    } else {
        MF = 0.0; net = 0.0; outw = 0.0;
    }
    out[0] = MF; out[1] = net; out[2] = outw;
    out[3] = (bc_z_bot == ORC_REFLECT) ? 2.0 * Lsum : Lsum;
}

/* Gravity variant: src/ConstantsAndFunctions.jl:691-804.
 * out = {M_F, M_F_antigravity, M_F_radial_net, M_F_radial_out, L_Lyα_tot} */
void orc_update_f_mf_l_grav(orc_lya *L, const orc_geom *G, double dt,
                            const double *fg_r, const double *fg_z,
                            double z_center, int bc_z_bot, double out[5])
{
    const int nr = G->n_r, nz = G->n_z;
    const double inv_dt = 1.0 / dt;
    for (int c = 0; c < nr * nz; ++c) {
        L->f_r[c] = L->drhou_r[c] * inv_dt;
        L->f_z[c] = L->drhou_z[c] * inv_dt;
    }
    double MF = 0.0, Lsum = 0.0;
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nr; ++i) {
            int c = j * nr + i;
            MF += G->dV[c] * sqrt(L->f_r[c] * L->f_r[c] + L->f_z[c] * L->f_z[c]);
            Lsum += FOUR_PI * G->dV[c] * L->js[c];
        }
    double ag = 0.0, net = 0.0, outw = 0.0;
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nr; ++i) {
            int c = j * nr + i;
            double gm = sqrt(fg_r[c] * fg_r[c] + fg_z[c] * fg_z[c]);
            double fa = (gm > 0.0) ? -(L->f_r[c] * fg_r[c] + L->f_z[c] * fg_z[c]) / gm : 0.0;
            ag += G->dV[c] * fa;
            double dr = G->rcom[i], dz = G->zcom[j] - z_center;
            double dist = sqrt(dr * dr + dz * dz);
            double frad;
            if (dist > 0.0) {
                double inv_d = 1.0 / dist;
                frad = L->f_r[c] * (dr * inv_d) + L->f_z[c] * (dz * inv_d);
            } else {
                frad = 0.0;
            }
            net += G->dV[c] * frad;
            outw += G->dV[c] * jmax(frad, 0.0);
        }
    if (Lsum > 0.0) {
        double s = 1.0 / (Lsum / C_LIGHT);
        MF = MF * s; ag = ag * s; net = net * s; outw = outw * s;
    } else {
        MF = 0.0; ag = 0.0; net = 0.0; outw = 0.0;
    }
    out[0] = MF; out[1] = ag; out[2] = net; out[3] = outw;
    out[4] = (bc_z_bot == ORC_REFLECT) ? 2.0 * Lsum : Lsum;
}

/* Compute_f_esc_Lyα_Cylindrical -> Compute_f_esc_Cylindrical:
 * src/ConstantsAndFunctions.jl:914-993 */
double orc_f_esc_lya(const orc_lya *L, const orc_geom *G, double L_allspace, int bc_z_bot)
{
    const int nn = L->n_nu, nr = G->n_r, nz = G->n_z;
    if (L_allspace <= 0.0) return 0.0;
    double Lsim = (bc_z_bot == ORC_REFLECT) ? 0.5 * L_allspace : L_allspace;
    double fesc = 0.0;
    for (int k = 0; k < nn; ++k) {
        double Lesc = 0.0;
        for (int j = 0; j < nz; ++j)
            Lesc += G->A_r[j * (nr + 1) + nr] * FOUR_PI * jmax(L->H_r[IDX(k, nr - 1, j)], 0.0);
        for (int i = 0; i < nr; ++i)
            Lesc += G->A_z[nz * nr + i] * FOUR_PI * jmax(L->H_z[IDX(k, i, nz - 1)], 0.0);
        if (bc_z_bot == ORC_OUTFLOW)
            for (int i = 0; i < nr; ++i)
                Lesc += G->A_z[i] * FOUR_PI * jmax(-L->H_z[IDX(k, i, 0)], 0.0);
        double fk = (Lsim > 0.0) ? Lesc / Lsim : 0.0;
        fesc += L->dnu[k] * fk;
    }
    return fesc;
}

/* Reset_Δρu!: src/ConstantsAndFunctions.jl:809-816 */
void orc_reset_drhou(orc_lya *L, int n_r, int n_z)
{
    for (int c = 0; c < n_r * n_z; ++c) { L->drhou_r[c] = 0.0; L->drhou_z[c] = 0.0; }
}

/* Update_e_Lyα!: src/ConstantsAndFunctions.jl:416-446 */
void orc_update_e_lya(orc_lya *L, int n_r, int n_z)
{
    const int nn = L->n_nu;
    for (int c = 0; c < n_r * n_z; ++c) {
        double s = 0.0;
        for (int k = 0; k < nn; ++k) s += L->dnu[k] * L->J[(size_t)c * nn + k];
        L->e_lya[c] = (FOUR_PI / C_LIGHT) * s;
    }
}

/* Edge spectra: src/SaveHDF5Data_StaticCloudTests.jl:152-200, 392-420.
 * raw/norm are [N_edge][n_nu]. */
int orc_edge_count(int n_r, int n_z) { return (n_z - 2) + 2 * (n_r - 3); }

void orc_edge_spectra(const orc_lya *L, int n_r, int n_z, double *raw, double *norm)
{
    const int nn = L->n_nu;
    int e = 0;
    int ne = orc_edge_count(n_r, n_z);
    int *ci = (int *)malloc(sizeof(int) * (size_t)(ne > 0 ? ne : 1) * 2);
    for (int j = 1; j <= n_z - 2; ++j) { ci[2 * e] = n_r - 2; ci[2 * e + 1] = j; ++e; }
    for (int i = 1; i <= n_r - 3; ++i) { ci[2 * e] = i; ci[2 * e + 1] = n_z - 2; ++e; }
    for (int i = 1; i <= n_r - 3; ++i) { ci[2 * e] = i; ci[2 * e + 1] = 1; ++e; }
    for (e = 0; e < ne; ++e) {
        const double *Jc = L->J + ((size_t)ci[2 * e + 1] * n_r + ci[2 * e]) * nn;
        double s = 0.0;
        for (int k = 0; k < nn; ++k) s += Jc[k] * L->dnu[k];
        for (int k = 0; k < nn; ++k) {
            raw[(size_t)e * nn + k] = Jc[k];
            norm[(size_t)e * nn + k] = (s > 0.0) ? Jc[k] / s : Jc[k];
        }
    }
    free(ci);
}

/* One pass of the RT-only static cloud time loop: src/StaticCloudTest.jl:357-392.
 * out = {M_F, M_F_radial_net, M_F_radial_out, L_Lyα, f_esc, kmin, kmax, 0} */
void orc_static_cloud_step(double c_tilde, double dt, orc_lya *L, const orc_geom *G,
                           const orc_vel *V, const double *T, double z_center,
                           int bc_z_bot, double out[8])
{
    double cdt = c_tilde * dt;
    int kmin, kmax;
    orc_advect_lya(cdt, L, G, V, ORC_REFLECT, ORC_OUTFLOW, bc_z_bot, ORC_OUTFLOW, &kmin, &kmax);
    orc_emission_absorption(L, G->n_r, G->n_z, c_tilde, dt);
    orc_frequency_update_all_cells(cdt, L, G->n_r, G->n_z, T, V, 1);
    double m[4];
    orc_update_f_mf_l(L, G, dt, z_center, bc_z_bot, m);
    double fesc = orc_f_esc_lya(L, G, m[3], bc_z_bot);
    orc_reset_drhou(L, G->n_r, G->n_z);
    out[0] = m[0]; out[1] = m[1]; out[2] = m[2]; out[3] = m[3];
    out[4] = fesc; out[5] = kmin; out[6] = kmax; out[7] = 0.0;
}

/* ---------------------------------------------------------------------- */
/* Setup                                                                  */
/* ---------------------------------------------------------------------- */

/* Uniform SetupCylindricalGrid (N_r_log = N_Z_log_side = 0): src/SpatialGrid.jl:606-913,
 * Get_*_Cylindrical! :26-169.  Julia LinRange element i is fma(t,b,fma(-t,a,a)), t=i/N;
 * with a = 0 that is (i/N)*b. */
void orc_cyl_grid_uniform(double rmax, double zmax, orc_geom *G)
{
    const int nr = G->n_r, nz = G->n_z;
    for (int i = 0; i <= nr; ++i) G->r[i] = ((double)i / (double)nr) * rmax;
    for (int j = 0; j <= nz; ++j) G->z[j] = ((double)j / (double)nz) * zmax;
    for (int i = 0; i < nr; ++i) {
        double a = G->r[i + 1], b = G->r[i];
        G->rcom[i] = (2.0 / 3.0) * (a * a * a - b * b * b) / (a * a - b * b);
        G->inv_r_cell[i] = 2.0 / (G->r[i + 1] + G->r[i]);
    }
    for (int j = 0; j < nz; ++j) G->zcom[j] = 0.5 * (G->z[j] + G->z[j + 1]);
    for (int j = 0; j < nz; ++j) {
        double dz = G->z[j + 1] - G->z[j];
        for (int i = 0; i < nr; ++i) {
            G->dV[j * nr + i] = PI_D * (G->r[i + 1] * G->r[i + 1] - G->r[i] * G->r[i]) * dz;
            G->inv_dV[j * nr + i] = 1.0 / G->dV[j * nr + i];
        }
        for (int i = 0; i <= nr; ++i) G->A_r[j * (nr + 1) + i] = 2.0 * PI_D * G->r[i] * dz;
    }
    for (int j = 0; j <= nz; ++j)
        for (int i = 0; i < nr; ++i)
            G->A_z[j * nr + i] = PI_D * (G->r[i + 1] * G->r[i + 1] - G->r[i] * G->r[i]);
    orc_grid_differences(nr, G->r, G->rcom, G->dr_L, G->dr_R, G->inv_dr_i_im1, G->inv_dr_ip1_i,
                         G->inv_dr_ip2_ip1, G->dr_inner, G->dr_outer);
    orc_grid_differences(nz, G->z, G->zcom, G->dz_L, G->dz_R, G->inv_dz_j_jm1, G->inv_dz_jp1_j,
                         G->inv_dz_jp2_jp1, G->dz_inner, G->dz_outer);
}

/* Compute_Grid_Differences!: src/SpatialGrid.jl:516-586.
 * inner/outer = {ΔL, ΔR, inv(i,i-1), inv(i+1,i), inv(i+2,i+1)} */
void orc_grid_differences(int n, const double *x, const double *xc,
                          double *dL, double *dR, double *inv_i_im1, double *inv_ip1_i,
                          double *inv_ip2_ip1, double inner[5], double outer[5])
{
    double c0 = x[0] - (xc[0] - x[0]);
    double cm1 = x[0] - (xc[1] - x[0]);
    double cNp1 = x[n] + (x[n] - xc[n - 1]);
    double cNp2 = x[n] + (x[n] - xc[n - 2]);
    inner[2] = 1.0 / (c0 - cm1);
    inner[3] = 1.0 / (xc[0] - c0);
    inner[4] = 1.0 / (xc[1] - xc[0]);
    inner[0] = x[0] - c0;
    inner[1] = xc[0] - x[0];
    for (int i = 0; i < n; ++i) {
        double cim1 = (i == 0) ? c0 : xc[i - 1];
        double cip1 = (i == n - 1) ? cNp1 : xc[i + 1];
        double cip2 = (i <= n - 3) ? xc[i + 2] : ((i == n - 2) ? cNp1 : cNp2);
        inv_i_im1[i] = 1.0 / (xc[i] - cim1);
        inv_ip1_i[i] = 1.0 / (cip1 - xc[i]);
        inv_ip2_ip1[i] = 1.0 / (cip2 - cip1);
        dL[i] = x[i + 1] - xc[i];
        dR[i] = cip1 - x[i + 1];
    }
    outer[2] = 1.0 / (xc[n - 1] - xc[n - 2]);
    outer[3] = 1.0 / (cNp1 - xc[n - 1]);
    outer[4] = 1.0 / (cNp2 - cNp1);
    outer[0] = x[n] - xc[n - 1];
    outer[1] = cNp1 - x[n];
}

/* Compute_b / Compute_ΔνD / Compute_a / Compute_σ0: src/ConstantsAndFunctions.jl:282-314 */
double orc_compute_b(double T) { return sqrt(2.0 * K_B / M_H) * sqrt(T); }
double orc_compute_dnuD(double T) { return orc_compute_b(T) * (NU_LYA / C_LIGHT); }
double orc_compute_a(double T)
{
    double a_pref = A_LYA / (FOUR_PI * NU_LYA * 1.0 / C_LIGHT);
    return a_pref / orc_compute_b(T);
}
double orc_compute_sigma0(double T)
{
    const double sqrt_pi = sqrt(PI_D);
    double s_pref = sqrt_pi * (E_CHARGE * E_CHARGE) * F_LU_LYA / (M_E * NU_LYA);
    return s_pref / orc_compute_b(T);
}

/* Frequency grid of GenerateLyαArrays: src/SetupLymanAlphaArrays.jl:40-104.
 * Writes ν (N_ν+1 interfaces) into x_out; returns N_ν+1, or -1 if cap is too small. */
int orc_lya_freq_grid(double Tmin, double q_fine, double q_wing, double *nu_out, int cap)
{
    double x_fine_3e4 = 5.0, x_max = 1e5;
    double dnu_min = (orc_compute_b(Tmin) / C_LIGHT) * NU_LYA;
    x_max = jmin(x_max, 0.99 * NU_LYA / dnu_min);
    x_max = jmin(x_max, (NU_LYC - NU_LYA) / dnu_min);
    double xfm = x_fine_3e4 * orc_compute_b(3e4) / orc_compute_b(Tmin);
    int npos = 0, maxpos = 100000;
    double *xp = (double *)malloc(sizeof(double) * (size_t)maxpos);
    xp[npos++] = 0.0;
    double dx = 1.0;
    while (xp[npos - 1] < x_max && npos < maxpos - 1) {
        double x = xp[npos - 1];
        double y = (x - xfm) / (0.1 * xfm);
        double S = (y < 0.0) ? 0.0 : ((y > 1.0) ? 1.0 : y);
        double q = q_fine + (q_wing - q_fine) * S;
        xp[npos] = xp[npos - 1] + dx;
        ++npos;
        dx *= q;
    }
    if (xp[npos - 1] > x_max) xp[npos - 1] = x_max;
    else if (xp[npos - 1] < x_max) xp[npos++] = x_max;
    int ntot = 2 * npos - 1;
    if (ntot > cap) { free(xp); return -1; }
    int m = 0;
    for (int i = npos - 1; i >= 1; --i) nu_out[m++] = NU_LYA + (-xp[i]) * dnu_min;
    for (int i = 0; i < npos; ++i) nu_out[m++] = NU_LYA + xp[i] * dnu_min;
    free(xp);
    return ntot;
}

/* EvaluateLUT: src/FastInterpolations.jl:83-112 */
double orc_lut_eval(const double *y, int nx, double xmin, double inv_dx, double x)
{
    double xt = (x - xmin) * inv_dx;
    long i = (long)floor(xt);
    if (i + 2 > nx) i = nx - 2;
    double d = (double)(i + 1) - xt;
    return d * y[i] + (1.0 - d) * y[i + 1];
}

/* H(x,a), ∫dx_H, Htophat: src/ConstantsAndFunctions.jl:227-279 with the fast
 * LUT evaluators of src/FastInterpolations.jl:203-274.
 * luts = [erf | dawson | exp(-x²) | x exp(-x²) | (1-2x²) exp(-x²)], each nx nodes. */
typedef struct { const double *t; int nx; double xmax, xmin, inv_dx; } vlut;

static double lut_dawson(const vlut *V, double x)
{
    if (fabs(x) >= V->xmax) {
        double ix = 1.0 / x, ix3 = ix * ix * ix;
        return 0.5 * ix + 0.25 * ix3;
    }
    return orc_lut_eval(V->t + (size_t)V->nx, V->nx, V->xmin, V->inv_dx, x);
}
static double lut_fn(const vlut *V, int which, double x)
{
    if (fabs(x) >= V->xmax) {
        if (which == 0) return copysign(1.0, x);
        return 0.0;
    }
    return orc_lut_eval(V->t + (size_t)which * V->nx, V->nx, V->xmin, V->inv_dx, x);
}
static double voigt_H(const vlut *V, double x, double a)
{
    const double sqrt_pi_over_2 = sqrt(PI_D) / 2.0;
    const double inv_sp2 = 1.0 / sqrt_pi_over_2;
    double ex = lut_fn(V, 2, x), daw = lut_dawson(V, x), st = lut_fn(V, 4, x);
    return ex + inv_sp2 * a * (2.0 * x * daw - 1.0) + (a * a) * st;
}
static double voigt_intH(const vlut *V, double x, double a)
{
    const double sqrt_pi_over_2 = sqrt(PI_D) / 2.0;
    const double inv_sp2 = 1.0 / sqrt_pi_over_2;
    double er = lut_fn(V, 0, x), daw = lut_dawson(V, x), st = lut_fn(V, 3, x);
    return sqrt_pi_over_2 * er - inv_sp2 * a * daw + (a * a) * st;
}

/* Update_Hvoigt!: src/ConstantsAndFunctions.jl:319-357 */
void orc_update_hvoigt(double *Hvoigt, int n_nu, int n_cells, const double *dnu_rel,
                       const double *dnuC_rel, const double *a, const double *inv_dnuD,
                       const double *luts, int nx, double lut_xmax)
{
    vlut V;
    V.t = luts; V.nx = nx; V.xmax = lut_xmax; V.xmin = -lut_xmax;
    V.inv_dx = 1.0 / ((lut_xmax - (-lut_xmax)) / (double)(nx - 1));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n_cells; ++c) {
        double ic = inv_dnuD[c], ac = a[c];
        for (int k = 0; k < n_nu; ++k) {
            double x = dnuC_rel[k] * ic, xmx = dnu_rel[k + 1] * ic, xmn = dnu_rel[k] * ic;
            double dx = xmx - xmn;
            double h;
            if (dx > 1e-5) h = (voigt_intH(&V, xmx, ac) - voigt_intH(&V, xmn, ac)) / dx;
            else h = voigt_H(&V, x, ac);
            Hvoigt[(size_t)c * n_nu + k] = h;
        }
    }
}

/* Update_ϕν!: src/ConstantsAndFunctions.jl:362-393 (kmin..kmax 0-based inclusive) */
void orc_update_phi(double *phi, int n_nu, int n_cells, const double *dnu,
                    const double *Hvoigt, int kmin, int kmax)
{
    for (int c = 0; c < n_cells; ++c) {
        double s = 0.0;
        for (int k = kmin; k <= kmax; ++k) s += dnu[k] * Hvoigt[(size_t)c * n_nu + k];
        double inv = 1.0 / s;
        for (int k = kmin; k <= kmax; ++k)
            phi[(size_t)c * n_nu + k] = Hvoigt[(size_t)c * n_nu + k] * inv;
    }
}

/* GetRHDCFLTimeStep: src/AdvectJandH.jl:29-68 */
double orc_cfl_dt(const orc_geom *G, double c_tilde, const double *u_r, const double *u_z, double cfl)
{
    double mn = INFINITY;
    for (int j = 0; j < G->n_z; ++j) {
        double dz = G->z[j + 1] - G->z[j];
        for (int i = 0; i < G->n_r; ++i) {
            double dr = G->r[i + 1] - G->r[i];
            double u = hypot(u_r[j * G->n_r + i], u_z[j * G->n_r + i]);
            double d = 0.5 * jmin(dr, dz) / (c_tilde + u);
            if (d < mn) mn = d;
        }
    }
    return cfl * mn;
}

/* GetHubbleFlow: src/StaticCloudTest.jl:34-59.
 * out = {u_r, u_Z, ∂u_r/∂r, ∂u_r/∂Z, ∂u_Z/∂r, ∂u_Z/∂Z, u_r/r} */
void orc_hubble_flow(double r, double z, double umax, double rcl, double zc, double out[7])
{
    double R = hypot(r, z - zc);
    if (R > rcl) { for (int i = 0; i < 7; ++i) out[i] = 0.0; return; }
    double H = umax / rcl;
    out[0] = H * r; out[1] = H * (z - zc);
    out[2] = H; out[3] = 0.0; out[4] = 0.0; out[5] = H; out[6] = H;
}

/* ConeOpticalDepthFactor: src/StaticCloudTest.jl:70-89 */
double orc_cone_factor(double mu, double mucone, double taucone, double taucl,
                       int smooth, double dmu)
{
    if (mucone >= 1.0) return 1.0;
    double fc = taucone / taucl;
    if (smooth) return fc + 0.5 * (1.0 - fc) * (1.0 - tanh((mu - mucone) / dmu));
    return (mu > mucone) ? fc : 1.0;
}


/* ------------------------------------------------------------------------
 * RHD time-step control and the Lyα absorption sub-step
 * ------------------------------------------------------------------------ */
#define X_HE      0.247
#define D_H_RATIO 2.53e-5
#define M_D       (2.014 * AMU)
#define EPS_RSL   4.0
#define TOL_ABS   0.1
#define J_FLOOR_ABS 1e-15

/* FindRSL_LyαRT, src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:616-722 (the Lyα
 * branch; X_over_mH from src/ConstantsAndFunctions.jl:124-128,167, ϵ_RSL from
 * PhotoThermoChemistry.jl:34).  rho_d is [n_z][n_r][3] (the reference's ρd[1:3,i,j]),
 * kappa_abs the three κd_abs_Lyα of silicate, graphite and PAH dust. */
double orc_find_rsl(double c_old, double c_min, const orc_geom *G, const double *u_r, const double *u_z,
                    const double *gamma_ad, const double *P, const double *rho, const double *rho_d,
                    const double *xHII, const double *xH2, const double *T, const double *pd,
                    const double *kappa_abs)
{
    const double X_D = D_H_RATIO * (1 - X_HE) * M_D / M_H;
    const double X_H = 1 - X_HE - X_D;
    const double X_over_mH = X_H / M_H;
    double cnew = 0.0;
    for (int j = 0; j < G->n_z; ++j) {
        double local = cnew;
        const double dZ = G->z[j + 1] - G->z[j];
        for (int i = 0; i < G->n_r; ++i) {
            const int c = j * G->n_r + i;
            const double dr = G->r[i + 1] - G->r[i];
            const double c_s = sqrt(gamma_ad[c] * P[c] / rho[c]);
            const double xHI = jmax(1 - xHII[c] - 2 * xH2[c], 0.0);
            const double nHI = xHI * X_over_mH * rho[c];
            const double alpha = nHI * orc_compute_sigma0(T[c]);
            const double dX = jmax(dr, dZ);
            const double tcell = alpha * dX;
            const double tcabs = (kappa_abs[0] * rho_d[3 * c] + kappa_abs[1] * rho_d[3 * c + 1] +
                                  kappa_abs[2] * rho_d[3 * c + 2]) * dX;
            const double at = orc_compute_a(T[c]) * tcell;
            const double at13 = pow(at, 1.0 / 3.0);
            const double fabs_ = 1 / (1 + at13 * tcabs + 0.6 * pd[c] * tcell);
            const double u = hypot(u_r[c], u_z[c]);
            const double cmin_cell = EPS_RSL * jmax(1, at13 * fabs_) * jmax(u, c_s);
            local = jmax(local, cmin_cell);
        }
        cnew = jmax(cnew, local);
    }
    cnew = jmin(jmax(cnew, c_min), C_LIGHT);
    if (cnew < c_old) cnew = jmax(cnew, 0.75 * c_old);
    if (cnew > c_old) cnew = jmin(cnew, 4.0 * c_old);
    return cnew;
}

/* The Lyα block of AbsorptionStep (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:1577-1670)
 * for every cell, with SplitRadPressure (:1017-1048) per bin, the rejection rule (tol_Abs,
 * e_Lyα_floor: :1051-1055) and RestoreLyαRejection (:980-990) on rejected cells.
 * dt [n_z][n_r] is each cell's sub-step; rho_d [n_z][n_r][3] may be NULL (no dust);
 * kappa_F the three κd_F_Lyα.  Reads J_old, H_old, e_lya; writes J, H (the old values on a
 * rejected cell).  redo [n_z][n_r]; out [12][n_z][n_r] = {Δe_Lyα, e_Lyα_new, Δρu_r/Z_Lyα,
 * Δρu_r/Z of gas, silicate, graphite, PAH}, all zero except Δe on a rejected cell. */
void orc_absorption_substep(orc_lya *L, int n_r, int n_z, const double *dt, double c_tilde,
                            const double *rho_d, const double *kappa_F, int *redo, double *out)
{
    const int nn = L->n_nu, nc = n_r * n_z;
    const double inv_ct = 1.0 / c_tilde;
    const double e_floor = (FOUR_PI / C_LIGHT) * J_FLOOR_ABS;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; ++c) {
        const double dts = dt[c];
        const double ct_js = c_tilde * L->js[c];
        const double ac = L->alpha_c[c];
        const double pd_a = L->pd[c] * L->alpha[c];
        const double aS = rho_d ? kappa_F[0] * rho_d[3 * c] : 0.0;
        const double aC = rho_d ? kappa_F[1] * rho_d[3 * c + 1] : 0.0;
        const double aP = rho_d ? kappa_F[2] * rho_d[3 * c + 2] : 0.0;
        const double dt_lya = inv_ct / (ac + pd_a);
        double *J = L->J + (size_t)c * nn, *Jo = L->J_old + (size_t)c * nn;
        double *Hr = L->H_r + (size_t)c * nn, *HZ = L->H_z + (size_t)c * nn;
        const double *Hro = L->H_r_old + (size_t)c * nn, *HZo = L->H_z_old + (size_t)c * nn;
        const double *Hv = L->Hvoigt + (size_t)c * nn, *phi = L->phi + (size_t)c * nn;
        const double *aF = L->alpha_F + (size_t)c * nn;
        double sJ = 0.0, mr = 0.0, mz = 0.0, g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < nn; ++k) {
            const double dnu = L->dnu[k];
            const double dJdt = ct_js * phi[k];
            const double a_abs = ac + pd_a * Hv[k];
            J[k] = jmax(absorption_update(Jo[k], dts, dJdt, a_abs, c_tilde), Jeps);
            const double aFk = aF[k];
            Hr[k] = absorption_update(Hro[k], dts, 0.0, aFk, c_tilde);
            HZ[k] = absorption_update(HZo[k], dts, 0.0, aFk, c_tilde);
            const double w = dnu * (FOUR_PI / C_LIGHT);
            const double dr_k = w * (inv_ct * (Hro[k] - Hr[k]));
            const double dz_k = w * (inv_ct * (HZo[k] - HZ[k]));
            mr += dr_k;
            mz += dz_k;
            const double a_gas = jmax(aFk - aS - aC - aP, 0.0);
            const double a_tot = a_gas + aS + aC + aP;
            double fg = 1.0, fS = 0.0, fC = 0.0, fP = 0.0;
            if (a_tot > 0.0) {
                const double inv = 1 / a_tot;
                fg = a_gas * inv; fS = aS * inv; fC = aC * inv; fP = aP * inv;
            }
            g[0] += fg * dr_k; g[1] += fg * dz_k;
            g[2] += fS * dr_k; g[3] += fS * dz_k;
            g[4] += fC * dr_k; g[5] += fC * dz_k;
            g[6] += fP * dr_k; g[7] += fP * dz_k;
            sJ += dnu * J[k];
        }
        const double e_new = (FOUR_PI / C_LIGHT) * sJ;
        const double e_old = L->e_lya[c];
        const double de = fabs(e_new - e_old) / (e_old + e_floor);
        const int rej = (de > TOL_ABS && dts > TOL_ABS * dt_lya) || !isfinite(de);
        redo[c] = rej;
        double v[12] = {de, e_new, mr, mz, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7]};
        if (rej) {
            for (int m = 1; m < 12; ++m) v[m] = 0.0;
            for (int k = 0; k < nn; ++k) { J[k] = Jo[k]; Hr[k] = Hro[k]; HZ[k] = HZo[k]; }
        }
        for (int m = 0; m < 12; ++m) out[(size_t)m * nc + c] = v[m];
    }
}


/* ------------------------------------------------------------------------
 * Lyα source terms of PrepareLyαSubCycling! (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2086-2155)
 * ------------------------------------------------------------------------ */
#define E_LYA_INV_4PI ((H_PLANCK * NU_LYA) * (1.0 / FOUR_PI))
#define SIGMA_T      6.6524587321e-25
#define SIGMA_RAMAN  2.1e-25
#define SIGMA_OH     4.07e-18
#define SIGMA_CH     0.05e-18
#define A_2P2S       8.78e-7
#define A_2S2P       1.597e-9
#define A_2GAMMA     8.2206
#define RHOD_IGNORE  (1.1 * (1e-13 * M_H))

/* Storey & Hummer (1995) probability of Lyα emission per recombination, the tables of
 * src/LymanAlphaEmission.jl:30-60 on log10 T (10 nodes) x log10 n_e (13 Case B / 9 Case A nodes). */
static const double SH95_T[10] = {5e2, 1e3, 3e3, 5e3, 7.5e3, 1e4, 1.25e4, 1.5e4, 2e4, 3e4};
static const double SH95_B[10][13] = {
    {0.80793637, 0.90215197, 0.98157401, 0.9970534, 0.99933853, 0.99934845, 0.99967896, 0.99948547, 0.99931146, 0.99951894, 0.9997498, 0.99864818, 0.99372263},
    {0.78439303, 0.87743841, 0.97423464, 0.99680114, 0.99960243, 0.99967816, 0.9997671, 0.99939133, 0.99922946, 0.99960973, 0.99937386, 0.99906812, 0.9954834},
    {0.74159266, 0.82809599, 0.95734434, 0.99359891, 0.99892162, 0.99927343, 0.99979471, 0.99938909, 0.99899055, 0.99925979, 0.99957955, 0.99923525, 0.99703571},
    {0.71946081, 0.80133228, 0.94581237, 0.99246182, 0.99848958, 0.99940329, 0.99949246, 0.99952048, 0.99981447, 0.99956726, 0.9993515, 0.99952024, 0.99692499},
    {0.70117195, 0.7788453, 0.93441344, 0.99075058, 0.99831301, 0.99960223, 0.99940949, 0.99946884, 0.99948297, 0.99947146, 0.99965311, 0.99918254, 0.99690311},
    {0.68808416, 0.76220739, 0.92503048, 0.9891268, 0.99828523, 0.9994074, 0.99944339, 0.99952601, 0.99932358, 0.99927528, 0.99938987, 0.99904232, 0.9949615},
    {0.6775581, 0.74868826, 0.9176273, 0.98784252, 0.99835238, 0.99903921, 0.99977064, 0.99954296, 0.99930529, 0.99954743, 0.99944997, 0.99874439, 0.9916493},
    {0.66862468, 0.73785038, 0.91033402, 0.98701405, 0.99817046, 0.99892625, 0.99953078, 0.99960556, 0.9993362, 0.99941558, 0.99944365, 0.99828039, 0.98817287},
    {0.65534642, 0.72080463, 0.89823705, 0.98497224, 0.99798854, 0.9988133, 0.99929092, 0.99909416, 0.99933199, 0.9995549, 0.99924446, 0.99764512, 0.97960152},
    {0.63566376, 0.69600394, 0.87952243, 0.98121463, 0.99716593, 0.99896144, 0.99905958, 0.9996252, 0.99953649, 0.99931002, 0.99872497, 0.99561256, 0.96391965}};
static const double SH95_A[10][9] = {
    {0.58042473, 0.59595911, 0.60335363, 0.58990186, 0.56623927, 0.53426043, 0.492528, 0.44534414, 0.40955195},
    {0.54712958, 0.56465163, 0.57999443, 0.57241623, 0.5541246, 0.5276139, 0.49052298, 0.44606687, 0.4101595},
    {0.48320245, 0.50151448, 0.52818381, 0.53008608, 0.52008701, 0.50298698, 0.47625175, 0.43924427, 0.40696841},
    {0.44933274, 0.46714191, 0.49852558, 0.50475148, 0.49815515, 0.48505756, 0.46319941, 0.43069186, 0.40125666},
    {0.42087234, 0.43810743, 0.47280049, 0.48268151, 0.47821493, 0.46794659, 0.4497736, 0.42097532, 0.39442497},
    {0.40023971, 0.41675887, 0.45373431, 0.46591319, 0.46300415, 0.45454038, 0.43876097, 0.41249415, 0.38810578},
    {0.3840314, 0.40007514, 0.43832241, 0.45257107, 0.45064954, 0.4434932, 0.42939783, 0.40511911, 0.38230403},
    {0.37046215, 0.38617473, 0.42546609, 0.44143608, 0.44049428, 0.43394168, 0.42127472, 0.39854551, 0.37684759},
    {0.34937867, 0.36397106, 0.40484801, 0.42349201, 0.42367822, 0.41860328, 0.40761256, 0.38699493, 0.36739625},
    {0.31960139, 0.33302781, 0.37528203, 0.39805701, 0.39984039, 0.39602958, 0.38718651, 0.36963656, 0.3518518}};

/* Bilinear interpolation on nodes xs[nx] x ys[ny] of f (row stride ny), at a point inside the grid. */
static double bilinear(const double *xs, int nx, const double *ys, int ny, const double *f, double x, double y)
{
    int i = 0, j = 0;
    while (i < nx - 2 && x > xs[i + 1]) ++i;
    while (j < ny - 2 && y > ys[j + 1]) ++j;
    const double t = (x - xs[i]) / (xs[i + 1] - xs[i]);
    const double u = (y - ys[j]) / (ys[j + 1] - ys[j]);
    const double f0 = f[i * ny + j] + u * (f[i * ny + j + 1] - f[i * ny + j]);
    const double f1 = f[(i + 1) * ny + j] + u * (f[(i + 1) * ny + j + 1] - f[(i + 1) * ny + j]);
    return f0 + t * (f1 - f0);
}

/* Prob_Lyα_SH95 (src/LymanAlphaEmission.jl:79-91): clamped to the table, flat beyond it. */
double orc_prob_lya_sh95(double T, double ne, int caseB)
{
    double lt[10], ln[13];
    for (int i = 0; i < 10; ++i) lt[i] = log10(SH95_T[i]);
    const int nn = caseB ? 13 : 9;
    for (int j = 0; j < nn; ++j) ln[j] = (double)(2 + j);   /* log10 n_e = 2, 3, ..., 14 (or 10) */
    const double x = jmin(jmax(log10(T), lt[0]), lt[9]);
    const double y = jmin(jmax(log10(ne), ln[0]), ln[nn - 1]);
    return bilinear(lt, 10, ln, nn, caseB ? &SH95_B[0][0] : &SH95_A[0][0], x, y);
}

/* krecHII_A / krecHII_B (src/PhotoChemistryAndCooling/GasChemistryRates.jl:166-204). */
double orc_krec_hii(double T, int caseB)
{
    T = jmin(jmax(T, 1.0), 1e9);
    const double lam = 315614.0 / T;
    if (caseB) return 2.753e-14 * pow(lam, 1.5) / pow(1 + pow(lam / 2.74, 0.407), 2.242);
    return 1.269e-13 * pow(lam, 1.503) / pow(1 + pow(lam / 0.522, 0.47), 1.923);
}

/* ∫dν_js_recLyα (src/LymanAlphaEmission.jl:105-124): Case B with the on-the-spot approximation. */
double orc_js_rec_lya(double T, double nHII, double ne, int ots)
{
    const double krec = orc_krec_hii(T, ots);
    const double P = orc_prob_lya_sh95(T, ne, ots);
    return P * krec * nHII * ne * E_LYA_INV_4PI;
}

/* ∫dν_js_collLyα (src/LymanAlphaEmission.jl:143-173). */
double orc_js_coll_lya(double T, double nHI, double ne)
{
    const double T_ryd = (H_PLANCK * NU_LYC) / K_B;
    const double T12 = (1 - 1.0 / 4.0) * T_ryd, T13 = (1 - 1.0 / 9.0) * T_ryd;
    const double T14 = (1 - 1.0 / 16.0) * T_ryd, T15 = (1 - 1.0 / 25.0) * T_ryd;
    T = jmax(T, 1.0);
    const double inv_T = 1.0 / T;
    const double T6 = jmin(T, 3e5) * 1e-6;
    const double T6s = T6 * T6, T6c = T6s * T6;
    const double Y12 = 0.267486 + 1.57257 * T6 - 6.44026 * T6s + 11.5401 * T6c;
    const double Y13 = 0.0940821 + 2.90748 * T6 - 6.69528 * T6s + 7.86419 * T6c;
    const double Y14 = 0.0415028 + 1.13673 * T6 - 3.97454 * T6s + 6.05941 * T6c;
    const double Y15 = 0.032277 + 0.40856 * T6 - 1.34604 * T6s + 2.07558 * T6c;
    const double q = (8.63e-6 / (2 * sqrt(T))) *
                     (Y12 * exp(-T12 * inv_T) + Y13 * exp(-T13 * inv_T) + Y14 * exp(-T14 * inv_T) +
                      Y15 * exp(-T15 * inv_T));
    return q * nHI * ne * E_LYA_INV_4PI;
}

/* pd_Lyα (src/LymanAlphaDestruction.jl:43-97). */
double orc_pd_lya(double nHII, double ne, double nHI, double nHeI, double nH2, double T, double z)
{
    T = jmin(jmax(T, 1.0), 1e5);
    const double T100 = T / 100.0;
    const double lT = log10(T), lT2 = lT * lT, lT3 = lT2 * lT, lT4 = lT3 * lT;
    const double G2p2s = 3.1e-6 * (1 + z), G2s2p = 3 * G2p2s;
    const double r25 = 2.5 / T;
    const double C_e = 5e-5 * exp(-(r25 * r25)) * ne / (1 + pow(T100, 0.45));
    const double C_HII = (6.98e-3 * (1 + 0.933 * lT) * exp(-pow(601 / T, 0.406)) / pow(T, 0.5) +
                          pow(10.0, -7.69 + 1.21 * lT + 0.455 * lT2 - 0.199 * lT3 + 0.0182 * lT4)) * nHII;
    const double C_HeI = 2.6e-9 * pow(T100, 0.315) * nHeI;
    const double C_H2 = 1.7e-9 * pow(T100, 0.3) * nH2;
    const double den = 8.9 * pow(T, -0.55) + 0.078 * pow(T, 0.35);
    const double C_HI = (4.1e-11 * exp(-pow(1.6 / T, 3.0 / 4.0) - 0.15 * T) + (1.6e-10 + 1.57e-12 * pow(T, 1.1)) / (den * den)) * nHI;
    const double C2s2p = C_e + C_HII + C_HeI + C_H2 + C_HI;
    const double C2p2s = (1.0 / 3.0) * C2s2p;
    const double R2p2s = (2.0 / 3.0) * A_2P2S + C2p2s + G2p2s;
    const double R2s2p = A_2S2P + C2s2p + G2s2p;
    const double pd_HI = R2p2s / ((A_LYA / A_2GAMMA) * (A_2GAMMA + R2s2p) + R2p2s);
    double pd_H2 = 0.0;
    if (nHI > 1e-20)
        pd_H2 = jmin((33.0 * exp(-14540.0 / T) + 13.0 * exp(-15564.0 / T)) / (46.0 + 1e3 * exp(-pow(1e3 / T, 1.5))),
                     1.3e-3) * (nH2 / nHI);
    return jmin(pd_HI + pd_H2, 1.0);
}

/* The per-cell source terms of PrepareLyαSubCycling! for every cell: js, α_c (dust + gas
 * absorption), α_c,F (dust + gas flux opacity, added to every bin of α_F), p_d.  chem is
 * [8][n_cells] = T, n_HI, n_H2, n_HII, n_HeI, n_e, n_OHx, n_CHx; rho_d [n_cells][3];
 * kappa = {κd_abs_Lyα, κd_F_Lyα} of silicate, graphite, PAH.  out [4][n_cells]. */
void orc_lya_sources(int n_cells, const double *chem, const double *rho_d, const double *kappa, double z, int ots,
                     int destruction, double *out)
{
    for (int c = 0; c < n_cells; ++c) {
        const double T = chem[c], nHI = chem[n_cells + c], nH2 = chem[2 * n_cells + c];
        const double nHII = chem[3 * n_cells + c], nHeI = chem[4 * n_cells + c], ne = chem[5 * n_cells + c];
        const double nOH = chem[6 * n_cells + c], nCH = chem[7 * n_cells + c];
        double rd[3];
        for (int s = 0; s < 3; ++s) rd[s] = (rho_d[3 * c + s] > RHOD_IGNORE) ? rho_d[3 * c + s] : 0.0;
        const double js = orc_js_rec_lya(T, nHII, ne, ots) + orc_js_coll_lya(T, nHI, ne);
        const double ac_gas = nH2 * SIGMA_RAMAN + nOH * SIGMA_OH + nCH * SIGMA_CH;
        const double acF_gas = ac_gas + ne * SIGMA_T;
        const double ac = kappa[0] * rd[0] + kappa[1] * rd[1] + kappa[2] * rd[2] + ac_gas;
        const double pd = destruction ? orc_pd_lya(nHII, ne, nHI, nHeI, nH2, T, z) : 0.0;
        const double acF = kappa[3] * rd[0] + kappa[4] * rd[1] + kappa[5] * rd[2] + acF_gas;
        out[c] = js;
        out[n_cells + c] = ac;
        out[2 * n_cells + c] = acF;
        out[3 * n_cells + c] = pd;
    }
}


/* Number of OpenMP threads the oracle's loops use (the CPU baseline's single-thread sample). */
#ifdef _OPENMP
#include <omp.h>
#endif
void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
