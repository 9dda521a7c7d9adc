// m1lya_internal.h -- kernel argument blocks and launch wrappers shared by
// the kernels (m1lya_kernels.hip) and the C ABI (m1lya_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "m1lya_device.h"

namespace m1 {

// Cell columns per wavefront in the column kernel (lane-per-column dgtsv).
constexpr int kColumnW = 4;
// Largest N_ν of the FP64 context (nine bins per lane, one column per wavefront).
constexpr int kMaxBinsFP64 = 576;
constexpr int kMaxBinsAdjoint = 256;

// LDS row stride of one column's tridiagonal arrays: >= N_ν - 2 rows, odd.
__host__ __device__ inline int col_stride(int nn)
{
    int n = nn - 2 > 1 ? nn - 2 : 1;
    return n | 1;
}

struct DevGeom {
    int nr, nz;
    const double* A_r;        // [nz][nr+1]
    const double* A_z;        // [nz+1][nr]
    const double* inv_dV;     // [nz][nr]
    const double* inv_r_cell; // [nr]
    const double* ivr;        // [(nr+1)][5] interval set of every r interface
    const double* ivz;        // [(nz+1)][5] interval set of every Z interface
};

struct DevVel {
    const double* ur_star;    // [nz][nr+1]
    const double* uz_star;    // [nz+1][nr]
    const double* dur_dr;
    const double* dur_dz;
    const double* duz_dr;
    const double* duz_dz;
    const double* ur_over_r;
};

struct AdvectArgs {
    int nn;
    DevGeom g;
    DevVel v;
    const double* J;
    const double* Hr;
    const double* Hz;
    const double* aF;
    double* J_out;
    double* Hr_out;
    double* Hz_out;
    double cdt;
    int bc_r_right, bc_z_bot, bc_z_top, hll;
    const int* krange;
    const int* halt;      // may be null; when *halt != 0 the launch does nothing
};

enum ColumnFlags {
    COL_EMIT = 1,     // LyαEmissionAbsorptionUpdate!
    COL_FREQ = 2,     // frequency factors computed in-kernel
    COL_STORE_F = 4,  // ... and stored to Fij_kp/k/km
    COL_SOLVE = 8,    // tridiagonal frequency solve
    COL_HCORR = 16,   // J floor + H rescale + causality
    COL_FUSED = 32,   // fused step: no *_old arrays, per-block reductions
    COL_CONV = 64     // fused step: also Σ Δν ΔV |J_new - J_start| (J_start read from J_old)
};

// Columns of a fused-step partial row after the 2 N_ν active-bin energies.
enum StepSums { SUM_MF = 0, SUM_L, SUM_NET, SUM_OUT, SUM_ESC, SUM_CONV, SUM_BAD, kStepSums };

// Bits of the context's device error flag.
enum ErrBits { ERR_SINGULAR = 1, ERR_NONFINITE = 2 };

struct ColumnArgs {
    int nn, nr, nz, flags;
    double* J;
    double* Hr;
    double* Hz;
    double* J_old;
    double* Hr_old;
    double* Hz_old;
    const double* Hv;
    const double* phi;
    const double* aF;
    double* Fkp;
    double* Fk;
    double* Fkm;
    const double* alpha;
    const double* alpha_c;
    const double* pd;
    const double* js;
    const double* T;
    const double* dnuD;
    const double* dur_dr;
    const double* dur_dz;
    const double* duz_dr;
    const double* duz_dz;
    const double* ur_over_r;
    double* drhou_r;
    double* drhou_z;
    double* f_r;
    double* f_z;
    const double* nu;
    const double* nuC;
    const double* dnu;
    const double* inv_dnu;
    const double* inv_nuC;
    const double* dV;
    const double* rcom;
    const double* zcom;
    const double* A_r;
    const double* A_z;
    double c_tilde, dt, dw, alpha_floor, z_center;
    int recoil, bc_z_bot;
    double* part;
    int pstride;
    int* err_flag;
    const int* halt;      // may be null; when *halt != 0 the launch does nothing
};

// Device-side control of a fused step: halt flag and count of executed steps.
struct StepCtl {
    int halt;             // set by k_step_final once ΔJ/J < tol; later launches return at once
    int steps;            // fused steps executed since the host last cleared it
};

struct MfArgs {
    int nr, nz;
    const double* drhou_r;
    const double* drhou_z;
    double* f_r;
    double* f_z;
    const double* js;
    const double* dV;
    const double* rcom;
    const double* zcom;
    const double* fg_r;   // may be null
    const double* fg_z;
    double dt, z_center;
    double* part;         // [nblocks][8]
};

// ---- reverse sweep of the fused step (m1lya_adjoint.hip) ----
struct AdjAdvArgs {
    int nn;
    DevGeom g;
    DevVel v;
    const double* J;      // advection input X_n
    const double* Hr;
    const double* Hz;
    const double* aF;
    const double* lJ1;    // cotangent of the advection output
    const double* lHr1;
    const double* lHz1;
    double* pJ;           // cotangent of each cell's flux-divergence sums
    double* pHr;
    double* pHz;
    double* oJ;           // cotangent of X_n
    double* oHr;
    double* oHz;
    double* gaF;          // += d/dα_F
    double cdt;
    int bc_r_right, bc_z_bot, bc_z_top, hll;
    const int* krange;
    double* fpart;        // reverse-mode interfaces: 14 planes of per-face input cotangents
    int mode;             // 0 interfaces in reverse mode; 1 dual-number passes per cell
};

// doubles of AdjAdvArgs::fpart for an N_r x N_Z x N_ν grid
inline size_t adj_face_part_size(int nn, int nr, int nz)
{
    return (size_t)14 * ((size_t)nz * (nr + 1) + (size_t)nr * (nz + 1)) * nn;
}

struct AdjColArgs {
    int nn, nr, nz, recoil;
    const double* J1;     // advection output (input of emission/absorption)
    const double* Hr1;
    const double* Hz1;
    const double* J4;     // step output
    const double* lJ4;    // cotangent of the step output
    const double* lHr4;
    const double* lHz4;
    double* lJ1;          // cotangent of the advection output
    double* lHr1;
    double* lHz1;
    const double* Hv;
    const double* phi;
    const double* aF;
    const double* alpha;
    const double* alpha_c;
    const double* pd;
    const double* js;
    const double* T;
    const double* dnuD;
    const double* dur_dr;
    const double* dur_dz;
    const double* duz_dr;
    const double* duz_dz;
    const double* ur_over_r;
    const double* nu;
    const double* nuC;
    const double* dnu;
    const double* inv_dnu;
    const double* inv_nuC;
    const double* dV;
    const double* rcom;
    const double* zcom;
    double c_tilde, dt, dw, alpha_floor, z_center;
    double mf_scale;      // c / L_sum (0 when there is no source)
    double w0, w1, w2;    // seeds of M_F, M_F_radial_net, M_F_radial_out at this step
    double* g_aF;         // += [n3]
    double* g_Hv;
    double* g_phi;
    double* g_alpha;      // += [ncell]
    double* g_alpha_c;
    double* g_js;
    double* g_dnuD;
    double* g_T;
    int mode;             // 0 frequency rows in reverse mode; 1 dual-number passes
};

void launch_adj_column(hipStream_t s, const AdjColArgs& a);
void launch_adj_advect(hipStream_t s, const AdjAdvArgs& a);
void launch_adj_spectrum(hipStream_t s, const double* Hr, const double* Hz, const double* A_r, const double* A_z,
                         int nn, int nr, int nz, int bc_z_bot, double inv_Lsim, double* S);
void launch_adj_spectrum_seed(hipStream_t s, const double* Hr, const double* Hz, const double* A_r,
                              const double* A_z, int nn, int nr, int nz, int bc_z_bot, double inv_Lsim,
                              const double* v, double* lHr, double* lHz);
void launch_adj_axpy_cell(hipStream_t s, double* y, const double* dV, double a, int n);

void launch_active_partial(hipStream_t s, const double* J, const double* dnu, const double* dV, int nn,
                           int ncells, int nblocks, double* part, int pstride);
void launch_reduce_cols(hipStream_t s, const double* part, int nrows, int stride, int ncols, double* out,
                        const int* halt = nullptr);
void launch_active_final(hipStream_t s, const double* red, int nn, int kcmin, int kcmax, int* krange);
void launch_advect(hipStream_t s, const AdvectArgs& a);
// Cell columns per wavefront of the column kernel for an N_ν-bin grid.
int column_width(int nn);
// Wavefronts per block of the column kernel for an N_ν-bin grid (the block shares its dgtsv phase).
int column_waves(int nn);
// Blocks of the column kernel for ngroups column groups: a fixed count (at most ngroups),
// so that the partition of the fused reductions, and with it their bits, does not depend
// on the device.  kColumnBlocks / column_waves(N_ν) blocks are one launch wave on MI355X
// (8 wavefronts per CU, 256 CUs).
constexpr int kColumnBlocks = 2048;
int column_blocks(int nn, int ngroups);
void launch_column(hipStream_t s, const ColumnArgs& a, int nblocks);
// Step scalars {M_F, net, out, L, f_esc, kmin, kmax, ΔJ/J}; ΔJ/J is 0 unless the column
// kernel ran with COL_CONV.  ctl may be null (no halt, no count); tol > 0 sets ctl->halt
// once ΔJ/J < tol.  A non-finite count sets ERR_NONFINITE in err_flag.
void launch_step_final(hipStream_t s, const double* red, int nn, int kcmin, int kcmax, int bc_z_bot,
                       int* krange, double* scalars, double* history, int* err_flag, StepCtl* ctl,
                       double tol, int conv);
void launch_mf_partial(hipStream_t s, const MfArgs& a, int nblocks);
void launch_fesc_partial(hipStream_t s, const double* Hr, const double* Hz, const double* A_r,
                         const double* A_z, const double* dnu, int nn, int nr, int nz, int bc_z_bot,
                         int nblocks, double* part);
void launch_e_lya(hipStream_t s, const double* J, const double* dnu, int nn, int ncells, double* e);
void launch_edge(hipStream_t s, const double* J, const double* dnu, int nn, int nr, int nz, double* raw,
                 double* norm);
void launch_nonfinite(hipStream_t s, const double* x, size_t n, int* flag);
void launch_selftest(hipStream_t s, int op, const double* a, const double* b, double* out, size_t n);

}  // namespace m1

namespace m1 {

// ---- device-side coefficient refresh (k_refresh in m1lya_kernels.hip) ----
struct LineConst {
    double b_prefact;       // b(T) = b_prefact √T
    double nu_over_c;       // ΔνD = b ν_Lyα / c
    double a_prefact;       // a = a_prefact / b
    double sigma0_prefact;  // σ0 = sigma0_prefact / b
};

struct RefreshArgs {
    int nn, ncell, kmin, kmax, phi_all_bins;
    const double* lut;      // [5][nx]: erf, dawson, exp(-x²), x exp(-x²), (1-2x²) exp(-x²)
    int nx;
    double xmax;
    LineConst lc;
    const double* dnu_rel;  // ν - ν_Lyα at the N_ν + 1 interfaces
    const double* dnuC_rel; // νC - ν_Lyα
    const double* dnu;
    const double* T;        // [ncell] input temperature
    const double* nHI;      // [ncell] input H I density
    const double* acF;      // [ncell] continuum flux opacity added to every bin of α_F (may be null)
    double* Hv;             // outputs
    double* phi;
    double* aF;
    double* alpha;
    double* dnuD;
};

void launch_refresh(hipStream_t s, const RefreshArgs& a);

// ---- RHD time-step control: GetRHDCFLTimeStep / FindRSL_LyαRT per-cell terms ----
struct RhdScanArgs {
    int nr, nz, mode;       // mode 0: min over cells of 0.5 min(Δr, ΔZ) / (c̃ + |u|); 1: max of the RSL floor
    const double* r;        // [nr+1] interfaces
    const double* Z;        // [nz+1]
    const double* u_r;      // [nz][nr] cell velocities
    const double* u_z;
    const double* gamma_ad; // mode 1 only
    const double* P;
    const double* rho;
    const double* rho_d;    // [nz][nr][3]
    const double* xHII;
    const double* xH2;
    const double* T;
    const double* pd;
    double kS, kG, kP;      // κd_abs_Lyα of silicate, graphite, PAH dust
    double c_tilde;
    LineConst lc;
    double X_over_mH;
};
// Writes the reduced value to out[0] (deterministic: min / max are exact); part holds nblocks doubles.
void launch_rhd_scan(hipStream_t s, const RhdScanArgs& a, int nblocks, double* part, double* out);

// ---- the Lyα block of the RHD absorption sub-step (AbsorptionStep) ----
struct AbsorbArgs {
    int nn, ncell;
    const double* dt;       // [ncell] sub-step of each cell
    double c_tilde;
    const double* J_old;
    const double* Hr_old;
    const double* Hz_old;
    double* J;
    double* Hr;
    double* Hz;
    const double* phi;
    const double* Hv;
    const double* aF;
    const double* alpha;
    const double* alpha_c;
    const double* pd;
    const double* js;
    const double* e_lya;    // e_Lyα before the sub-step
    const double* dnu;
    const double* rho_d;    // [ncell][3] or null
    double kS, kG, kP;      // κd_F_Lyα of silicate, graphite, PAH dust
    int* redo;              // [ncell]
    double* out;            // [12][ncell]
};
void launch_absorb(hipStream_t s, const AbsorbArgs& a);

// ---- Lyα source terms of PrepareLyαSubCycling! from the chemical state ----
struct SourceArgs {
    int ncell, ots, destruction;
    double z;
    const double* chem;     // [8][ncell]: T, n_HI, n_H2, n_HII, n_HeI, n_e, n_OHx, n_CHx
    const double* rho_d;    // [ncell][3] or null
    double kappa[6];        // κd_abs_Lyα, κd_F_Lyα of silicate, graphite, PAH
    double* js;             // outputs [ncell]
    double* alpha_c;
    double* acF;
    double* pd;
    double* T;
    double* nHI;
};
void launch_sources(hipStream_t s, const SourceArgs& a);

}  // namespace m1
