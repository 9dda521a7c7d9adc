/*
 * m1lya.h -- C ABI of libm1lya, the MI355X (gfx950) multi-frequency M1 Lya
 * radiative-transfer step of LYDION.
 *
 * The library is the drop-in behind the Julia operator API that
 * src/StaticCloudTest.jl and src/RHDSolver_*.jl call (olofnebrin/LYDION):
 * every entry point below names the reference function it replaces.
 *
 * Conventions
 *  - Every function returns an int status: M1LYA_OK (0) or a negative code.
 *    m1lya_last_error() returns the message of the last failure.  No C++
 *    exception crosses this boundary.
 *  - Host arrays use the reference (Julia column-major) layout, so a Julia
 *    Array{Float64,3} of size (N_ν, N_r, N_Z) is passed as-is:
 *      spectral fields  [N_ν, N_r, N_Z]   (k fastest)
 *      cell fields      [N_r, N_Z]
 *      u_r_star, A_r    [N_r+1, N_Z]
 *      u_Z_star, A_Z    [N_r, N_Z+1]
 *  - Bin indices (kcore_min/kmax, active-bin ranges) are 0-based here; the
 *    Julia shim subtracts 1 from LyαArrays.kcore_min/kcore_max.
 *  - A context is not re-entrant: one host thread at a time.  All work is
 *    queued on the context's HIP stream; functions that return host data
 *    synchronise that stream.  Host pointers are never retained.
 *  - State stays resident in device memory between calls; upload/download
 *    it at save or checkpoint points.
 */
#ifndef M1LYA_H
#define M1LYA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M1LYA_ABI_VERSION 1

typedef struct m1lya_ctx m1lya_ctx;

enum m1lya_status {
    M1LYA_OK = 0,
    M1LYA_ERR_ARG = -1,        /* bad argument (NULL pointer, wrong count, bad enum) */
    M1LYA_ERR_DIMS = -2,       /* unsupported dimensions (N_r < 3, N_Z < 3, N_ν < 3;
                                  N_ν > 576 for a context, > 256 for a batch)         */
    M1LYA_ERR_HIP = -3,        /* HIP runtime error (allocation, launch)            */
    M1LYA_ERR_NONFINITE = -4,  /* a fused step wrote a non-finite J, H_r or H_Z, or its
                                  tridiagonal system was singular                      */
    M1LYA_ERR_STATE = -5,      /* geometry / frequency grid not set                  */
    M1LYA_ERR_UNSUPPORTED = -6 /* option not implemented                             */
};

/* Symbol BCs of the reference (:Reflect, :Outflow). */
enum m1lya_bc { M1LYA_REFLECT = 0, M1LYA_OUTFLOW = 1 };

/* LyαArrays{FluxMethod}: Val{:GLF} or Val{:HLL} (src/SetupLymanAlphaArrays.jl:220-224). */
enum m1lya_flux { M1LYA_GLF = 0, M1LYA_HLL = 1 };

/* Arithmetic of the step. */
enum m1lya_precision { M1LYA_FP64 = 0 };

/* Device-resident fields, named after the LyαArrays / solver arrays they mirror. */
enum m1lya_field {
    /* [N_ν, N_r, N_Z] */
    M1LYA_J = 0,          /* J_Lyα        */
    M1LYA_J_OLD,          /* J_Lyα_old    */
    M1LYA_H_R,            /* H_r_Lyα      */
    M1LYA_H_Z,            /* H_Z_Lyα      */
    M1LYA_H_R_OLD,        /* H_r_Lyα_old  */
    M1LYA_H_Z_OLD,        /* H_Z_Lyα_old  */
    M1LYA_ALPHA_F,        /* α_F_Lyα      */
    M1LYA_HVOIGT,         /* Hvoigt       */
    M1LYA_PHI_NU,         /* ϕ_ν          */
    M1LYA_FIJ_KP,         /* Fij_kp       */
    M1LYA_FIJ_K,          /* Fij_k        */
    M1LYA_FIJ_KM,         /* Fij_km       */
    /* [N_r, N_Z] */
    M1LYA_ALPHA_LYA,      /* α_Lyα        */
    M1LYA_ALPHA_C,        /* αc_Lyα       */
    M1LYA_PD,             /* pd           */
    M1LYA_JS,             /* ∫dν_js_Lyα   */
    M1LYA_T,              /* gas temperature T                       */
    M1LYA_DNU_D,          /* ΔνD          */
    M1LYA_DUR_DR,         /* ∂u_r/∂r      */
    M1LYA_DUR_DZ,         /* ∂u_r/∂Z      */
    M1LYA_DUZ_DR,         /* ∂u_Z/∂r      */
    M1LYA_DUZ_DZ,         /* ∂u_Z/∂Z      */
    M1LYA_UR_OVER_R,      /* ⟨u_r/r⟩      */
    M1LYA_DRHOU_R,        /* Δρu_Lyα_r    */
    M1LYA_DRHOU_Z,        /* Δρu_Lyα_Z    */
    M1LYA_F_LYA_R,        /* f_Lyα_r      */
    M1LYA_F_LYA_Z,        /* f_Lyα_Z      */
    M1LYA_E_LYA,          /* e_Lyα        */
    /* [N_r+1, N_Z] */
    M1LYA_UR_STAR,        /* u_r_star     */
    /* [N_r, N_Z+1] */
    M1LYA_UZ_STAR,        /* u_Z_star     */
    M1LYA_NFIELDS
};

/* GeometryCylindrical{Float64} (src/Structs.jl:650-740), the fields the M1
 * step reads.  Interval arrays have N entries (N = N_r or N_Z) and follow
 * Compute_Grid_Differences! (src/SpatialGrid.jl:516-586); the *_inner and
 * *_outer arrays are the {ΔL, ΔR, inv(i,i-1), inv(i+1,i), inv(i+2,i+1)}
 * values of the first and last interface. */
typedef struct {
    const double *r, *Z;            /* [N_r+1], [N_Z+1]  */
    const double *rCoM, *ZCoM;      /* [N_r], [N_Z]      */
    const double *dV, *inv_dV;      /* [N_r, N_Z]        */
    const double *A_r;              /* [N_r+1, N_Z]      */
    const double *A_Z;              /* [N_r, N_Z+1]      */
    const double *inv_r_cell;       /* [N_r]             */
    const double *dr_L, *dr_R, *inv_dr_i_im1, *inv_dr_ip1_i, *inv_dr_ip2_ip1;
    double dr_inner[5], dr_outer[5];
    const double *dZ_L, *dZ_R, *inv_dZ_j_jm1, *inv_dZ_jp1_j, *inv_dZ_jp2_jp1;
    double dZ_inner[5], dZ_outer[5];
} m1lya_geometry;

/* Frequency grid and settings of LyαArrays (src/Structs.jl:23-124). */
typedef struct {
    const double *nu;          /* ν, [N_ν+1] interfaces  */
    const double *nuC;         /* νC, [N_ν]              */
    const double *dnu;         /* Δν, [N_ν]              */
    const double *inv_dnu;     /* inv_Δν, [N_ν]          */
    const double *inv_nuC;     /* inv_νC, [N_ν]          */
    int kcore_min, kcore_max;  /* 0-based, inclusive     */
    int recoil;                /* Recoil                 */
    int flux_method;           /* enum m1lya_flux        */
    double alpha_floor;        /* α_Lyα_floor            */
} m1lya_freq_grid;

/* ---- context --------------------------------------------------------- */
int  m1lya_abi_version(void);
/* Creates a context for an N_ν x N_r x N_Z cylindrical problem on the current
 * HIP device (GenerateLyαArrays, src/SetupLymanAlphaArrays.jl:17-239, state
 * allocation; all state starts at zero like the reference's zeros()). */
int  m1lya_create(int n_nu, int n_r, int n_z, int precision, m1lya_ctx **out);
void m1lya_destroy(m1lya_ctx *ctx);
/* Copies the last error message (ctx may be NULL: last create() failure). */
int  m1lya_last_error(const m1lya_ctx *ctx, char *buf, size_t len);
int  m1lya_set_geometry(m1lya_ctx *ctx, const m1lya_geometry *geometry);
int  m1lya_set_freq_grid(m1lya_ctx *ctx, const m1lya_freq_grid *grid);
/* count must equal the field's element count. */
int  m1lya_upload(m1lya_ctx *ctx, int field, const double *host, size_t count);
int  m1lya_download(m1lya_ctx *ctx, int field, double *host, size_t count);
int  m1lya_fill(m1lya_ctx *ctx, int field, double value);
int  m1lya_synchronize(m1lya_ctx *ctx);

/* ---- operator entry points (drop-ins for the reference functions) ----- */
/* Velocity / temperature arrays may be NULL: the device-resident field is used. */

/* FindActiveBins, src/UpdateLyaJandH.jl:47-97 (0-based kmin/kmax). */
int m1lya_find_active_bins(m1lya_ctx *ctx, int *kmin, int *kmax);

/* AdvectLyα!(c̃Δt, Lyα, geometry, u_r_star, u_Z_star, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r,
 * ∂u_Z_∂Z, u_r_over_r, BC_r_left, BC_r_right, BC_Z_bot, BC_Z_top),
 * src/UpdateLyaJandH.jl:107-171 -> AdvectJandH_Cylindrical!, src/AdvectJandH.jl:526-1390.
 * BC_r_left must be M1LYA_REFLECT (cylindrical axis). */
int m1lya_advect_lya(m1lya_ctx *ctx, double c_dt,
                     const double *u_r_star, const double *u_z_star,
                     const double *dur_dr, const double *dur_dz,
                     const double *duz_dr, const double *duz_dz, const double *ur_over_r,
                     int bc_r_left, int bc_r_right, int bc_z_bot, int bc_z_top);

/* LyαEmissionAbsorptionUpdate!(J, J_old, H_r, H_Z, H_r_old, H_Z_old, ∫dν_js, ϕ_ν,
 * Hvoigt, α_Lyα, αc, α_F, pd, Δρu_r, Δρu_Z, Δν, c̃, Δt),
 * src/ConstantsAndFunctions.jl:483-573.  All arrays are the device-resident fields. */
int m1lya_emission_absorption_update(m1lya_ctx *ctx, double c_tilde, double dt);

/* UpdateFreqFactors!(Lyα, T, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r),
 * src/FreqFactors.jl:142-278: fills the Fij_kp / Fij_k / Fij_km fields. */
int m1lya_update_freq_factors(m1lya_ctx *ctx, const double *T,
                              const double *dur_dr, const double *dur_dz,
                              const double *duz_dr, const double *duz_dz, const double *ur_over_r);

/* FrequencyUpdateJLyα_AllCells!(c̃Δt, Lyα, T, ∂u..., u_r_over_r, UpdateFreqFactors),
 * src/UpdateLyaJandH.jl:231-391 (TridiagonalSolve -> LAPACK gtsv, src/TridiagonalSolver.jl:33-58). */
int m1lya_frequency_update_all_cells(m1lya_ctx *ctx, double c_dt, const double *T,
                                     const double *dur_dr, const double *dur_dz,
                                     const double *duz_dr, const double *duz_dz,
                                     const double *ur_over_r, int update_freq_factors);

/* Update_f_Lyα_MF_L_Lyα!(f_Lyα_r, f_Lyα_Z, Δρu_r, Δρu_Z, ∫dν_js, ΔV, rCoM, ZCoM, Δt,
 * Z_center, BC_Z_bot), src/ConstantsAndFunctions.jl:588-674.
 * out = {M_F, M_F_radial_net, M_F_radial_out, L_Lyα_tot}; fills f_Lyα_r/Z. */
int m1lya_update_f_mf_l(m1lya_ctx *ctx, double dt, double z_center, int bc_z_bot, double out[4]);

/* Gravity variant, src/ConstantsAndFunctions.jl:691-804 (f_grav arrays [N_r, N_Z]).
 * out = {M_F, M_F_antigravity, M_F_radial_net, M_F_radial_out, L_Lyα_tot}. */
int m1lya_update_f_mf_l_gravity(m1lya_ctx *ctx, double dt, const double *f_grav_r,
                                const double *f_grav_z, double z_center, int bc_z_bot,
                                double out[5]);

/* Compute_f_esc_Lyα_Cylindrical(H_r, H_Z, A_r, A_Z, Δν, L_Lyα_allspace, BC_Z_bot),
 * src/ConstantsAndFunctions.jl:969-993. */
int m1lya_compute_f_esc(m1lya_ctx *ctx, double L_allspace, int bc_z_bot, double *f_esc);

/* Reset_Δρu!, src/ConstantsAndFunctions.jl:809-816. */
int m1lya_reset_drhou(m1lya_ctx *ctx);

/* Update_e_Lyα!, src/ConstantsAndFunctions.jl:416-446 (fills M1LYA_E_LYA). */
int m1lya_update_e_lya(m1lya_ctx *ctx);

/* Edge spectra written by SaveTimeStepHDF5, src/SaveHDF5Data_StaticCloudTests.jl:152-200,
 * 392-420.  raw/norm are [N_edge, N_ν] row-major (edge slowest), N_edge = m1lya_edge_count(). */
int m1lya_edge_count(const m1lya_ctx *ctx);
int m1lya_edge_spectra(m1lya_ctx *ctx, double *raw, double *norm, size_t count);

/* ---- fused step ------------------------------------------------------ */
/* One pass of the RT-only time loop body, src/StaticCloudTest.jl:357-392:
 * AdvectLyα! -> LyαEmissionAbsorptionUpdate! -> FrequencyUpdateJLyα_AllCells! ->
 * Update_f_Lyα_MF_L_Lyα! -> Compute_f_esc_Lyα_Cylindrical -> Reset_Δρu!,
 * fused into device kernels with no host round trip.  The velocity, T and
 * coefficient fields are the device-resident ones.  BC_r_left is the axis
 * (always reflecting).  out (may be NULL) = {M_F, M_F_radial_net,
 * M_F_radial_out, L_Lyα, f_esc, kmin, kmax, 0}; out != NULL synchronises.
 * Every fused step checks that the J, H_r and H_Z it writes are finite.  A step
 * queued without synchronisation (out == NULL) reports M1LYA_ERR_NONFINITE at the
 * next call that synchronises the context: m1lya_step with out, m1lya_run_steps,
 * m1lya_run_until_steady, m1lya_synchronize or m1lya_download (which still fills
 * its host buffer). */
int m1lya_step(m1lya_ctx *ctx, double c_tilde, double dt, double z_center,
               int bc_r_right, int bc_z_bot, int bc_z_top, double out[8]);

/* n_steps fused steps without host synchronisation.  history (may be NULL)
 * receives n_steps rows of the 8 step scalars. */
int m1lya_run_steps(m1lya_ctx *ctx, int n_steps, double c_tilde, double dt, double z_center,
                    int bc_r_right, int bc_z_bot, int bc_z_top, double *history);

/* Fused steps until the state is steady: stops after the first step whose
 *   ΔJ/J = Σ Δν ΔV |J_new - J_start| / Σ Δν ΔV max(J_new - Jeps, 0)
 * is below tol, or after max_steps.  The reference runs a fixed horizon
 * (src/StaticCloudTest.jl:317-322,357) and has no such stop; this is the
 * "iterated to steady state (ΔJ/J < 1e-8)" option of the RT-only static cloud.
 * history (may be NULL) receives *steps_done rows of the step scalars, whose
 * last entry is that step's ΔJ/J.  The stop is decided on the device: steps
 * are queued in chunks and those after the stop return without work. */
int m1lya_run_until_steady(m1lya_ctx *ctx, int max_steps, double tol, double c_tilde, double dt,
                           double z_center, int bc_r_right, int bc_z_bot, int bc_z_top,
                           double *history, int *steps_done);

/* ---- device-side coefficient refresh -------------------------------------- */
typedef struct {
    double nu_lya;          /* ν_Lyα [Hz] */
    double b_prefact;       /* b(T) = b_prefact √T                (Compute_b) */
    double nu_over_c;       /* ΔνD = b ν_Lyα / c                  (Compute_ΔνD) */
    double a_prefact;       /* a = a_prefact / b                  (Compute_a) */
    double sigma0_prefact;  /* σ0 = sigma0_prefact / b            (Compute_σ0) */
} m1lya_line_constants;

/* The Voigt lookup tables of CreateVoigtLUTs (src/FastInterpolations.jl:26-80): luts is
 * [5][nx], the node values of erf, dawson, exp(-x²), x exp(-x²) and (1-2x²) exp(-x²) on
 * x_i = -xmax + i·2 xmax/(nx-1).  Copied to the device once. */
int m1lya_set_voigt_lut(m1lya_ctx *ctx, const double *luts, int nx, double xmax, const m1lya_line_constants *lc);

/* PrepareLyαSubCycling! (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2090-2155) with
 * Update_Hvoigt_And_ϕν_cell! (:2039-2082) on the device for every cell: from T and n_HI
 * [N_Z][N_r] it recomputes ΔνD, Hvoigt (the bin average Htophat of the Voigt profile,
 * src/ConstantsAndFunctions.jl:265-279), ϕ_ν, α_Lyα = n_HI σ0(T) and α_F = α_Lyα Hvoigt + α_c,F.
 * The thermochemistry terms are inputs: alpha_c (continuum absorption α_c), alpha_c_F
 * (continuum flux opacity, added to every bin of α_F), js (∫dν js) and pd; each may be NULL
 * (α_c, js and pd keep their device values, α_c,F = 0).  phi_all_bins = 1 normalises ϕ_ν over
 * all bins (the RHD form), 0 over the core bins (Update_ϕν!, src/ConstantsAndFunctions.jl:362-393). */
int m1lya_prepare_subcycling(m1lya_ctx *ctx, const double *T, const double *n_HI, const double *alpha_c,
                             const double *alpha_c_F, const double *js, const double *pd, int phi_all_bins);

/* ---- RHD time-step control and the Lyα absorption sub-step ---------------- */
/* GetRHDCFLTimeStep(r, Z, c̃, u_r, u_Z, CFL), src/AdvectJandH.jl:29-68: dt = CFL min over cells
 * of 0.5 min(Δr, ΔZ) / (c̃ + |u|), a device reduction.  u_r, u_z [N_r, N_Z] cell velocities;
 * the cell interfaces r, Z come from m1lya_set_geometry. */
int m1lya_cfl_time_step(m1lya_ctx *ctx, double c_tilde, const double *u_r, const double *u_z, double cfl,
                        double *dt);

/* Inputs of FindRSL_LyαRT: hydro, chemistry and dust state, [N_r, N_Z] each (rho_d
 * [3, N_r, N_Z]: silicate, graphite, PAH dust densities, species fastest, as Dust ρd). */
typedef struct {
    const double *u_r, *u_z, *gamma_ad, *P, *rho, *rho_d, *xHII, *xH2, *T;
    double kappa_abs[3];       /* κd_abs_Lyα of silicate, graphite, PAH dust (DustState) */
    m1lya_line_constants lc;   /* b(T), a(T) and σ0(T) prefactors */
} m1lya_rsl_inputs;

/* FindRSL_LyαRT(c̃_Lyα, c̃_min, Lyα, Hydro, Chem, Dust, Geom),
 * src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:616-722: the largest per-cell RSL
 * floor ϵ_RSL max(1, (aτ)^{1/3} f_abs) max(|u|, c_s) as a device reduction, clamped to
 * [c̃_min, c] and to [0.75, 4] times c_old.  p_d is the device-resident field. */
int m1lya_find_rsl(m1lya_ctx *ctx, double c_old, double c_min, const m1lya_rsl_inputs *in, double *c_new);

/* The Lyα block of AbsorptionStep, src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:1577-1670,
 * for every cell at once: from J_old, H_r_old, H_Z_old (device fields) over each cell's
 * sub-step (dt_cell [N_r, N_Z], or dt for every cell when dt_cell is NULL) it writes
 * J = max(backward Euler, Jeps), H_r, H_Z, and splits the momentum each bin deposits between
 * gas and the dust species by flux opacity (SplitRadPressure, :1017-1048; rho_d [3, N_r, N_Z]
 * may be NULL, kappa_F = the three κd_F_Lyα).  A cell whose e_Lyα (device field M1LYA_E_LYA,
 * before the sub-step) changes by more than tol_Abs = 0.1 while dt > 0.1 / (c̃ (αc + p_d α)), or
 * not finitely, is rejected (RejectAbsorption, :996) and keeps its old J and H
 * (RestoreLyαRejection, :980-990).  Outputs (each may be NULL): redo [N_r, N_Z] (1 = rejected);
 * cell_out [12][N_r, N_Z] = {Δe_Lyα, e_Lyα_new, Δρu_r_Lyα, Δρu_Z_Lyα, then r and Z of the
 * gas, silicate, graphite and PAH shares}, zero but Δe_Lyα on a rejected cell.  e_Lyα itself is
 * left to the caller, as in the reference (stored once the whole sub-step is accepted). */
int m1lya_absorption_substep(m1lya_ctx *ctx, const double *dt_cell, double dt, double c_tilde,
                             const double *rho_d, const double *kappa_F, int *redo, double *cell_out);

/* PrepareLyαSubCycling! (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2086-2155) for every cell
 * from the chemical state, all on the device: ∫dν js from recombinations (∫dν_js_recLyα,
 * src/LymanAlphaEmission.jl:105-124; Case B and the on-the-spot Lyα probability when ots != 0)
 * and collisional excitation (∫dν_js_collLyα, :143-173); p_d (pd_Lyα, src/LymanAlphaDestruction.jl:43-97,
 * at redshift z; 0 when destruction == 0); the continuum absorption α_c and flux opacity α_c,F of
 * dust (densities below 1.1 ρd_floor ignored) and gas (Raman, OH, CH; Thomson in α_c,F); then the
 * refresh of m1lya_prepare_subcycling (ΔνD, Hvoigt, ϕ_ν, α_Lyα = n_HI σ0(T), α_F = α_Lyα Hvoigt + α_c,F).
 * chem [8][N_r, N_Z] = T, n_HI, n_H2, n_HII, n_HeI, n_e, n_OHx, n_CHx; rho_d [3, N_r, N_Z] (may be
 * NULL); kappa = {κd_abs_Lyα, κd_F_Lyα} of silicate, graphite, PAH.  Needs m1lya_set_voigt_lut. */
int m1lya_prepare_subcycling_chem(m1lya_ctx *ctx, const double *chem, const double *rho_d, const double *kappa,
                                  double z, int ots, int destruction, int phi_all_bins);

/* ---- reverse-mode gradient of the fused step ------------------------------
 * The reference has no adjoint; this entry point is what a ChainRules `rrule`
 * around m1lya_run_steps calls.  It runs n_steps fused steps from the
 * context's current state (Δρu starting at zero) and returns the gradient of
 *
 *   Obj = Σ_s Σ_{q<3} mf_seed[3 s + q] · (M_F, M_F_radial_net, M_F_radial_out)_q of step s
 *       + Σ_k spec_seed[k] · S_k of the final state,
 *
 * S_k = 4π Σ_{outer faces} A max(H_out, 0) / L_sim, the escaping spectrum whose
 * Δν-weighted sum is f_esc (Compute_f_esc_Lyα_Cylindrical,
 * src/ConstantsAndFunctions.jl:969-993).  mf_seed [n_steps][3] and spec_seed
 * [N_ν] may be NULL (no such term).  The steps are those of m1lya_run_steps;
 * on return the context holds the state after n_steps and history (may be
 * NULL) the [n_steps][8] step scalars.  Outputs (each may be NULL):
 *   grad_state [3][N_Z][N_r][N_ν]  dObj/d(J, H_r, H_Z) of the initial state;
 *   grad_cell  [5][N_Z][N_r]       dObj/d(α_Lyα, α_c, ∫js, ΔνD, T), where the T entry
 *                                  is the direct (recoil) dependence only;
 *   grad_bin   [3][N_Z][N_r][N_ν]  dObj/d(α_F, Hvoigt, ϕ_ν).
 * The active-bin range of each step is held at its forward value.  At kinks
 * (limiters, floors, clamps) the derivative of the branch taken is returned.
 * Memory: states are checkpointed every checkpoint_every steps (0: ceil(√n))
 * and one segment is replayed at a time: about (n/S + 2 S + 12) state-sized
 * FP64 arrays of N_ν N_r N_Z each (S = checkpoint interval).  N_ν <= 256
 * (M1LYA_ERR_DIMS beyond). */
int m1lya_adjoint_run(m1lya_ctx *ctx, int n_steps, double c_tilde, double dt, double z_center,
                      int bc_r_right, int bc_z_bot, int bc_z_top, const double *mf_seed,
                      const double *spec_seed, int checkpoint_every, double *objective, double *history,
                      double *grad_state, double *grad_cell, double *grad_bin);

/* ---- mixed-precision batched step ---------------------------------------
 * A batch advances many clouds of uneven size (N_r, N_Z >= 3, 3 <= N_ν <= 256),
 * optical depth and frequency resolution by the same fused step as m1lya_step
 * (the RT-only loop body of src/StaticCloudTest.jl:357-392), one launch per step
 * for the whole batch.  Precision: J, H_r, H_Z, α_F, Hvoigt and ϕ_ν are stored
 * in FP32 (J and H divided by a per-cloud scale J0); the flux sweep and its M1
 * closure run in FP32, with the cell-centred closure in FP64 near free
 * streaming (f > 0.9); emission/absorption, the frequency factors and the
 * tridiagonal frequency solve run in FP64; M_F, L, f_esc are accumulated with
 * compensated FP64 sums.  A batch of one cloud is the single-cloud FP32 step.
 * Differences from the FP64 context: the J floor is 1e-30·J0 instead of Jeps;
 * Δρu starts each step at zero (the step resets it, as m1lya_step does); the
 * frequency solve is a cyclic reduction instead of the pivoted gtsv.
 * Field ids are enum m1lya_field; a batch holds J, H_R, H_Z, ALPHA_F, HVOIGT,
 * PHI_NU, ALPHA_LYA, ALPHA_C, PD, JS, T, DNU_D, the velocity gradients,
 * UR_OVER_R, UR_STAR, UZ_STAR (upload) and J, H_R, H_Z, F_LYA_R, F_LYA_Z
 * (download).  Velocity fields default to zero. */
typedef struct m1lya_batch m1lya_batch;

int  m1lya_batch_create(int n_clouds, const int *n_nu, const int *n_r, const int *n_z, m1lya_batch **out);
void m1lya_batch_destroy(m1lya_batch *b);
int  m1lya_batch_last_error(const m1lya_batch *b, char *buf, size_t len);
int  m1lya_batch_count(const m1lya_batch *b);
/* Per-cloud setup, as m1lya_set_geometry / m1lya_set_freq_grid. */
int  m1lya_batch_set_geometry(m1lya_batch *b, int cloud, const m1lya_geometry *geometry);
int  m1lya_batch_set_freq_grid(m1lya_batch *b, int cloud, const m1lya_freq_grid *grid);
/* J0 (default 1): J and H are stored as J/J0, H/J0.  A resident state keeps its
 * physical value when J0 changes.  Choose J0 near the largest J of the cloud. */
int  m1lya_batch_set_scale(m1lya_batch *b, int cloud, double J0);
int  m1lya_batch_get_scale(const m1lya_batch *b, int cloud, double *J0);
int  m1lya_batch_upload(m1lya_batch *b, int cloud, int field, const double *host, size_t count);
int  m1lya_batch_download(m1lya_batch *b, int cloud, int field, double *host, size_t count);
/* c̃, Δt, Z_center and the boundary conditions of one cloud's steps. */
/* The device-side coefficient refresh of one cloud, as m1lya_set_voigt_lut and
 * m1lya_prepare_subcycling; the refreshed Hvoigt, ϕ_ν and α_F are stored in FP32. */
int  m1lya_batch_set_voigt_lut(m1lya_batch *b, const double *luts, int nx, double xmax,
                               const m1lya_line_constants *lc);
int  m1lya_batch_prepare_subcycling(m1lya_batch *b, int cloud, const double *T, const double *n_HI,
                                    const double *alpha_c, const double *alpha_c_F, const double *js,
                                    const double *pd, int phi_all_bins);
int  m1lya_batch_set_step(m1lya_batch *b, int cloud, double c_tilde, double dt, double z_center,
                          int bc_r_right, int bc_z_bot, int bc_z_top);
/* n_steps steps of every cloud.  history (may be NULL) receives
 * [n_steps][n_clouds][8] step scalars laid out as m1lya_step's out, with ΔJ/J (as
 * m1lya_run_until_steady defines it) as the last entry.  Returns
 * M1LYA_ERR_NONFINITE, naming the first cloud and step, if a step wrote a
 * non-finite J, H_r or H_Z. */
int  m1lya_batch_run_steps(m1lya_batch *b, int n_steps, double *history);
/* Steps every cloud until each one's ΔJ/J (as m1lya_run_until_steady defines it, in the
 * cloud's FP32 state) has fallen below tol, or max_steps.  history (may be NULL) receives
 * [*steps_done][n_clouds][8] step scalars (the last entry of each row is ΔJ/J); steady
 * (may be NULL) receives [n_clouds] the step count after which each cloud was first below
 * tol (0: never).  Clouds that are already steady keep stepping until the last one is.
 * The FP32 state's rounding sets a floor of about 1e-7 under ΔJ/J, so tol should be
 * above it. */
int  m1lya_batch_run_until_steady(m1lya_batch *b, int max_steps, double tol, double *history,
                                  int *steady, int *steps_done);
/* [n_clouds][8] scalars of the last step. */
int  m1lya_batch_scalars(m1lya_batch *b, double *out);
/* [n_clouds][2] active-bin ranges the next step uses (0-based, inclusive). */
int  m1lya_batch_active_bins(m1lya_batch *b, int *out);
/* Tile shape: cells per wavefront segment along r (1..64), wavefronts per block (1..8).
 * The state after a step does not depend on the tile shape. */
int  m1lya_batch_set_tiles(m1lya_batch *b, int seg_len, int waves_per_block);
int  m1lya_batch_get_tiles(const m1lya_batch *b, int *seg_len, int *waves_per_block);
/* Times each candidate tile shape over trial_steps steps, in two passes over the
 * candidates, and keeps the fastest; the state is restored afterwards.  Outputs may be NULL. */
int  m1lya_batch_autotune(m1lya_batch *b, int trial_steps, int *seg_len, int *waves_per_block, double *best_ms);
/* on != 0: steps replay a captured hipGraph (one per state-buffer parity). */
int  m1lya_batch_set_graph(m1lya_batch *b, int on);
int  m1lya_batch_synchronize(m1lya_batch *b);
/* on = 1: every launch of a step is bracketed by timing events (m1lya_batch_kernel_stats);
 * on = 2: only the step kernels k_step and k_step_pair (clouds of N_ν <= 32), the step's
 * dominant work (two events per launch instead of around every launch);
 * 0: none.  Graph replay is suspended while profiling. */
int  m1lya_batch_set_profiling(m1lya_batch *b, int on);
int  m1lya_batch_kernel_stats_count(m1lya_batch *b);
int  m1lya_batch_kernel_stats(m1lya_batch *b, int index, char *name, size_t name_len,
                              long *launches, double *total_ms);
int  m1lya_batch_reset_kernel_stats(m1lya_batch *b);

/* How m1lya_adjoint_run differentiates the interfaces of the advection.  0 (default):
 * each interface once in reverse mode, its input cotangents gathered per cell.
 * 1: per cell, dual-number passes through the <= 8 interfaces whose stencil holds it
 * (16 passes; kept to cross-check mode 0).  The two agree to rounding.
 * No counterpart in the reference, which has no adjoint. */
int m1lya_set_adjoint_mode(m1lya_ctx *ctx, int mode);

/* ---- allocation-free entry point ------------------------------------------
 * No counterpart in the reference: FindActiveBins (src/UpdateLyaJandH.jl:47-97, as
 * m1lya_find_active_bins) for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  J [N_ν, N_r, N_Z],
 * dnu [N_ν] and dV [N_r, N_Z] are device pointers the caller owns; kmin/kmax are host
 * outputs (0-based).  The partial sums use a fixed device array of the library (4 MiB per
 * GPU, kept until the process ends), so N_ν is limited to 256 (M1LYA_ERR_DIMS beyond).
 * Runs on the null stream and synchronises it; no device memory is allocated.  Calls that
 * overlap on one GPU share that array and race.  The result is bit-identical to
 * m1lya_find_active_bins on the same state.  m1lya_noalloc_last_error returns the message
 * of its last failure on the calling thread. */
int m1lya_find_active_bins_noalloc(const double *J, const double *dnu, const double *dV, int n_nu, int n_r,
                                   int n_z, int kcore_min, int kcore_max, int *kmin, int *kmax);
int m1lya_noalloc_last_error(char *buf, size_t len);

/* ---- instrumentation ------------------------------------------------- */
/* on = 1: every kernel launch is bracketed by HIP events on the context's
 * stream; on = 2: only the column kernel of the fused step (its dominant
 * kernel), so that the events add little to a timed run; 0: off.
 * m1lya_kernel_stats() reports per-kernel launch counts and total device
 * milliseconds (synchronises). */
int m1lya_set_profiling(m1lya_ctx *ctx, int on);
int m1lya_kernel_stats_count(m1lya_ctx *ctx);
int m1lya_kernel_stats(m1lya_ctx *ctx, int index, char *name, size_t name_len,
                       long *launches, double *total_ms);
int m1lya_reset_kernel_stats(m1lya_ctx *ctx);

/* Runs n IEEE operations on the device for the bit-exactness self-test:
 * op 0: a/b, 1: sqrt(a), 2: the library's exp(a), 3: a plain copy out = a
 * (one kernel reading and writing n doubles; calibrates HBM byte counters). */
int m1lya_selftest_math(int op, const double *a, const double *b, double *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
