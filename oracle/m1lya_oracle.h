/*
 * m1lya_oracle.h -- CPU FP64 restatement of LYDION's multi-frequency M1 Lya step.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library (lydion_amd/libm1lya.so) is compared
 * with this oracle: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product path may
 * link, import or call it.
 *
 * Every function follows the reference Julia source line by line, in the
 * written operation order, compiled with -ffp-contract=off.  Citations are
 * file:line into the olofnebrin/LYDION repository.
 *
 * Memory layout is the reference's: Julia column-major [k,i,j] with k
 * (frequency) fastest == C row-major [j][i][k].  All indices here are 0-based.
 *
 *   cell arrays     [n_z][n_r][n_nu]
 *   r-face arrays   [n_z][n_r+1][n_nu]     (face f = Julia interface r[f+1])
 *   z-face arrays   [n_z+1][n_r][n_nu]
 *   2-D cell arrays [n_z][n_r];  A_r/u_r_star [n_z][n_r+1];  A_z/u_z_star [n_z+1][n_r]
 */
#ifndef M1LYA_ORACLE_H
#define M1LYA_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_REFLECT = 0, ORC_OUTFLOW = 1 };

/* GeometryCylindrical{Float64} (src/Structs.jl:650-740), M1-relevant fields. */
typedef struct {
    int n_r, n_z;
    double *r, *z;            /* [n_r+1], [n_z+1] interfaces            */
    double *rcom, *zcom;      /* [n_r], [n_z] centres of mass           */
    double *dV, *inv_dV;      /* [n_z][n_r]                             */
    double *A_r;              /* [n_z][n_r+1]                           */
    double *A_z;              /* [n_z+1][n_r]                           */
    double *inv_r_cell;       /* [n_r]  <1/r>                           */
    /* interval arrays [n_r]: dL, dR, inv(i,i-1), inv(i+1,i), inv(i+2,i+1) */
    double *dr_L, *dr_R, *inv_dr_i_im1, *inv_dr_ip1_i, *inv_dr_ip2_ip1;
    double dr_inner[5], dr_outer[5];  /* same order, for faces 1/2 and N+1/2 */
    double *dz_L, *dz_R, *inv_dz_j_jm1, *inv_dz_jp1_j, *inv_dz_jp2_jp1;
    double dz_inner[5], dz_outer[5];
} orc_geom;

/* LyαArrays{FluxMethod} (src/Structs.jl:23-124). */
typedef struct {
    int n_nu;
    int kcore_min, kcore_max;   /* 0-based, inclusive                       */
    int recoil;                 /* RECOIL_ON                                */
    int flux_hll;               /* 0 = Val(:GLF), 1 = Val(:HLL)             */
    double alpha_floor;         /* α_Lyα_floor                              */
    double *nu;                 /* [n_nu+1]                                 */
    double *nuC, *dnu, *inv_dnu, *inv_nuC;   /* [n_nu]                      */
    double *J, *J_old, *H_r, *H_z, *H_r_old, *H_z_old;   /* cell arrays     */
    double *Kphi_ip, *Kphi_im;                           /* cell arrays     */
    double *F_r, *F_r_Hr, *F_r_HZ;                       /* r-face arrays   */
    double *F_z, *F_z_Hr, *F_z_HZ;                       /* z-face arrays   */
    double *alpha_F, *Hvoigt, *phi;                      /* cell arrays     */
    double *Fkp, *Fk, *Fkm;                              /* cell arrays     */
    double *alpha, *alpha_c, *pd, *js;                   /* [n_z][n_r]      */
    double *dnuD;                                        /* [n_z][n_r]      */
    double *drhou_r, *drhou_z, *f_r, *f_z, *e_lya;       /* [n_z][n_r]      */
} orc_lya;

/* Gas velocity field passed to the advection / frequency operators. */
typedef struct {
    double *u_r_star;   /* [n_z][n_r+1] */
    double *u_z_star;   /* [n_z+1][n_r] */
    double *dur_dr, *dur_dz, *duz_dr, *duz_dz, *ur_over_r;   /* [n_z][n_r] */
} orc_vel;

/* --- math primitives -------------------------------------------------- */
double orc_exp(double x);
double orc_minmod(double a, double b);
double orc_chi_L84(double f);
void   orc_D_edd(double H_r, double H_z, double J, double out[3]);
void   orc_D_edd_from_f_n(double f, double nr, double nz, double out[3]);
void   orc_hll_eigenvalues(double f, double mu, double out[2]);
void   orc_flux_glf(const double in[16], double out[3]);
void   orc_flux_hll(const double in[16], const double l[4], double out[3]);
void   orc_reconstruct_interface(const double in[19], double out[14]);
double orc_Kphiphi_inner_axis(double J_axis, double HZ_axis);

/* --- tridiagonal ------------------------------------------------------- */
int    orc_gtsv(int n, double *dl, double *d, double *du, double *b);
void   orc_thomas(int n, const double *a, const double *b, const double *c,
                  const double *d, double *cp, double *dp, double *x);

/* --- hot-path operators ------------------------------------------------ */
void orc_find_active_bins(const orc_lya *L, const orc_geom *G, int *kmin, int *kmax);
void orc_advect_jandh_cyl(double cdt, orc_lya *L, const orc_geom *G, const orc_vel *V,
                          int update_jold, int bc_r_right, int bc_z_bot, int bc_z_top,
                          int kmin, int kmax);
void orc_advect_lya(double cdt, orc_lya *L, const orc_geom *G, const orc_vel *V,
                    int bc_r_left, int bc_r_right, int bc_z_bot, int bc_z_top,
                    int *kmin_out, int *kmax_out);
void orc_emission_absorption(orc_lya *L, int n_r, int n_z, double c_tilde, double dt);
void orc_update_freq_factors(orc_lya *L, int n_r, int n_z, const double *T, const orc_vel *V);
void orc_frequency_update_all_cells(double cdt, orc_lya *L, int n_r, int n_z,
                                    const double *T, const orc_vel *V, int update_ff);
void orc_update_f_mf_l(orc_lya *L, const orc_geom *G, double dt, double z_center,
                       int bc_z_bot, double out[4]);
void orc_update_f_mf_l_grav(orc_lya *L, const orc_geom *G, double dt,
                            const double *f_grav_r, const double *f_grav_z,
                            double z_center, int bc_z_bot, double out[5]);
double orc_f_esc_lya(const orc_lya *L, const orc_geom *G, double L_allspace, int bc_z_bot);
void orc_reset_drhou(orc_lya *L, int n_r, int n_z);
void orc_update_e_lya(orc_lya *L, int n_r, int n_z);
int  orc_edge_count(int n_r, int n_z);
void orc_edge_spectra(const orc_lya *L, int n_r, int n_z, double *raw, double *norm);

/* One time step of the RT-only static-cloud loop (src/StaticCloudTest.jl:357-392). */
void orc_static_cloud_step(double c_tilde, double dt, orc_lya *L, const orc_geom *G,
                           const orc_vel *V, const double *T, double z_center,
                           int bc_z_bot, double out[8]);

/* --- setup (inputs of the hot path) ------------------------------------ */
void orc_cyl_grid_uniform(double rmax, double zmax, orc_geom *G);
void orc_grid_differences(int n, const double *x, const double *xcom,
                          double *dL, double *dR, double *inv_i_im1, double *inv_ip1_i,
                          double *inv_ip2_ip1, double inner[5], double outer[5]);
int  orc_lya_freq_grid(double Tmin, double q_fine, double q_wing, double *x_out, int cap);
double orc_compute_b(double T);
double orc_compute_dnuD(double T);
double orc_compute_a(double T);
double orc_compute_sigma0(double T);
double orc_lut_eval(const double *y, int nx, double xmin, double inv_dx, double x);
void orc_update_hvoigt(double *Hvoigt, int n_nu, int n_cells, const double *dnu_rel,
                       const double *dnuC_rel, const double *a, const double *inv_dnuD,
                       const double *luts /* 5 tables of nx */, int nx, double lut_xmax);
void orc_update_phi(double *phi, int n_nu, int n_cells, const double *dnu,
                    const double *Hvoigt, int kmin, int kmax);
double orc_cfl_dt(const orc_geom *G, double c_tilde, const double *u_r, const double *u_z, double cfl);
double orc_find_rsl(double c_old, double c_min, const orc_geom *G, const double *u_r, const double *u_z,
                    const double *gamma_ad, const double *P, const double *rho, const double *rho_d,
                    const double *xHII, const double *xH2, const double *T, const double *pd,
                    const double *kappa_abs);
void orc_absorption_substep(orc_lya *L, int n_r, int n_z, const double *dt, double c_tilde,
                            const double *rho_d, const double *kappa_F, int *redo, double *out);
void orc_set_threads(int n);
double orc_prob_lya_sh95(double T, double ne, int caseB);
double orc_krec_hii(double T, int caseB);
double orc_js_rec_lya(double T, double nHII, double ne, int ots);
double orc_js_coll_lya(double T, double nHI, double ne);
double orc_pd_lya(double nHII, double ne, double nHI, double nHeI, double nH2, double T, double z);
void orc_lya_sources(int n_cells, const double *chem, const double *rho_d, const double *kappa, double z, int ots,
                     int destruction, double *out);
void orc_hubble_flow(double r, double z, double umax, double rcl, double zc, double out[7]);
double orc_cone_factor(double mu, double mucone, double taucone, double taucl,
                       int smooth, double dmu);

#ifdef __cplusplus
}
#endif
#endif
