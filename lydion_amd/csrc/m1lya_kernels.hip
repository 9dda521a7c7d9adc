// m1lya_kernels.hip -- HIP kernels of the FP64 M1 Lya step for gfx950.
//
// Layout: spectral arrays are [j][i][k] (k = frequency bin fastest), exactly the
// Julia column-major [k,i,j] of the reference, so a wavefront of 64 lanes on
// consecutive k reads 512 contiguous bytes per FP64 field.
#include "m1lya_internal.h"

namespace m1 {

// ------------------------------------------------------------------------
// Wave helpers (64-wide wavefronts)
// ------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------------
// FindActiveBins partial sums: src/UpdateLyaJandH.jl:61-90.
// Block b sums over its cell range, per bin k:
//   e1[k]  = Σ Δν[k] * max(J - Jeps, 0) * ΔV
//   e10[k] = Σ Δν[k] * max(J - 10 Jeps, 0) * ΔV
// part[b][0..nn) = e1, part[b][nn..2nn) = e10 (row stride = pstride).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_active_partial(const double* __restrict__ J,
                                                         const double* __restrict__ dnu,
                                                         const double* __restrict__ dV, int nn,
                                                         int ncells, int cells_per_block,
                                                         double* __restrict__ part, int pstride)
{
    __shared__ double s1[256], s2[256];
    const int kw = ((nn + 63) / 64) * 64 > 256 ? 256 : ((nn + 63) / 64) * 64;
    const int t = threadIdx.x;
    const int sub = t / kw, nsub = 256 / kw;
    const int c0 = blockIdx.x * cells_per_block;
    const int c1 = min(c0 + cells_per_block, ncells);
    for (int kb = 0; kb < nn; kb += kw) {
        const int k = kb + (t % kw);
        double a1 = 0.0, a10 = 0.0;
        if (k < nn && sub < nsub) {
            const double d = dnu[k];
            for (int c = c0 + sub; c < c1; c += nsub) {
                double Jv = J[(size_t)c * nn + k];
                double v = dV[c];
                a1 += d * dmax(Jv - kJeps, 0.0) * v;
                a10 += d * dmax(Jv - 10 * kJeps, 0.0) * v;
            }
        }
        s1[t] = a1; s2[t] = a10;
        __syncthreads();
        if (t < kw && k < nn) {
            double r1 = 0.0, r10 = 0.0;
            for (int s = 0; s < nsub; ++s) { r1 += s1[s * kw + t]; r10 += s2[s * kw + t]; }
            part[(size_t)blockIdx.x * pstride + k] = r1;
            part[(size_t)blockIdx.x * pstride + nn + k] = r10;
        }
        __syncthreads();
    }
}

// Column sums of a [nrows][stride] partial array: out[col] = Σ_rows part[row][col].
// One block per column, fixed-shape tree -> deterministic.
__global__ __launch_bounds__(256) void k_reduce_cols(const double* __restrict__ part, int nrows,
                                                      int stride, double* __restrict__ out)
{
    __shared__ double s[256];
    const int col = blockIdx.x;
    double a = 0.0;
    for (int r = threadIdx.x; r < nrows; r += 256) a += part[(size_t)r * stride + col];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[col] = s[0];
}

// Active-bin decision (src/UpdateLyaJandH.jl:73-96) from reduced e1/e10.
__device__ void active_range(const double* red, int nn, int kcore_min, int kcore_max, int* krange)
{
    double Etot = 0.0;
    for (int k = 0; k < nn; ++k) Etot += red[k];
    int kmn = nn - 1, kmx = 0;
    for (int k = 0; k < nn; ++k) {
        if (red[nn + k] > kEpsActive * Etot) {
            if (k < kmn) kmn = k;
            if (k > kmx) kmx = k;
        }
    }
    if (kcore_max > kmx) kmx = kcore_max;
    if (kcore_min < kmn) kmn = kcore_min;
    krange[0] = kmn;
    krange[1] = kmx;
}

__global__ void k_active_final(const double* __restrict__ red, int nn, int kcore_min, int kcore_max,
                               int* __restrict__ krange)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) active_range(red, nn, kcore_min, kcore_max, krange);
}

// ------------------------------------------------------------------------
// Ghost-aware stencil loads for the r and Z sweeps.  The ghost values are the
// ones the reference builds at each boundary:
//   r axis (i < 0): mirror, H_r negated                  src/AdvectJandH.jl:666-673
//   r outer (i >= N_r): :Reflect mirror / :Outflow copy with H_r = max(H_r,0)  :853-871
//   Z bottom (j < 0): :Reflect mirror / :Outflow copy with H_Z = min(H_Z,0)    :1000-1018
//   Z top (j >= N_Z): :Reflect mirror / :Outflow copy with H_Z = max(H_Z,0)    :1185-1203
// ------------------------------------------------------------------------
struct SpecIn {
    const double* J;
    const double* Hr;
    const double* Hz;
};

__device__ __forceinline__ void load_r(const SpecIn& s, int nn, int nr, int j, int k, int c, int bc_right,
                                       double& J, double& Hr, double& Hz)
{
    if (c < 0) {
        size_t idx = ((size_t)j * nr + (-c - 1)) * nn + k;
        J = s.J[idx]; Hr = -s.Hr[idx]; Hz = s.Hz[idx];
    } else if (c < nr) {
        size_t idx = ((size_t)j * nr + c) * nn + k;
        J = s.J[idx]; Hr = s.Hr[idx]; Hz = s.Hz[idx];
    } else if (bc_right == BC_REFLECT) {
        size_t idx = ((size_t)j * nr + (2 * nr - 1 - c)) * nn + k;
        J = s.J[idx]; Hr = -s.Hr[idx]; Hz = s.Hz[idx];
    } else {
        size_t idx = ((size_t)j * nr + (nr - 1)) * nn + k;
        J = s.J[idx]; Hr = dmax(s.Hr[idx], 0.0); Hz = s.Hz[idx];
    }
}

__device__ __forceinline__ void load_z(const SpecIn& s, int nn, int nr, int nz, int i, int k, int c,
                                       int bc_bot, int bc_top, double& J, double& Hr, double& Hz)
{
    if (c < 0) {
        if (bc_bot == BC_REFLECT) {
            size_t idx = ((size_t)(-c - 1) * nr + i) * nn + k;
            J = s.J[idx]; Hr = s.Hr[idx]; Hz = -s.Hz[idx];
        } else {
            size_t idx = (size_t)i * nn + k;
            J = s.J[idx]; Hr = s.Hr[idx]; Hz = dmin(s.Hz[idx], 0.0);
        }
    } else if (c < nz) {
        size_t idx = ((size_t)c * nr + i) * nn + k;
        J = s.J[idx]; Hr = s.Hr[idx]; Hz = s.Hz[idx];
    } else if (bc_top == BC_REFLECT) {
        size_t idx = ((size_t)(2 * nz - 1 - c) * nr + i) * nn + k;
        J = s.J[idx]; Hr = s.Hr[idx]; Hz = -s.Hz[idx];
    } else {
        size_t idx = ((size_t)(nz - 1) * nr + i) * nn + k;
        J = s.J[idx]; Hr = s.Hr[idx]; Hz = dmax(s.Hz[idx], 0.0);
    }
}

// ------------------------------------------------------------------------
// AdvectJandH_Cylindrical!: src/AdvectJandH.jl:526-1390.
// A block is 64 bins x BI cells along r x BJ cells along Z, one wavefront per
// cell.  Each interface is evaluated once: a wavefront evaluates its cell's left
// r face and bottom Z face, the block's last column and row (or the cloud's last
// cell) also the right / top face, and the results go through LDS to the two
// cells that share them.  Then each thread applies its cell update (:1302-1382).
// Reads the state, writes the new state to separate buffers (all fluxes of a
// step are computed from the same old state, as in the reference).  Bins outside
// [kmin, kmax] are copied.  A face's value does not depend on which cell's
// wavefront evaluates it, so the result is the one of a per-cell evaluation.
// ------------------------------------------------------------------------
template <int BI, int BJ>
__global__ __launch_bounds__(64 * BI * BJ) void k_advect(AdvectArgs a)
{
    // r faces of the block's rows: F_J, F_Hr, F_HZ, K_ϕϕ left, K_ϕϕ right (the inner-axis
    // value at f = 0); Z faces of its columns: F_J, F_Hr, F_HZ
    __shared__ double rF[BJ][BI + 1][5][64];
    __shared__ double zF[BJ + 1][BI][3][64];
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    const int lane = threadIdx.x, il = threadIdx.y, jl = threadIdx.z;
    const int k = blockIdx.x * 64 + lane;
    const int i = blockIdx.y * BI + il;
    const int j = blockIdx.z * BJ + jl;
    const bool cell_ok = k < nn && i < nr && j < nz;
    const int kmin = a.krange[0], kmax = a.krange[1];
    const bool act = cell_ok && k >= kmin && k <= kmax;
    const SpecIn s{a.J, a.Hr, a.Hz};

    auto r_face = [&](int f, int slot) {
        Stencil st;
        const int cl = f - 1;
        const int cells[4] = {cl, cl + 1, cl + 2, cl - 1};
#pragma unroll
        for (int m = 0; m < 4; ++m) load_r(s, nn, nr, j, k, cells[m], a.bc_r_right, st.J[m], st.Hr[m], st.Hz[m]);
        const double aFL = a.aF[((size_t)j * nr + max(f - 1, 0)) * nn + k];
        const double aFR = a.aF[((size_t)j * nr + min(f, nr - 1)) * nn + k];
        const double u = a.v.ur_star[(size_t)j * (nr + 1) + f] * kInvC;
        const FaceResult r = interface_flux(0, a.hll, st, a.g.ivr + (size_t)f * 5, aFL, aFR, u);
        double* o = &rF[jl][slot][0][lane];
        o[0] = r.F.J; o[64] = r.F.Hr; o[128] = r.F.Hz;
        o[192] = r.kphi_L;
        o[256] = (f == 0) ? kphi_inner_axis(r.J_R, r.Hz_R) : r.kphi_R;
    };
    auto z_face = [&](int g, int slot) {
        Stencil st;
        const int cl = g - 1;
        const int cells[4] = {cl, cl + 1, cl + 2, cl - 1};
#pragma unroll
        for (int m = 0; m < 4; ++m)
            load_z(s, nn, nr, nz, i, k, cells[m], a.bc_z_bot, a.bc_z_top, st.J[m], st.Hr[m], st.Hz[m]);
        const double aFL = a.aF[((size_t)max(g - 1, 0) * nr + i) * nn + k];
        const double aFR = a.aF[((size_t)min(g, nz - 1) * nr + i) * nn + k];
        const double u = a.v.uz_star[(size_t)g * nr + i] * kInvC;
        const FaceResult r = interface_flux(1, a.hll, st, a.g.ivz + (size_t)g * 5, aFL, aFR, u);
        double* o = &zF[slot][il][0][lane];
        o[0] = r.F.J; o[64] = r.F.Hr; o[128] = r.F.Hz;
    };
    if (act) {
        r_face(i, il);
        if (il == BI - 1 || i == nr - 1) r_face(i + 1, il + 1);
        z_face(j, jl);
        if (jl == BJ - 1 || j == nz - 1) z_face(j + 1, jl + 1);
    }
    __syncthreads();
    if (!cell_ok) return;
    const size_t idx = ((size_t)j * nr + i) * nn + k;
    if (!act) {
        a.J_out[idx] = s.J[idx];
        a.Hr_out[idx] = s.Hr[idx];
        a.Hz_out[idx] = s.Hz[idx];
        return;
    }
    const double FrL0 = rF[jl][il][0][lane], FrL1 = rF[jl][il][1][lane], FrL2 = rF[jl][il][2][lane];
    const double FrR0 = rF[jl][il + 1][0][lane], FrR1 = rF[jl][il + 1][1][lane], FrR2 = rF[jl][il + 1][2][lane];
    const double FzL0 = zF[jl][il][0][lane], FzL1 = zF[jl][il][1][lane], FzL2 = zF[jl][il][2][lane];
    const double FzR0 = zF[jl + 1][il][0][lane], FzR1 = zF[jl + 1][il][1][lane], FzR2 = zF[jl + 1][il][2][lane];
    const double kphi_im = rF[jl][il][4][lane];
    const double kphi_ip = rF[jl][il + 1][3][lane];

    // ---- cell update: src/AdvectJandH.jl:1319-1379 ----
    const int c = j * nr + i;
    const double Ar_L = a.g.A_r[(size_t)j * (nr + 1) + i], Ar_R = a.g.A_r[(size_t)j * (nr + 1) + i + 1];
    const double Az_L = a.g.A_z[(size_t)j * nr + i], Az_R = a.g.A_z[(size_t)(j + 1) * nr + i];
    const double cdt = a.cdt;
    const double cdt_dV = cdt * a.g.inv_dV[c];
    const double cdt_over_c = cdt * kInvC;
    const double durdr = a.v.dur_dr[c], durdz = a.v.dur_dz[c], duzdr = a.v.duz_dr[c];
    const double duzdz = a.v.duz_dz[c], urr = a.v.ur_over_r[c];
    const double Js = s.J[idx], Hrs = s.Hr[idx], Hzs = s.Hz[idx];
    Edd D = D_edd(Hrs, Hzs, Js);
    double Jn = Js - cdt_dV * ((Ar_R * FrR0 - Ar_L * FrL0) + (Az_R * FzR0 - Az_L * FzL0));
    Jn = dmax(Jn, kJeps);
    double Hzn = Hzs - cdt_dV * ((Ar_R * FrR2 - Ar_L * FrL2) + (Az_R * FzR2 - Az_L * FzL2));
    double Hrn = Hrs - cdt_dV * ((Ar_R * FrR1 - Ar_L * FrL1) + (Az_R * FzR1 - Az_L * FzL1));
    double Ddu = D_colon_grad_u(D, durdr, durdz, duzdr, duzdz, urr);
    double Kav = 0.5 * (kphi_im + kphi_ip);
    Hrn += cdt * Kav * a.g.inv_r_cell[i];
    double dHr = -cdt_over_c * (Hrs * durdr + Hzs * durdz);
    double dHz = -cdt_over_c * (Hrs * duzdr + Hzs * duzdz);
    Hrn = Hrn + dHr;
    Hzn = Hzn + dHz;
    Jn = dmax(Jn - cdt_over_c * Ddu * Js, kJeps);
    // EnforceCausalReducedFlux: src/AdvectJandH.jl:478-493
    double Jc = dmax(Jn, kJeps);
    double Hm2 = Hrn * Hrn + Hzn * Hzn;
    if (Hm2 > Jc * Jc) {
        double fac = Jc / sqrt(Hm2);
        Hrn *= fac;
        Hzn *= fac;
    }
    a.J_out[idx] = Jn;
    a.Hr_out[idx] = Hrn;
    a.Hz_out[idx] = Hzn;
}

// ------------------------------------------------------------------------
// LAPACK dgtsv (NRHS = 1, published reference algorithm: elimination with
// partial pivoting by row interchange) on one LDS-resident column, as called by
// TridiagonalSolve (src/TridiagonalSolver.jl:54).  Row i of each array is at
// [i]; the solution overwrites b.  The current and next rows are carried in
// registers, so the dependent chain is arithmetic only; every value stored
// is the one LAPACK would store, computed by the same operations.
// ------------------------------------------------------------------------
__device__ int gtsv_col(int n, double* __restrict__ dl, double* __restrict__ d, double* __restrict__ du,
                        double* __restrict__ b)
{
    double d_i = d[0], b_i = b[0];
    double du_i = (n > 1) ? du[0] : 0.0;
    for (int i = 0; i < n - 1; ++i) {
        const bool last = (i == n - 2);
        const double dl_i = dl[i];
        double d_n = d[i + 1], b_n = b[i + 1];
        double du_n = last ? 0.0 : du[i + 1];
        if (fabs(d_i) >= fabs(dl_i)) {
            if (d_i == 0.0) return i + 1;
            double fact = dl_i / d_i;
            d_n = d_n - fact * du_i;
            b_n = b_n - fact * b_i;
            if (!last) dl[i] = 0.0;
        } else {
            double fact = d_i / dl_i;
            double new_d_i = dl_i;
            double temp = d_n;
            d_n = du_i - fact * temp;
            if (!last) {
                dl[i] = du_n;             // fill-in of the second super-diagonal
                du_n = -fact * du_n;
            }
            du_i = temp;
            double tb = b_i;
            b_i = b_n;
            b_n = tb - fact * b_n;
            d_i = new_d_i;
        }
        d[i] = d_i; du[i] = du_i; b[i] = b_i;
        d_i = d_n; b_i = b_n; du_i = du_n;
    }
    if (d_i == 0.0) return n;
    d[n - 1] = d_i;
    double x1 = b_i / d_i;
    b[n - 1] = x1;
    if (n == 1) return 0;
    double x0 = (b[n - 2] - du[n - 2] * x1) / d[n - 2];
    b[n - 2] = x0;
    for (int i = n - 3; i >= 0; --i) {
        double x = (b[i] - du[i] * x0 - dl[i] * x1) / d[i];
        b[i] = x;
        x1 = x0;
        x0 = x;
    }
    return 0;
}

// ------------------------------------------------------------------------
// Column kernel: one wavefront owns W cell columns at a time.
//   P1 (lanes over k): emission/absorption (src/ConstantsAndFunctions.jl:483-573),
//       frequency factors (src/FreqFactors.jl:142-278), tridiagonal rows to LDS
//       (src/UpdateLyaJandH.jl:178-227);
//   P2 (lane per column): dgtsv on the LDS column;
//   P3 (lanes over k): J floor, H rescale and causality, old copies
//       (src/UpdateLyaJandH.jl:335-379), plus the fused step's reductions.
// LDS per block: 4 arrays x (N_ν-2) rows x W columns (column index fastest,
// so the lane-per-column sweep is bank-conflict free).
// ------------------------------------------------------------------------
template <int W, int KPL>
__global__ __launch_bounds__(64) void k_column(ColumnArgs a)
{
    extern __shared__ double lds[];
    const int nn = a.nn, nr = a.nr, nz = a.nz;
    const int n = nn - 2;
    const int lane = threadIdx.x;
    const int flags = a.flags;
    const bool EMIT = flags & COL_EMIT, FREQ = flags & COL_FREQ, STOREF = flags & COL_STORE_F;
    const bool SOLVE = flags & COL_SOLVE, HCORR = flags & COL_HCORR, FUSED = flags & COL_FUSED;
    // [column][row] with an odd row stride: conflict-free for both the lanes-over-k
    // writes (consecutive rows) and the lane-per-column sweep (stride S, S odd).
    const int S = col_stride(nn);
    double* sDL = lds;
    double* sD = lds + (size_t)S * W;
    double* sDU = lds + (size_t)2 * S * W;
    double* sB = lds + (size_t)3 * S * W;
    const int ncells = nr * nz;
    const int ngroups = (ncells + W - 1) / W;
    const double ct = a.c_tilde, dt = a.dt, dw = a.dw;
    const double inv_ct = 1.0 / ct;
    const double fpc_ict = (kFourPi / kC) * inv_ct;
    const double inv_dt = 1.0 / dt;

    double e1[KPL], e10[KPL];
#pragma unroll
    for (int kk = 0; kk < KPL; ++kk) { e1[kk] = 0.0; e10[kk] = 0.0; }
    double acc_mf = 0.0, acc_L = 0.0, acc_net = 0.0, acc_out = 0.0, acc_esc = 0.0;

    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        double Jo[W][KPL], Hr[W][KPL], Hz[W][KPL];
        // ================= P1 =================
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int c = grp * W + w;
            if (c >= ncells) continue;
            const size_t base = (size_t)c * nn;
            double sr = 0.0, sz = 0.0;
            double al_pd = 0.0, js_ct = 0.0, ac = 0.0;
            if (EMIT) {
                ac = a.alpha_c[c];
                al_pd = a.alpha[c] * a.pd[c];
                js_ct = a.js[c] * ct;
            }
            double Jcur[KPL];
#pragma unroll
            for (int kk = 0; kk < KPL; ++kk) {
                const int k = lane + 64 * kk;
                if (k >= nn) continue;
                const size_t idx = base + k;
                double jc = a.J[idx];
                double jo = FUSED ? jc : a.J_old[idx];
                double hr = a.Hr[idx], hz = a.Hz[idx];
                if (EMIT) {
                    // AbsorptionUpdateJorH (src/ConstantsAndFunctions.jl:451-462)
                    double aJ = ac + al_pd * a.Hv[idx];
                    double dJdt = js_ct * a.phi[idx];
                    double jn = (jo + dJdt * dt) / (1.0 + aJ * ct * dt);
                    double aH = a.aF[idx];
                    double hrn = (hr + 0.0 * dt) / (1.0 + aH * ct * dt);
                    double hzn = (hz + 0.0 * dt) / (1.0 + aH * ct * dt);
                    const double dk = a.dnu[k];
                    sr += dk * (hr - hrn);
                    sz += dk * (hz - hzn);
                    jo = jn; jc = jn; hr = hrn; hz = hzn;
                }
                Jo[w][kk] = jo; Hr[w][kk] = hr; Hz[w][kk] = hz; Jcur[kk] = jc;
            }
            if (EMIT) {
                sr = wave_sum(sr);
                sz = wave_sum(sz);
                if (!FUSED) {
                    if (lane == 0) {
                        a.drhou_r[c] += fpc_ict * sr;
                        a.drhou_z[c] += fpc_ict * sz;
                    }
                } else {
                    // Update_f_Lyα_MF_L_Lyα! per-cell terms (src/ConstantsAndFunctions.jl:612-660)
                    const int i = c % nr, j = c / nr;
                    double dr_u = a.drhou_r[c] + fpc_ict * sr;
                    double dz_u = a.drhou_z[c] + fpc_ict * sz;
                    double fr = dr_u * inv_dt, fz = dz_u * inv_dt;
                    if (lane == 0) {
                        a.f_r[c] = fr; a.f_z[c] = fz;
                        a.drhou_r[c] = 0.0; a.drhou_z[c] = 0.0;
                    }
                    const double dV = a.dV[c];
                    acc_mf += dV * sqrt(fr * fr + fz * fz);
                    acc_L += kFourPi * dV * a.js[c];
                    double dr = a.rcom[i], dz = a.zcom[j] - a.z_center;
                    double dist = sqrt(dr * dr + dz * dz);
                    double frad = 0.0;
                    if (dist > 0.0) {
                        double inv_d = 1.0 / dist;
                        frad = fr * (dr * inv_d) + fz * (dz * inv_d);
                    }
                    acc_net += dV * frad;
                    acc_out += dV * dmax(frad, 0.0);
                }
            }
            if (FREQ || SOLVE) {
                const double dnuD = a.dnuD ? a.dnuD[c] : 0.0;
                const double PREFACT = FREQ ? 0.5 * dmax(a.alpha[c], a.alpha_floor) * (dnuD * dnuD) : 0.0;
                const double RECOIL_0 = FREQ ? kHoverKB / a.T[c] : 0.0;
                const double durdr = FREQ ? a.dur_dr[c] : 0.0, durdz = FREQ ? a.dur_dz[c] : 0.0;
                const double duzdr = FREQ ? a.duz_dr[c] : 0.0, duzdz = FREQ ? a.duz_dz[c] : 0.0;
                const double urr = FREQ ? a.ur_over_r[c] : 0.0;
#pragma unroll
                for (int kk = 0; kk < KPL; ++kk) {
                    const int k = lane + 64 * kk;
                    if (k < 1 || k > nn - 2) continue;
                    const size_t idx = base + k;
                    double Fkp, Fk, Fkm;
                    if (FREQ) {
                        const double* nu = a.nu;
                        const double* nuC = a.nuC;
                        double nuCk = nuC[k], nuCkp = nuC[k + 1], nuCkm = nuC[k - 1];
                        double dkp = nuCkp - nu[k + 1];
                        double dk = nu[k + 1] - nuCk;
                        double dkm = nu[k] - nuCkm;
                        double inv_dnu_k = a.inv_dnu[k];
                        double two_k = 2.0 * a.inv_nuC[k];
                        double two_kp = 2.0 * a.inv_nuC[k + 1];
                        double two_km = 2.0 * a.inv_nuC[k - 1];
                        Edd D = D_edd(Hr[w][kk], Hz[w][kk], Jcur[kk]);
                        double Ddu = D_colon_grad_u(D, durdr, durdz, duzdr, duzdz, urr);
                        double VEL = Ddu * kInvC;
                        double Kkp = PREFACT * a.Hv[idx + 1], Kk = PREFACT * a.Hv[idx];
                        double Kkm = PREFACT * a.Hv[idx - 1];
                        double Ckp = VEL * nuCkp, Ck = VEL * nuCk, Ckm = VEL * nuCkm;
                        if (a.recoil) {
                            Ckp += Kkp * (RECOIL_0 - two_kp);
                            Ck += Kk * (RECOIL_0 - two_k);
                            Ckm += Kkm * (RECOIL_0 - two_km);
                        }
                        double qkp = dkp * (1.0 / Kkp), qk = dk * (1.0 / Kk), qkm = dkm * (1.0 / Kkm);
                        double Pkp = dclamp(Ckp * qkp, -kPmax, kPmax);
                        double Pk = dclamp(Ck * qk, -kPmax, kPmax);
                        double Pkm = dclamp(Ckm * qkm, -kPmax, kPmax);
                        double oRkp, xRkp, oLk, xLk, oRk, xRk, oLkm, xLkm;
                        one_minus_phi_R(Pkp, oRkp, xRkp);
                        one_minus_phi_L(Pk, oLk, xLk);
                        one_minus_phi_R(Pk, oRk, xRk);
                        one_minus_phi_L(Pkm, oLkm, xLkm);
                        double den_kp = inv_dnu_k * (1.0 / (xLk * qk - xRkp * qkp));
                        Fkp = oRkp * den_kp;
                        double den_km = inv_dnu_k * (1.0 / (xLkm * qkm - xRk * qk));
                        Fkm = oLkm * den_km;
                        Fk = oLk * den_kp + oRk * den_km;
                        if (STOREF) { a.Fkp[idx] = Fkp; a.Fk[idx] = Fk; a.Fkm[idx] = Fkm; }
                    } else {
                        Fkp = a.Fkp[idx]; Fk = a.Fk[idx]; Fkm = a.Fkm[idx];
                    }
                    if (SOLVE) {
                        // rows of FrequencyUpdateJLyα! (src/UpdateLyaJandH.jl:200-215) in gtsv form:
                        // dl <- A[2:end], du <- C[1:end-1] (src/TridiagonalSolver.jl:45-46)
                        const int row = k - 1;
                        sD[(size_t)w * S + row] = -(Fk + dw);
                        sB[(size_t)w * S + row] = -dw * Jo[w][kk];
                        if (row >= 1) sDL[(size_t)w * S + row - 1] = Fkm;
                        if (row <= n - 2) sDU[(size_t)w * S + row] = Fkp;
                    }
                }
            }
        }
        // ================= P2 =================
        if (SOLVE) {
            __syncthreads();
            const int c = grp * W + lane;
            if (lane < W && c < ncells && n > 0) {
                int info = gtsv_col(n, sDL + (size_t)lane * S, sD + (size_t)lane * S, sDU + (size_t)lane * S,
                                    sB + (size_t)lane * S);
                if (info != 0) atomicMax(a.err_flag, 1);
            }
            __syncthreads();
        }
        // ================= P3 =================
        if (!(EMIT || SOLVE || HCORR)) continue;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int c = grp * W + w;
            if (c >= ncells) continue;
            const size_t base = (size_t)c * nn;
            const int i = c % nr, j = c / nr;
            const double dV = a.dV ? a.dV[c] : 0.0;
            double esc = 0.0;
#pragma unroll
            for (int kk = 0; kk < KPL; ++kk) {
                const int k = lane + 64 * kk;
                if (k >= nn) continue;
                const size_t idx = base + k;
                double jn = Jo[w][kk];
                double hr = Hr[w][kk], hz = Hz[w][kk];
                if (SOLVE) {
                    if (k >= 1 && k <= nn - 2) jn = sB[(size_t)w * S + (k - 1)];
                    else if (!FUSED) jn = a.J[idx];
                }
                if (HCORR) {
                    double Jk = dmax(jn, kJeps);
                    double inv_Jold = 1.0 / dmax(Jo[w][kk], kJeps);
                    double ratio = Jk * inv_Jold;
                    hr = hr * ratio;
                    hz = hz * ratio;
                    double Hm2 = hr * hr + hz * hz;
                    if (Hm2 > Jk * Jk) {
                        double s = Jk / sqrt(Hm2);
                        hr = hr * s;
                        hz = hz * s;
                    }
                    jn = Jk;
                }
                a.J[idx] = jn; a.Hr[idx] = hr; a.Hz[idx] = hz;
                if (!FUSED) { a.J_old[idx] = jn; a.Hr_old[idx] = hr; a.Hz_old[idx] = hz; }
                if (FUSED) {
                    const double dk = a.dnu[k];
                    e1[kk] += dk * dmax(jn - kJeps, 0.0) * dV;
                    e10[kk] += dk * dmax(jn - 10 * kJeps, 0.0) * dV;
                    // Compute_f_esc_Cylindrical terms (src/ConstantsAndFunctions.jl:937-953)
                    if (i == nr - 1) esc += dk * (a.A_r[(size_t)j * (nr + 1) + nr] * kFourPi * dmax(hr, 0.0));
                    if (j == nz - 1) esc += dk * (a.A_z[(size_t)nz * nr + i] * kFourPi * dmax(hz, 0.0));
                    if (j == 0 && a.bc_z_bot == BC_OUTFLOW) esc += dk * (a.A_z[i] * kFourPi * dmax(-hz, 0.0));
                }
            }
            if (FUSED && (i == nr - 1 || j == nz - 1 || (j == 0 && a.bc_z_bot == BC_OUTFLOW)))
                acc_esc += wave_sum(esc);
        }
    }
    if (FUSED) {
        double* prow = a.part + (size_t)blockIdx.x * a.pstride;
#pragma unroll
        for (int kk = 0; kk < KPL; ++kk) {
            const int k = lane + 64 * kk;
            if (k < nn) { prow[k] = e1[kk]; prow[nn + k] = e10[kk]; }
        }
        if (lane == 0) {
            prow[2 * nn + 0] = acc_mf;
            prow[2 * nn + 1] = acc_L;
            prow[2 * nn + 2] = acc_net;
            prow[2 * nn + 3] = acc_out;
            prow[2 * nn + 4] = acc_esc;
        }
    }
}

// Step scalars from the reduced partials of the fused column kernel:
// Update_f_Lyα_MF_L_Lyα! normalisation (src/ConstantsAndFunctions.jl:663-673),
// Compute_f_esc_Lyα_Cylindrical (:969-993) and FindActiveBins for the next step.
__global__ void k_step_final(const double* __restrict__ red, int nn, int kcore_min, int kcore_max,
                             int bc_z_bot, int* __restrict__ krange, double* __restrict__ scalars,
                             double* __restrict__ history)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int kmin_used = krange[0], kmax_used = krange[1];
    double MF = red[2 * nn + 0], L = red[2 * nn + 1], net = red[2 * nn + 2], out = red[2 * nn + 3];
    double esc = red[2 * nn + 4];
    if (L > 0.0) {
        double s = 1.0 / (L / kC);
        MF = MF * s; net = net * s; out = out * s;
    } else {
        MF = 0.0; net = 0.0; out = 0.0;
    }
    double Ltot = (bc_z_bot == BC_REFLECT) ? 2.0 * L : L;
    double fesc = 0.0;
    if (Ltot > 0.0) {
        double Lsim = (bc_z_bot == BC_REFLECT) ? 0.5 * Ltot : Ltot;
        fesc = (Lsim > 0.0) ? esc / Lsim : 0.0;
    }
    double v[8] = {MF, net, out, Ltot, fesc, (double)kmin_used, (double)kmax_used, 0.0};
    for (int m = 0; m < 8; ++m) scalars[m] = v[m];
    if (history)
        for (int m = 0; m < 8; ++m) history[m] = v[m];
    active_range(red, nn, kcore_min, kcore_max, krange);
}

// ------------------------------------------------------------------------
// Operator-level reductions
// ------------------------------------------------------------------------
// Update_f_Lyα_MF_L_Lyα! (and the gravity variant), per-block partials:
// part[b] = {Σ ΔV|f|, Σ 4π ΔV js, Σ ΔV f_rad, Σ ΔV max(f_rad,0), Σ ΔV f_antigrav}
__global__ __launch_bounds__(256) void k_mf_partial(MfArgs a)
{
    __shared__ double s[5][256];
    const int nr = a.nr, nz = a.nz, ncells = nr * nz;
    double acc[5] = {0, 0, 0, 0, 0};
    const double inv_dt = 1.0 / a.dt;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < ncells; c += gridDim.x * 256) {
        const int i = c % nr, j = c / nr;
        double fr = a.drhou_r[c] * inv_dt, fz = a.drhou_z[c] * inv_dt;
        a.f_r[c] = fr; a.f_z[c] = fz;
        const double dV = a.dV[c];
        acc[0] += dV * sqrt(fr * fr + fz * fz);
        acc[1] += kFourPi * dV * a.js[c];
        double dr = a.rcom[i], dz = a.zcom[j] - a.z_center;
        double dist = sqrt(dr * dr + dz * dz);
        double frad = 0.0;
        if (dist > 0.0) {
            double inv_d = 1.0 / dist;
            frad = fr * (dr * inv_d) + fz * (dz * inv_d);
        }
        acc[2] += dV * frad;
        acc[3] += dV * dmax(frad, 0.0);
        if (a.fg_r) {
            double gr = a.fg_r[c], gz = a.fg_z[c];
            double gm = sqrt(gr * gr + gz * gz);
            double fa = (gm > 0.0) ? -(fr * gr + fz * gz) / gm : 0.0;
            acc[4] += dV * fa;
        }
    }
    for (int m = 0; m < 5; ++m) s[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w)
            for (int m = 0; m < 5; ++m) s[m][threadIdx.x] += s[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < 5; ++m) a.part[(size_t)blockIdx.x * 8 + m] = s[m][0];
}

// Compute_f_esc_Cylindrical leak sums, per block: Σ_k Δν_k Σ_faces A 4π max(±H, 0).
__global__ __launch_bounds__(256) void k_fesc_partial(const double* __restrict__ Hr,
                                                       const double* __restrict__ Hz,
                                                       const double* __restrict__ A_r,
                                                       const double* __restrict__ A_z,
                                                       const double* __restrict__ dnu, int nn, int nr,
                                                       int nz, int bc_z_bot, double* __restrict__ part)
{
    __shared__ double s[256];
    const int nfaces = nz + nr + (bc_z_bot == BC_OUTFLOW ? nr : 0);
    const long total = (long)nfaces * nn;
    double acc = 0.0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
        const int k = (int)(t % nn);
        const int f = (int)(t / nn);
        double term;
        if (f < nz) {
            term = A_r[(size_t)f * (nr + 1) + nr] * kFourPi * dmax(Hr[((size_t)f * nr + nr - 1) * nn + k], 0.0);
        } else if (f < nz + nr) {
            const int i = f - nz;
            term = A_z[(size_t)nz * nr + i] * kFourPi * dmax(Hz[((size_t)(nz - 1) * nr + i) * nn + k], 0.0);
        } else {
            const int i = f - nz - nr;
            term = A_z[i] * kFourPi * dmax(-Hz[(size_t)i * nn + k], 0.0);
        }
        acc += dnu[k] * term;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Update_e_Lyα!: one wavefront per cell, e = (4π/c) Σ_k Δν_k J_k.
__global__ __launch_bounds__(256) void k_e_lya(const double* __restrict__ J, const double* __restrict__ dnu,
                                                int nn, int ncells, double* __restrict__ e)
{
    const int c = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (c >= ncells) return;
    double s = 0.0;
    for (int k = lane; k < nn; k += 64) s += dnu[k] * J[(size_t)c * nn + k];
    s = wave_sum(s);
    if (lane == 0) e[c] = (kFourPi / kC) * s;
}

// Edge spectra: raw J of the edge cells and J / Σ_k J_k Δν_k.
__global__ __launch_bounds__(64) void k_edge(const double* __restrict__ J, const double* __restrict__ dnu,
                                              int nn, int nr, int nz, double* __restrict__ raw,
                                              double* __restrict__ norm)
{
    const int e = blockIdx.x;
    const int lane = threadIdx.x;
    int i, j;
    const int nright = nz - 2, ntop = nr - 3;
    if (e < nright) { i = nr - 2; j = 1 + e; }
    else if (e < nright + ntop) { i = 1 + (e - nright); j = nz - 2; }
    else { i = 1 + (e - nright - ntop); j = 1; }
    const double* Jc = J + ((size_t)j * nr + i) * nn;
    double s = 0.0;
    for (int k = lane; k < nn; k += 64) s += Jc[k] * dnu[k];
    s = wave_sum(s);
    for (int k = lane; k < nn; k += 64) {
        raw[(size_t)e * nn + k] = Jc[k];
        norm[(size_t)e * nn + k] = (s > 0.0) ? Jc[k] / s : Jc[k];
    }
}

// Non-finite scan of a state array: sets *flag when a NaN/Inf is found.
__global__ __launch_bounds__(256) void k_nonfinite(const double* __restrict__ x, size_t n, int* flag)
{
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) {
        double v = x[t];
        if (!(v - v == 0.0)) atomicMax(flag, 1);
    }
}

__global__ void k_selftest(int op, const double* a, const double* b, double* out, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (op == 0) out[t] = a[t] / b[t];
    else if (op == 1) out[t] = sqrt(a[t]);
    else if (op == 2) out[t] = exp_det(a[t]);
    else out[t] = a[t];
}

// ------------------------------------------------------------------------
// Launch wrappers
// ------------------------------------------------------------------------
void launch_active_partial(hipStream_t s, const double* J, const double* dnu, const double* dV, int nn,
                           int ncells, int nblocks, double* part, int pstride)
{
    int cpb = (ncells + nblocks - 1) / nblocks;
    hipLaunchKernelGGL(k_active_partial, dim3(nblocks), dim3(256), 0, s, J, dnu, dV, nn, ncells, cpb, part,
                       pstride);
}

void launch_reduce_cols(hipStream_t s, const double* part, int nrows, int stride, int ncols, double* out)
{
    hipLaunchKernelGGL(k_reduce_cols, dim3(ncols), dim3(256), 0, s, part, nrows, stride, out);
}

void launch_active_final(hipStream_t s, const double* red, int nn, int kcmin, int kcmax, int* krange)
{
    hipLaunchKernelGGL(k_active_final, dim3(1), dim3(64), 0, s, red, nn, kcmin, kcmax, krange);
}

void launch_advect(hipStream_t s, const AdvectArgs& a)
{
    // 2 x 4 cells: 2.75 face evaluations per cell instead of 4, 46 KB of LDS, two blocks per CU
    constexpr int BI = 2, BJ = 4;
    dim3 grid((a.nn + 63) / 64, (a.g.nr + BI - 1) / BI, (a.g.nz + BJ - 1) / BJ);
    hipLaunchKernelGGL((k_advect<BI, BJ>), grid, dim3(64, BI, BJ), 0, s, a);
}

template <int W, int KPL>
static void launch_column_t(hipStream_t s, const ColumnArgs& a, int nblocks)
{
    size_t lds = 4 * (size_t)col_stride(a.nn) * W * sizeof(double);
    hipLaunchKernelGGL((k_column<W, KPL>), dim3(nblocks), dim3(64), lds, s, a);
}

int column_width() { return kColumnW; }

// A grid of exactly one wave of resident blocks: every block loops over the same number of
// column groups and they all start together.  A larger grid left a partial second wave of
// blocks starting late (2560 blocks against 2048 resident ones on the 256x256x128 cone
// cloud: 1.15 ms against 0.81 ms for the kernel).
int column_blocks(int nn, int ngroups)
{
    const size_t lds = 4 * (size_t)col_stride(nn) * kColumnW * sizeof(double);
    int per_cu = 0;
    hipError_t e;
    switch ((nn + 63) / 64) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_column<kColumnW, 1>, 64, lds); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_column<kColumnW, 2>, 64, lds); break;
    case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_column<kColumnW, 3>, 64, lds); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_column<kColumnW, 4>, 64, lds); break;
    }
    if (e != hipSuccess || per_cu < 1) per_cu = 8;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess || cus < 1)
        cus = 256;
    const long nb = (long)per_cu * cus;
    return (int)std::max<long>(1, std::min<long>(ngroups, nb));
}

void launch_column(hipStream_t s, const ColumnArgs& a, int nblocks)
{
    const int kpl = (a.nn + 63) / 64;
    switch (kpl) {
    case 1: launch_column_t<kColumnW, 1>(s, a, nblocks); break;
    case 2: launch_column_t<kColumnW, 2>(s, a, nblocks); break;
    case 3: launch_column_t<kColumnW, 3>(s, a, nblocks); break;
    default: launch_column_t<kColumnW, 4>(s, a, nblocks); break;
    }
}

void launch_step_final(hipStream_t s, const double* red, int nn, int kcmin, int kcmax, int bc_z_bot,
                       int* krange, double* scalars, double* history)
{
    hipLaunchKernelGGL(k_step_final, dim3(1), dim3(64), 0, s, red, nn, kcmin, kcmax, bc_z_bot, krange, scalars,
                       history);
}

void launch_mf_partial(hipStream_t s, const MfArgs& a, int nblocks)
{
    hipLaunchKernelGGL(k_mf_partial, dim3(nblocks), dim3(256), 0, s, a);
}

void launch_fesc_partial(hipStream_t s, const double* Hr, const double* Hz, const double* A_r,
                         const double* A_z, const double* dnu, int nn, int nr, int nz, int bc_z_bot,
                         int nblocks, double* part)
{
    hipLaunchKernelGGL(k_fesc_partial, dim3(nblocks), dim3(256), 0, s, Hr, Hz, A_r, A_z, dnu, nn, nr, nz,
                       bc_z_bot, part);
}

void launch_e_lya(hipStream_t s, const double* J, const double* dnu, int nn, int ncells, double* e)
{
    hipLaunchKernelGGL(k_e_lya, dim3((ncells + 3) / 4), dim3(256), 0, s, J, dnu, nn, ncells, e);
}

void launch_edge(hipStream_t s, const double* J, const double* dnu, int nn, int nr, int nz, double* raw,
                 double* norm)
{
    int ne = (nz - 2) + 2 * (nr - 3);
    if (ne > 0) hipLaunchKernelGGL(k_edge, dim3(ne), dim3(64), 0, s, J, dnu, nn, nr, nz, raw, norm);
}

void launch_nonfinite(hipStream_t s, const double* x, size_t n, int* flag)
{
    hipLaunchKernelGGL(k_nonfinite, dim3(1024), dim3(256), 0, s, x, n, flag);
}

void launch_selftest(hipStream_t s, int op, const double* a, const double* b, double* out, size_t n)
{
    hipLaunchKernelGGL(k_selftest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, op, a, b, out, n);
}

}  // namespace m1

namespace m1 {

// ---------------------------------------------------------------------------
// Device-side coefficient refresh: the temperature- and density-dependent part of
// PrepareLyαSubCycling! (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2090-2155)
// with Update_Hvoigt_And_ϕν_cell! (:2039-2082).  Per cell: a(T), ΔνD(T), the bin-averaged
// Voigt profile Htophat (src/ConstantsAndFunctions.jl:227-279) from the lookup tables of
// CreateVoigtLUTs with the evaluators of src/FastInterpolations.jl:83-274, ϕ_ν normalised
// over all bins (the RHD form) or over the core bins (Update_ϕν!,
// src/ConstantsAndFunctions.jl:362-393), α = n_HI σ0(T), α_F = α Hvoigt + α_c,F.
// The formulas are evaluated in the order the host setup evaluates them, so the
// tables agree with the host's bit for bit.  One wavefront per cell, lanes over bins;
// the normalisation sum runs on lane 0 in bin order.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ double lut_eval(const double* y, int nx, double xmin, double inv_dx, double x)
{
    const double xt = (x - xmin) * inv_dx;
    long i = (long)floor(xt);
    if (i + 2 > nx) i = nx - 2;
    const double d = (double)(i + 1) - xt;
    return d * y[i] + (1.0 - d) * y[i + 1];
}

struct VLut {
    const double* t;
    int nx;
    double xmax, xmin, inv_dx;
    __device__ double fn(int which, double x) const
    {
        if (fabs(x) >= xmax) {
            if (which == 0) return copysign(1.0, x);
            if (which == 1) {
                const double ix = 1.0 / x;
                return 0.5 * ix + 0.25 * (ix * ix * ix);
            }
            return 0.0;
        }
        return lut_eval(t + (size_t)which * nx, nx, xmin, inv_dx, x);
    }
};

// H(x,a) and its integral, Smith et al. 2018 eqs. 12-13 (src/ConstantsAndFunctions.jl:227-256)
__device__ double voigt_H(const VLut& V, double x, double a)
{
    const double sp2 = sqrt(kPi) / 2.0;
    const double isp2 = 1.0 / sp2;
    return V.fn(2, x) + isp2 * a * (2.0 * x * V.fn(1, x) - 1.0) + (a * a) * V.fn(4, x);
}
__device__ double voigt_intH(const VLut& V, double x, double a)
{
    const double sp2 = sqrt(kPi) / 2.0;
    const double isp2 = 1.0 / sp2;
    return sp2 * V.fn(0, x) - isp2 * a * V.fn(1, x) + (a * a) * V.fn(3, x);
}

}  // namespace

__global__ __launch_bounds__(64) void k_refresh(RefreshArgs a)
{
    extern __shared__ double sHv[];
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int nn = a.nn;
    const size_t base = (size_t)c * nn;
    VLut V;
    V.t = a.lut; V.nx = a.nx; V.xmax = a.xmax; V.xmin = -a.xmax;
    V.inv_dx = 1.0 / ((a.xmax - (-a.xmax)) / (double)(a.nx - 1));
    const double T = a.T[c];
    const double b = a.lc.b_prefact * sqrt(T);
    const double dnuD = b * a.lc.nu_over_c;
    const double ic = 1.0 / dnuD;
    const double av = a.lc.a_prefact / b;
    const double alpha = a.nHI[c] * (a.lc.sigma0_prefact / b);
    const double acF = a.acF ? a.acF[c] : 0.0;
    for (int k = lane; k < nn; k += 64) {
        const double x = a.dnuC_rel[k] * ic, xmx = a.dnu_rel[k + 1] * ic, xmn = a.dnu_rel[k] * ic;
        const double dx = xmx - xmn;
        const double h = (dx > 1e-5) ? (voigt_intH(V, xmx, av) - voigt_intH(V, xmn, av)) / dx : voigt_H(V, x, av);
        sHv[k] = h;
        a.Hv[base + k] = h;
        a.aF[base + k] = alpha * h + acF;
    }
    __syncthreads();
    const int k0 = a.phi_all_bins ? 0 : a.kmin, k1 = a.phi_all_bins ? nn - 1 : a.kmax;
    if (lane == 0) {
        double s = 0.0;
        for (int k = k0; k <= k1; ++k) s += a.dnu[k] * sHv[k];
        sHv[nn] = 1.0 / s;
        a.alpha[c] = alpha;
        a.dnuD[c] = dnuD;
    }
    __syncthreads();
    const double inv = sHv[nn];
    for (int k = lane; k < nn; k += 64) a.phi[base + k] = (k >= k0 && k <= k1) ? sHv[k] * inv : 0.0;
}

void launch_refresh(hipStream_t s, const RefreshArgs& a)
{
    hipLaunchKernelGGL(k_refresh, dim3(a.ncell), dim3(64), (size_t)(a.nn + 1) * sizeof(double), s, a);
}

}  // namespace m1
