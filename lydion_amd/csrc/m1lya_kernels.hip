// m1lya_kernels.hip -- HIP kernels of the FP64 M1 Lya step for gfx950.
//
// Layout: spectral arrays are [j][i][k] (k = frequency bin fastest), exactly the
// Julia column-major [k,i,j] of the reference, so a wavefront of 64 lanes on
// consecutive k reads 512 contiguous bytes per FP64 field.
#include <algorithm>

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
                                                      int stride, double* __restrict__ out,
                                                      const int* __restrict__ halt)
{
    __shared__ double s[256];
    if (halt && *halt) return;
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
// E_tot = Σ_k e1[k], summed in bin order
__device__ double active_energy(const double* red, int nn)
{
    double Etot = 0.0;
    for (int k = 0; k < nn; ++k) Etot += red[k];
    return Etot;
}

__device__ void active_range(const double* red, int nn, int kcore_min, int kcore_max, int* krange, double Etot)
{
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

// The reduced row (2 N_ν bin sums and the step sums) copied into LDS by the whole block, so that
// the serial tail that follows reads it at LDS latency instead of one global load per bin.
constexpr int kRedMax = 2 * kMaxBinsFP64 + kStepSums;

__device__ __forceinline__ void stage_red(double* sred, const double* red, int n)
{
    for (int t = threadIdx.x; t < n; t += blockDim.x) sred[t] = red[t];
    __syncthreads();
}

__global__ void k_active_final(const double* __restrict__ red, int nn, int kcore_min, int kcore_max,
                               int* __restrict__ krange)
{
    __shared__ double sred[kRedMax];
    stage_red(sred, red, 2 * nn);
    if (threadIdx.x == 0 && blockIdx.x == 0) active_range(sred, nn, kcore_min, kcore_max, krange, active_energy(sred, nn));
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

// Cell update of AdvectJandH_Cylindrical! (src/AdvectJandH.jl:1319-1382) from the fluxes of the
// cell's four faces (J, H_r, H_Z components; r faces left/right, Z faces bottom/top) and the
// K_ϕϕ of the two r faces' sides that touch the cell.  Writes the new state of bin k.
__device__ __forceinline__ void advect_cell_update(const AdvectArgs& a, int i, int j, size_t idx, double Js,
                                                   double Hrs, double Hzs, const double (&FrL)[3],
                                                   const double (&FrR)[3], const double (&FzL)[3],
                                                   const double (&FzR)[3], double kphi_im, double kphi_ip)
{
    const int nr = a.g.nr;
    const int c = j * nr + i;
    const double Ar_L = a.g.A_r[(size_t)j * (nr + 1) + i], Ar_R = a.g.A_r[(size_t)j * (nr + 1) + i + 1];
    const double Az_L = a.g.A_z[(size_t)j * nr + i], Az_R = a.g.A_z[(size_t)(j + 1) * nr + i];
    const double cdt = a.cdt;
    const double cdt_dV = cdt * a.g.inv_dV[c];
    const double cdt_over_c = cdt * kInvC;
    const double durdr = a.v.dur_dr[c], durdz = a.v.dur_dz[c], duzdr = a.v.duz_dr[c];
    const double duzdz = a.v.duz_dz[c], urr = a.v.ur_over_r[c];
    Edd D = D_edd(Hrs, Hzs, Js);
    double Jn = Js - cdt_dV * ((Ar_R * FrR[0] - Ar_L * FrL[0]) + (Az_R * FzR[0] - Az_L * FzL[0]));
    Jn = dmax(Jn, kJeps);
    double Hzn = Hzs - cdt_dV * ((Ar_R * FrR[2] - Ar_L * FrL[2]) + (Az_R * FzR[2] - Az_L * FzL[2]));
    double Hrn = Hrs - cdt_dV * ((Ar_R * FrR[1] - Ar_L * FrL[1]) + (Az_R * FzR[1] - Az_L * FzL[1]));
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
// AdvectJandH_Cylindrical!: src/AdvectJandH.jl:526-1390, as a march along r.
// A block is WPB rows (one wavefront each) x one segment of seg_len cells along r
// x 64 bins (lanes).  Each wavefront keeps the r window (cells i-1 .. i+2 of its
// row) in registers and evaluates one r face per cell; the left face of a cell
// is the right face of the one before it.  The Z faces of the block's rows go
// through LDS: every wavefront publishes its cell's state, wavefront w evaluates
// face j0 + w and wavefront 0 also the block's top face.  That is 1 + 1/seg_len
// r faces and 1 + 1/WPB Z faces per cell.  All fluxes of a step are computed from
// the same old state (read from the input buffers, written to separate ones), as
// in the reference; bins outside [kmin, kmax] are copied.  A face's value does not
// depend on which wavefront evaluates it, so the result is the one of a per-cell
// evaluation.
// ------------------------------------------------------------------------
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_advect(AdvectArgs a, int seg_len)
{
    // Z stencil rows j0-2 .. j0+WPB+1 of the current column (J, H_r, H_Z, ghost-mapped) and the
    // Z faces j0 .. j0+WPB (F_J, F_Hr, F_HZ)
    __shared__ double zrow[WPB + 4][3][64];
    __shared__ double zflx[WPB + 1][3][64];
    if (a.halt && *a.halt) return;
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int k = blockIdx.x * 64 + lane;
    const int i0 = blockIdx.y * seg_len, i1 = min(i0 + seg_len, nr);
    const int j0 = blockIdx.z * WPB, j = j0 + w;
    const bool row_ok = j < nz;
    const bool kin = k < nn;
    const int kmin = a.krange[0], kmax = a.krange[1];
    const bool act = kin && k >= kmin && k <= kmax;
    const SpecIn s{a.J, a.Hr, a.Hz};

    // r window of row j: cells (i-1, i, i+1, i+2) in slots 0..3; before the march, cells
    // i0-2, i0-1, i0 in slots 1..3, so that the first shift and load make (i0-2 .. i0+1)
    double wJ[4] = {0.0, 0.0, 0.0, 0.0}, wHr[4] = {0.0, 0.0, 0.0, 0.0}, wHz[4] = {0.0, 0.0, 0.0, 0.0};
    if (act && row_ok) {
#pragma unroll
        for (int m = 1; m < 4; ++m) load_r(s, nn, nr, j, k, i0 - 3 + m, a.bc_r_right, wJ[m], wHr[m], wHz[m]);
    }
    // left r face of cell i: its flux and the K_ϕϕ of its right side
    double FL[3] = {0.0, 0.0, 0.0}, KL = 0.0;
    for (int i = i0 - 1; i < i1; ++i) {
        double FR[3] = {0.0, 0.0, 0.0}, KRl = 0.0, KRr = 0.0;
        if (act && row_ok) {
#pragma unroll
            for (int m = 0; m < 3; ++m) { wJ[m] = wJ[m + 1]; wHr[m] = wHr[m + 1]; wHz[m] = wHz[m + 1]; }
            load_r(s, nn, nr, j, k, i + 2, a.bc_r_right, wJ[3], wHr[3], wHz[3]);
            // face f = i + 1 between cells i and i+1; stencil (i, i+1, i+2, i-1)
            const int f = i + 1;
            Stencil st;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int slot = (m + 1) & 3;
                st.J[m] = wJ[slot]; st.Hr[m] = wHr[slot]; st.Hz[m] = wHz[slot];
            }
            const double aFL = a.aF[((size_t)j * nr + max(f - 1, 0)) * nn + k];
            const double aFR = a.aF[((size_t)j * nr + min(f, nr - 1)) * nn + k];
            const double u = a.v.ur_star[(size_t)j * (nr + 1) + f] * kInvC;
            const FaceResult r = interface_flux(0, a.hll, st, a.g.ivr + (size_t)f * 5, aFL, aFR, u);
            FR[0] = r.F.J; FR[1] = r.F.Hr; FR[2] = r.F.Hz;
            KRl = r.kphi_L;
            KRr = (f == 0) ? kphi_inner_axis(r.J_R, r.Hz_R) : r.kphi_R;
        }
        if (i < i0) {
#pragma unroll
            for (int m = 0; m < 3; ++m) FL[m] = FR[m];
            KL = KRr;
            continue;
        }
        // ---- Z faces of column i through LDS ----
        if (act) {
            if (row_ok) {
                zrow[w + 2][0][lane] = wJ[1]; zrow[w + 2][1][lane] = wHr[1]; zrow[w + 2][2][lane] = wHz[1];
            } else {
                double vj, vr, vz;   // a ghost row above the grid (rows past N_Z + 1 feed no needed face)
                load_z(s, nn, nr, nz, i, k, min(j, nz + 1), a.bc_z_bot, a.bc_z_top, vj, vr, vz);
                zrow[w + 2][0][lane] = vj; zrow[w + 2][1][lane] = vr; zrow[w + 2][2][lane] = vz;
            }
            for (int h = w; h < 4; h += WPB) {
                const int slot = (h < 2) ? h : WPB + h;
                double vj, vr, vz;
                load_z(s, nn, nr, nz, i, k, min(j0 - 2 + slot, nz + 1), a.bc_z_bot, a.bc_z_top, vj, vr, vz);
                zrow[slot][0][lane] = vj; zrow[slot][1][lane] = vr; zrow[slot][2][lane] = vz;
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll 1
            for (int it = 0; it < 2; ++it) {
                if (it && w != 0) break;
                const int fz = it ? WPB : w;    // the face's slot: face g = j0 + fz
                const int g = j0 + fz;
                if (g > nz) continue;
                Stencil st;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int slot = fz + ((m + 1) & 3);   // rows (g-1, g, g+1, g-2)
                    st.J[m] = zrow[slot][0][lane]; st.Hr[m] = zrow[slot][1][lane]; st.Hz[m] = zrow[slot][2][lane];
                }
                const double aFL = a.aF[((size_t)max(g - 1, 0) * nr + i) * nn + k];
                const double aFR = a.aF[((size_t)min(g, nz - 1) * nr + i) * nn + k];
                const double u = a.v.uz_star[(size_t)g * nr + i] * kInvC;
                const FaceResult r = interface_flux(1, a.hll, st, a.g.ivz + (size_t)g * 5, aFL, aFR, u);
                zflx[fz][0][lane] = r.F.J; zflx[fz][1][lane] = r.F.Hr; zflx[fz][2][lane] = r.F.Hz;
            }
        }
        __syncthreads();
        if (row_ok && kin) {
            const size_t idx = ((size_t)j * nr + i) * nn + k;
            if (!act) {
                a.J_out[idx] = s.J[idx];
                a.Hr_out[idx] = s.Hr[idx];
                a.Hz_out[idx] = s.Hz[idx];
            } else {
                const double FzL[3] = {zflx[w][0][lane], zflx[w][1][lane], zflx[w][2][lane]};
                const double FzR[3] = {zflx[w + 1][0][lane], zflx[w + 1][1][lane], zflx[w + 1][2][lane]};
                advect_cell_update(a, i, j, idx, wJ[1], wHr[1], wHz[1], FL, FR, FzL, FzR, KL, KRl);
            }
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) FL[m] = FR[m];
        KL = KRr;
    }
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
    // The inputs of row i+1 are read one row ahead of their use, so the LDS latency
    // overlaps the previous row's division instead of lengthening the chain.  No row
    // step writes what a later read takes (it writes dl, d, du, b at row i only).
    double dl_c = 0.0, d_c = 0.0, b_c = 0.0, du_c = 0.0;
    if (n > 1) { dl_c = dl[0]; d_c = d[1]; b_c = b[1]; du_c = du[min(1, n - 2)]; }
    int info = 0;
    for (int i = 0; i < n - 1; ++i) {
        const bool last = (i == n - 2);
        const double dl_i = dl_c, d_n = d_c, b_n = b_c, du_n = last ? 0.0 : du_c;
        {
            // unconditional reads, clamped to the column (what a clamped read returns is
            // never used), so that nothing waits for them before the next row step
            const int r1 = min(i + 1, n - 2), r2 = min(i + 2, n - 1), r3 = max(min(i + 2, n - 2), 0);
            dl_c = dl[r1]; d_c = d[r2]; b_c = b[r2]; du_c = du[r3];
        }
        __builtin_amdgcn_sched_barrier(0);   // keep the reads ahead of the row step
        // Both branches of the row step as selects around one division: without an
        // interchange fact = dl_i/d_i, row i stays and row i+1 loses fact·row i; with
        // one (|d_i| < |dl_i|) fact = d_i/dl_i, rows i and i+1 swap places, row i+1
        // fills in dl[i] (the second super-diagonal) and becomes old row i - fact·row i+1.
        // Each value is formed by the same operations as in LAPACK's two branches.
        const bool swap = !(fabs(d_i) >= fabs(dl_i));
        if (!swap && d_i == 0.0 && info == 0) info = i + 1;
        const double fact = (swap ? d_i : dl_i) / (swap ? dl_i : d_i);
        const double st_d = swap ? dl_i : d_i, st_du = swap ? d_n : du_i, st_b = swap ? b_n : b_i;
        const double nd = (swap ? du_i : d_n) - fact * (swap ? d_n : du_i);
        const double nb = (swap ? b_i : b_n) - fact * (swap ? b_n : b_i);
        if (!last) {
            dl[i] = swap ? du_n : 0.0;
            du_i = swap ? -fact * du_n : du_n;
        } else {
            du_i = du_n;
        }
        d[i] = st_d; du[i] = st_du; b[i] = st_b;
        d_i = nd; b_i = nb;
    }
    if (info != 0) return info;
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
// Column kernel: a block of NW wavefronts owns W·NW cell columns at a time, W per wavefront.
//   P1 (lanes over k): emission/absorption (src/ConstantsAndFunctions.jl:483-573),
//       frequency factors (src/FreqFactors.jl:142-278), tridiagonal rows to LDS
//       (src/UpdateLyaJandH.jl:178-227);
//   P2 (lane per column): dgtsv on the LDS columns, all W·NW of them by the block's first
//       wavefront: the elimination is a serial chain per column, so the issue cost of P2 is
//       one wavefront instruction per row step for as many columns as lanes take part;
//   P3 (lanes over k): J floor, H rescale and causality, old copies
//       (src/UpdateLyaJandH.jl:335-379), plus the fused step's reductions.
// LDS per block: 4 arrays x (N_ν-2) rows x W·NW columns (column index fastest, odd row
// stride, so the lane-per-column sweep is bank-conflict free).
// ------------------------------------------------------------------------
// kFused: the instantiation for the fused step's phases (emission, frequency factors, solve,
// H correction and reductions, kFusedFlags), with the phase tests resolved at compile time;
// only COL_CONV is read from the arguments.
constexpr int kFusedFlags = COL_EMIT | COL_FREQ | COL_SOLVE | COL_HCORR | COL_FUSED;

template <int W, int KPL, int NW, bool kFused>
__global__ __launch_bounds__(64 * NW) void k_column(ColumnArgs a)
{
    extern __shared__ double lds[];
    const int nn = a.nn, nr = a.nr, nz = a.nz;
    const int n = nn - 2;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int WB = W * NW;   // columns of the block
    const int flags = kFused ? (kFusedFlags | (a.flags & COL_CONV)) : a.flags;
    const bool EMIT = flags & COL_EMIT, FREQ = flags & COL_FREQ, STOREF = flags & COL_STORE_F;
    const bool SOLVE = flags & COL_SOLVE, HCORR = flags & COL_HCORR, FUSED = flags & COL_FUSED;
    const bool CONV = flags & COL_CONV;
    if (a.halt && *a.halt) return;
    // [column][row] with an odd row stride: conflict-free for both the lanes-over-k
    // writes (consecutive rows) and the lane-per-column sweep (stride S, S odd).
    const int S = col_stride(nn);
    double* sDL = lds;
    double* sD = lds + (size_t)S * WB;
    double* sDU = lds + (size_t)2 * S * WB;
    double* sB = lds + (size_t)3 * S * WB;
    const int ncells = nr * nz;
    const int ngroups = (ncells + WB - 1) / WB;
    const double ct = a.c_tilde, dt = a.dt, dw = a.dw;
    const double inv_ct = 1.0 / ct;
    const double fpc_ict = (kFourPi / kC) * inv_ct;
    const double inv_dt = 1.0 / dt;

    double e1[KPL], e10[KPL];
#pragma unroll
    for (int kk = 0; kk < KPL; ++kk) { e1[kk] = 0.0; e10[kk] = 0.0; }
    double acc_mf = 0.0, acc_L = 0.0, acc_net = 0.0, acc_out = 0.0, acc_esc = 0.0;
    double conv = 0.0;   // per lane: Σ Δν ΔV |J_new - J_start| (COL_CONV)
    int bad = 0;         // per lane: non-finite outputs of the fused step

    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        double Jo[W][KPL], Hr[W][KPL], Hz[W][KPL];
        // ================= P1 =================
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int lc = wave * W + w;   // the column's index in the block's LDS
            const int c = grp * WB + lc;
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
                // With no velocity gradient D:∇u is +0 whatever D is (D finite; its rr, zz and
                // ϕϕ entries are >= 0), so the cell's bins skip the closure
                const bool still = durdr == 0.0 && durdz == 0.0 && duzdr == 0.0 && duzdz == 0.0 && urr == 0.0;
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
                        double VEL = 0.0;
                        if (!still) {
                            Edd D = D_edd(Hr[w][kk], Hz[w][kk], Jcur[kk]);
                            double Ddu = D_colon_grad_u(D, durdr, durdz, duzdr, duzdz, urr);
                            VEL = Ddu * kInvC;
                        }
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
                        sD[(size_t)lc * S + row] = -(Fk + dw);
                        sB[(size_t)lc * S + row] = -dw * Jo[w][kk];
                        if (row >= 1) sDL[(size_t)lc * S + row - 1] = Fkm;
                        if (row <= n - 2) sDU[(size_t)lc * S + row] = Fkp;
                    }
                }
            }
        }
        // ================= P2 =================
        if (SOLVE) {
            __syncthreads();
            const int c = grp * WB + lane;
            if (wave == 0) {
                // The solve is one dependent chain per column; the other block on the CU keeps
                // its SIMD busy meanwhile.  At raised priority the chain's instructions issue as
                // soon as they are ready, and the other wavefronts fill the cycles in between.
                __builtin_amdgcn_s_setprio(3);
                if (lane < WB && c < ncells && n > 0) {
                    int info = gtsv_col(n, sDL + (size_t)lane * S, sD + (size_t)lane * S, sDU + (size_t)lane * S,
                                        sB + (size_t)lane * S);
                    if (info != 0) atomicMax(a.err_flag, 1);
                }
                __builtin_amdgcn_s_setprio(0);
            }
            __syncthreads();
        }
        // ================= P3 =================
        if (!(EMIT || SOLVE || HCORR)) continue;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int lc = wave * W + w;
            const int c = grp * WB + lc;
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
                    if (k >= 1 && k <= nn - 2) jn = sB[(size_t)lc * S + (k - 1)];
                    else if (!FUSED) jn = a.J[idx];
                }
                // the floor below would hide a NaN in J (max(NaN, Jeps) = Jeps): test first
                if (FUSED) bad += !(__builtin_isfinite(jn) && __builtin_isfinite(hr) && __builtin_isfinite(hz));
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
                    if (CONV) conv += dk * fabs(jn - a.J_old[idx]) * dV;
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
        // the block's partial row: each wavefront's values into LDS, summed in wavefront order
        const int ncol = 2 * nn + kStepSums;
        __syncthreads();
        double* wrow = lds + (size_t)wave * ncol;
#pragma unroll
        for (int kk = 0; kk < KPL; ++kk) {
            const int k = lane + 64 * kk;
            if (k < nn) { wrow[k] = e1[kk]; wrow[nn + k] = e10[kk]; }
        }
        conv = wave_sum(conv);
        const double nbad = wave_sum((double)bad);
        if (lane == 0) {
            wrow[2 * nn + SUM_MF] = acc_mf;
            wrow[2 * nn + SUM_L] = acc_L;
            wrow[2 * nn + SUM_NET] = acc_net;
            wrow[2 * nn + SUM_OUT] = acc_out;
            wrow[2 * nn + SUM_ESC] = acc_esc;
            wrow[2 * nn + SUM_CONV] = conv;
            wrow[2 * nn + SUM_BAD] = nbad;
        }
        __syncthreads();
        double* prow = a.part + (size_t)blockIdx.x * a.pstride;
        for (int t = threadIdx.x; t < ncol; t += 64 * NW) {
            double v = lds[t];
#pragma unroll
            for (int w = 1; w < NW; ++w) v += lds[(size_t)w * ncol + t];
            prow[t] = v;
        }
    }
}

// Step scalars from the reduced partials of the fused column kernel:
// Update_f_Lyα_MF_L_Lyα! normalisation (src/ConstantsAndFunctions.jl:663-673),
// Compute_f_esc_Lyα_Cylindrical (:969-993) and FindActiveBins for the next step.
__global__ void k_step_final(const double* __restrict__ red_g, int nn, int kcore_min, int kcore_max,
                             int bc_z_bot, int* __restrict__ krange, double* __restrict__ scalars,
                             double* __restrict__ history, int* __restrict__ err_flag, StepCtl* __restrict__ ctl,
                             double tol, int conv)
{
    __shared__ double sred[kRedMax];
    if (ctl && ctl->halt) return;
    stage_red(sred, red_g, 2 * nn + kStepSums);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double* red = sred;
    const int kmin_used = krange[0], kmax_used = krange[1];
    double MF = red[2 * nn + SUM_MF], L = red[2 * nn + SUM_L], net = red[2 * nn + SUM_NET];
    double out = red[2 * nn + SUM_OUT];
    double esc = red[2 * nn + SUM_ESC];
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
    // ΔJ/J = Σ Δν ΔV |J_new - J_start| / E_tot, E_tot = Σ Δν ΔV max(J_new - Jeps, 0) (the
    // energy FindActiveBins measures, src/UpdateLyaJandH.jl:60-66)
    const double Etot = active_energy(red, nn);
    double rel = 0.0;
    if (conv) rel = (Etot > 0.0) ? red[2 * nn + SUM_CONV] / Etot : 0.0;
    if (red[2 * nn + SUM_BAD] != 0.0) atomicOr(err_flag, (int)ERR_NONFINITE);
    double v[8] = {MF, net, out, Ltot, fesc, (double)kmin_used, (double)kmax_used, rel};
    for (int m = 0; m < 8; ++m) scalars[m] = v[m];
    if (history)
        for (int m = 0; m < 8; ++m) history[m] = v[m];
    active_range(red, nn, kcore_min, kcore_max, krange, Etot);
    if (ctl) {
        ctl->steps += 1;
        if (tol > 0.0 && conv && rel < tol) ctl->halt = 1;
    }
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

void launch_reduce_cols(hipStream_t s, const double* part, int nrows, int stride, int ncols, double* out,
                        const int* halt)
{
    hipLaunchKernelGGL(k_reduce_cols, dim3(ncols), dim3(256), 0, s, part, nrows, stride, out, halt);
}

void launch_active_final(hipStream_t s, const double* red, int nn, int kcmin, int kcmax, int* krange)
{
    hipLaunchKernelGGL(k_active_final, dim3(1), dim3(64), 0, s, red, nn, kcmin, kcmax, krange);
}

void launch_advect(hipStream_t s, const AdvectArgs& a)
{
    // Rows of 8 wavefronts marching over segments of 16 cells: 1.0625 r faces and 1.125 Z faces
    // per cell, 32 KB of LDS per block.  Segments are halved while the grid holds fewer blocks
    // than MI355X keeps resident (2 per CU at 123 VGPRs, 512 in all): on the 128x128x128
    // expanding cloud 8-cell segments take 0.110 ms against 0.120 ms for 16 (256 blocks), on
    // the 256x256x128 cone cloud 16 and 8 are within 1%.
    constexpr int WPB = 8, kResident = 512;
    const int chunks = (a.nn + 63) / 64, rows = (a.g.nz + WPB - 1) / WPB;
    int seg = 16;
    while (seg > 4 && chunks * ((a.g.nr + seg - 1) / seg) * rows < kResident) seg /= 2;
    dim3 grid(chunks, (a.g.nr + seg - 1) / seg, rows);
    hipLaunchKernelGGL((k_advect<WPB>), grid, dim3(64 * WPB), 0, s, a, seg);
}

template <int W, int KPL, int NW>
static void launch_column_t(hipStream_t s, const ColumnArgs& a, int nblocks)
{
    const size_t rows = 4 * (size_t)col_stride(a.nn) * W * NW;
    const size_t sums = (size_t)NW * (2 * a.nn + kStepSums);
    const size_t lds = std::max(rows, sums) * sizeof(double);
    if ((a.flags & ~COL_CONV) == kFusedFlags)
        hipLaunchKernelGGL((k_column<W, KPL, NW, true>), dim3(nblocks), dim3(64 * NW), lds, s, a);
    else
        hipLaunchKernelGGL((k_column<W, KPL, NW, false>), dim3(nblocks), dim3(64 * NW), lds, s, a);
}

// Up to 256 bins (four per lane) four cell columns share a wavefront; beyond, up to 576 bins
// (nine per lane, the reference's q = 1.03 grid of 542 bins), one column per wavefront keeps the
// per-lane registers and the LDS rows within bounds.
int column_width(int nn) { return nn <= 256 ? kColumnW : 1; }

// Wavefronts per block of the column kernel: the block's W·NW columns fill 64 KB of LDS up to
// 128 bins (16 columns), 64 KB up to 256 bins (8 columns); wider grids keep one wavefront
// (one column of up to 18 KB).  Two blocks per CU then hold 8 wavefronts.
int column_waves(int nn) { return nn <= 128 ? 4 : (nn <= 256 ? 2 : 1); }

// A grid of exactly one wave of resident blocks: every block loops over the same number of
// column groups and they all start together.  A larger grid left a partial second wave of
// blocks starting late (2560 blocks against 2048 resident ones on the 256x256x128 cone
// cloud: 1.15 ms against 0.81 ms for the kernel).  The count is fixed per block shape, so the
// partition of the fused reductions does not depend on the device.
int column_blocks(int nn, int ngroups)
{
    return std::max(1, std::min(ngroups, kColumnBlocks / column_waves(nn)));
}

void launch_column(hipStream_t s, const ColumnArgs& a, int nblocks)
{
    const int kpl = (a.nn + 63) / 64;
    switch (kpl) {
    case 1: launch_column_t<kColumnW, 1, 4>(s, a, nblocks); break;
    case 2: launch_column_t<kColumnW, 2, 4>(s, a, nblocks); break;
    case 3: launch_column_t<kColumnW, 3, 2>(s, a, nblocks); break;
    case 4: launch_column_t<kColumnW, 4, 2>(s, a, nblocks); break;
    case 5: launch_column_t<1, 5, 1>(s, a, nblocks); break;
    case 6: launch_column_t<1, 6, 1>(s, a, nblocks); break;
    case 7: launch_column_t<1, 7, 1>(s, a, nblocks); break;
    case 8: launch_column_t<1, 8, 1>(s, a, nblocks); break;
    default: launch_column_t<1, 9, 1>(s, a, nblocks); break;
    }
}

void launch_step_final(hipStream_t s, const double* red, int nn, int kcmin, int kcmax, int bc_z_bot,
                       int* krange, double* scalars, double* history, int* err_flag, StepCtl* ctl,
                       double tol, int conv)
{
    hipLaunchKernelGGL(k_step_final, dim3(1), dim3(64), 0, s, red, nn, kcmin, kcmax, bc_z_bot, krange, scalars,
                       history, err_flag, ctl, tol, conv);
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

namespace m1 {

// ------------------------------------------------------------------------
// RHD time-step control: one value per cell, reduced by min (CFL) or max (RSL).
//   mode 0, GetRHDCFLTimeStep (src/AdvectJandH.jl:29-68): 0.5 min(Δr, ΔZ) / (c̃ + |u|)
//   mode 1, FindRSL_LyαRT (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:655-705):
//           ϵ_RSL max(1, (aτ)^{1/3} f_abs) max(|u|, c_s)
// ------------------------------------------------------------------------
__device__ double rhd_cell_value(const RhdScanArgs& a, int c)
{
    const int i = c % a.nr, j = c / a.nr;
    const double dr = a.r[i + 1] - a.r[i];
    const double dZ = a.Z[j + 1] - a.Z[j];
    const double u = hypot(a.u_r[c], a.u_z[c]);
    if (a.mode == 0) return 0.5 * dmin(dr, dZ) / (a.c_tilde + u);
    const double c_s = sqrt(a.gamma_ad[c] * a.P[c] / a.rho[c]);
    const double xHI = dmax(1 - a.xHII[c] - 2 * a.xH2[c], 0.0);
    const double nHI = xHI * a.X_over_mH * a.rho[c];
    const double T = a.T[c];
    const double b = a.lc.b_prefact * sqrt(T);
    const double alpha = nHI * (a.lc.sigma0_prefact / b);
    const double dX = dmax(dr, dZ);
    const double tcell = alpha * dX;
    const double tcabs = (a.kS * a.rho_d[3 * c] + a.kG * a.rho_d[3 * c + 1] + a.kP * a.rho_d[3 * c + 2]) * dX;
    const double at = (a.lc.a_prefact / b) * tcell;
    const double at13 = pow(at, 1.0 / 3.0);
    const double fabs_ = 1 / (1 + at13 * tcabs + 0.6 * a.pd[c] * tcell);
    return 4.0 * dmax(1, at13 * fabs_) * dmax(u, c_s);   // ϵ_RSL = 4 (PhotoThermoChemistry.jl:34)
}

__global__ __launch_bounds__(256) void k_rhd_scan(RhdScanArgs a, double* __restrict__ part)
{
    __shared__ double s[256];
    const int n = a.nr * a.nz;
    const bool mx = a.mode == 1;
    double v = mx ? 0.0 : INFINITY;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
        const double x = rhd_cell_value(a, c);
        v = mx ? dmax(v, x) : dmin(v, x);
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = mx ? dmax(s[threadIdx.x], s[threadIdx.x + w])
                                                 : dmin(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_rhd_scan_final(const double* __restrict__ part, int nb, int mode,
                                                         double* __restrict__ out)
{
    __shared__ double s[256];
    const bool mx = mode == 1;
    double v = mx ? 0.0 : INFINITY;
    for (int b = threadIdx.x; b < nb; b += 256) v = mx ? dmax(v, part[b]) : dmin(v, part[b]);
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = mx ? dmax(s[threadIdx.x], s[threadIdx.x + w])
                                                 : dmin(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

void launch_rhd_scan(hipStream_t s, const RhdScanArgs& a, int nblocks, double* part, double* out)
{
    hipLaunchKernelGGL(k_rhd_scan, dim3(nblocks), dim3(256), 0, s, a, part);
    hipLaunchKernelGGL(k_rhd_scan_final, dim3(1), dim3(256), 0, s, part, nblocks, a.mode, out);
}

// ------------------------------------------------------------------------
// The Lyα block of AbsorptionStep (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:1577-1670):
// one wavefront per cell, lanes over the bins.  Backward-Euler J (floored at Jeps) and H
// over the cell's sub-step, the momentum each bin deposits split between gas and the
// three dust species by flux opacity (SplitRadPressure, :1017-1048), the would-be e_Lyα and
// the rejection rule (tol_Abs = 0.1, e_Lyα_floor = (4π/c) 1e-15: :1051-1055); a rejected
// cell keeps its old J and H (RestoreLyαRejection, :980-990).
// ------------------------------------------------------------------------
template <int KPL>
__global__ __launch_bounds__(64) void k_absorb(AbsorbArgs a)
{
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= a.ncell) return;
    const int nn = a.nn;
    const double ct = a.c_tilde, inv_ct = 1.0 / ct;
    const double dts = a.dt[c];
    const double ct_js = ct * a.js[c];
    const double ac = a.alpha_c[c];
    const double pd_a = a.pd[c] * a.alpha[c];
    const double aS = a.rho_d ? a.kS * a.rho_d[3 * c] : 0.0;
    const double aG = a.rho_d ? a.kG * a.rho_d[3 * c + 1] : 0.0;
    const double aP = a.rho_d ? a.kP * a.rho_d[3 * c + 2] : 0.0;
    const double dt_lya = inv_ct / (ac + pd_a);
    const size_t base = (size_t)c * nn;
    double Jn[KPL], Hrn[KPL], Hzn[KPL];
    double sJ = 0.0, mr = 0.0, mz = 0.0, g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KPL; ++kk) {
        const int k = lane + 64 * kk;
        Jn[kk] = 0.0; Hrn[kk] = 0.0; Hzn[kk] = 0.0;
        if (k >= nn) continue;
        const size_t idx = base + k;
        const double dnu = a.dnu[k];
        const double dJdt = ct_js * a.phi[idx];
        const double a_abs = ac + pd_a * a.Hv[idx];
        const double Jo = a.J_old[idx], Hro = a.Hr_old[idx], Hzo = a.Hz_old[idx];
        // AbsorptionUpdateJorH (src/ConstantsAndFunctions.jl:451-462)
        Jn[kk] = dmax((Jo + dJdt * dts) / (1 + a_abs * ct * dts), kJeps);
        const double aFk = a.aF[idx];
        Hrn[kk] = (Hro + 0.0 * dts) / (1 + aFk * ct * dts);
        Hzn[kk] = (Hzo + 0.0 * dts) / (1 + aFk * ct * dts);
        const double w = dnu * (kFourPi / kC);
        const double dr_k = w * (inv_ct * (Hro - Hrn[kk]));
        const double dz_k = w * (inv_ct * (Hzo - Hzn[kk]));
        mr += dr_k;
        mz += dz_k;
        const double a_gas = dmax(aFk - aS - aG - aP, 0.0);
        const double a_tot = a_gas + aS + aG + aP;
        double fg = 1.0, fS = 0.0, fG = 0.0, fP = 0.0;
        if (a_tot > 0.0) {
            const double inv = 1 / a_tot;
            fg = a_gas * inv; fS = aS * inv; fG = aG * inv; fP = aP * inv;
        }
        g[0] += fg * dr_k; g[1] += fg * dz_k;
        g[2] += fS * dr_k; g[3] += fS * dz_k;
        g[4] += fG * dr_k; g[5] += fG * dz_k;
        g[6] += fP * dr_k; g[7] += fP * dz_k;
        sJ += dnu * Jn[kk];
    }
    sJ = wave_sum(sJ);
    mr = wave_sum(mr);
    mz = wave_sum(mz);
#pragma unroll
    for (int m = 0; m < 8; ++m) g[m] = wave_sum(g[m]);
    const double e_new = (kFourPi / kC) * sJ;
    const double e_old = a.e_lya[c];
    const double de = fabs(e_new - e_old) / (e_old + (kFourPi / kC) * 1e-15);
    const bool rej = (de > 0.1 && dts > 0.1 * dt_lya) || !__builtin_isfinite(de);
#pragma unroll
    for (int kk = 0; kk < KPL; ++kk) {
        const int k = lane + 64 * kk;
        if (k >= nn) continue;
        const size_t idx = base + k;
        a.J[idx] = rej ? a.J_old[idx] : Jn[kk];
        a.Hr[idx] = rej ? a.Hr_old[idx] : Hrn[kk];
        a.Hz[idx] = rej ? a.Hz_old[idx] : Hzn[kk];
    }
    if (lane == 0) {
        const double v[12] = {de, e_new, mr, mz, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7]};
        a.redo[c] = rej ? 1 : 0;
        for (int m = 0; m < 12; ++m) a.out[(size_t)m * a.ncell + c] = (rej && m > 0) ? 0.0 : v[m];
    }
}

void launch_absorb(hipStream_t s, const AbsorbArgs& a)
{
    const int kpl = (a.nn + 63) / 64;
    switch (kpl) {
    case 1: hipLaunchKernelGGL(k_absorb<1>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_absorb<2>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_absorb<3>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_absorb<4>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_absorb<5>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_absorb<6>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 7: hipLaunchKernelGGL(k_absorb<7>, dim3(a.ncell), dim3(64), 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_absorb<8>, dim3(a.ncell), dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL(k_absorb<9>, dim3(a.ncell), dim3(64), 0, s, a); break;
    }
}

}  // namespace m1

namespace m1 {

// ------------------------------------------------------------------------
// Lyα source terms of PrepareLyαSubCycling! (src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2086-2155):
// the frequency-integrated emissivity from recombinations (src/LymanAlphaEmission.jl:105-124) and
// collisional excitation (:143-173), the destruction probability (src/LymanAlphaDestruction.jl:43-97)
// and the dust + gas continuum absorption and flux opacities, one thread per cell.
// ------------------------------------------------------------------------
// Storey & Hummer (1995) probability of Lyα emission per recombination (src/LymanAlphaEmission.jl:30-60),
// T = 5e2 .. 3e4 K (10 nodes) by n_e = 1e2 .. 1e14 cm^-3 (Case B, 13 nodes) or 1e2 .. 1e10 (Case A, 9 nodes).
__constant__ double SH95_T[10] = {5e2, 1e3, 3e3, 5e3, 7.5e3, 1e4, 1.25e4, 1.5e4, 2e4, 3e4};
__constant__ double SH95_B[10][13] = {
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
__constant__ double SH95_A[10][9] = {
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


__device__ double sh95(double T, double ne, bool caseB)
{
    const int nn = caseB ? 13 : 9;
    const double* f = caseB ? &SH95_B[0][0] : &SH95_A[0][0];
    const double x = dmin(dmax(log10(T), log10(SH95_T[0])), log10(SH95_T[9]));
    const double y = dmin(dmax(log10(ne), 2.0), (double)(nn + 1));
    int i = 0, j = 0;
    while (i < 8 && x > log10(SH95_T[i + 1])) ++i;
    while (j < nn - 2 && y > (double)(j + 3)) ++j;
    const double xi = log10(SH95_T[i]), xi1 = log10(SH95_T[i + 1]);
    const double t = (x - xi) / (xi1 - xi);
    const double u = (y - (double)(j + 2)) / ((double)(j + 3) - (double)(j + 2));
    const double f0 = f[i * nn + j] + u * (f[i * nn + j + 1] - f[i * nn + j]);
    const double f1 = f[(i + 1) * nn + j] + u * (f[(i + 1) * nn + j + 1] - f[(i + 1) * nn + j]);
    return f0 + t * (f1 - f0);
}

__device__ double krec_hii(double T, bool caseB)
{
    T = dmin(dmax(T, 1.0), 1e9);
    const double lam = 315614.0 / T;
    if (caseB) return 2.753e-14 * pow(lam, 1.5) / pow(1 + pow(lam / 2.74, 0.407), 2.242);
    return 1.269e-13 * pow(lam, 1.503) / pow(1 + pow(lam / 0.522, 0.47), 1.923);
}

__global__ __launch_bounds__(256) void k_sources(SourceArgs a)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.ncell) return;
    const int n = a.ncell;
    const double T = a.chem[c], nHI = a.chem[n + c], nH2 = a.chem[2 * n + c], nHII = a.chem[3 * n + c];
    const double nHeI = a.chem[4 * n + c], ne = a.chem[5 * n + c], nOH = a.chem[6 * n + c], nCH = a.chem[7 * n + c];
    constexpr double kE = (6.626069573e-27 * 2.466068e15) * (1.0 / kFourPi);   // E_Lyα / 4π
    // recombinations (Case B with the on-the-spot approximation, else Case A)
    const double js_rec = sh95(T, ne, a.ots) * krec_hii(T, a.ots) * nHII * ne * kE;
    // collisional excitation
    constexpr double T_ryd = (6.626069573e-27 * 3.289842e15) / 1.380648813e-16;
    constexpr double T12 = (1 - 1.0 / 4.0) * T_ryd, T13 = (1 - 1.0 / 9.0) * T_ryd;
    constexpr double T14 = (1 - 1.0 / 16.0) * T_ryd, T15 = (1 - 1.0 / 25.0) * T_ryd;
    const double Tc = dmax(T, 1.0);
    const double inv_T = 1.0 / Tc;
    const double T6 = dmin(Tc, 3e5) * 1e-6;
    const double T6s = T6 * T6, T6c = T6s * T6;
    const double Y12 = 0.267486 + 1.57257 * T6 - 6.44026 * T6s + 11.5401 * T6c;
    const double Y13 = 0.0940821 + 2.90748 * T6 - 6.69528 * T6s + 7.86419 * T6c;
    const double Y14 = 0.0415028 + 1.13673 * T6 - 3.97454 * T6s + 6.05941 * T6c;
    const double Y15 = 0.032277 + 0.40856 * T6 - 1.34604 * T6s + 2.07558 * T6c;
    const double q = (8.63e-6 / (2 * sqrt(Tc))) *
                     (Y12 * exp(-T12 * inv_T) + Y13 * exp(-T13 * inv_T) + Y14 * exp(-T14 * inv_T) +
                      Y15 * exp(-T15 * inv_T));
    const double js = js_rec + q * nHI * ne * kE;
    // destruction probability
    double pd = 0.0;
    if (a.destruction) {
        const double Tp = dmin(dmax(T, 1.0), 1e5);
        const double T100 = Tp / 100.0;
        const double lT = log10(Tp), lT2 = lT * lT, lT3 = lT2 * lT, lT4 = lT3 * lT;
        const double G2p2s = 3.1e-6 * (1 + a.z), G2s2p = 3 * G2p2s;
        const double r25 = 2.5 / Tp;
        const double C_e = 5e-5 * exp(-(r25 * r25)) * ne / (1 + pow(T100, 0.45));
        const double C_HII = (6.98e-3 * (1 + 0.933 * lT) * exp(-pow(601 / Tp, 0.406)) / pow(Tp, 0.5) +
                              pow(10.0, -7.69 + 1.21 * lT + 0.455 * lT2 - 0.199 * lT3 + 0.0182 * lT4)) * nHII;
        const double C_HeI = 2.6e-9 * pow(T100, 0.315) * nHeI;
        const double C_H2 = 1.7e-9 * pow(T100, 0.3) * nH2;
        const double den = 8.9 * pow(Tp, -0.55) + 0.078 * pow(Tp, 0.35);
        const double C_HI = (4.1e-11 * exp(-pow(1.6 / Tp, 3.0 / 4.0) - 0.15 * Tp) +
                             (1.6e-10 + 1.57e-12 * pow(Tp, 1.1)) / (den * den)) * nHI;
        const double C2s2p = C_e + C_HII + C_HeI + C_H2 + C_HI;
        const double C2p2s = (1.0 / 3.0) * C2s2p;
        const double R2p2s = (2.0 / 3.0) * 8.78e-7 + C2p2s + G2p2s;
        const double R2s2p = 1.597e-9 + C2s2p + G2s2p;
        const double pd_HI = R2p2s / ((6.265e8 / 8.2206) * (8.2206 + R2s2p) + R2p2s);
        double pd_H2 = 0.0;
        if (nHI > 1e-20)
            pd_H2 = dmin((33.0 * exp(-14540.0 / Tp) + 13.0 * exp(-15564.0 / Tp)) /
                             (46.0 + 1e3 * exp(-pow(1e3 / Tp, 1.5))),
                         1.3e-3) * (nH2 / nHI);
        pd = dmin(pd_HI + pd_H2, 1.0);
    }
    // continuum opacities: dust below 1.1 ρd_floor is ignored
    constexpr double rd_ignore = 1.1 * (1e-13 * (1.0080 * 1.660539069e-24));
    double rd[3] = {0.0, 0.0, 0.0};
    if (a.rho_d)
        for (int s = 0; s < 3; ++s) rd[s] = (a.rho_d[3 * c + s] > rd_ignore) ? a.rho_d[3 * c + s] : 0.0;
    const double ac_gas = nH2 * 2.1e-25 + nOH * 4.07e-18 + nCH * 0.05e-18;
    const double acF_gas = ac_gas + ne * 6.6524587321e-25;
    a.js[c] = js;
    a.alpha_c[c] = a.kappa[0] * rd[0] + a.kappa[1] * rd[1] + a.kappa[2] * rd[2] + ac_gas;
    a.acF[c] = a.kappa[3] * rd[0] + a.kappa[4] * rd[1] + a.kappa[5] * rd[2] + acF_gas;
    a.pd[c] = pd;
    a.T[c] = T;
    a.nHI[c] = nHI;
}

void launch_sources(hipStream_t s, const SourceArgs& a)
{
    hipLaunchKernelGGL(k_sources, dim3((a.ncell + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace m1
