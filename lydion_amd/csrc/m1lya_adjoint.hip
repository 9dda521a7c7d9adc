// m1lya_adjoint.hip -- reverse sweep of the FP64 fused step (m1lya_adjoint_run).
//
// The step is X_{n+1} = C(E(A(X_n))): advection A (AdvectJandH_Cylindrical!,
// src/AdvectJandH.jl:526-1390), emission/absorption E
// (LyαEmissionAbsorptionUpdate!, src/ConstantsAndFunctions.jl:483-573), and the
// frequency update C (FrequencyUpdateJLyα_AllCells!, src/UpdateLyaJandH.jl:231-391)
// with its flux correction; the scalars are Update_f_Lyα_MF_L_Lyα!
// (src/ConstantsAndFunctions.jl:588-674) and the escaping spectrum of
// Compute_f_esc_Lyα_Cylindrical (:969-993).  Given the cotangent of X_{n+1},
// three kernels produce the cotangent of X_n and accumulate those of the
// coefficient fields:
//
//   k_adj_column    one wavefront per cell, lanes over bins: E and C in
//                   reverse.  Recomputes E(X1) and the frequency-coupling
//                   rows, solves the transposed tridiagonal system, contracts
//                   the row cotangents with the local derivatives of each row
//                   (dual-number passes over its inputs), and reverses the
//                   flux correction and emission.
//   k_adj_cell      one thread per (bin, cell): the cell update of A in
//                   reverse (floors, causality limit, aberration, D:∇u and
//                   K_ϕϕ terms); writes the cotangent of each cell's
//                   flux-divergence sums and the cell's local cotangent.
//   k_adj_gather    one thread per (bin, cell): adds, over the eight
//                   interfaces whose stencil holds the cell, the derivative of
//                   the interface outputs along the cell's J, H_r, H_Z and α_F
//                   contracted with the interface cotangents (which come from
//                   the flux-divergence cotangents of the two cells it
//                   separates).  Gather form: no atomics, deterministic.
//
// Bins outside the step's active range pass through A unchanged.  The active
// range itself is piecewise constant in the inputs and is held fixed.
#include <hip/hip_runtime.h>

#include "m1lya_adjoint.h"
#include "m1lya_adjoint_face.h"
#include "m1lya_internal.h"

namespace m1 {

using m1a::Dl;
using m1a::Dl2;

namespace {

// Tangent d on input `which` (0 J, 1 H_r, 2 H_Z, 3 α_F of the target cell) when it is
// the direction `var` asks for: Dl carries direction var; Dl2 the pair (2 var, 2 var + 1).
__device__ __forceinline__ void seedv(double&, int, int, double) {}
__device__ __forceinline__ void seedv(Dl& x, int which, int var, double d)
{
    if (which == var) x.d = d;
}
__device__ __forceinline__ void seedv(Dl2& x, int which, int var, double d)
{
    if (which == 2 * var) x.d0 = d;
    else if (which == 2 * var + 1) x.d1 = d;
}

// Ghost-aware stencil load along r (the values of load_r in m1lya_kernels.hip),
// with the tangent of the slot along variable `var` (0 J, 1 H_r, 2 H_Z, else
// none) of cell `tc` of the same row.
template <class T>
__device__ __forceinline__ void tload_r(const double* J, const double* Hr, const double* Hz, int nn, int nr, int j,
                                        int k, int c, int bc_right, int tc, int var, T& sJ, T& sHr, T& sHz)
{
    int src;
    double hr_fac = 1.0;
    bool clip = false;
    if (c < 0) { src = -c - 1; hr_fac = -1.0; }
    else if (c < nr) { src = c; }
    else if (bc_right == BC_REFLECT) { src = 2 * nr - 1 - c; hr_fac = -1.0; }
    else { src = nr - 1; clip = true; }
    const size_t idx = ((size_t)j * nr + src) * nn + k;
    const double vJ = J[idx], vHr = Hr[idx], vHz = Hz[idx];
    sJ = T(vJ);
    sHr = T(clip ? dmax(vHr, 0.0) : hr_fac * vHr);
    sHz = T(vHz);
    if (src == tc) {
        seedv(sJ, 0, var, 1.0);
        seedv(sHr, 1, var, clip ? ((0.0 > vHr) ? 0.0 : 1.0) : hr_fac);
        seedv(sHz, 2, var, 1.0);
    }
}

// Same along Z (load_z), tc = row index of the target cell in column i.
template <class T>
__device__ __forceinline__ void tload_z(const double* J, const double* Hr, const double* Hz, int nn, int nr, int nz,
                                        int i, int k, int c, int bc_bot, int bc_top, int tc, int var, T& sJ,
                                        T& sHr, T& sHz)
{
    int src;
    double hz_fac = 1.0;
    int clip = 0;   // 1: max(H_Z, 0) (top outflow), 2: min(H_Z, 0) (bottom outflow)
    if (c < 0) {
        if (bc_bot == BC_REFLECT) { src = -c - 1; hz_fac = -1.0; }
        else { src = 0; clip = 2; }
    } else if (c < nz) {
        src = c;
    } else if (bc_top == BC_REFLECT) {
        src = 2 * nz - 1 - c; hz_fac = -1.0;
    } else {
        src = nz - 1; clip = 1;
    }
    const size_t idx = ((size_t)src * nr + i) * nn + k;
    const double vJ = J[idx], vHr = Hr[idx], vHz = Hz[idx];
    sJ = T(vJ);
    sHr = T(vHr);
    double hz = hz_fac * vHz, dhz = hz_fac;
    if (clip == 1) { hz = dmax(vHz, 0.0); dhz = (0.0 > vHz) ? 0.0 : 1.0; }
    if (clip == 2) { hz = dmin(vHz, 0.0); dhz = (0.0 < vHz) ? 0.0 : 1.0; }
    sHz = T(hz);
    if (src == tc) {
        seedv(sJ, 0, var, 1.0);
        seedv(sHr, 1, var, 1.0);
        seedv(sHz, 2, var, dhz);
    }
}

// Interface f of row j (r, dir 0) or column i (Z, dir 1) at bin k, with the
// tangent along `var` of target cell tc (var 3: α_F of tc).
template <class T>
__device__ m1a::FaceT<T> face_at(const AdjAdvArgs& a, int dir, int f, int i, int j, int k, int tc, int var)
{
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    m1a::StencilT<T> st;
    const int cells[4] = {f - 1, f, f + 1, f - 2};
    const double* iv;
    double u;
    T aFL, aFR;
    if (dir == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            tload_r(a.J, a.Hr, a.Hz, nn, nr, j, k, cells[m], a.bc_r_right, tc, var, st.J[m], st.Hr[m], st.Hz[m]);
        iv = a.g.ivr + (size_t)f * 5;
        const int cl = max(f - 1, 0), cr = min(f, nr - 1);
        aFL = T(a.aF[((size_t)j * nr + cl) * nn + k]);
        aFR = T(a.aF[((size_t)j * nr + cr) * nn + k]);
        if (cl == tc) seedv(aFL, 3, var, 1.0);
        if (cr == tc) seedv(aFR, 3, var, 1.0);
        u = a.v.ur_star[(size_t)j * (nr + 1) + f] * kInvC;
    } else {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            tload_z(a.J, a.Hr, a.Hz, nn, nr, nz, i, k, cells[m], a.bc_z_bot, a.bc_z_top, tc, var, st.J[m],
                    st.Hr[m], st.Hz[m]);
        iv = a.g.ivz + (size_t)f * 5;
        const int cl = max(f - 1, 0), cr = min(f, nz - 1);
        aFL = T(a.aF[((size_t)cl * nr + i) * nn + k]);
        aFR = T(a.aF[((size_t)cr * nr + i) * nn + k]);
        if (cl == tc) seedv(aFL, 3, var, 1.0);
        if (cr == tc) seedv(aFR, 3, var, 1.0);
        u = a.v.uz_star[(size_t)f * nr + i] * kInvC;
    }
    return m1a::tface<T>(dir, a.hll, dir == 0 && f == 0, st, iv, aFL, aFR, u);
}

}  // namespace

// ---------------------------------------------------------------------------
// Cell update of AdvectJandH_Cylindrical! (src/AdvectJandH.jl:1302-1382) in reverse.
// ---------------------------------------------------------------------------
// Block of 64 bins x BI x BJ cells, a wavefront per cell; the forward values of each interface
// are evaluated once per block and reach the two cells that share it through LDS.
template <int BI, int BJ>
__global__ __launch_bounds__(64 * BI * BJ) void k_adj_cell(AdjAdvArgs a)
{
    __shared__ double rF[BJ][BI + 1][5][64];   // F_J, F_Hr, F_HZ, K_ϕϕ left, K_ϕϕ right
    __shared__ double zF[BJ + 1][BI][3][64];
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    const int lane = threadIdx.x, il = threadIdx.y, jl = threadIdx.z;
    const int k = blockIdx.x * 64 + lane;
    const int i = blockIdx.y * BI + il;
    const int j = blockIdx.z * BJ + jl;
    const bool cell_ok = k < nn && i < nr && j < nz;
    const bool act = cell_ok && k >= a.krange[0] && k <= a.krange[1];
    if (act) {
        auto put_r = [&](int f, int slot) {
            const m1a::FaceT<double> r = face_at<double>(a, 0, f, i, j, k, -1, -1);
            double* o = &rF[jl][slot][0][lane];
            o[0] = r.FJ; o[64] = r.FHr; o[128] = r.FHz; o[192] = r.kphi_L; o[256] = r.kphi_R;
        };
        auto put_z = [&](int g, int slot) {
            const m1a::FaceT<double> r = face_at<double>(a, 1, g, i, j, k, -1, -1);
            double* o = &zF[slot][il][0][lane];
            o[0] = r.FJ; o[64] = r.FHr; o[128] = r.FHz;
        };
        put_r(i, il);
        if (il == BI - 1 || i == nr - 1) put_r(i + 1, il + 1);
        put_z(j, jl);
        if (jl == BJ - 1 || j == nz - 1) put_z(j + 1, jl + 1);
    }
    __syncthreads();
    if (!cell_ok) return;
    const size_t idx = ((size_t)j * nr + i) * nn + k;
    const double lJ3 = a.lJ1[idx], lHr3 = a.lHr1[idx], lHz3 = a.lHz1[idx];
    if (!act) {
        a.pJ[idx] = 0.0; a.pHr[idx] = 0.0; a.pHz[idx] = 0.0;
        a.oJ[idx] = lJ3; a.oHr[idx] = lHr3; a.oHz[idx] = lHz3;
        return;
    }
    double Fr[2][3], Fz[2][3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        Fr[0][m] = rF[jl][il][m][lane];
        Fr[1][m] = rF[jl][il + 1][m][lane];
        Fz[0][m] = zF[jl][il][m][lane];
        Fz[1][m] = zF[jl + 1][il][m][lane];
    }
    const double kphi_im = rF[jl][il][4][lane], kphi_ip = rF[jl][il + 1][3][lane];
    const int c = j * nr + i;
    const double Ar_L = a.g.A_r[(size_t)j * (nr + 1) + i], Ar_R = a.g.A_r[(size_t)j * (nr + 1) + i + 1];
    const double Az_L = a.g.A_z[(size_t)j * nr + i], Az_R = a.g.A_z[(size_t)(j + 1) * nr + i];
    const double cdt = a.cdt;
    const double cdt_dV = cdt * a.g.inv_dV[c];
    const double cdt_over_c = cdt * kInvC;
    const double durdr = a.v.dur_dr[c], durdz = a.v.dur_dz[c], duzdr = a.v.duz_dr[c];
    const double duzdz = a.v.duz_dz[c], urr = a.v.ur_over_r[c];
    const double Js = a.J[idx], Hrs = a.Hr[idx], Hzs = a.Hz[idx];
    Edd D = D_edd(Hrs, Hzs, Js);
    const double Ddu = D_colon_grad_u(D, durdr, durdz, duzdr, duzdz, urr);
    const double Jn0 = Js - cdt_dV * ((Ar_R * Fr[1][0] - Ar_L * Fr[0][0]) + (Az_R * Fz[1][0] - Az_L * Fz[0][0]));
    const double Jn1 = dmax(Jn0, kJeps);
    const double Hzn0 = Hzs - cdt_dV * ((Ar_R * Fr[1][2] - Ar_L * Fr[0][2]) + (Az_R * Fz[1][2] - Az_L * Fz[0][2]));
    const double Hrn0 = Hrs - cdt_dV * ((Ar_R * Fr[1][1] - Ar_L * Fr[0][1]) + (Az_R * Fz[1][1] - Az_L * Fz[0][1]));
    const double Kav = 0.5 * (kphi_im + kphi_ip);
    const double Hrn1 = Hrn0 + cdt * Kav * a.g.inv_r_cell[i];
    const double Hrn2 = Hrn1 + (-cdt_over_c * (Hrs * durdr + Hzs * durdz));
    const double Hzn2 = Hzn0 + (-cdt_over_c * (Hrs * duzdr + Hzs * duzdz));
    const double Jm = Jn1 - cdt_over_c * Ddu * Js;
    const double Jn2 = dmax(Jm, kJeps);

    // EnforceCausalReducedFlux in reverse
    double lHrn2 = lHr3, lHzn2 = lHz3, lJn2 = lJ3;
    const double Hm2 = Hrn2 * Hrn2 + Hzn2 * Hzn2;
    if (Hm2 > Jn2 * Jn2) {
        const double hm = sqrt(Hm2);
        const double fac = Jn2 / hm;
        const double dot = Hrn2 * lHr3 + Hzn2 * lHz3;
        const double g = Jn2 * dot / (hm * hm * hm);
        lHrn2 = fac * lHr3 - g * Hrn2;
        lHzn2 = fac * lHz3 - g * Hzn2;
        lJn2 += dot / hm;
    }
    const double lJm = (Jm >= kJeps) ? lJn2 : 0.0;
    const double lDdu = -lJm * cdt_over_c * Js;
    double lJs = -lJm * cdt_over_c * Ddu;
    const double lJn0 = (Jn0 >= kJeps) ? lJm : 0.0;
    // aberration ΔH (src/AdvectJandH.jl:300-312)
    double lHrs = -cdt_over_c * (lHrn2 * durdr + lHzn2 * duzdr);
    double lHzs = -cdt_over_c * (lHrn2 * durdz + lHzn2 * duzdz);
    // identity parts of the updates
    lJs += lJn0;
    lHrs += lHrn2;
    lHzs += lHzn2;
    // D:∇u of the cell's own closure
    if (lDdu != 0.0) {
#pragma unroll 1
        for (int var = 0; var < 3; ++var) {
            Dl tJ(Js, var == 0 ? 1.0 : 0.0), tHr(Hrs, var == 1 ? 1.0 : 0.0), tHz(Hzs, var == 2 ? 1.0 : 0.0);
            m1a::EddT<Dl> Dd = m1a::tD_edd(tHr, tHz, tJ);
            const double d = m1a::tD_colon_grad_u(Dd, durdr, durdz, duzdr, duzdz, urr).d;
            if (var == 0) lJs += lDdu * d;
            else if (var == 1) lHrs += lDdu * d;
            else lHzs += lDdu * d;
        }
    }
    a.pJ[idx] = lJn0;
    a.pHr[idx] = lHrn2;   // the K_ϕϕ term and the flux divergence enter H_r with the same cotangent
    a.pHz[idx] = lHzn2;
    a.oJ[idx] = lJs;
    a.oHr[idx] = lHrs;
    a.oHz[idx] = lHzs;
}

// ---------------------------------------------------------------------------
// Interface contributions gathered per cell.
// ---------------------------------------------------------------------------
// two waves per SIMD: the Dl2 passes need about 280 VGPRs unbounded, 256 with a few spills
__global__ __launch_bounds__(256, 2) void k_adj_gather(AdjAdvArgs a)
{
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    const int j = blockIdx.z;
    if (k >= nn || i >= nr) return;
    if (k < a.krange[0] || k > a.krange[1]) return;
    const size_t idx = ((size_t)j * nr + i) * nn + k;
    const double cdt = a.cdt;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int t = 0; t < 8; ++t) {
        const int dir = t < 4 ? 0 : 1;
        const int f = dir == 0 ? i - 1 + t : j - 1 + (t - 4);
        const int nf = dir == 0 ? nr : nz;
        if (f < 0 || f > nf) continue;
        // cotangents of the interface outputs from the two cells it separates
        double lF[3] = {0.0, 0.0, 0.0}, lkL = 0.0, lkR = 0.0;
        if (f >= 1) {
            const int cl = dir == 0 ? j * nr + (f - 1) : (f - 1) * nr + i;
            const double A = dir == 0 ? a.g.A_r[(size_t)j * (nr + 1) + f] : a.g.A_z[(size_t)f * nr + i];
            const double s = -cdt * a.g.inv_dV[cl] * A;
            const size_t q = (size_t)cl * nn + k;
            lF[0] += s * a.pJ[q]; lF[1] += s * a.pHr[q]; lF[2] += s * a.pHz[q];
            if (dir == 0) lkL = 0.5 * cdt * a.g.inv_r_cell[f - 1] * a.pHr[q];
        }
        if (f <= nf - 1) {
            const int cr = dir == 0 ? j * nr + f : f * nr + i;
            const double A = dir == 0 ? a.g.A_r[(size_t)j * (nr + 1) + f] : a.g.A_z[(size_t)f * nr + i];
            const double s = cdt * a.g.inv_dV[cr] * A;
            const size_t q = (size_t)cr * nn + k;
            lF[0] += s * a.pJ[q]; lF[1] += s * a.pHr[q]; lF[2] += s * a.pHz[q];
            if (dir == 0) lkR = 0.5 * cdt * a.g.inv_r_cell[f] * a.pHr[q];
        }
        if (lF[0] == 0.0 && lF[1] == 0.0 && lF[2] == 0.0 && lkL == 0.0 && lkR == 0.0) continue;
        const int tc = dir == 0 ? i : j;
        // two directions per pass: (J, H_r), then (H_Z, α_F)
#pragma unroll 1
        for (int vp = 0; vp < 2; ++vp) {
            m1a::FaceT<Dl2> r = face_at<Dl2>(a, dir, f, i, j, k, tc, vp);
            acc[2 * vp] += lF[0] * r.FJ.d0 + lF[1] * r.FHr.d0 + lF[2] * r.FHz.d0 + lkL * r.kphi_L.d0 + lkR * r.kphi_R.d0;
            acc[2 * vp + 1] +=
                lF[0] * r.FJ.d1 + lF[1] * r.FHr.d1 + lF[2] * r.FHz.d1 + lkL * r.kphi_L.d1 + lkR * r.kphi_R.d1;
        }
    }
    a.oJ[idx] += acc[0];
    a.oHr[idx] += acc[1];
    a.oHz[idx] += acc[2];
    a.gaF[idx] += acc[3];
}

// ---------------------------------------------------------------------------
// Interface contributions through reverse-mode interfaces.  k_adj_face: thread per
// (bin, interface); the interface cotangents come from the two cells it separates,
// m1a::face_vjp turns them into cotangents of the interface's 14 inputs (J, H_r, H_Z
// of its four stencil slots, α_F on its two sides), stored as planes of fpart.
// k_adj_gather_rev: thread per (bin, cell); sums, over the <= 8 interfaces whose
// stencil holds the cell, the slots that read it, through the ghost map of each
// boundary (mirror: -1 on the normal flux; outflow clip: 0 or 1), and the α_F sides
// that are the cell.  Same contributions as k_adj_gather's dual-number passes, in a
// fixed order, without atomics.
// ---------------------------------------------------------------------------
namespace {

struct FacePlanes {
    size_t base, plane;   // plane p of interface index q at fpart[base + p * plane + q * nn + k]
};

__device__ __forceinline__ FacePlanes face_planes(int dir, int nn, int nr, int nz)
{
    const size_t nfr = (size_t)nz * (nr + 1) * nn, nfz = (size_t)nr * (nz + 1) * nn;
    return dir == 0 ? FacePlanes{0, nfr} : FacePlanes{14 * nfr, nfz};
}

// Ghost map of stencil cell c along one direction (as tload_r / tload_z): the source
// cell, and in fac the derivative of the slot's normal flux component (H_r along r,
// H_Z along Z) with respect to the source's; J and the other component map with 1.
// vnorm: the source's normal component, which decides the outflow clips.
__device__ __forceinline__ int ghost_src(int c, int n, int dir, int low_reflect, int high_reflect, double vnorm,
                                         double& fac)
{
    fac = 1.0;
    if (c < 0) {
        if (dir == 0 || low_reflect) { fac = -1.0; return -c - 1; }
        fac = (0.0 < vnorm) ? 0.0 : 1.0;   // min(H_Z, 0) at an open bottom
        return 0;
    }
    if (c < n) return c;
    if (high_reflect) { fac = -1.0; return 2 * n - 1 - c; }
    fac = (0.0 > vnorm) ? 0.0 : 1.0;       // max(H, 0) at an open outer / top boundary
    return n - 1;
}

}  // namespace

template <int DIR>
__global__ __launch_bounds__(256, 2) void k_adj_face(AdjAdvArgs a)
{
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    const int k = blockIdx.x * 64 + threadIdx.x;
    // DIR 0: interface f = 0..N_r of row j; DIR 1: interface f = 0..N_Z of column i
    const int f = DIR == 0 ? (int)(blockIdx.y * 4 + threadIdx.y) : (int)blockIdx.z;
    const int i = DIR == 0 ? 0 : (int)(blockIdx.y * 4 + threadIdx.y);
    const int j = DIR == 0 ? (int)blockIdx.z : 0;
    const int nf = DIR == 0 ? nr : nz;
    if (k >= nn || f > nf || (DIR == 1 && i >= nr)) return;
    if (k < a.krange[0] || k > a.krange[1]) return;
    const double cdt = a.cdt;
    double lF[3] = {0.0, 0.0, 0.0}, lkL = 0.0, lkR = 0.0;
    if (f >= 1) {
        const int cl = DIR == 0 ? j * nr + (f - 1) : (f - 1) * nr + i;
        const double A = DIR == 0 ? a.g.A_r[(size_t)j * (nr + 1) + f] : a.g.A_z[(size_t)f * nr + i];
        const double s = -cdt * a.g.inv_dV[cl] * A;
        const size_t q = (size_t)cl * nn + k;
        lF[0] += s * a.pJ[q]; lF[1] += s * a.pHr[q]; lF[2] += s * a.pHz[q];
        if (DIR == 0) lkL = 0.5 * cdt * a.g.inv_r_cell[f - 1] * a.pHr[q];
    }
    if (f <= nf - 1) {
        const int cr = DIR == 0 ? j * nr + f : f * nr + i;
        const double A = DIR == 0 ? a.g.A_r[(size_t)j * (nr + 1) + f] : a.g.A_z[(size_t)f * nr + i];
        const double s = cdt * a.g.inv_dV[cr] * A;
        const size_t q = (size_t)cr * nn + k;
        lF[0] += s * a.pJ[q]; lF[1] += s * a.pHr[q]; lF[2] += s * a.pHz[q];
        if (DIR == 0) lkR = 0.5 * cdt * a.g.inv_r_cell[f] * a.pHr[q];
    }
    const FacePlanes P = face_planes(DIR, nn, nr, nz);
    const size_t q0 = P.base + (size_t)(DIR == 0 ? j * (nr + 1) + f : f * nr + i) * nn + k;
    m1a::FaceGrad g;
    if (lF[0] == 0.0 && lF[1] == 0.0 && lF[2] == 0.0 && lkL == 0.0 && lkR == 0.0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) { g.J[m] = 0.0; g.Hr[m] = 0.0; g.Hz[m] = 0.0; }
        g.aF_L = 0.0; g.aF_R = 0.0;
    } else {
        double J[4], Hr[4], Hz[4];
        const int cells[4] = {f - 1, f, f + 1, f - 2};
        double aFL, aFR, u;
        const double* iv;
        if (DIR == 0) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
                tload_r<double>(a.J, a.Hr, a.Hz, nn, nr, j, k, cells[m], a.bc_r_right, -1, -1, J[m], Hr[m], Hz[m]);
            iv = a.g.ivr + (size_t)f * 5;
            aFL = a.aF[((size_t)j * nr + max(f - 1, 0)) * nn + k];
            aFR = a.aF[((size_t)j * nr + min(f, nr - 1)) * nn + k];
            u = a.v.ur_star[(size_t)j * (nr + 1) + f] * kInvC;
        } else {
#pragma unroll
            for (int m = 0; m < 4; ++m)
                tload_z<double>(a.J, a.Hr, a.Hz, nn, nr, nz, i, k, cells[m], a.bc_z_bot, a.bc_z_top, -1, -1, J[m],
                                Hr[m], Hz[m]);
            iv = a.g.ivz + (size_t)f * 5;
            aFL = a.aF[((size_t)max(f - 1, 0) * nr + i) * nn + k];
            aFR = a.aF[((size_t)min(f, nz - 1) * nr + i) * nn + k];
            u = a.v.uz_star[(size_t)f * nr + i] * kInvC;
        }
        g = m1a::face_vjp(DIR, a.hll, DIR == 0 && f == 0, J, Hr, Hz, iv, aFL, aFR, u, lF[0], lF[1], lF[2], lkL, lkR);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        a.fpart[q0 + (size_t)(3 * m) * P.plane] = g.J[m];
        a.fpart[q0 + (size_t)(3 * m + 1) * P.plane] = g.Hr[m];
        a.fpart[q0 + (size_t)(3 * m + 2) * P.plane] = g.Hz[m];
    }
    a.fpart[q0 + (size_t)12 * P.plane] = g.aF_L;
    a.fpart[q0 + (size_t)13 * P.plane] = g.aF_R;
}

__global__ __launch_bounds__(256) void k_adj_gather_rev(AdjAdvArgs a)
{
    const int nn = a.nn, nr = a.g.nr, nz = a.g.nz;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    const int j = blockIdx.z;
    if (k >= nn || i >= nr) return;
    if (k < a.krange[0] || k > a.krange[1]) return;
    const size_t idx = ((size_t)j * nr + i) * nn + k;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int dir = 0; dir < 2; ++dir) {
        const FacePlanes P = face_planes(dir, nn, nr, nz);
        const int tc = dir == 0 ? i : j;
        const int n = dir == 0 ? nr : nz;
        const int low_reflect = dir == 0 ? 1 : (a.bc_z_bot == BC_REFLECT);
        const int high_reflect = dir == 0 ? (a.bc_r_right == BC_REFLECT) : (a.bc_z_top == BC_REFLECT);
        const double vnorm = dir == 0 ? a.Hr[idx] : a.Hz[idx];
#pragma unroll 1
        for (int t = 0; t < 4; ++t) {
            const int f = tc - 1 + t;
            if (f < 0 || f > n) continue;
            const size_t q0 = P.base + (size_t)(dir == 0 ? j * (nr + 1) + f : f * nr + i) * nn + k;
            const int cells[4] = {f - 1, f, f + 1, f - 2};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                double fac;
                const int src = ghost_src(cells[m], n, dir, low_reflect, high_reflect, vnorm, fac);
                if (src != tc) continue;
                acc[0] += a.fpart[q0 + (size_t)(3 * m) * P.plane];
                const double gr = a.fpart[q0 + (size_t)(3 * m + 1) * P.plane];
                const double gz = a.fpart[q0 + (size_t)(3 * m + 2) * P.plane];
                if (dir == 0) { acc[1] += fac * gr; acc[2] += gz; }
                else { acc[1] += gr; acc[2] += fac * gz; }
            }
            if (max(f - 1, 0) == tc) acc[3] += a.fpart[q0 + (size_t)12 * P.plane];
            if (min(f, n - 1) == tc) acc[3] += a.fpart[q0 + (size_t)13 * P.plane];
        }
    }
    a.oJ[idx] += acc[0];
    a.oHr[idx] += acc[1];
    a.oHz[idx] += acc[2];
    a.gaF[idx] += acc[3];
}

// ---------------------------------------------------------------------------
// Emission/absorption and the frequency update in reverse: one wavefront per
// cell, lanes over bins; the two tridiagonal sweeps (the forward system for x,
// its transpose for mu) run on lane 0, interleaved.
// ---------------------------------------------------------------------------
namespace {
// Tridiagonal solve of n <= 128 rows (row r: A[r] x[r-1] + B[r] x[r] + C[r] x[r+1] = d[r], an
// M-matrix, no pivoting) by one wavefront: lane l holds rows 2l, 2l+1; one cyclic-reduction level
// folds the odd rows into the even ones, parallel cyclic reduction over the 64 lanes solves the
// even rows, back substitution gives the odd ones.  transposed: solve M^T instead (sub-diagonal
// C[r-1], super-diagonal A[r+1]).  rhs(r) gives d[r]; x receives the solution (lane-own rows,
// written after every read, so x may alias the storage rhs reads).
template <class Rhs>
__device__ __forceinline__ void tri_pcr(int n, const double* A, const double* B, const double* C, Rhs rhs,
                                        bool transposed, double* x)
{
    const int lane = threadIdx.x & 63;
    double ra[2], rb[2], rc[2], rd[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = 2 * lane + q;
        if (r < n) {
            ra[q] = (r == 0) ? 0.0 : (transposed ? C[r - 1] : A[r]);
            rc[q] = (r == n - 1) ? 0.0 : (transposed ? A[r + 1] : C[r]);
            rb[q] = B[r];
            rd[q] = rhs(r);
        } else {
            ra[q] = 0.0; rb[q] = 1.0; rc[q] = 0.0; rd[q] = 0.0;
        }
    }
    // odd row into the even rows: x1 = ud - ua x0 - uc x0(next lane)
    const double ib1 = 1.0 / rb[1];
    const double ua = ra[1] * ib1, uc = rc[1] * ib1, ud = rd[1] * ib1;
    double pua = __shfl_up(ua, 1, 64), puc = __shfl_up(uc, 1, 64), pud = __shfl_up(ud, 1, 64);
    if (lane == 0) { pua = 0.0; puc = 0.0; pud = 0.0; }
    double b = rb[0] - ra[0] * puc - rc[0] * ua;
    double d = rd[0] - ra[0] * pud - rc[0] * ud;
    double a = -ra[0] * pua;
    double c = -rc[0] * uc;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const double ib = 1.0 / b;
        const double u = a * ib, v = c * ib, w = d * ib;
        double ub = __shfl_up(u, s, 64), vb = __shfl_up(v, s, 64), wb = __shfl_up(w, s, 64);
        double uf = __shfl_down(u, s, 64), vf = __shfl_down(v, s, 64), wf = __shfl_down(w, s, 64);
        if (lane < s) { ub = 0.0; vb = 0.0; wb = 0.0; }
        if (lane + s >= 64) { uf = 0.0; vf = 0.0; wf = 0.0; }
        const double nb = b - a * vb - c * uf;
        const double nd = d - a * wb - c * wf;
        a = -a * ub;
        c = -c * vf;
        b = nb;
        d = nd;
    }
    const double x0 = d / b;
    double x0n = __shfl_down(x0, 1, 64);
    if (lane == 63) x0n = 0.0;
    const double x1 = ud - ua * x0 - uc * x0n;
    if (2 * lane < n) x[2 * lane] = x0;
    if (2 * lane + 1 < n) x[2 * lane + 1] = x1;
}

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
}  // namespace

__global__ __launch_bounds__(64) void k_adj_column(AdjColArgs a)
{
    extern __shared__ double lds[];
    const int nn = a.nn, nr = a.nr;
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int i = c % nr, j = c / nr;
    const size_t base = (size_t)c * nn;
    double* X2J = lds;
    double* X2r = lds + nn;
    double* X2z = lds + 2 * nn;
    double* rA = lds + 3 * nn;
    double* rB = lds + 4 * nn;
    double* rC = lds + 5 * nn;
    double* x = lds + 6 * nn;
    double* mu = lds + 7 * nn;
    double* cp = lds + 8 * nn;
    double* cq = lds + 9 * nn;
    double* lJ = lds + 10 * nn;    // cotangent of E's output, then of the advection output
    double* lHr = lds + 11 * nn;
    double* lHz = lds + 12 * nn;
    double* gm = lds + 13 * nn;    // row k's derivative contributions to Hvoigt at k-1, k, k+1
    double* g0 = lds + 14 * nn;
    double* gp = lds + 15 * nn;
    const double ct = a.c_tilde, dt = a.dt, dw = a.dw;
    const double fpc_ict = (kFourPi / kC) * (1.0 / ct);
    const double ac = a.alpha_c[c], al = a.alpha[c], pd = a.pd[c], jsv = a.js[c];
    const double* Hv = a.Hv + base;
    const double* phi = a.phi + base;
    const double* aF = a.aF + base;

    // ---- E forward (values) and the momentum deposit Δρu ----
    double sr = 0.0, sz = 0.0;
    for (int k = lane; k < nn; k += 64) {
        const double aJ = ac + al * pd * Hv[k];
        const double dJdt = jsv * ct * phi[k];
        X2J[k] = (a.J1[base + k] + dJdt * dt) / (1.0 + aJ * ct * dt);
        const double denH = 1.0 + aF[k] * ct * dt;
        const double h1r = a.Hr1[base + k], h1z = a.Hz1[base + k];
        X2r[k] = h1r / denH;
        X2z[k] = h1z / denH;
        sr += a.dnu[k] * (h1r - X2r[k]);
        sz += a.dnu[k] * (h1z - X2z[k]);
        gm[k] = 0.0; g0[k] = 0.0; gp[k] = 0.0;
    }
    sr = wsum(sr);
    sz = wsum(sz);
    // ---- cotangent of Δρu from the seeds of M_F, M_F_radial_net, M_F_radial_out ----
    double lDr = 0.0, lDz = 0.0;
    if (a.mf_scale != 0.0) {
        const double fr = fpc_ict * sr / dt, fz = fpc_ict * sz / dt;
        const double s = a.mf_scale * a.dV[c];
        double lfr = 0.0, lfz = 0.0;
        const double fm = sqrt(fr * fr + fz * fz);
        if (fm > 0.0) { lfr += a.w0 * s * fr / fm; lfz += a.w0 * s * fz / fm; }
        const double dr = a.rcom[i], dz = a.zcom[j] - a.z_center;
        const double dist = sqrt(dr * dr + dz * dz);
        if (dist > 0.0) {
            const double er = dr / dist, ez = dz / dist;
            const double frad = fr * er + fz * ez;
            const double w = a.w1 + ((frad > 0.0) ? a.w2 : 0.0);
            lfr += w * s * er;
            lfz += w * s * ez;
        }
        lDr = lfr / dt;
        lDz = lfz / dt;
    }

    const int n = nn - 2;
    const double dnuD = a.dnuD[c];
    const double amax = dmax(al, a.alpha_floor);
    const double PREFACT = 0.5 * amax * (dnuD * dnuD);
    const double Tc = a.T[c];
    const double RECOIL_0 = kHoverKB / Tc;
    const double durdr = a.dur_dr[c], durdz = a.dur_dz[c], duzdr = a.duz_dr[c], duzdz = a.duz_dz[c];
    const double urr = a.ur_over_r[c];
    auto row_const = [&](int k) {
        m1a::RowConst rc;
        rc.nuCk = a.nuC[k]; rc.nuCkp = a.nuC[k + 1]; rc.nuCkm = a.nuC[k - 1];
        rc.dkp = rc.nuCkp - a.nu[k + 1];
        rc.dk = a.nu[k + 1] - rc.nuCk;
        rc.dkm = a.nu[k] - rc.nuCkm;
        rc.inv_dnu_k = a.inv_dnu[k];
        rc.two_inv_k = 2.0 * a.inv_nuC[k];
        rc.two_inv_kp = 2.0 * a.inv_nuC[k + 1];
        rc.two_inv_km = 2.0 * a.inv_nuC[k - 1];
        rc.recoil = a.recoil;
        return rc;
    };
    __syncthreads();

    // ---- C forward: the rows of the coupling matrix ----
    for (int k = lane + 1; k <= nn - 2; k += 64) {
        Edd D = D_edd(X2r[k], X2z[k], X2J[k]);
        const double VEL = D_colon_grad_u(D, durdr, durdz, duzdr, duzdz, urr) * kInvC;
        m1a::RowT<double> row = m1a::tfreq_row<double>(row_const(k), Hv[k - 1], Hv[k], Hv[k + 1], PREFACT,
                                                        RECOIL_0, VEL);
        rA[k - 1] = row.Fkm;
        rB[k - 1] = -(row.Fk + dw);
        rC[k - 1] = row.Fkp;
    }
    // ---- flux correction in reverse (src/UpdateLyaJandH.jl:335-379) ----
    for (int k = lane; k < nn; k += 64) {
        const double J4 = a.J4[base + k];
        const double J2 = X2J[k], Hr2 = X2r[k], Hz2 = X2z[k];
        const double inv_J2 = 1.0 / dmax(J2, kJeps);
        const double ratio = J4 * inv_J2;
        const double hr = Hr2 * ratio, hz = Hz2 * ratio;
        double lJ4 = a.lJ4[base + k], lhr = a.lHr4[base + k], lhz = a.lHz4[base + k];
        const double Hm2 = hr * hr + hz * hz;
        if (Hm2 > J4 * J4) {
            const double hm = sqrt(Hm2);
            const double s = J4 / hm;
            const double dot = hr * lhr + hz * lhz;
            const double g = J4 * dot / (hm * hm * hm);
            lJ4 += dot / hm;
            const double nhr = s * lhr - g * hr, nhz = s * lhz - g * hz;
            lhr = nhr;
            lhz = nhz;
        }
        const double lr = Hr2 * lhr + Hz2 * lhz;
        lJ4 += lr * inv_J2;
        double lJ2 = (J2 >= kJeps) ? -lr * ratio * inv_J2 : 0.0;
        const double lJ3 = (J4 > kJeps) ? lJ4 : 0.0;
        if (k == 0 || k == nn - 1) lJ2 += lJ3;   // bins outside the solve keep E's value
        else mu[k - 1] = lJ3;                      // right-hand side of the transposed solve
        lJ[k] = lJ2;
        lHr[k] = ratio * lhr;
        lHz[k] = ratio * lhz;
    }
    __syncthreads();

    // ---- the forward system M x = -dw J2 and M^T mu = lJ3 (M is an M-matrix): cyclic
    // reduction over the wavefront's lanes up to 128 rows, else Thomas on lane 0 ----
    if (n >= 1 && n <= 128) {
        tri_pcr(n, rA, rB, rC, [&](int r) { return -dw * X2J[r + 1]; }, false, x);
        tri_pcr(n, rA, rB, rC, [&](int r) { return mu[r]; }, true, mu);
    } else if (lane == 0 && n >= 1) {
        double cpr = rC[0] / rB[0], xr = (-dw * X2J[1]) / rB[0];
        double cqr = (n > 1 ? rA[1] : 0.0) / rB[0], mur = mu[0] / rB[0];
        cp[0] = cpr; x[0] = xr; cq[0] = cqr; mu[0] = mur;
        for (int r = 1; r < n; ++r) {
            const double m = rB[r] - rA[r] * cpr;
            const double mt = rB[r] - rC[r - 1] * cqr;
            cpr = rC[r] / m;
            xr = ((-dw * X2J[r + 1]) - rA[r] * xr) / m;
            cqr = (r + 1 < n ? rA[r + 1] : 0.0) / mt;
            mur = (mu[r] - rC[r - 1] * mur) / mt;
            cp[r] = cpr; x[r] = xr; cq[r] = cqr; mu[r] = mur;
        }
        for (int r = n - 2; r >= 0; --r) {
            xr = x[r] - cp[r] * xr;
            mur = mu[r] - cq[r] * mur;
            x[r] = xr;
            mu[r] = mur;
        }
    }
    __syncthreads();

    // ---- matrix cotangents contracted with each row's local derivatives ----
    double lPREF = 0.0, lREC = 0.0;
    for (int k = lane + 1; k <= nn - 2; k += 64) {
        const int r = k - 1;
        const double m_r = mu[r];
        lJ[k] += -dw * m_r;   // b = -dw J2
        const double lFkm = (r >= 1) ? -m_r * x[r - 1] : 0.0;
        const double lFk = m_r * x[r];                       // B = -(Fk + dw)
        const double lFkp = (r <= n - 2) ? -m_r * x[r + 1] : 0.0;
        if (lFkm == 0.0 && lFk == 0.0 && lFkp == 0.0) continue;
        Edd D = D_edd(X2r[k], X2z[k], X2J[k]);
        const double VEL = D_colon_grad_u(D, durdr, durdz, duzdr, duzdz, urr) * kInvC;
        const m1a::RowConst rc = row_const(k);
        // derivatives along the six inputs: reverse mode, or (mode 1) two per dual pass
        double g[6];
        if (a.mode == 0) {
            m1a::freq_row_vjp(rc, Hv[k - 1], Hv[k], Hv[k + 1], PREFACT, RECOIL_0, VEL, lFkp, lFk, lFkm, g);
        } else {
#pragma unroll 1
        for (int dp = 0; dp < 3; ++dp) {
            auto in = [&](double v, int d) { return Dl2(v, d == 2 * dp ? 1.0 : 0.0, d == 2 * dp + 1 ? 1.0 : 0.0); };
            const Dl2 hm = in(Hv[k - 1], 0), h0 = in(Hv[k], 1), hp = in(Hv[k + 1], 2);
            const Dl2 pf = in(PREFACT, 3), rec = in(RECOIL_0, 4), vel = in(VEL, 5);
            m1a::RowT<Dl2> row = m1a::tfreq_row<Dl2>(rc, hm, h0, hp, pf, rec, vel);
            g[2 * dp] = lFkp * row.Fkp.d0 + lFk * row.Fk.d0 + lFkm * row.Fkm.d0;
            g[2 * dp + 1] = lFkp * row.Fkp.d1 + lFk * row.Fk.d1 + lFkm * row.Fkm.d1;
        }
        }
        if (!a.recoil) g[4] = 0.0;
        gm[k] = g[0];
        g0[k] = g[1];
        gp[k] = g[2];
        lPREF += g[3];
        lREC += g[4];
        const double lDdu = g[5] * kInvC;
        if (lDdu != 0.0) {
            // D:∇u along (J, H_r) in one pass, then along H_Z
            const Dl2 tJ(X2J[k], 1.0, 0.0), tHr(X2r[k], 0.0, 1.0), tHz(X2z[k]);
            const Dl2 dd2 = m1a::tD_colon_grad_u(m1a::tD_edd(tHr, tHz, tJ), durdr, durdz, duzdr, duzdz, urr);
            lJ[k] += lDdu * dd2.d0;
            lHr[k] += lDdu * dd2.d1;
            const Dl uJ(X2J[k]), uHr(X2r[k]), uHz(X2z[k], 1.0);
            lHz[k] += lDdu * m1a::tD_colon_grad_u(m1a::tD_edd(uHr, uHz, uJ), durdr, durdz, duzdr, duzdz, urr).d;
        }
    }
    __syncthreads();

    // ---- E in reverse (src/ConstantsAndFunctions.jl:451-573) ----
    double lal = 0.0, lac = 0.0, ljs = 0.0;
    for (int k = lane; k < nn; k += 64) {
        const double aJ = ac + al * pd * Hv[k];
        const double den = 1.0 + aJ * ct * dt;
        const double lJ2 = lJ[k];
        a.lJ1[base + k] = lJ2 / den;
        const double q = lJ2 * dt * ct / den;
        a.g_phi[base + k] += q * jsv;
        ljs += q * phi[k];
        const double laJ = -q * X2J[k];
        lac += laJ;
        lal += laJ * pd * Hv[k];
        // Hvoigt: emission's α p_d term plus the rows k-1, k, k+1 of the coupling matrix
        double gh = laJ * al * pd + g0[k];
        if (k >= 1) gh += gp[k - 1];
        if (k + 1 < nn) gh += gm[k + 1];
        a.g_Hv[base + k] += gh;
        const double denH = 1.0 + aF[k] * ct * dt;
        const double dr_k = lDr * fpc_ict * a.dnu[k], dz_k = lDz * fpc_ict * a.dnu[k];
        const double lHr2 = lHr[k], lHz2 = lHz[k];
        a.lHr1[base + k] = lHr2 / denH + dr_k * (1.0 - 1.0 / denH);
        a.lHz1[base + k] = lHz2 / denH + dz_k * (1.0 - 1.0 / denH);
        a.g_aF[base + k] += -((lHr2 - dr_k) * X2r[k] + (lHz2 - dz_k) * X2z[k]) * ct * dt / denH;
    }
    lal = wsum(lal);
    lac = wsum(lac);
    ljs = wsum(ljs);
    lPREF = wsum(lPREF);
    lREC = wsum(lREC);
    if (lane == 0) {
        // PREFACT = ½ max(α, α_floor) ΔνD², RECOIL_0 = h/(k_B T)
        a.g_alpha[c] += lal + ((a.alpha_floor > al) ? 0.0 : lPREF * 0.5 * (dnuD * dnuD));
        a.g_alpha_c[c] += lac;
        a.g_js[c] += ljs;
        a.g_dnuD[c] += lPREF * amax * dnuD;
        a.g_T[c] += -lREC * kHoverKB / (Tc * Tc);
    }
}

// Escaping luminosity per bin over L_sim (the integrand of f_esc,
// src/ConstantsAndFunctions.jl:914-993): one block per bin.
__global__ __launch_bounds__(256) void k_adj_spectrum(const double* Hr, const double* Hz, const double* A_r,
                                                      const double* A_z, int nn, int nr, int nz, int bc_z_bot,
                                                      double inv_Lsim, double* S)
{
    __shared__ double sh[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int t = threadIdx.x; t < nz + 2 * nr; t += blockDim.x) {
        if (t < nz) {
            s += A_r[(size_t)t * (nr + 1) + nr] * kFourPi * dmax(Hr[((size_t)t * nr + nr - 1) * nn + k], 0.0);
        } else if (t < nz + nr) {
            const int i = t - nz;
            s += A_z[(size_t)nz * nr + i] * kFourPi * dmax(Hz[((size_t)(nz - 1) * nr + i) * nn + k], 0.0);
        } else if (bc_z_bot == BC_OUTFLOW) {
            const int i = t - nz - nr;
            s += A_z[i] * kFourPi * dmax(-Hz[(size_t)i * nn + k], 0.0);
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) S[k] = sh[0] * inv_Lsim;
}

// Seeds Σ_k v_k S_k of the final state into the state cotangent.
__global__ __launch_bounds__(256) void k_adj_spectrum_seed(const double* Hr, const double* Hz, const double* A_r,
                                                           const double* A_z, int nn, int nr, int nz, int bc_z_bot,
                                                           double inv_Lsim, const double* v, double* lHr,
                                                           double* lHz)
{
    const int k = blockIdx.x;
    for (int t = threadIdx.x; t < nz + 2 * nr; t += blockDim.x) {
        const double w = v[k] * kFourPi * inv_Lsim;
        if (t < nz) {
            const size_t q = ((size_t)t * nr + nr - 1) * nn + k;
            if (Hr[q] > 0.0) lHr[q] += w * A_r[(size_t)t * (nr + 1) + nr];
        } else if (t < nz + nr) {
            const int i = t - nz;
            const size_t q = ((size_t)(nz - 1) * nr + i) * nn + k;
            if (Hz[q] > 0.0) lHz[q] += w * A_z[(size_t)nz * nr + i];
        } else if (bc_z_bot == BC_OUTFLOW && nz > 1) {
            const int i = t - nz - nr;
            const size_t q = (size_t)i * nn + k;
            if (-Hz[q] > 0.0) lHz[q] -= w * A_z[i];
        }
    }
}

// y[c] += s * dV[c]
__global__ void k_adj_axpy_cell(double* y, const double* dV, double s, int n)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) y[c] += s * dV[c];
}

void launch_adj_column(hipStream_t s, const AdjColArgs& a)
{
    const int ncell = a.nr * a.nz;
    const size_t lds = 16 * (size_t)a.nn * sizeof(double);
    hipLaunchKernelGGL(k_adj_column, dim3(ncell), dim3(64), lds, s, a);
}

void launch_adj_advect(hipStream_t s, const AdjAdvArgs& a)
{
    dim3 grid((a.nn + 63) / 64, (a.g.nr + 3) / 4, a.g.nz);
    constexpr int BI = 2, BJ = 4;   // 2.75 interface evaluations per cell
    hipLaunchKernelGGL((k_adj_cell<BI, BJ>), dim3((a.nn + 63) / 64, (a.g.nr + BI - 1) / BI, (a.g.nz + BJ - 1) / BJ),
                       dim3(64, BI, BJ), 0, s, a);
    if (a.mode == 1) {
        hipLaunchKernelGGL(k_adj_gather, grid, dim3(64, 4), 0, s, a);
        return;
    }
    hipLaunchKernelGGL(k_adj_face<0>, dim3((a.nn + 63) / 64, (a.g.nr + 1 + 3) / 4, a.g.nz), dim3(64, 4), 0, s, a);
    hipLaunchKernelGGL(k_adj_face<1>, dim3((a.nn + 63) / 64, (a.g.nr + 3) / 4, a.g.nz + 1), dim3(64, 4), 0, s, a);
    hipLaunchKernelGGL(k_adj_gather_rev, grid, dim3(64, 4), 0, s, a);
}

void launch_adj_spectrum(hipStream_t s, const double* Hr, const double* Hz, const double* A_r, const double* A_z,
                         int nn, int nr, int nz, int bc_z_bot, double inv_Lsim, double* S)
{
    hipLaunchKernelGGL(k_adj_spectrum, dim3(nn), dim3(256), 0, s, Hr, Hz, A_r, A_z, nn, nr, nz, bc_z_bot, inv_Lsim,
                       S);
}

void launch_adj_spectrum_seed(hipStream_t s, const double* Hr, const double* Hz, const double* A_r,
                              const double* A_z, int nn, int nr, int nz, int bc_z_bot, double inv_Lsim,
                              const double* v, double* lHr, double* lHz)
{
    hipLaunchKernelGGL(k_adj_spectrum_seed, dim3(nn), dim3(256), 0, s, Hr, Hz, A_r, A_z, nn, nr, nz, bc_z_bot,
                       inv_Lsim, v, lHr, lHz);
}

void launch_adj_axpy_cell(hipStream_t s, double* y, const double* dV, double a, int n)
{
    hipLaunchKernelGGL(k_adj_axpy_cell, dim3((n + 255) / 256), dim3(256), 0, s, y, dV, a, n);
}

}  // namespace m1
