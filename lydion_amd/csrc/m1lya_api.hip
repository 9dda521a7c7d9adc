// m1lya_api.hip -- the C ABI of libm1lya (include/m1lya.h): context, device
// memory, the operator entry points that replace the reference Julia
// functions, and the fused time step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/m1lya.h"
#include "m1lya_internal.h"

using namespace m1;

namespace {

thread_local std::string g_last_error;

struct KStat {
    long launches = 0;
    double total_ms = 0.0;
};

struct Pending {
    std::string name;
    hipEvent_t e0, e1;
};

}  // namespace

struct m1lya_ctx {
    int nn = 0, nr = 0, nz = 0, precision = 0;
    size_t ncell = 0, n3 = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    // spectral state: cur set and the second set ("old" arrays / ping-pong target)
    double *J = nullptr, *Hr = nullptr, *Hz = nullptr;
    double *Jo = nullptr, *Hro = nullptr, *Hzo = nullptr;
    bool old_stale = false;   // after fused steps the *_old arrays equal the current state
    bool fij_stale = false;   // Fij_* not maintained by fused steps
    bool krange_valid = false;
    int adj_mode = 0;                      // adjoint interfaces: 0 reverse mode, 1 dual-number passes

    double *aF = nullptr, *Hv = nullptr, *phi = nullptr;
    double *Fkp = nullptr, *Fk = nullptr, *Fkm = nullptr;

    double *alpha = nullptr, *alpha_c = nullptr, *pd = nullptr, *js = nullptr, *T = nullptr, *dnuD = nullptr;
    double *dur_dr = nullptr, *dur_dz = nullptr, *duz_dr = nullptr, *duz_dz = nullptr, *ur_over_r = nullptr;
    double *drhou_r = nullptr, *drhou_z = nullptr, *f_r = nullptr, *f_z = nullptr, *e_lya = nullptr;
    double *ur_star = nullptr, *uz_star = nullptr;

    // geometry
    double *A_r = nullptr, *A_z = nullptr, *dV = nullptr, *inv_dV = nullptr, *inv_r_cell = nullptr;
    double *rcom = nullptr, *zcom = nullptr, *ivr = nullptr, *ivz = nullptr;
    double *r_if = nullptr, *z_if = nullptr;   // cell interfaces (RHD time-step control)
    bool geom_set = false, interfaces_set = false;

    // RHD sub-step and time-step control: staging of the host inputs (allocated on first use)
    double* rhd = nullptr;                 // [16][ncell]
    int* rhd_redo = nullptr;               // [ncell]

    // device-side coefficient refresh: Voigt tables, line constants, scratch inputs
    double *lut = nullptr, *dnu_rel = nullptr, *dnuC_rel = nullptr, *nHI = nullptr, *acF = nullptr;
    int lut_nx = 0;
    double lut_xmax = 0.0;
    LineConst lc{};

    // frequency grid
    double *nu = nullptr, *nuC = nullptr, *dnu = nullptr, *inv_dnu = nullptr, *inv_nuC = nullptr;
    int kcore_min = 0, kcore_max = 0, recoil = 0, flux = 0;
    double alpha_floor = 0.0;
    bool grid_set = false;

    // reductions and scratch
    int col_blocks = 0, act_blocks = 0, pstride = 0, prows = 0;
    double *part = nullptr, *red = nullptr, *scalars = nullptr;
    int *krange = nullptr, *err_flag = nullptr;
    StepCtl* ctl = nullptr;                // halt flag and executed-step count (m1lya_run_until_steady)
    double *edge_buf = nullptr;
    double *fg_r = nullptr, *fg_z = nullptr;
    double *history = nullptr;
    size_t history_cap = 0;

    std::vector<void*> allocs;
    std::string err;

    int profiling = 0;   // 0 off, 1 every launch, 2 the fused step's column kernel only
    std::map<std::string, KStat> stats;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};

namespace {

struct Fail {
    int code;
};

void fail(m1lya_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    g_last_error = msg;
    throw Fail{code};
}

void hipchk(m1lya_ctx* c, hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(c, M1LYA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

double* dalloc(m1lya_ctx* c, size_t n)
{
    void* p = nullptr;
    hipchk(c, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)), "hipMalloc");
    c->allocs.push_back(p);
    hipchk(c, hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(double)), "hipMemset");
    return (double*)p;
}

int* ialloc(m1lya_ctx* c, size_t n)
{
    void* p = nullptr;
    hipchk(c, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(int)), "hipMalloc");
    c->allocs.push_back(p);
    hipchk(c, hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(int)), "hipMemset");
    return (int*)p;
}

hipEvent_t get_event(m1lya_ctx* c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipchk(c, hipEventCreate(&e), "hipEventCreate");
    return e;
}

// Brackets one kernel launch with events on the context stream.
struct Prof {
    m1lya_ctx* c;
    const char* name;
    hipEvent_t e0{}, e1{};
    bool on;
    // dominant: the launch is bracketed at profiling level 2 as well as at level 1
    Prof(m1lya_ctx* c_, const char* n, bool dominant = false)
        : c(c_), name(n), on(c_->profiling == 1 || (c_->profiling == 2 && dominant))
    {
        if (on) {
            e0 = get_event(c);
            e1 = get_event(c);
            hipchk(c, hipEventRecord(e0, c->stream), "hipEventRecord");
        }
    }
    void done()
    {
        hipchk(c, hipGetLastError(), name);
        if (on) {
            hipchk(c, hipEventRecord(e1, c->stream), "hipEventRecord");
            c->pending.push_back(Pending{name, e0, e1});
        }
    }
};

void collect_stats(m1lya_ctx* c)
{
    if (c->pending.empty()) return;
    hipchk(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (auto& p : c->pending) {
        float ms = 0.f;
        hipchk(c, hipEventElapsedTime(&ms, p.e0, p.e1), "hipEventElapsedTime");
        KStat& s = c->stats[p.name];
        s.launches += 1;
        s.total_ms += ms;
        c->pool.push_back(p.e0);
        c->pool.push_back(p.e1);
    }
    c->pending.clear();
}

struct FieldInfo {
    double* ptr;
    size_t count;
};

FieldInfo field_info(m1lya_ctx* c, int f, bool for_download)
{
    const size_t n3 = c->n3, nc = c->ncell;
    const bool stale = c->old_stale && for_download;
    switch (f) {
    case M1LYA_J: return {c->J, n3};
    case M1LYA_J_OLD: return {stale ? c->J : c->Jo, n3};
    case M1LYA_H_R: return {c->Hr, n3};
    case M1LYA_H_Z: return {c->Hz, n3};
    case M1LYA_H_R_OLD: return {stale ? c->Hr : c->Hro, n3};
    case M1LYA_H_Z_OLD: return {stale ? c->Hz : c->Hzo, n3};
    case M1LYA_ALPHA_F: return {c->aF, n3};
    case M1LYA_HVOIGT: return {c->Hv, n3};
    case M1LYA_PHI_NU: return {c->phi, n3};
    case M1LYA_FIJ_KP: return {c->Fkp, n3};
    case M1LYA_FIJ_K: return {c->Fk, n3};
    case M1LYA_FIJ_KM: return {c->Fkm, n3};
    case M1LYA_ALPHA_LYA: return {c->alpha, nc};
    case M1LYA_ALPHA_C: return {c->alpha_c, nc};
    case M1LYA_PD: return {c->pd, nc};
    case M1LYA_JS: return {c->js, nc};
    case M1LYA_T: return {c->T, nc};
    case M1LYA_DNU_D: return {c->dnuD, nc};
    case M1LYA_DUR_DR: return {c->dur_dr, nc};
    case M1LYA_DUR_DZ: return {c->dur_dz, nc};
    case M1LYA_DUZ_DR: return {c->duz_dr, nc};
    case M1LYA_DUZ_DZ: return {c->duz_dz, nc};
    case M1LYA_UR_OVER_R: return {c->ur_over_r, nc};
    case M1LYA_DRHOU_R: return {c->drhou_r, nc};
    case M1LYA_DRHOU_Z: return {c->drhou_z, nc};
    case M1LYA_F_LYA_R: return {c->f_r, nc};
    case M1LYA_F_LYA_Z: return {c->f_z, nc};
    case M1LYA_E_LYA: return {c->e_lya, nc};
    case M1LYA_UR_STAR: return {c->ur_star, (size_t)(c->nr + 1) * c->nz};
    case M1LYA_UZ_STAR: return {c->uz_star, (size_t)c->nr * (c->nz + 1)};
    default: fail(c, M1LYA_ERR_ARG, "unknown field id " + std::to_string(f));
    }
    return {nullptr, 0};
}

// After fused steps the *_old arrays hold a stale state; the reference
// invariant at step boundaries is J_old == J, H_old == H.  Materialise it
// before an operator-level call reads or writes the old arrays.
void materialize_old(m1lya_ctx* c)
{
    if (!c->old_stale) return;
    size_t b = c->n3 * sizeof(double);
    hipchk(c, hipMemcpyAsync(c->Jo, c->J, b, hipMemcpyDeviceToDevice, c->stream), "copy J_old");
    hipchk(c, hipMemcpyAsync(c->Hro, c->Hr, b, hipMemcpyDeviceToDevice, c->stream), "copy H_r_old");
    hipchk(c, hipMemcpyAsync(c->Hzo, c->Hz, b, hipMemcpyDeviceToDevice, c->stream), "copy H_Z_old");
    c->old_stale = false;
}

void require_setup(m1lya_ctx* c)
{
    if (!c->geom_set) fail(c, M1LYA_ERR_STATE, "geometry not set (m1lya_set_geometry)");
    if (!c->grid_set) fail(c, M1LYA_ERR_STATE, "frequency grid not set (m1lya_set_freq_grid)");
}

void check_bc(m1lya_ctx* c, int bc, const char* which)
{
    if (bc != M1LYA_REFLECT && bc != M1LYA_OUTFLOW)
        fail(c, M1LYA_ERR_ARG, std::string("invalid boundary condition for ") + which);
}

void upload_opt(m1lya_ctx* c, double* dst, const double* src, size_t n)
{
    if (src) hipchk(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload");
}

void upload_velocity(m1lya_ctx* c, const double* dur_dr, const double* dur_dz, const double* duz_dr,
                     const double* duz_dz, const double* ur_over_r)
{
    upload_opt(c, c->dur_dr, dur_dr, c->ncell);
    upload_opt(c, c->dur_dz, dur_dz, c->ncell);
    upload_opt(c, c->duz_dr, duz_dr, c->ncell);
    upload_opt(c, c->duz_dz, duz_dz, c->ncell);
    upload_opt(c, c->ur_over_r, ur_over_r, c->ncell);
}

void compute_krange(m1lya_ctx* c)
{
    Prof p1(c, "active_partial");
    launch_active_partial(c->stream, c->J, c->dnu, c->dV, c->nn, (int)c->ncell, c->act_blocks, c->part,
                          c->pstride);
    p1.done();
    Prof p2(c, "reduce_cols");
    launch_reduce_cols(c->stream, c->part, c->act_blocks, c->pstride, 2 * c->nn, c->red);
    p2.done();
    Prof p3(c, "active_final");
    launch_active_final(c->stream, c->red, c->nn, c->kcore_min, c->kcore_max, c->krange);
    p3.done();
    c->krange_valid = true;
}

AdvectArgs advect_args(m1lya_ctx* c, double cdt, int bc_r_right, int bc_z_bot, int bc_z_top,
                       const int* halt = nullptr)
{
    AdvectArgs a;
    a.nn = c->nn;
    a.g = DevGeom{c->nr, c->nz, c->A_r, c->A_z, c->inv_dV, c->inv_r_cell, c->ivr, c->ivz};
    a.v = DevVel{c->ur_star, c->uz_star, c->dur_dr, c->dur_dz, c->duz_dr, c->duz_dz, c->ur_over_r};
    a.J = c->J; a.Hr = c->Hr; a.Hz = c->Hz; a.aF = c->aF;
    a.J_out = c->Jo; a.Hr_out = c->Hro; a.Hz_out = c->Hzo;
    a.cdt = cdt;
    a.bc_r_right = bc_r_right; a.bc_z_bot = bc_z_bot; a.bc_z_top = bc_z_top;
    a.hll = c->flux;
    a.krange = c->krange;
    a.halt = halt;
    return a;
}

// Advection into the second buffer set, then swap roles.
void run_advect(m1lya_ctx* c, double cdt, int bc_r_right, int bc_z_bot, int bc_z_top, const int* halt = nullptr)
{
    AdvectArgs a = advect_args(c, cdt, bc_r_right, bc_z_bot, bc_z_top, halt);
    Prof p(c, "advect");
    launch_advect(c->stream, a);
    p.done();
    std::swap(c->J, c->Jo);
    std::swap(c->Hr, c->Hro);
    std::swap(c->Hz, c->Hzo);
}

ColumnArgs column_args(m1lya_ctx* c, int flags)
{
    ColumnArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nn = c->nn; a.nr = c->nr; a.nz = c->nz; a.flags = flags;
    a.J = c->J; a.Hr = c->Hr; a.Hz = c->Hz;
    a.J_old = c->Jo; a.Hr_old = c->Hro; a.Hz_old = c->Hzo;
    a.Hv = c->Hv; a.phi = c->phi; a.aF = c->aF;
    a.Fkp = c->Fkp; a.Fk = c->Fk; a.Fkm = c->Fkm;
    a.alpha = c->alpha; a.alpha_c = c->alpha_c; a.pd = c->pd; a.js = c->js; a.T = c->T; a.dnuD = c->dnuD;
    a.dur_dr = c->dur_dr; a.dur_dz = c->dur_dz; a.duz_dr = c->duz_dr; a.duz_dz = c->duz_dz;
    a.ur_over_r = c->ur_over_r;
    a.drhou_r = c->drhou_r; a.drhou_z = c->drhou_z; a.f_r = c->f_r; a.f_z = c->f_z;
    a.nu = c->nu; a.nuC = c->nuC; a.dnu = c->dnu; a.inv_dnu = c->inv_dnu; a.inv_nuC = c->inv_nuC;
    a.dV = c->dV; a.rcom = c->rcom; a.zcom = c->zcom; a.A_r = c->A_r; a.A_z = c->A_z;
    a.alpha_floor = c->alpha_floor; a.recoil = c->recoil;
    a.c_tilde = 1.0; a.dt = 1.0; a.dw = 0.0;
    a.part = c->part; a.pstride = c->pstride; a.err_flag = c->err_flag;
    a.halt = nullptr;
    return a;
}

// Reports (and clears) what the device error flag has collected since the last check:
// a singular tridiagonal system (gtsv INFO > 0) or a non-finite J, H_r or H_Z written by a
// fused step.  Steps queued without a host synchronisation report here, at the next call
// that synchronises the context (m1lya_step with out, m1lya_run_steps, m1lya_synchronize,
// m1lya_download).
void check_err_flag(m1lya_ctx* c, const char* what)
{
    int h = 0;
    hipchk(c, hipMemcpyAsync(&h, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream), "flag");
    hipchk(c, hipStreamSynchronize(c->stream), "sync");
    if (h) {
        hipchk(c, hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream), "flag reset");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        std::string msg = std::string(what) + ":";
        if (h & ERR_NONFINITE) msg += " non-finite J, H_r or H_Z in the state after a fused step;";
        if (h & ERR_SINGULAR) msg += " singular tridiagonal system (gtsv INFO > 0);";
        msg.pop_back();
        fail(c, M1LYA_ERR_NONFINITE, msg);
    }
}

// Options of a fused step run by m1lya_run_until_steady: measure ΔJ/J, halt below tol.
struct StepRun {
    bool conv = false;
    double tol = 0.0;
    StepCtl* ctl = nullptr;
};

// tape_X1 (may be null): receives the advection output (J, H_r, H_Z), as replay_step records it
void fused_step(m1lya_ctx* c, double c_tilde, double dt, double z_center, int bc_r_right, int bc_z_bot,
                int bc_z_top, double* history_row, double* tape_X1 = nullptr, const StepRun& run = StepRun{})
{
    if (!c->krange_valid) compute_krange(c);
    const double cdt = c_tilde * dt;
    const int* halt = run.ctl ? &run.ctl->halt : nullptr;
    run_advect(c, cdt, bc_r_right, bc_z_bot, bc_z_top, halt);
    if (tape_X1) {
        const size_t b = c->n3 * sizeof(double);
        hipchk(c, hipMemcpyAsync(tape_X1, c->J, b, hipMemcpyDeviceToDevice, c->stream), "tape");
        hipchk(c, hipMemcpyAsync(tape_X1 + c->n3, c->Hr, b, hipMemcpyDeviceToDevice, c->stream), "tape");
        hipchk(c, hipMemcpyAsync(tape_X1 + 2 * c->n3, c->Hz, b, hipMemcpyDeviceToDevice, c->stream), "tape");
    }
    // in a fused step the *_old arrays are not read as such: J_old is the step's input state
    ColumnArgs a = column_args(c, COL_EMIT | COL_FREQ | COL_SOLVE | COL_HCORR | COL_FUSED | (run.conv ? COL_CONV : 0));
    a.c_tilde = c_tilde; a.dt = dt; a.dw = 1.0 / cdt; a.z_center = z_center; a.bc_z_bot = bc_z_bot;
    a.halt = halt;
    Prof p(c, "column", true);
    launch_column(c->stream, a, c->col_blocks);
    p.done();
    Prof p2(c, "reduce_cols");
    launch_reduce_cols(c->stream, c->part, c->col_blocks, c->pstride, 2 * c->nn + kStepSums, c->red, halt);
    p2.done();
    Prof p3(c, "step_final");
    launch_step_final(c->stream, c->red, c->nn, c->kcore_min, c->kcore_max, bc_z_bot, c->krange, c->scalars,
                      history_row, c->err_flag, run.ctl, run.tol, run.conv ? 1 : 0);
    p3.done();
    c->old_stale = true;
    c->fij_stale = true;
    c->krange_valid = true;
}

// Device buffers of one adjoint run; freed (after the stream drains) on any exit.
struct AdjMem {
    m1lya_ctx* c;
    std::vector<void*> ptrs;
    explicit AdjMem(m1lya_ctx* c_) : c(c_) {}
    AdjMem(const AdjMem&) = delete;
    AdjMem& operator=(const AdjMem&) = delete;
    double* get(size_t n)
    {
        void* p = nullptr;
        hipchk(c, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)), "hipMalloc (adjoint)");
        ptrs.push_back(p);
        return (double*)p;
    }
    int* geti(size_t n)
    {
        void* p = nullptr;
        hipchk(c, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(int)), "hipMalloc (adjoint)");
        ptrs.push_back(p);
        return (int*)p;
    }
    ~AdjMem()
    {
        (void)hipStreamSynchronize(c->stream);
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// The fused step with the advection output copied to X1 (J, H_r, H_Z at
// X1, X1 + n3, X1 + 2 n3).  The active range must already be in c->krange.
void replay_step(m1lya_ctx* c, double c_tilde, double dt, double z_center, int bc_r_right, int bc_z_bot,
                 int bc_z_top, double* X1)
{
    const double cdt = c_tilde * dt;
    run_advect(c, cdt, bc_r_right, bc_z_bot, bc_z_top);
    const size_t b = c->n3 * sizeof(double);
    hipchk(c, hipMemcpyAsync(X1, c->J, b, hipMemcpyDeviceToDevice, c->stream), "tape");
    hipchk(c, hipMemcpyAsync(X1 + c->n3, c->Hr, b, hipMemcpyDeviceToDevice, c->stream), "tape");
    hipchk(c, hipMemcpyAsync(X1 + 2 * c->n3, c->Hz, b, hipMemcpyDeviceToDevice, c->stream), "tape");
    ColumnArgs a = column_args(c, COL_EMIT | COL_FREQ | COL_SOLVE | COL_HCORR | COL_FUSED);
    a.c_tilde = c_tilde; a.dt = dt; a.dw = 1.0 / cdt; a.z_center = z_center; a.bc_z_bot = bc_z_bot;
    Prof p(c, "column");
    launch_column(c->stream, a, c->col_blocks);
    p.done();
    c->old_stale = true;
    c->fij_stale = true;
}

template <class F>
int guarded(m1lya_ctx* c, F&& f)
{
    if (!c) {
        g_last_error = "null context";
        return M1LYA_ERR_ARG;
    }
    try {
        f();
        return M1LYA_OK;
    } catch (const Fail& e) {
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        g_last_error = e.what();
        return M1LYA_ERR_HIP;
    }
}

}  // namespace

extern "C" {

int m1lya_abi_version(void) { return M1LYA_ABI_VERSION; }

int m1lya_create(int n_nu, int n_r, int n_z, int precision, m1lya_ctx** out)
{
    if (!out) {
        g_last_error = "out is NULL";
        return M1LYA_ERR_ARG;
    }
    *out = nullptr;
    if (n_nu < 3 || n_r < 3 || n_z < 3) {
        g_last_error = "need N_ν >= 3, N_r >= 3, N_Z >= 3 (src/SpatialGrid.jl:629-630)";
        return M1LYA_ERR_DIMS;
    }
    if (n_nu > kMaxBinsFP64) {
        g_last_error = "N_ν > " + std::to_string(kMaxBinsFP64) + " is not supported by the column kernel";
        return M1LYA_ERR_DIMS;
    }
    if (precision != M1LYA_FP64) {
        g_last_error = "unsupported precision";
        return M1LYA_ERR_UNSUPPORTED;
    }
    m1lya_ctx* c = new m1lya_ctx();
    int rc = guarded(c, [&] {
        c->nn = n_nu; c->nr = n_r; c->nz = n_z; c->precision = precision;
        c->ncell = (size_t)n_r * n_z;
        c->n3 = c->ncell * n_nu;
        hipchk(c, hipGetDevice(&c->device), "hipGetDevice");
        hipchk(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        const size_t n3 = c->n3, nc = c->ncell;
        c->J = dalloc(c, n3); c->Hr = dalloc(c, n3); c->Hz = dalloc(c, n3);
        c->Jo = dalloc(c, n3); c->Hro = dalloc(c, n3); c->Hzo = dalloc(c, n3);
        c->aF = dalloc(c, n3); c->Hv = dalloc(c, n3); c->phi = dalloc(c, n3);
        c->Fkp = dalloc(c, n3); c->Fk = dalloc(c, n3); c->Fkm = dalloc(c, n3);
        double** cells[] = {&c->alpha, &c->alpha_c, &c->pd, &c->js, &c->T, &c->dnuD, &c->dur_dr, &c->dur_dz,
                            &c->duz_dr, &c->duz_dz, &c->ur_over_r, &c->drhou_r, &c->drhou_z, &c->f_r,
                            &c->f_z, &c->e_lya, &c->dV, &c->inv_dV, &c->fg_r, &c->fg_z};
        for (double** p : cells) *p = dalloc(c, nc);
        c->ur_star = dalloc(c, (size_t)(n_r + 1) * n_z);
        c->uz_star = dalloc(c, (size_t)n_r * (n_z + 1));
        c->A_r = dalloc(c, (size_t)(n_r + 1) * n_z);
        c->A_z = dalloc(c, (size_t)n_r * (n_z + 1));
        c->inv_r_cell = dalloc(c, n_r); c->rcom = dalloc(c, n_r); c->zcom = dalloc(c, n_z);
        c->ivr = dalloc(c, (size_t)(n_r + 1) * 5); c->ivz = dalloc(c, (size_t)(n_z + 1) * 5);
        c->nu = dalloc(c, n_nu + 1); c->nuC = dalloc(c, n_nu); c->dnu = dalloc(c, n_nu);
        c->inv_dnu = dalloc(c, n_nu); c->inv_nuC = dalloc(c, n_nu);
        const int W = column_width(n_nu) * column_waves(n_nu);   // columns per block
        const int ngroups = (int)((nc + W - 1) / W);
        c->col_blocks = column_blocks(n_nu, ngroups);
        c->act_blocks = (int)std::min<size_t>(nc, 1024);
        c->pstride = 2 * n_nu + 8;
        c->prows = std::max(c->col_blocks, c->act_blocks);
        c->part = dalloc(c, (size_t)c->prows * c->pstride);
        c->red = dalloc(c, c->pstride);
        c->scalars = dalloc(c, 8);
        c->krange = ialloc(c, 2);
        c->err_flag = ialloc(c, 1);
        c->ctl = (StepCtl*)ialloc(c, sizeof(StepCtl) / sizeof(int));
        int ne = (n_z - 2) + 2 * (n_r - 3);
        c->edge_buf = dalloc(c, (size_t)std::max(ne, 1) * n_nu * 2);
        hipchk(c, hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
    if (rc != M1LYA_OK) {
        m1lya_destroy(c);
        return rc;
    }
    *out = c;
    return M1LYA_OK;
}

void m1lya_destroy(m1lya_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) {
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->history) (void)hipFree(c->history);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int m1lya_last_error(const m1lya_ctx* c, char* buf, size_t len)
{
    if (!buf || len == 0) return M1LYA_ERR_ARG;
    const std::string& s = c ? c->err : g_last_error;
    std::snprintf(buf, len, "%s", s.c_str());
    return M1LYA_OK;
}

int m1lya_set_geometry(m1lya_ctx* c, const m1lya_geometry* g)
{
    return guarded(c, [&] {
        if (!g || !g->dV || !g->inv_dV || !g->A_r || !g->A_Z || !g->inv_r_cell || !g->rCoM || !g->ZCoM ||
            !g->dr_L || !g->dr_R || !g->inv_dr_i_im1 || !g->inv_dr_ip1_i || !g->inv_dr_ip2_ip1 || !g->dZ_L ||
            !g->dZ_R || !g->inv_dZ_j_jm1 || !g->inv_dZ_jp1_j || !g->inv_dZ_jp2_jp1)
            fail(c, M1LYA_ERR_ARG, "geometry has a NULL array");
        const int nr = c->nr, nz = c->nz;
        auto up = [&](double* d, const double* h, size_t n) {
            hipchk(c, hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice), "geometry upload");
        };
        up(c->dV, g->dV, c->ncell);
        up(c->inv_dV, g->inv_dV, c->ncell);
        up(c->A_r, g->A_r, (size_t)(nr + 1) * nz);
        up(c->A_z, g->A_Z, (size_t)nr * (nz + 1));
        up(c->inv_r_cell, g->inv_r_cell, nr);
        up(c->rcom, g->rCoM, nr);
        up(c->zcom, g->ZCoM, nz);
        if (g->r && g->Z) {
            if (!c->r_if) {
                c->r_if = dalloc(c, (size_t)nr + 1);
                c->z_if = dalloc(c, (size_t)nz + 1);
            }
            up(c->r_if, g->r, (size_t)nr + 1);
            up(c->z_if, g->Z, (size_t)nz + 1);
            c->interfaces_set = true;
        }
        // One interval set per interface: face 0 = inner boundary, face N = outer
        // boundary, face f in 1..N-1 uses the interior arrays at f-1
        // (src/AdvectJandH.jl:632-639, 803-812, 1092-1098, 1105-1114).
        auto pack = [&](int n, const double* dL, const double* dR, const double* a, const double* b,
                        const double* cc, const double* inner, const double* outer, double* dst) {
            std::vector<double> h((size_t)(n + 1) * 5);
            for (int m = 0; m < 5; ++m) { h[m] = inner[m]; h[(size_t)n * 5 + m] = outer[m]; }
            for (int f = 1; f < n; ++f) {
                h[(size_t)f * 5 + 0] = dL[f - 1];
                h[(size_t)f * 5 + 1] = dR[f - 1];
                h[(size_t)f * 5 + 2] = a[f - 1];
                h[(size_t)f * 5 + 3] = b[f - 1];
                h[(size_t)f * 5 + 4] = cc[f - 1];
            }
            up(dst, h.data(), h.size());
        };
        pack(nr, g->dr_L, g->dr_R, g->inv_dr_i_im1, g->inv_dr_ip1_i, g->inv_dr_ip2_ip1, g->dr_inner, g->dr_outer,
             c->ivr);
        pack(nz, g->dZ_L, g->dZ_R, g->inv_dZ_j_jm1, g->inv_dZ_jp1_j, g->inv_dZ_jp2_jp1, g->dZ_inner, g->dZ_outer,
             c->ivz);
        c->geom_set = true;
        c->krange_valid = false;
    });
}

int m1lya_set_freq_grid(m1lya_ctx* c, const m1lya_freq_grid* g)
{
    return guarded(c, [&] {
        if (!g || !g->nu || !g->nuC || !g->dnu || !g->inv_dnu || !g->inv_nuC)
            fail(c, M1LYA_ERR_ARG, "frequency grid has a NULL array");
        if (g->kcore_min < 0 || g->kcore_max >= c->nn || g->kcore_min > g->kcore_max)
            fail(c, M1LYA_ERR_ARG, "kcore range out of bounds");
        if (g->flux_method != M1LYA_GLF && g->flux_method != M1LYA_HLL)
            fail(c, M1LYA_ERR_ARG, "unknown flux method");
        const int nn = c->nn;
        auto up = [&](double* d, const double* h, size_t n) {
            hipchk(c, hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice), "grid upload");
        };
        up(c->nu, g->nu, nn + 1);
        up(c->nuC, g->nuC, nn);
        up(c->dnu, g->dnu, nn);
        up(c->inv_dnu, g->inv_dnu, nn);
        up(c->inv_nuC, g->inv_nuC, nn);
        c->kcore_min = g->kcore_min; c->kcore_max = g->kcore_max;
        c->recoil = g->recoil ? 1 : 0;
        c->flux = g->flux_method;
        c->alpha_floor = g->alpha_floor;
        c->grid_set = true;
        c->krange_valid = false;
    });
}

int m1lya_upload(m1lya_ctx* c, int field, const double* host, size_t count)
{
    return guarded(c, [&] {
        if (!host) fail(c, M1LYA_ERR_ARG, "host pointer is NULL");
        FieldInfo fi = field_info(c, field, false);
        if (count != fi.count)
            fail(c, M1LYA_ERR_ARG, "field " + std::to_string(field) + ": count " + std::to_string(count) +
                                       " != " + std::to_string(fi.count));
        if (field <= M1LYA_H_Z_OLD) materialize_old(c);
        if (field >= M1LYA_FIJ_KP && field <= M1LYA_FIJ_KM) c->fij_stale = false;
        hipchk(c, hipMemcpyAsync(fi.ptr, host, count * sizeof(double), hipMemcpyHostToDevice, c->stream),
               "upload");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        if (field == M1LYA_J) c->krange_valid = false;
    });
}

int m1lya_download(m1lya_ctx* c, int field, double* host, size_t count)
{
    return guarded(c, [&] {
        if (!host) fail(c, M1LYA_ERR_ARG, "host pointer is NULL");
        if (field >= M1LYA_FIJ_KP && field <= M1LYA_FIJ_KM && c->fij_stale)
            fail(c, M1LYA_ERR_STATE, "Fij_* are not maintained by the fused step; call "
                                     "m1lya_update_freq_factors first");
        FieldInfo fi = field_info(c, field, true);
        if (count != fi.count)
            fail(c, M1LYA_ERR_ARG, "field " + std::to_string(field) + ": count " + std::to_string(count) +
                                       " != " + std::to_string(fi.count));
        hipchk(c, hipMemcpyAsync(host, fi.ptr, count * sizeof(double), hipMemcpyDeviceToHost, c->stream),
               "download");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        check_err_flag(c, "step");
    });
}

int m1lya_fill(m1lya_ctx* c, int field, double value)
{
    return guarded(c, [&] {
        FieldInfo fi = field_info(c, field, false);
        if (field <= M1LYA_H_Z_OLD) materialize_old(c);
        std::vector<double> h(fi.count, value);
        hipchk(c, hipMemcpyAsync(fi.ptr, h.data(), fi.count * sizeof(double), hipMemcpyHostToDevice, c->stream),
               "fill");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        if (field == M1LYA_J) c->krange_valid = false;
    });
}

int m1lya_synchronize(m1lya_ctx* c)
{
    return guarded(c, [&] {
        hipchk(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        check_err_flag(c, "step");
    });
}

int m1lya_find_active_bins(m1lya_ctx* c, int* kmin, int* kmax)
{
    return guarded(c, [&] {
        require_setup(c);
        compute_krange(c);
        int h[2];
        hipchk(c, hipMemcpyAsync(h, c->krange, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream), "krange");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        if (kmin) *kmin = h[0];
        if (kmax) *kmax = h[1];
    });
}

int m1lya_advect_lya(m1lya_ctx* c, double c_dt, const double* u_r_star, const double* u_z_star,
                     const double* dur_dr, const double* dur_dz, const double* duz_dr, const double* duz_dz,
                     const double* ur_over_r, int bc_r_left, int bc_r_right, int bc_z_bot, int bc_z_top)
{
    return guarded(c, [&] {
        require_setup(c);
        if (bc_r_left != M1LYA_REFLECT)
            fail(c, M1LYA_ERR_ARG, "cylindrical geometry: the axis boundary must be M1LYA_REFLECT");
        check_bc(c, bc_r_right, "BC_r_right");
        check_bc(c, bc_z_bot, "BC_Z_bot");
        check_bc(c, bc_z_top, "BC_Z_top");
        materialize_old(c);
        upload_opt(c, c->ur_star, u_r_star, (size_t)(c->nr + 1) * c->nz);
        upload_opt(c, c->uz_star, u_z_star, (size_t)c->nr * (c->nz + 1));
        upload_velocity(c, dur_dr, dur_dz, duz_dr, duz_dz, ur_over_r);
        compute_krange(c);   // FindActiveBins (src/UpdateLyaJandH.jl:153)
        run_advect(c, c_dt, bc_r_right, bc_z_bot, bc_z_top);
        // J_old = J (src/AdvectJandH.jl:1385-1387); H_old = H (src/UpdateLyaJandH.jl:167-168)
        size_t b = c->n3 * sizeof(double);
        hipchk(c, hipMemcpyAsync(c->Jo, c->J, b, hipMemcpyDeviceToDevice, c->stream), "J_old");
        hipchk(c, hipMemcpyAsync(c->Hro, c->Hr, b, hipMemcpyDeviceToDevice, c->stream), "H_r_old");
        hipchk(c, hipMemcpyAsync(c->Hzo, c->Hz, b, hipMemcpyDeviceToDevice, c->stream), "H_Z_old");
        c->krange_valid = false;
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_emission_absorption_update(m1lya_ctx* c, double c_tilde, double dt)
{
    return guarded(c, [&] {
        require_setup(c);
        materialize_old(c);
        ColumnArgs a = column_args(c, COL_EMIT);
        a.c_tilde = c_tilde; a.dt = dt;
        Prof p(c, "column_emit");
        launch_column(c->stream, a, c->col_blocks);
        p.done();
        c->krange_valid = false;
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_update_freq_factors(m1lya_ctx* c, const double* T, const double* dur_dr, const double* dur_dz,
                              const double* duz_dr, const double* duz_dz, const double* ur_over_r)
{
    return guarded(c, [&] {
        require_setup(c);
        upload_opt(c, c->T, T, c->ncell);
        upload_velocity(c, dur_dr, dur_dz, duz_dr, duz_dz, ur_over_r);
        ColumnArgs a = column_args(c, COL_FREQ | COL_STORE_F);
        Prof p(c, "column_freq_factors");
        launch_column(c->stream, a, c->col_blocks);
        p.done();
        c->fij_stale = false;
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_frequency_update_all_cells(m1lya_ctx* c, double c_dt, const double* T, const double* dur_dr,
                                     const double* dur_dz, const double* duz_dr, const double* duz_dz,
                                     const double* ur_over_r, int update_freq_factors)
{
    return guarded(c, [&] {
        require_setup(c);
        if (!update_freq_factors && c->fij_stale)
            fail(c, M1LYA_ERR_STATE, "Fij_* are stale (fused steps); call with update_freq_factors = 1");
        materialize_old(c);
        upload_opt(c, c->T, T, c->ncell);
        upload_velocity(c, dur_dr, dur_dz, duz_dr, duz_dz, ur_over_r);
        int flags = COL_SOLVE | COL_HCORR | (update_freq_factors ? (COL_FREQ | COL_STORE_F) : 0);
        ColumnArgs a = column_args(c, flags);
        a.dw = 1.0 / c_dt;
        Prof p(c, "column_freq_update");
        launch_column(c->stream, a, c->col_blocks);
        p.done();
        if (update_freq_factors) c->fij_stale = false;
        c->krange_valid = false;
        check_err_flag(c, "FrequencyUpdateJLyα_AllCells");
    });
}

static void mf_common(m1lya_ctx* c, double dt, const double* fgr, const double* fgz, double z_center,
                      double* sums)
{
    MfArgs a;
    a.nr = c->nr; a.nz = c->nz;
    a.drhou_r = c->drhou_r; a.drhou_z = c->drhou_z; a.f_r = c->f_r; a.f_z = c->f_z;
    a.js = c->js; a.dV = c->dV; a.rcom = c->rcom; a.zcom = c->zcom;
    a.fg_r = fgr; a.fg_z = fgz; a.dt = dt; a.z_center = z_center; a.part = c->part;
    int nb = (int)std::min<size_t>((c->ncell + 255) / 256, 512);
    Prof p(c, "mf_partial");
    launch_mf_partial(c->stream, a, nb);
    p.done();
    launch_reduce_cols(c->stream, c->part, nb, 8, 5, c->red);
    hipchk(c, hipGetLastError(), "reduce");
    hipchk(c, hipMemcpyAsync(sums, c->red, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "sums");
    hipchk(c, hipStreamSynchronize(c->stream), "sync");
}

int m1lya_update_f_mf_l(m1lya_ctx* c, double dt, double z_center, int bc_z_bot, double out[4])
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_z_bot, "BC_Z_bot");
        if (!out) fail(c, M1LYA_ERR_ARG, "out is NULL");
        double s[5];
        mf_common(c, dt, nullptr, nullptr, z_center, s);
        // normalisation: src/ConstantsAndFunctions.jl:663-673
        double MF = s[0], L = s[1], net = s[2], ow = s[3];
        if (L > 0.0) {
            double f = 1.0 / (L / kC);
            MF = MF * f; net = net * f; ow = ow * f;
        } else {
            MF = 0.0; net = 0.0; ow = 0.0;
        }
        out[0] = MF; out[1] = net; out[2] = ow;
        out[3] = (bc_z_bot == M1LYA_REFLECT) ? 2.0 * L : L;
    });
}

int m1lya_update_f_mf_l_gravity(m1lya_ctx* c, double dt, const double* f_grav_r, const double* f_grav_z,
                                double z_center, int bc_z_bot, double out[5])
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_z_bot, "BC_Z_bot");
        if (!out || !f_grav_r || !f_grav_z) fail(c, M1LYA_ERR_ARG, "NULL argument");
        upload_opt(c, c->fg_r, f_grav_r, c->ncell);
        upload_opt(c, c->fg_z, f_grav_z, c->ncell);
        double s[5];
        mf_common(c, dt, c->fg_r, c->fg_z, z_center, s);
        double MF = s[0], L = s[1], net = s[2], ow = s[3], ag = s[4];
        if (L > 0.0) {
            double f = 1.0 / (L / kC);
            MF = MF * f; ag = ag * f; net = net * f; ow = ow * f;
        } else {
            MF = 0.0; ag = 0.0; net = 0.0; ow = 0.0;
        }
        out[0] = MF; out[1] = ag; out[2] = net; out[3] = ow;
        out[4] = (bc_z_bot == M1LYA_REFLECT) ? 2.0 * L : L;
    });
}

int m1lya_compute_f_esc(m1lya_ctx* c, double L_allspace, int bc_z_bot, double* f_esc)
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_z_bot, "BC_Z_bot");
        if (!f_esc) fail(c, M1LYA_ERR_ARG, "f_esc is NULL");
        if (L_allspace <= 0.0) {
            *f_esc = 0.0;
            return;
        }
        int nb = 256;
        Prof p(c, "fesc_partial");
        launch_fesc_partial(c->stream, c->Hr, c->Hz, c->A_r, c->A_z, c->dnu, c->nn, c->nr, c->nz, bc_z_bot, nb,
                            c->part);
        p.done();
        launch_reduce_cols(c->stream, c->part, nb, 1, 1, c->red);
        double s = 0.0;
        hipchk(c, hipMemcpyAsync(&s, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream), "sum");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        double Lsim = (bc_z_bot == M1LYA_REFLECT) ? 0.5 * L_allspace : L_allspace;
        *f_esc = (Lsim > 0.0) ? s / Lsim : 0.0;
    });
}

int m1lya_reset_drhou(m1lya_ctx* c)
{
    return guarded(c, [&] {
        hipchk(c, hipMemsetAsync(c->drhou_r, 0, c->ncell * sizeof(double), c->stream), "memset");
        hipchk(c, hipMemsetAsync(c->drhou_z, 0, c->ncell * sizeof(double), c->stream), "memset");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_update_e_lya(m1lya_ctx* c)
{
    return guarded(c, [&] {
        require_setup(c);
        Prof p(c, "e_lya");
        launch_e_lya(c->stream, c->J, c->dnu, c->nn, (int)c->ncell, c->e_lya);
        p.done();
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_edge_count(const m1lya_ctx* c)
{
    if (!c) return 0;
    return (c->nz - 2) + 2 * (c->nr - 3);
}

int m1lya_edge_spectra(m1lya_ctx* c, double* raw, double* norm, size_t count)
{
    return guarded(c, [&] {
        require_setup(c);
        size_t ne = (size_t)m1lya_edge_count(c);
        if (!raw || !norm || count != ne * c->nn) fail(c, M1LYA_ERR_ARG, "edge spectra: bad buffers/count");
        double* draw = c->edge_buf;
        double* dnorm = c->edge_buf + ne * c->nn;
        launch_edge(c->stream, c->J, c->dnu, c->nn, c->nr, c->nz, draw, dnorm);
        hipchk(c, hipGetLastError(), "edge");
        hipchk(c, hipMemcpyAsync(raw, draw, count * sizeof(double), hipMemcpyDeviceToHost, c->stream), "raw");
        hipchk(c, hipMemcpyAsync(norm, dnorm, count * sizeof(double), hipMemcpyDeviceToHost, c->stream), "norm");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_step(m1lya_ctx* c, double c_tilde, double dt, double z_center, int bc_r_right, int bc_z_bot,
               int bc_z_top, double out[8])
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_r_right, "BC_r_right");
        check_bc(c, bc_z_bot, "BC_Z_bot");
        check_bc(c, bc_z_top, "BC_Z_top");
        if (!(c_tilde > 0.0) || !(dt > 0.0)) fail(c, M1LYA_ERR_ARG, "c_tilde and dt must be > 0");
        fused_step(c, c_tilde, dt, z_center, bc_r_right, bc_z_bot, bc_z_top, nullptr);
        if (out) {
            hipchk(c, hipMemcpyAsync(out, c->scalars, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                   "scalars");
            check_err_flag(c, "step");
        }
    });
}

int m1lya_run_steps(m1lya_ctx* c, int n_steps, double c_tilde, double dt, double z_center, int bc_r_right,
                    int bc_z_bot, int bc_z_top, double* history)
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_r_right, "BC_r_right");
        check_bc(c, bc_z_bot, "BC_Z_bot");
        check_bc(c, bc_z_top, "BC_Z_top");
        if (n_steps < 0) fail(c, M1LYA_ERR_ARG, "n_steps < 0");
        if (!(c_tilde > 0.0) || !(dt > 0.0)) fail(c, M1LYA_ERR_ARG, "c_tilde and dt must be > 0");
        if (history && (size_t)n_steps * 8 > c->history_cap) {
            if (c->history) hipchk(c, hipFree(c->history), "hipFree");
            c->history = nullptr;
            hipchk(c, hipMalloc(&c->history, (size_t)n_steps * 8 * sizeof(double)), "hipMalloc");
            c->history_cap = (size_t)n_steps * 8;
        }
        for (int s = 0; s < n_steps; ++s)
            fused_step(c, c_tilde, dt, z_center, bc_r_right, bc_z_bot, bc_z_top,
                       history ? c->history + (size_t)s * 8 : nullptr);
        if (history && n_steps > 0)
            hipchk(c, hipMemcpyAsync(history, c->history, (size_t)n_steps * 8 * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream),
                   "history");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        check_err_flag(c, "run_steps");
    });
}

int m1lya_run_until_steady(m1lya_ctx* c, int max_steps, double tol, double c_tilde, double dt, double z_center,
                           int bc_r_right, int bc_z_bot, int bc_z_top, double* history, int* steps_done)
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_r_right, "BC_r_right");
        check_bc(c, bc_z_bot, "BC_Z_bot");
        check_bc(c, bc_z_top, "BC_Z_top");
        if (max_steps < 0) fail(c, M1LYA_ERR_ARG, "max_steps < 0");
        if (!(tol > 0.0)) fail(c, M1LYA_ERR_ARG, "tol must be > 0");
        if (!(c_tilde > 0.0) || !(dt > 0.0)) fail(c, M1LYA_ERR_ARG, "c_tilde and dt must be > 0");
        // Steps are queued in chunks; after the step whose ΔJ/J falls below tol, k_step_final
        // sets the halt flag and the rest of the chunk's launches return at once.
        const int chunk = 256;
        if ((size_t)chunk * 8 > c->history_cap) {
            if (c->history) hipchk(c, hipFree(c->history), "hipFree");
            c->history = nullptr;
            hipchk(c, hipMalloc(&c->history, (size_t)chunk * 8 * sizeof(double)), "hipMalloc");
            c->history_cap = (size_t)chunk * 8;
        }
        StepRun run;
        run.conv = true;
        run.tol = tol;
        run.ctl = c->ctl;
        int done = 0;
        bool halted = false;
        while (done < max_steps && !halted) {
            const int n = std::min(chunk, max_steps - done);
            const StepCtl zero{0, 0};
            hipchk(c, hipMemcpyAsync(c->ctl, &zero, sizeof(StepCtl), hipMemcpyHostToDevice, c->stream), "ctl");
            for (int s = 0; s < n; ++s)
                fused_step(c, c_tilde, dt, z_center, bc_r_right, bc_z_bot, bc_z_top, c->history + (size_t)s * 8,
                           nullptr, run);
            StepCtl h{0, 0};
            hipchk(c, hipMemcpyAsync(&h, c->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream), "ctl");
            hipchk(c, hipStreamSynchronize(c->stream), "sync");
            // every launched step swapped the buffer roles on the host; the halted ones did not run
            if ((n - h.steps) & 1) {
                std::swap(c->J, c->Jo);
                std::swap(c->Hr, c->Hro);
                std::swap(c->Hz, c->Hzo);
            }
            if (history && h.steps > 0)
                hipchk(c, hipMemcpyAsync(history + (size_t)done * 8, c->history, (size_t)h.steps * 8 * sizeof(double),
                                         hipMemcpyDeviceToHost, c->stream),
                       "history");
            done += h.steps;
            halted = h.halt != 0;
            check_err_flag(c, "run_until_steady");
        }
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        if (steps_done) *steps_done = done;
    });
}

}  // extern "C"

namespace {

double* rhd_stage(m1lya_ctx* c, int slot, const double* host, size_t n)
{
    if (!c->rhd) {
        c->rhd = dalloc(c, (size_t)16 * c->ncell);
        c->rhd_redo = ialloc(c, c->ncell);
    }
    double* d = c->rhd + (size_t)slot * c->ncell;
    if (host) hipchk(c, hipMemcpyAsync(d, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload");
    return d;
}

double rhd_scan(m1lya_ctx* c, RhdScanArgs& a)
{
    if (!c->interfaces_set) fail(c, M1LYA_ERR_STATE, "cell interfaces r, Z not set (m1lya_set_geometry)");
    a.nr = c->nr;
    a.nz = c->nz;
    a.r = c->r_if;
    a.Z = c->z_if;
    const int nb = (int)std::min<size_t>(std::max<size_t>((c->ncell + 255) / 256, 1), 1024);
    Prof p(c, a.mode ? "rsl_scan" : "cfl_scan");
    launch_rhd_scan(c->stream, a, nb, c->part, c->red);
    p.done();
    double v = 0.0;
    hipchk(c, hipMemcpyAsync(&v, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream), "scan");
    hipchk(c, hipStreamSynchronize(c->stream), "sync");
    return v;
}

}  // namespace

extern "C" {

int m1lya_cfl_time_step(m1lya_ctx* c, double c_tilde, const double* u_r, const double* u_z, double cfl, double* dt)
{
    return guarded(c, [&] {
        if (!u_r || !u_z || !dt) fail(c, M1LYA_ERR_ARG, "NULL argument");
        if (!(c_tilde > 0.0)) fail(c, M1LYA_ERR_ARG, "c_tilde must be > 0");
        RhdScanArgs a{};
        a.mode = 0;
        a.c_tilde = c_tilde;
        a.u_r = rhd_stage(c, 0, u_r, c->ncell);
        a.u_z = rhd_stage(c, 1, u_z, c->ncell);
        *dt = cfl * rhd_scan(c, a);
    });
}

int m1lya_find_rsl(m1lya_ctx* c, double c_old, double c_min, const m1lya_rsl_inputs* in, double* c_new)
{
    return guarded(c, [&] {
        if (!in || !c_new || !in->u_r || !in->u_z || !in->gamma_ad || !in->P || !in->rho || !in->rho_d ||
            !in->xHII || !in->xH2 || !in->T)
            fail(c, M1LYA_ERR_ARG, "NULL argument");
        const size_t n = c->ncell;
        RhdScanArgs a{};
        a.mode = 1;
        a.u_r = rhd_stage(c, 0, in->u_r, n);
        a.u_z = rhd_stage(c, 1, in->u_z, n);
        a.gamma_ad = rhd_stage(c, 2, in->gamma_ad, n);
        a.P = rhd_stage(c, 3, in->P, n);
        a.rho = rhd_stage(c, 4, in->rho, n);
        a.xHII = rhd_stage(c, 5, in->xHII, n);
        a.xH2 = rhd_stage(c, 6, in->xH2, n);
        a.T = rhd_stage(c, 7, in->T, n);
        a.rho_d = rhd_stage(c, 8, in->rho_d, 3 * n);
        a.pd = c->pd;
        a.kS = in->kappa_abs[0];
        a.kG = in->kappa_abs[1];
        a.kP = in->kappa_abs[2];
        a.lc = LineConst{in->lc.b_prefact, in->lc.nu_over_c, in->lc.a_prefact, in->lc.sigma0_prefact};
        // X / m_H of src/ConstantsAndFunctions.jl:124-128,167
        const double Y = 0.247, D_H = 2.53e-5, amu = 1.660539069e-24;
        const double mH = 1.0080 * amu, mD = 2.014 * amu;
        const double X_D = D_H * (1 - Y) * mD / mH;
        const double X = 1 - Y - X_D;
        a.X_over_mH = X / mH;
        double v = rhd_scan(c, a);
        // clamp and rate limits of src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:706-721
        v = std::min(std::max(v, c_min), 2.99792458e10);
        if (v < c_old) v = std::max(v, 0.75 * c_old);
        if (v > c_old) v = std::min(v, 4.0 * c_old);
        *c_new = v;
    });
}

int m1lya_absorption_substep(m1lya_ctx* c, const double* dt_cell, double dt, double c_tilde, const double* rho_d,
                             const double* kappa_F, int* redo, double* cell_out)
{
    return guarded(c, [&] {
        require_setup(c);
        if (!(c_tilde > 0.0)) fail(c, M1LYA_ERR_ARG, "c_tilde must be > 0");
        if (!dt_cell && !(dt > 0.0)) fail(c, M1LYA_ERR_ARG, "dt must be > 0 when dt_cell is NULL");
        if (rho_d && !kappa_F) fail(c, M1LYA_ERR_ARG, "kappa_F is NULL");
        materialize_old(c);
        const size_t n = c->ncell;
        double* d_dt = rhd_stage(c, 0, dt_cell, n);
        if (!dt_cell) {
            std::vector<double> h(n, dt);
            hipchk(c, hipMemcpyAsync(d_dt, h.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream), "dt");
            hipchk(c, hipStreamSynchronize(c->stream), "sync");
        }
        AbsorbArgs a{};
        a.nn = c->nn;
        a.ncell = (int)n;
        a.dt = d_dt;
        a.c_tilde = c_tilde;
        a.J_old = c->Jo; a.Hr_old = c->Hro; a.Hz_old = c->Hzo;
        a.J = c->J; a.Hr = c->Hr; a.Hz = c->Hz;
        a.phi = c->phi; a.Hv = c->Hv; a.aF = c->aF;
        a.alpha = c->alpha; a.alpha_c = c->alpha_c; a.pd = c->pd; a.js = c->js; a.e_lya = c->e_lya;
        a.dnu = c->dnu;
        a.rho_d = rho_d ? rhd_stage(c, 1, rho_d, 3 * n) : nullptr;
        if (kappa_F) { a.kS = kappa_F[0]; a.kG = kappa_F[1]; a.kP = kappa_F[2]; }
        a.redo = c->rhd_redo;
        a.out = c->rhd + 4 * n;
        Prof p(c, "absorb");
        launch_absorb(c->stream, a);
        p.done();
        if (redo) hipchk(c, hipMemcpyAsync(redo, c->rhd_redo, n * sizeof(int), hipMemcpyDeviceToHost, c->stream), "redo");
        if (cell_out)
            hipchk(c, hipMemcpyAsync(cell_out, a.out, 12 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                   "cell_out");
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        c->krange_valid = false;
    });
}

int m1lya_set_voigt_lut(m1lya_ctx* c, const double* luts, int nx, double xmax, const m1lya_line_constants* lc)
{
    return guarded(c, [&] {
        require_setup(c);
        if (!luts || !lc || nx < 2 || !(xmax > 0.0)) fail(c, M1LYA_ERR_ARG, "set_voigt_lut: bad arguments");
        if (c->lut && c->lut_nx != nx) fail(c, M1LYA_ERR_ARG, "set_voigt_lut: table size changed");
        if (!c->lut) {
            c->lut = dalloc(c, (size_t)5 * nx);
            c->dnu_rel = dalloc(c, (size_t)c->nn + 1);
            c->dnuC_rel = dalloc(c, (size_t)c->nn);
            c->nHI = dalloc(c, c->ncell);
            c->acF = dalloc(c, c->ncell);
        }
        hipchk(c, hipMemcpy(c->lut, luts, (size_t)5 * nx * sizeof(double), hipMemcpyHostToDevice), "lut");
        // ν - ν_Lyα at the interfaces and centres, formed as the host setup forms them
        std::vector<double> nu(c->nn + 1), nuC(c->nn), dr(c->nn + 1), dcr(c->nn);
        hipchk(c, hipMemcpy(nu.data(), c->nu, nu.size() * sizeof(double), hipMemcpyDeviceToHost), "nu");
        hipchk(c, hipMemcpy(nuC.data(), c->nuC, nuC.size() * sizeof(double), hipMemcpyDeviceToHost), "nuC");
        for (int k = 0; k <= c->nn; ++k) dr[k] = nu[k] - lc->nu_lya;
        for (int k = 0; k < c->nn; ++k) dcr[k] = nuC[k] - lc->nu_lya;
        hipchk(c, hipMemcpy(c->dnu_rel, dr.data(), dr.size() * sizeof(double), hipMemcpyHostToDevice), "dnu_rel");
        hipchk(c, hipMemcpy(c->dnuC_rel, dcr.data(), dcr.size() * sizeof(double), hipMemcpyHostToDevice), "dnuC_rel");
        c->lut_nx = nx;
        c->lut_xmax = xmax;
        c->lc = LineConst{lc->b_prefact, lc->nu_over_c, lc->a_prefact, lc->sigma0_prefact};
    });
}

int m1lya_prepare_subcycling(m1lya_ctx* c, const double* T, const double* n_HI, const double* alpha_c,
                             const double* alpha_c_F, const double* js, const double* pd, int phi_all_bins)
{
    return guarded(c, [&] {
        require_setup(c);
        if (!c->lut) fail(c, M1LYA_ERR_STATE, "prepare_subcycling: Voigt tables not set (m1lya_set_voigt_lut)");
        if (!T || !n_HI) fail(c, M1LYA_ERR_ARG, "prepare_subcycling: T and n_HI are required");
        const size_t b = c->ncell * sizeof(double);
        hipStream_t st = c->stream;
        hipchk(c, hipMemcpyAsync(c->T, T, b, hipMemcpyHostToDevice, st), "T");
        hipchk(c, hipMemcpyAsync(c->nHI, n_HI, b, hipMemcpyHostToDevice, st), "n_HI");
        if (alpha_c) hipchk(c, hipMemcpyAsync(c->alpha_c, alpha_c, b, hipMemcpyHostToDevice, st), "alpha_c");
        if (alpha_c_F) hipchk(c, hipMemcpyAsync(c->acF, alpha_c_F, b, hipMemcpyHostToDevice, st), "alpha_c_F");
        if (js) hipchk(c, hipMemcpyAsync(c->js, js, b, hipMemcpyHostToDevice, st), "js");
        if (pd) hipchk(c, hipMemcpyAsync(c->pd, pd, b, hipMemcpyHostToDevice, st), "pd");
        RefreshArgs r;
        r.nn = c->nn; r.ncell = (int)c->ncell; r.kmin = c->kcore_min; r.kmax = c->kcore_max;
        r.phi_all_bins = phi_all_bins ? 1 : 0;
        r.lut = c->lut; r.nx = c->lut_nx; r.xmax = c->lut_xmax; r.lc = c->lc;
        r.dnu_rel = c->dnu_rel; r.dnuC_rel = c->dnuC_rel; r.dnu = c->dnu;
        r.T = c->T; r.nHI = c->nHI; r.acF = alpha_c_F ? c->acF : nullptr;
        r.Hv = c->Hv; r.phi = c->phi; r.aF = c->aF; r.alpha = c->alpha; r.dnuD = c->dnuD;
        Prof p(c, "refresh");
        launch_refresh(st, r);
        p.done();
        c->fij_stale = true;
        hipchk(c, hipStreamSynchronize(st), "sync");
    });
}

int m1lya_prepare_subcycling_chem(m1lya_ctx* c, const double* chem, const double* rho_d, const double* kappa,
                                  double z, int ots, int destruction, int phi_all_bins)
{
    return guarded(c, [&] {
        require_setup(c);
        if (!c->lut) fail(c, M1LYA_ERR_STATE, "prepare_subcycling_chem: Voigt tables not set (m1lya_set_voigt_lut)");
        if (!chem || !kappa) fail(c, M1LYA_ERR_ARG, "prepare_subcycling_chem: chem and kappa are required");
        const size_t n = c->ncell;
        SourceArgs a{};
        a.ncell = (int)n;
        a.ots = ots ? 1 : 0;
        a.destruction = destruction ? 1 : 0;
        a.z = z;
        a.chem = rhd_stage(c, 0, chem, 8 * n);
        a.rho_d = rho_d ? rhd_stage(c, 8, rho_d, 3 * n) : nullptr;
        for (int m = 0; m < 6; ++m) a.kappa[m] = kappa[m];
        a.js = c->js; a.alpha_c = c->alpha_c; a.acF = c->acF; a.pd = c->pd; a.T = c->T; a.nHI = c->nHI;
        Prof p0(c, "sources");
        launch_sources(c->stream, a);
        p0.done();
        RefreshArgs r;
        r.nn = c->nn; r.ncell = (int)n; r.kmin = c->kcore_min; r.kmax = c->kcore_max;
        r.phi_all_bins = phi_all_bins ? 1 : 0;
        r.lut = c->lut; r.nx = c->lut_nx; r.xmax = c->lut_xmax; r.lc = c->lc;
        r.dnu_rel = c->dnu_rel; r.dnuC_rel = c->dnuC_rel; r.dnu = c->dnu;
        r.T = c->T; r.nHI = c->nHI; r.acF = c->acF;
        r.Hv = c->Hv; r.phi = c->phi; r.aF = c->aF; r.alpha = c->alpha; r.dnuD = c->dnuD;
        Prof p(c, "refresh");
        launch_refresh(c->stream, r);
        p.done();
        c->fij_stale = true;
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
    });
}

int m1lya_adjoint_run(m1lya_ctx* c, int n_steps, double c_tilde, double dt, double z_center, int bc_r_right,
                      int bc_z_bot, int bc_z_top, const double* mf_seed, const double* spec_seed,
                      int checkpoint_every, double* objective, double* history, double* grad_state,
                      double* grad_cell, double* grad_bin)
{
    return guarded(c, [&] {
        require_setup(c);
        check_bc(c, bc_r_right, "BC_r_right");
        check_bc(c, bc_z_bot, "BC_Z_bot");
        check_bc(c, bc_z_top, "BC_Z_top");
        if (n_steps < 1) fail(c, M1LYA_ERR_ARG, "n_steps < 1");
        if (checkpoint_every < 0) fail(c, M1LYA_ERR_ARG, "checkpoint_every < 0");
        if (!(c_tilde > 0.0) || !(dt > 0.0)) fail(c, M1LYA_ERR_ARG, "c_tilde and dt must be > 0");
        // the reverse column kernel keeps 16 rows of N_ν doubles in LDS (32 KiB at 256 bins)
        if (c->nn > kMaxBinsAdjoint)
            fail(c, M1LYA_ERR_DIMS, "the adjoint supports N_ν <= " + std::to_string(kMaxBinsAdjoint));
        const size_t n3 = c->n3, nc = c->ncell, B3 = 3 * n3;
        const int nn = c->nn, nr = c->nr, nz = c->nz;
        // √N checkpointing: states every S steps; a segment of S steps is replayed with its tape
        int S = checkpoint_every > 0 ? checkpoint_every : (int)std::ceil(std::sqrt((double)n_steps));
        S = std::max(1, std::min(S, n_steps));
        const int nck = (n_steps + S - 1) / S;
        AdjMem mem(c);
        double* ck = mem.get((size_t)nck * B3);
        double* tape = mem.get((size_t)2 * S * B3);   // per step of a segment: X_n, then the advection output
        double* fin = mem.get(B3);
        int* kr = mem.geti((size_t)2 * n_steps + 2);
        double* hist = mem.get((size_t)8 * n_steps);
        double* lam = mem.get(B3);     // cotangent of the state
        double* lam1 = mem.get(B3);    // cotangent of the advection output
        double* pre = mem.get(B3);
        double* gbin = mem.get(3 * n3);
        double* gcell = mem.get(5 * nc);
        double* fpart = c->adj_mode == 0 ? mem.get(adj_face_part_size(nn, nr, nz)) : nullptr;
        double* dv = mem.get(2 * (size_t)nn);
        hipStream_t st = c->stream;
        auto zero = [&](double* p, size_t n) {
            hipchk(c, hipMemsetAsync(p, 0, n * sizeof(double), st), "hipMemsetAsync");
        };
        auto copy3_from_state = [&](double* dst) {
            const size_t b = n3 * sizeof(double);
            hipchk(c, hipMemcpyAsync(dst, c->J, b, hipMemcpyDeviceToDevice, st), "copy state");
            hipchk(c, hipMemcpyAsync(dst + n3, c->Hr, b, hipMemcpyDeviceToDevice, st), "copy state");
            hipchk(c, hipMemcpyAsync(dst + 2 * n3, c->Hz, b, hipMemcpyDeviceToDevice, st), "copy state");
        };
        auto copy3_to_state = [&](const double* src) {
            const size_t b = n3 * sizeof(double);
            hipchk(c, hipMemcpyAsync(c->J, src, b, hipMemcpyDeviceToDevice, st), "restore state");
            hipchk(c, hipMemcpyAsync(c->Hr, src + n3, b, hipMemcpyDeviceToDevice, st), "restore state");
            hipchk(c, hipMemcpyAsync(c->Hz, src + 2 * n3, b, hipMemcpyDeviceToDevice, st), "restore state");
        };
        zero(lam, B3);
        zero(gbin, 3 * n3);
        zero(gcell, 5 * nc);
        zero(c->drhou_r, nc);
        zero(c->drhou_z, nc);

        // ---- forward sweep: checkpoints, active ranges, scalars; the last segment's tape ----
        if (!c->krange_valid) compute_krange(c);
        const int s_last = (nck - 1) * S;   // first step of the last segment, which is not replayed
        for (int s = 0; s < n_steps; ++s) {
            if (s % S == 0) copy3_from_state(ck + (size_t)(s / S) * B3);
            hipchk(c, hipMemcpyAsync(kr + 2 * s, c->krange, 2 * sizeof(int), hipMemcpyDeviceToDevice, st), "krange");
            double* tX1 = nullptr;
            if (s >= s_last) {
                double* Xn = tape + (size_t)2 * (s - s_last) * B3;
                copy3_from_state(Xn);
                tX1 = Xn + B3;
            }
            fused_step(c, c_tilde, dt, z_center, bc_r_right, bc_z_bot, bc_z_top, hist + (size_t)8 * s, tX1);
        }
        copy3_from_state(fin);
        hipchk(c, hipMemcpyAsync(kr + 2 * n_steps, c->krange, 2 * sizeof(int), hipMemcpyDeviceToDevice, st),
               "krange");
        std::vector<double> hh((size_t)8 * n_steps);
        hipchk(c, hipMemcpyAsync(hh.data(), hist, hh.size() * sizeof(double), hipMemcpyDeviceToHost, st),
               "history");
        check_err_flag(c, "adjoint forward sweep");

        // L_sum = 4π Σ ΔV js; the history's L_Lyα doubles it for a reflecting bottom
        const double Lsum = hh[3] / (bc_z_bot == M1LYA_REFLECT ? 2.0 : 1.0);
        const double mf_scale = Lsum > 0.0 ? kC / Lsum : 0.0;
        double obj = 0.0;
        if (mf_seed)
            for (int s = 0; s < n_steps; ++s)
                for (int q = 0; q < 3; ++q) obj += mf_seed[3 * s + q] * hh[(size_t)8 * s + q];
        if (spec_seed && Lsum > 0.0) {
            std::vector<double> Sk(nn);
            hipchk(c, hipMemcpyAsync(dv, spec_seed, nn * sizeof(double), hipMemcpyHostToDevice, st), "seed");
            launch_adj_spectrum(st, fin + n3, fin + 2 * n3, c->A_r, c->A_z, nn, nr, nz, bc_z_bot, 1.0 / Lsum,
                                dv + nn);
            hipchk(c, hipGetLastError(), "spectrum");
            hipchk(c, hipMemcpyAsync(Sk.data(), dv + nn, nn * sizeof(double), hipMemcpyDeviceToHost, st), "S");
            hipchk(c, hipStreamSynchronize(st), "sync");
            for (int k = 0; k < nn; ++k) obj += spec_seed[k] * Sk[k];
            launch_adj_spectrum_seed(st, fin + n3, fin + 2 * n3, c->A_r, c->A_z, nn, nr, nz, bc_z_bot, 1.0 / Lsum,
                                     dv, lam + n3, lam + 2 * n3);
            hipchk(c, hipGetLastError(), "spectrum seed");
        }

        // ---- reverse sweep, segment by segment ----
        for (int m = nck - 1; m >= 0; --m) {
            const int s0 = m * S, s1 = std::min(s0 + S, n_steps);
            if (m != nck - 1) {   // the last segment was taped by the forward sweep
                copy3_to_state(ck + (size_t)m * B3);
                for (int s = s0; s < s1; ++s) {
                    double* Xn = tape + (size_t)2 * (s - s0) * B3;
                    copy3_from_state(Xn);
                    hipchk(c, hipMemcpyAsync(c->krange, kr + 2 * s, 2 * sizeof(int), hipMemcpyDeviceToDevice, st),
                           "krange");
                    replay_step(c, c_tilde, dt, z_center, bc_r_right, bc_z_bot, bc_z_top, Xn + B3);
                }
            }
            for (int s = s1 - 1; s >= s0; --s) {
                const double* Xn = tape + (size_t)2 * (s - s0) * B3;
                const double* X1 = Xn + B3;
                const double* Xn1 = (s + 1 < s1) ? tape + (size_t)2 * (s + 1 - s0) * B3
                                                 : (s1 == n_steps ? fin : ck + (size_t)(m + 1) * B3);
                AdjColArgs ca;
                std::memset(&ca, 0, sizeof(ca));
                ca.nn = nn; ca.nr = nr; ca.nz = nz; ca.recoil = c->recoil;
                ca.J1 = X1; ca.Hr1 = X1 + n3; ca.Hz1 = X1 + 2 * n3;
                ca.J4 = Xn1;
                ca.lJ4 = lam; ca.lHr4 = lam + n3; ca.lHz4 = lam + 2 * n3;
                ca.lJ1 = lam1; ca.lHr1 = lam1 + n3; ca.lHz1 = lam1 + 2 * n3;
                ca.Hv = c->Hv; ca.phi = c->phi; ca.aF = c->aF;
                ca.alpha = c->alpha; ca.alpha_c = c->alpha_c; ca.pd = c->pd; ca.js = c->js; ca.T = c->T;
                ca.dnuD = c->dnuD;
                ca.dur_dr = c->dur_dr; ca.dur_dz = c->dur_dz; ca.duz_dr = c->duz_dr; ca.duz_dz = c->duz_dz;
                ca.ur_over_r = c->ur_over_r;
                ca.nu = c->nu; ca.nuC = c->nuC; ca.dnu = c->dnu; ca.inv_dnu = c->inv_dnu; ca.inv_nuC = c->inv_nuC;
                ca.dV = c->dV; ca.rcom = c->rcom; ca.zcom = c->zcom;
                ca.c_tilde = c_tilde; ca.dt = dt; ca.dw = 1.0 / (c_tilde * dt);
                ca.alpha_floor = c->alpha_floor; ca.z_center = z_center;
                ca.mf_scale = mf_scale;
                if (mf_seed) { ca.w0 = mf_seed[3 * s]; ca.w1 = mf_seed[3 * s + 1]; ca.w2 = mf_seed[3 * s + 2]; }
                ca.g_aF = gbin; ca.g_Hv = gbin + n3; ca.g_phi = gbin + 2 * n3;
                ca.g_alpha = gcell; ca.g_alpha_c = gcell + nc; ca.g_js = gcell + 2 * nc;
                ca.g_dnuD = gcell + 3 * nc; ca.g_T = gcell + 4 * nc;
                ca.mode = c->adj_mode;
                {
                    Prof p(c, "adj_column");
                    launch_adj_column(st, ca);
                    p.done();
                }
                AdjAdvArgs aa;
                aa.nn = nn;
                aa.g = DevGeom{nr, nz, c->A_r, c->A_z, c->inv_dV, c->inv_r_cell, c->ivr, c->ivz};
                aa.v = DevVel{c->ur_star, c->uz_star, c->dur_dr, c->dur_dz, c->duz_dr, c->duz_dz, c->ur_over_r};
                aa.J = Xn; aa.Hr = Xn + n3; aa.Hz = Xn + 2 * n3; aa.aF = c->aF;
                aa.lJ1 = lam1; aa.lHr1 = lam1 + n3; aa.lHz1 = lam1 + 2 * n3;
                aa.pJ = pre; aa.pHr = pre + n3; aa.pHz = pre + 2 * n3;
                aa.oJ = lam; aa.oHr = lam + n3; aa.oHz = lam + 2 * n3;
                aa.gaF = gbin;
                aa.cdt = c_tilde * dt;
                aa.bc_r_right = bc_r_right; aa.bc_z_bot = bc_z_bot; aa.bc_z_top = bc_z_top; aa.hll = c->flux;
                aa.krange = kr + 2 * s;
                aa.fpart = fpart;
                aa.mode = c->adj_mode;
                {
                    Prof p(c, "adj_advect");
                    launch_adj_advect(st, aa);
                    p.done();
                }
            }
        }
        // the context ends in the state after n_steps, as m1lya_run_steps leaves it
        copy3_to_state(fin);
        hipchk(c, hipMemcpyAsync(c->krange, kr + 2 * n_steps, 2 * sizeof(int), hipMemcpyDeviceToDevice, st),
               "krange");
        c->krange_valid = true;
        c->old_stale = true;
        c->fij_stale = true;
        // M_F and S carry 1/L_sum, L_sum = 4π Σ ΔV js
        if (Lsum > 0.0 && obj != 0.0) launch_adj_axpy_cell(st, gcell + 2 * nc, c->dV, -obj / Lsum * kFourPi, (int)nc);
        if (objective) *objective = obj;
        if (history) std::memcpy(history, hh.data(), hh.size() * sizeof(double));
        if (grad_state)
            hipchk(c, hipMemcpyAsync(grad_state, lam, B3 * sizeof(double), hipMemcpyDeviceToHost, st), "grad");
        if (grad_cell)
            hipchk(c, hipMemcpyAsync(grad_cell, gcell, 5 * nc * sizeof(double), hipMemcpyDeviceToHost, st), "grad");
        if (grad_bin)
            hipchk(c, hipMemcpyAsync(grad_bin, gbin, 3 * n3 * sizeof(double), hipMemcpyDeviceToHost, st), "grad");
        hipchk(c, hipStreamSynchronize(st), "sync");
        check_err_flag(c, "adjoint");
    });
}

int m1lya_set_adjoint_mode(m1lya_ctx* c, int mode)
{
    return guarded(c, [&] {
        if (mode != 0 && mode != 1) fail(c, M1LYA_ERR_ARG, "adjoint mode must be 0 or 1");
        c->adj_mode = mode;
    });
}

int m1lya_set_profiling(m1lya_ctx* c, int on)
{
    return guarded(c, [&] {
        collect_stats(c);
        c->profiling = (on == 2) ? 2 : (on != 0 ? 1 : 0);
    });
}

int m1lya_kernel_stats_count(m1lya_ctx* c)
{
    if (!c) return 0;
    int n = 0;
    int rc = guarded(c, [&] {
        collect_stats(c);
        n = (int)c->stats.size();
    });
    return rc == M1LYA_OK ? n : rc;
}

int m1lya_kernel_stats(m1lya_ctx* c, int index, char* name, size_t name_len, long* launches, double* total_ms)
{
    return guarded(c, [&] {
        collect_stats(c);
        if (index < 0 || index >= (int)c->stats.size()) fail(c, M1LYA_ERR_ARG, "stats index out of range");
        auto it = c->stats.begin();
        std::advance(it, index);
        if (name && name_len) std::snprintf(name, name_len, "%s", it->first.c_str());
        if (launches) *launches = it->second.launches;
        if (total_ms) *total_ms = it->second.total_ms;
    });
}

int m1lya_reset_kernel_stats(m1lya_ctx* c)
{
    return guarded(c, [&] {
        collect_stats(c);
        c->stats.clear();
    });
}

int m1lya_selftest_math(int op, const double* a, const double* b, double* out, size_t n)
{
    if (!a || !out || (op == 0 && !b) || op < 0 || op > 3) {
        g_last_error = "selftest: bad arguments";
        return M1LYA_ERR_ARG;
    }
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(double);
    hipError_t e = hipMalloc(&da, bytes);
    if (e == hipSuccess) e = hipMalloc(&db, bytes);
    if (e == hipSuccess) e = hipMalloc(&dout, bytes);
    if (e == hipSuccess) e = hipMemcpy(da, a, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && b) e = hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && n > 0) {
        launch_selftest(nullptr, op, da, db, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) {
        g_last_error = std::string("selftest: ") + hipGetErrorString(e);
        return M1LYA_ERR_HIP;
    }
    return M1LYA_OK;
}

}  // extern "C"
