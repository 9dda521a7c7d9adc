"""Benchmark of the multi-frequency M1 Lyα step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|fp64] [--workload cone|sweep]

Workload: the 256×256 r-Z cloud with 128 frequency bins (τ0 = 1e8, T = 1e4 K,
polar cone of half-angle π/4, τ_abs = 1, point source of radius 0.5 cells) in a
seeded, developed radiation field with every bin active.  A step is one pass
of the RT-only time-loop body (advection, emission/absorption, frequency
diffusion solve, M_F / f_esc reductions) over the whole cloud.

Default (--precision fp32): the mixed-precision batched step (FP32 state, flux
sweep and closure; FP64 emission, frequency factors and tridiagonal solve;
compensated reductions) for a batch of this one cloud, tile shape chosen by the
library's autotuner before the timed region.  Beside it the FP64 plain step
(bit-identical to the reference restatement) is timed on the same state, and
the accuracy block states the FP32 step's error in M_F, f_esc and the
frequency-integrated J against the FP64 step after the same number of steps.

--workload sweep: a batch of 64 seeded static clouds of uneven grid size
(32..256 per side), optical depth (τ0 1e5..1e9), temperature, dust and
frequency resolution (32..128 bins), stepped by one batched launch per step,
beside the same clouds run one after another through the FP64 plain step.

Metric: cell-frequency updates per second = N_r·N_Z·N_ν·K / wall time, state
resident in HBM.  The JSON line carries the roofline of the dominant kernel
(algorithmic bytes per launch / its mean duration from HIP events on the
library's stream during the timed region) and the CPU oracle (the FP64
restatement of the reference) timed on the host cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-frequency update of each kernel:
#   advect (FP64): reads J, H_r, H_Z, α_F, writes J, H_r, H_Z                     -> 7 x 8 B
#   column (FP64): reads J, H_r, H_Z, ϕ_ν, Hvoigt, α_F, writes J, H_r, H_Z          -> 9 x 8 B
#   k_step (FP32 state, one fused pass): reads J, H_r, H_Z, α_F, Hvoigt, ϕ_ν,
#                                         writes J, H_r, H_Z                       -> 9 x 4 B
ALGO_BYTES = {"advect": 7 * 8, "column": 9 * 8, "k_step": 9 * 4}
STEP_ALGO_BYTES = {"fp64": 72, "fp32": 36}
METRIC = "Lyα M1 cell-freq updates/s + %HBM peak; FP32/batched vs FP64 at M_F err≤1e-4"
WORKLOAD = ("cone cloud 256x256 r-Z, 128 bins, tau0=1e8, T=1e4 K, cone half-angle pi/4, tau_abs=1, "
            "point source of radius 0.5 cells")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None, help="timed steps (default 100; rhd: 200 substeps)")
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    p.add_argument("--workload", choices=("cone", "sweep", "adjoint", "rhd"), default="cone",
                   help="cone: the 256x256x128 cone cloud; sweep: the seeded batch of 64 uneven clouds; "
                        "adjoint: gradient of M_F over the steps of the 128x128x128 expanding cloud; "
                        "rhd: 200 FP32 substeps of the 256x512x128 RHD-like state, with the FP64 gradient of the "
                        "integrated radial force")
    p.add_argument("--clouds", type=int, default=64, help="sweep: number of clouds")
    p.add_argument("--n-r", type=int, default=256)
    p.add_argument("--n-z", type=int, default=256)
    p.add_argument("--n-nu", type=int, default=128)
    p.add_argument("--accuracy-steps", type=int, default=50)
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU-oracle sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run a plain 512 MiB device copy (calibrates rocprofv3 FETCH/WRITE_SIZE)")
    return p.parse_args()


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (profiles/)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(cloud_factory, seconds: float, single_seconds: float = 5.0):
    """FP64 oracle (C restatement of the reference, gcc -O3 -march=x86-64-v3, OpenMP over cells)
    on the host cores, and on one core for a shorter sample."""
    from oracle.oracle import OracleState, load

    from lydion_amd.static_cloud import developed_state
    cloud = developed_state(cloud_factory())
    o = OracleState(cloud)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    upd = cloud.lya.n_nu * cloud.lya.n_r * cloud.lya.n_z

    def sample(limit):
        o.step()  # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            o.step()
            n += 1
            el = time.perf_counter() - t0
            if el >= limit:
                return n, el

    n, el = sample(seconds)
    load().orc_set_threads(1)
    n1, el1 = sample(single_seconds)
    load().orc_set_threads(threads)
    return {"value": upd * n / el, "unit": "cell-freq updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 steps of the same {cloud.lya.n_r}x{cloud.lya.n_z}x{cloud.lya.n_nu} workload "
                      f"({el:.1f} s, OMP_NUM_THREADS={threads})",
            "single_thread": {"value": upd * n1 / el1, "cores": 1,
                              "sample": f"{n1} FP64 steps on one thread ({el1:.1f} s)"}}


def copy_calibration():
    """One plain copy kernel of 512 MiB in and 512 MiB out (the library's selftest op 3)."""
    from lydion_amd import m1lya as M
    n = 512 * 1024 * 1024 // 8
    M.selftest_math(3, np.ones(n))


def run_fp64(cloud, steps, warmup):
    """Timed FP64 plain steps; returns wall time and per-kernel event totals.  The timed region
    brackets only the column kernel (the dominant one) with HIP events; an untimed pass of up
    to 20 steps with every launch bracketed gives the other kernels' share."""
    from lydion_amd import m1lya as M
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    args = (cloud.c_tilde, cloud.dt, cloud.Zc)
    kw = dict(bc_z_bot=cloud.bc_z_bot, history=False)
    M.run_steps(ctx, warmup, *args, **kw)
    ctx.synchronize()
    ctx.reset_kernel_stats()
    ctx.set_profiling(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    M.run_steps(ctx, steps, *args, **kw)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    nb = min(steps, 20)
    ctx.reset_kernel_stats()
    ctx.set_profiling(1)
    M.run_steps(ctx, nb, *args, **kw)
    for name, (launches, total_ms) in ctx.kernel_stats().items():
        if name not in stats:   # per-step share scaled to the timed steps
            stats[name] = (launches * steps // nb, total_ms * steps / nb)
    ctx.set_profiling(False)
    if not np.all(np.isfinite(ctx.download("J"))):
        raise RuntimeError("non-finite J after the FP64 steps")
    ctx.close()
    return wall, stats


def run_fp32(clouds, steps, warmup, autotune):
    """Timed batched FP32 steps of a batch of clouds.  The timed region launches the
    kernels eagerly with HIP events around each k_step launch (the dominant kernel; no
    host synchronisation).  An untimed pass of up to 20 steps with every launch bracketed
    gives the small reduction kernels' share, and a last pass replays the captured hipGraph."""
    from lydion_amd.batch import M1LyaBatch
    b = M1LyaBatch(clouds)
    tiles = b.autotune(trial_steps=3) if autotune else (*b.tiles(), None)
    b.run_steps(warmup, history=False)
    b.reset_kernel_stats()
    b.set_profiling(2)
    b.synchronize()
    t0 = time.perf_counter()
    b.run_steps(steps, history=False)
    b.synchronize()
    wall = time.perf_counter() - t0
    stats = b.kernel_stats()
    nb = min(steps, 20)
    b.reset_kernel_stats()
    b.set_profiling(1)
    b.run_steps(nb, history=False)
    for name, (launches, total_ms) in b.kernel_stats().items():
        if name not in stats:   # per-step share scaled to the timed steps
            stats[name] = (launches * steps // nb, total_ms * steps / nb)
    b.set_profiling(False)
    b.set_graph(True)
    b.run_steps(2, history=False)
    b.synchronize()
    t1 = time.perf_counter()
    b.run_steps(steps, history=False)
    b.synchronize()
    graph_wall = time.perf_counter() - t1
    for idx in range(len(clouds)):
        if not np.all(np.isfinite(b.download(idx, "J"))):
            raise RuntimeError(f"non-finite J in cloud {idx} after the FP32 steps")
    b.close()
    return wall, stats, graph_wall, tiles


def accuracy(cloud, nsteps):
    """FP32 batched vs FP64 plain from the same state after nsteps steps: M_F and f_esc at every
    step, the frequency-integrated J, and the emergent (escaping) spectrum of the last step."""
    from lydion_amd import adjoint as A
    from lydion_amd import m1lya as M
    from lydion_amd.batch import M1LyaBatch
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    h64 = M.run_steps(ctx, nsteps, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
    b = M1LyaBatch([cloud])
    h32 = b.run_steps(nsteps)[:, 0, :]
    rel = lambda a, r: float(np.max(np.abs(a - r) / np.maximum(np.abs(r), 1e-300)))
    w = cloud.lya["dnu"][None, None, :] * cloud.geometry.dV[:, :, None]
    J64, J32 = ctx.download("J"), b.download(0, "J")
    L_sim = float(h64[-1, 3]) * (0.5 if cloud.bc_z_bot == 0 else 1.0)
    S64 = A.escaping_spectrum(cloud.geometry, ctx.download("H_r"), ctx.download("H_z"), L_sim, cloud.bc_z_bot)
    S32 = A.escaping_spectrum(cloud.geometry, b.download(0, "H_r"), b.download(0, "H_z"), L_sim, cloud.bc_z_bot)
    big = S64 > 1e-3 * S64.max()
    out = {"steps": nsteps, "M_F_max_rel_err": rel(h32[:, 0], h64[:, 0]),
           "f_esc_max_rel_err": rel(h32[:, 4], h64[:, 4]),
           "J_energy_rel_err": float(np.sum(np.abs(J32 - J64) * w) / np.sum(np.abs(J64) * w)),
           "spectrum_max_rel_err": rel(S32[big], S64[big]),
           "spectrum_bins_compared": int(big.sum()),
           "M_F_fp64_last": float(h64[-1, 0]), "f_esc_fp64_last": float(h64[-1, 4])}
    ctx.close()
    b.close()
    return out


CUS, CLOCK_HZ = 256, 2.4e9   # MI355X: compute units, peak engine clock


def load_valu(kernel: str, workload: str):
    """VALU wave-instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU, profiles/)."""
    path = os.path.join(ROOT, "profiles", f"sq_counters_{workload}.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["SQ_INSTS_VALU"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(stats, dom, upd_per_step, steps, workload):
    """The dominant kernel against the HBM roofline.  Beside it, the share of its duration that
    issuing its VALU instructions takes (one wave64 VALU instruction per CU per cycle): where
    that share exceeds the HBM fraction, instruction issue, not bandwidth, limits the kernel.
    For the batched step, k_step and k_step_pair (clouds of N_ν <= 32, launched before k_step in
    the same step) are taken together: their summed time per step against all the step's units."""
    parts = [k for k in ([dom, "k_step_pair"] if dom == "k_step" else [dom]) if k in stats]
    total_ms = sum(stats[k][1] for k in parts)
    avg_s = total_ms * 1e-3 / steps
    achieved = ALGO_BYTES[dom] * upd_per_step / avg_s / 1e9
    valus = [load_valu(k, workload) for k in parts]
    valu = None if any(v is None for v in valus) else sum(valus)
    issue = None if valu is None else valu / (CUS * CLOCK_HZ) / avg_s
    frac = achieved / HBM_PEAK_GBS
    traffics = [load_traffic(k, workload) for k in parts]
    traffic = None if any(t is None for t in traffics) else sum(traffics)
    return {"bound": "hbm", "kernel": "+".join(parts), "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(frac, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": ALGO_BYTES[dom] * upd_per_step,
            "avg_launch_ms": round(1e3 * avg_s, 4),
            "valu_issue_frac": None if issue is None else round(issue, 4),
            "valu_insts_per_update": None if valu is None else round(valu / upd_per_step, 3),
            "limiter": None if issue is None else ("valu-issue" if issue > frac else "hbm"),
            "kernels_ms_per_step": {k: round(v[1] / steps, 4) for k, v in sorted(stats.items())}}


SWEEP_WORKLOAD = ("sweep of {n} seeded static clouds (cloud i from PCG64({seed}+i)): N_r, N_Z uniform in "
                  "{{32,64,128,256}}, tau0 log-uniform in [1e5,1e9], T log-uniform in [1e2,1e4] K, tau_abs uniform "
                  "in [0,10], N_nu uniform in {{32,64,128}}, point or uniform source; one batched launch per step")


def cpu_baseline_sweep(clouds, seconds: float):
    """The FP64 oracle stepping the sweep's clouds in order, two steps per cloud per pass,
    passes repeated until the time bound."""
    from oracle.oracle import OracleState
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    done, el, visits = 0, 0.0, 0
    while el < seconds:
        for cloud in clouds:
            o = OracleState(cloud)
            t0 = time.perf_counter()
            o.step()
            o.step()
            el += time.perf_counter() - t0
            done += 2 * cloud.lya.n_nu * cloud.lya.n_r * cloud.lya.n_z
            visits += 1
            if el >= seconds:
                break
    return {"value": done / el, "unit": "cell-freq updates/s", "cores": threads, "kind": "port",
            "sample": f"{visits} visits of 2 FP64 steps over the sweep's clouds in order ({el:.1f} s, "
                      f"OMP_NUM_THREADS={threads})"}


def main_sweep(args):
    """The batched sweep: FP32 batch of all clouds vs the same clouds run one after another
    through the FP64 plain step."""
    from lydion_amd import m1lya as M
    from lydion_amd import workloads as W
    from lydion_amd.batch import M1LyaBatch
    from lydion_amd.static_cloud import DEFAULT_SEED, developed_state

    clouds = [developed_state(c) for c in W.sweep_clouds(args.clouds, DEFAULT_SEED)]
    upd = W.updates_per_step(clouds)
    wall, stats, graph_wall, tiles = run_fp32(clouds, args.steps, args.warmup, not args.no_autotune)
    value = upd * args.steps / wall

    # the same clouds, one FP64 plain run after another (each timed alone, summed)
    steps64 = max(5, args.steps // 10)
    wall64 = 0.0
    for cloud in clouds:
        w, _ = run_fp64(cloud, steps64, 2)
        wall64 += w
    seq64 = upd * steps64 / wall64

    # accuracy of the batch against each cloud's FP64 step after the same steps
    n_acc = args.accuracy_steps
    b = M1LyaBatch(clouds)
    h32 = b.run_steps(n_acc)
    b.close()
    err_mf, err_fesc = 0.0, 0.0
    for idx, cloud in enumerate(clouds):
        ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
        h64 = M.run_steps(ctx, n_acc, cloud.c_tilde, cloud.dt, cloud.Zc, bc_z_bot=cloud.bc_z_bot)
        ctx.close()
        rel = lambda a, r: float(np.max(np.abs(a - r) / np.maximum(np.abs(r), 1e-300)))
        err_mf = max(err_mf, rel(h32[:, idx, 0], h64[:, 0]))
        err_fesc = max(err_fesc, rel(h32[:, idx, 4], h64[:, 4]))

    if args.copy_calibration:
        copy_calibration()
    step_gbs = STEP_ALGO_BYTES["fp32"] * value / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline_sweep(clouds, args.cpu_seconds)
    line = {
        "metric": METRIC, "value": value, "unit": "cell-freq updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded developed radiation field per cloud, every bin active)",
        "config": {"workload": SWEEP_WORKLOAD.format(n=len(clouds), seed=DEFAULT_SEED), "clouds": len(clouds),
                   "updates_per_step": upd, "precision": "fp32", "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                   "step_algo_bytes_per_update": STEP_ALGO_BYTES["fp32"]},
        "roofline": roofline(stats, "k_step", upd, args.steps, f"sweep{len(clouds)}_fp32"),
        "cpu_baseline": cpu,
        "fp64_sequential": {"value": seq64, "ms_per_step": 1e3 * wall64 / steps64, "steps": steps64,
                            "note": "each cloud through its own FP64 context, one after another"},
        "speedup_vs_fp64_sequential": value / seq64,
        "graph_ms_per_step": 1e3 * graph_wall / args.steps,
        "tiles": {"seg_len": tiles[0], "waves_per_block": tiles[1]},
        "accuracy": {"steps": n_acc, "M_F_max_rel_err": err_mf, "f_esc_max_rel_err": err_fesc,
                     "over": f"all {len(clouds)} clouds"},
    }
    print(json.dumps(line))


ADJ_WORKLOAD = ("expanding cloud 128x128 r-Z, 128 bins, tau0=1e7, T=1e4 K, Hubble outflow u_max=200 km/s, "
                "lognormal density (sigma_ln=0.5, P(k)~k^-11/3), point source; gradient of the per-step M_F "
                "over {n} FP64 steps, states checkpointed every ceil(sqrt(n)) steps")


def main_adjoint(args):
    """Forward FP64 steps vs the gradient run (forward sweep + segment replays + reverse sweep)."""
    from lydion_amd import adjoint as A
    from lydion_amd import m1lya as M
    from lydion_amd import workloads as W
    from lydion_amd.static_cloud import developed_state

    cloud = developed_state(W.expanding_cloud(128, 128, 128, tau0=1e7, u_max=200.0, sigma_ln=0.5))
    n = args.steps
    upd = cloud.lya.n_nu * cloud.lya.n_r * cloud.lya.n_z
    run = dict(bc_z_bot=cloud.bc_z_bot)
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    M.run_steps(ctx, args.warmup, cloud.c_tilde, cloud.dt, cloud.Zc, history=False, **run)
    ctx.synchronize()
    t0 = time.perf_counter()
    M.run_steps(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, history=False, **run)
    ctx.synchronize()
    fwd = time.perf_counter() - t0
    ctx.close()

    mf = np.zeros((n, 3))
    mf[:, 0] = 1.0                 # Σ_s M_F(s)
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    A.adjoint_run(ctx, 2, cloud.c_tilde, cloud.dt, cloud.Zc, mf[:2], **run)   # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    res = A.adjoint_run(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, mf, **run)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # the kernels' shares from an untimed second run with every launch bracketed by HIP events
    ctx.reset_kernel_stats()
    ctx.set_profiling(True)
    A.adjoint_run(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, mf, **run)
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    ctx.close()
    if not all(np.all(np.isfinite(g)) for g in res.grad_state.values()):
        raise RuntimeError("non-finite gradient")
    line = {
        "metric": METRIC, "value": upd * n / wall, "unit": "cell-freq updates/s (gradient run)", "n_gpus": 1,
        "steps": n, "warmup": args.warmup, "ms_per_step": 1e3 * wall / n, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lognormal density, developed radiation field)",
        "config": {"workload": ADJ_WORKLOAD.format(n=n), "checkpoint_every": int(np.ceil(np.sqrt(n)))},
        "forward_ms_per_step": 1e3 * fwd / n,
        "gradient_over_forward": wall / fwd,
        "kernels_ms_per_step": {k: round(v[1] / n, 4) for k, v in sorted(stats.items())},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


RHD_WORKLOAD = ("RHD-like state 256x512 r-Z (open at both Z ends), 128 bins: lognormal n_H (mean 1e3 cm^-3, "
                "sigma_ln=1, P(k)~k^-11/3), T = clip(1e4 K (n/n_min)^-s, 1e2, 1e4) with s={s:.3f}, Mach-2 turbulent "
                "velocity, dust alpha_c = kappa_abs(Lya) x 0.01 x 0.1 x rho, central point source of 1e5 Lsun; "
                "{n} substeps")


def main_rhd(args):
    """FP32 batched substeps of the RHD-like state, the FP64 step beside them, and the FP64
    gradient of the integrated radial Lyα force Σ_s M_F_radial_net(s) w.r.t. n_H and L."""
    from lydion_amd import adjoint as A
    from lydion_amd import m1lya as M
    from lydion_amd import workloads as W
    from lydion_amd.static_cloud import developed_state

    cloud = developed_state(W.rhd_like_cloud())
    n = args.steps
    if args.copy_calibration:
        copy_calibration()
    upd = cloud.lya.n_nu * cloud.lya.n_r * cloud.lya.n_z
    wall, stats, graph_wall, tiles = run_fp32([cloud], n, args.warmup, not args.no_autotune)
    value = upd * n / wall
    steps64 = max(5, n // 10)
    wall64, stats64 = run_fp64(cloud, steps64, 2)
    acc = accuracy(cloud, args.accuracy_steps)

    mf = np.zeros((n, 3))
    mf[:, 1] = 1.0                   # Σ_s M_F_radial_net(s): the integrated radial force
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    t0 = time.perf_counter()
    res = A.adjoint_run(ctx, n, cloud.c_tilde, cloud.dt, cloud.Zc, mf, bc_z_bot=cloud.bc_z_bot)
    adj_wall = time.perf_counter() - t0
    ctx.close()
    g = A.physical_gradients(cloud, res, with_T=False)
    step_gbs = STEP_ALGO_BYTES["fp32"] * value / 1e9
    line = {
        "metric": METRIC, "value": value, "unit": "cell-freq updates/s", "n_gpus": 1, "steps": n,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / n, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded turbulent RHD-like state, developed radiation field)",
        "config": {"workload": RHD_WORKLOAD.format(s=cloud.params["s_T"], n=n), "grid": "256x512", "n_nu": 128,
                   "precision": "fp32", "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                   "step_algo_bytes_per_update": STEP_ALGO_BYTES["fp32"]},
        "roofline": roofline(stats, "k_step", upd, n, "rhd_256x512x128_fp32"),
        "cpu_baseline": None,
        "fp64_plain": {"value": upd * steps64 / wall64, "ms_per_step": 1e3 * wall64 / steps64, "steps": steps64},
        "speedup_vs_fp64": value / (upd * steps64 / wall64),
        "graph_ms_per_step": 1e3 * graph_wall / n,
        "tiles": {"seg_len": tiles[0], "waves_per_block": tiles[1]},
        "accuracy": acc,
        "adjoint": {"objective": "sum over substeps of M_F_radial_net", "steps": n, "wall_s": adj_wall,
                    "ms_per_step": 1e3 * adj_wall / n, "dObj_dlnL": g["L"] * cloud.params["L_lya"],
                    "dObj_dnH_abs_max": float(np.max(np.abs(g["n_H"]))),
                    "checkpoint_every": int(np.ceil(np.sqrt(n)))},
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.steps is None:
        args.steps = 200 if args.workload == "rhd" else 100
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of a cloud is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint":
        main_adjoint(args)
        return
    if args.workload == "rhd":
        main_rhd(args)
        return
    if args.workload == "sweep":
        if args.precision != "fp32":
            print("bench.py: the sweep workload measures the FP32 batch (with the FP64 sequential runs beside it)",
                  file=sys.stderr)
            sys.exit(2)
        main_sweep(args)
        return

    import lydion_amd as L
    from lydion_amd.static_cloud import developed_state

    def factory():
        return L.cone_dust_cloud_256x256x128(n_r=args.n_r, n_z=args.n_z, n_nu=args.n_nu)

    cloud = developed_state(factory())
    lya = cloud.lya
    upd_per_step = lya.n_nu * lya.n_r * lya.n_z
    grid = f"{lya.n_r}x{lya.n_z}x{lya.n_nu}"

    fp64_steps = max(10, args.steps // 3) if args.precision == "fp32" else args.steps
    wall64, stats64 = run_fp64(cloud, fp64_steps, args.warmup)
    fp64 = {"value": upd_per_step * fp64_steps / wall64, "ms_per_step": 1e3 * wall64 / fp64_steps,
            "steps": fp64_steps,
            "roofline": roofline(stats64, max((k for k in stats64 if k in ALGO_BYTES), key=lambda k: stats64[k][1]),
                                 upd_per_step, fp64_steps, f"cone_cloud_{grid}_fp64")}

    if args.precision == "fp32":
        wall, stats, graph_wall, tiles = run_fp32([cloud], args.steps, args.warmup, not args.no_autotune)
        value = upd_per_step * args.steps / wall
        dom = "k_step"
        workload = f"cone_cloud_{grid}_fp32"
        extra = {"fp64_plain": fp64, "speedup_vs_fp64": value / fp64["value"],
                 "graph_ms_per_step": 1e3 * graph_wall / args.steps,
                 "tiles": {"seg_len": tiles[0], "waves_per_block": tiles[1]},
                 "accuracy": accuracy(cloud, args.accuracy_steps)}
        dtype = "f32+f64"
    else:
        wall, stats, value = wall64, stats64, fp64["value"]
        dom = max((k for k in stats if k in ALGO_BYTES), key=lambda k: stats[k][1])
        workload = f"cone_cloud_{grid}_fp64"
        extra = {}
        dtype = "f64"
    steps = args.steps if args.precision == "fp32" else fp64_steps
    if args.copy_calibration:
        copy_calibration()
    step_gbs = STEP_ALGO_BYTES[args.precision] * value / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline(factory, args.cpu_seconds)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-freq updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded developed radiation field, every bin active)",
        "config": {"workload": WORKLOAD + ("; FP32 batched step (batch of 1)" if args.precision == "fp32"
                                           else "; FP64 plain step"),
                   "grid": f"{lya.n_r}x{lya.n_z}", "n_nu": lya.n_nu, "precision": args.precision,
                   "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                   "step_algo_bytes_per_update": STEP_ALGO_BYTES[args.precision]},
        "roofline": roofline(stats, dom, upd_per_step, steps, workload),
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
