"""Benchmark of the multi-frequency M1 Lyα step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 256×256 r-Z cloud with 128 frequency bins (τ0 = 1e8, T = 1e4 K,
polar cone of half-angle π/4, τ_abs = 1, point source of radius 0.5 cells) in a
seeded, developed radiation field with every bin active.  A step is one pass
of the RT-only time-loop body (advection, emission/absorption, frequency
diffusion solve, M_F / f_esc reductions) over the whole cloud, in FP64.

Metric: cell-frequency updates per second = N_r·N_Z·N_ν·K / wall time, with the
state resident in HBM.  The JSON line also carries the roofline of the dominant
kernel (algorithmic bytes per launch / its mean duration from HIP events on
the library's stream during the timed region) and the CPU oracle (the FP64
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
# Algorithmic HBM bytes per cell-frequency update of each kernel (FP64 words):
#   advect: reads J, H_r, H_Z, α_F, writes J, H_r, H_Z               -> 7 words
#   column: reads J, H_r, H_Z, ϕ_ν, Hvoigt, α_F, writes J, H_r, H_Z    -> 9 words
ALGO_BYTES = {"advect": 7 * 8, "column": 9 * 8}
STEP_ALGO_BYTES = 72   # one fused step: J, H_r, H_Z read+write, α_F, Hvoigt, ϕ_ν read (FP64)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--n-r", type=int, default=256)
    p.add_argument("--n-z", type=int, default=256)
    p.add_argument("--n-nu", type=int, default=128)
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


def cpu_baseline(cloud_factory, seconds: float):
    """FP64 oracle (C restatement of the reference, OpenMP over cells) on the host cores."""
    from oracle.oracle import OracleState

    from lydion_amd.static_cloud import developed_state
    cloud = developed_state(cloud_factory())
    o = OracleState(cloud)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o.step()  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        o.step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    upd = cloud.lya.n_nu * cloud.lya.n_r * cloud.lya.n_z
    return {"value": upd * n / el, "unit": "cell-freq updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same {cloud.lya.n_r}x{cloud.lya.n_z}x{cloud.lya.n_nu} workload "
                      f"({el:.1f} s, OMP_NUM_THREADS={threads})"}


def copy_calibration():
    """One plain copy kernel of 512 MiB in and 512 MiB out (the library's selftest op 3)."""
    from lydion_amd import m1lya as M
    n = 512 * 1024 * 1024 // 8
    M.selftest_math(3, np.ones(n))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of a cloud is a design note only)", file=sys.stderr)
        sys.exit(2)

    import lydion_amd as L
    from lydion_amd import m1lya as M
    from lydion_amd.static_cloud import developed_state

    def factory():
        return L.cone_dust_cloud_256x256x128(n_r=args.n_r, n_z=args.n_z, n_nu=args.n_nu)

    cloud = developed_state(factory())
    ctx = M.M1LyaContext(cloud.lya, cloud.geometry)
    lya = cloud.lya
    upd_per_step = lya.n_nu * lya.n_r * lya.n_z
    step_args = (cloud.c_tilde, cloud.dt, cloud.Zc)
    kw = dict(bc_z_bot=cloud.bc_z_bot, history=False)

    M.run_steps(ctx, args.warmup, *step_args, **kw)
    ctx.synchronize()
    ctx.reset_kernel_stats()
    ctx.set_profiling(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    M.run_steps(ctx, args.steps, *step_args, **kw)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    J = ctx.download("J")
    if not np.all(np.isfinite(J)):
        raise RuntimeError("non-finite J after the timed steps")
    if args.copy_calibration:
        copy_calibration()

    value = upd_per_step * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    dom = max((k for k in stats if k in ALGO_BYTES), key=lambda k: stats[k][1])
    launches, total_ms = stats[dom]
    avg_s = total_ms * 1e-3 / launches
    achieved = ALGO_BYTES[dom] * upd_per_step / avg_s / 1e9
    workload = f"cone_cloud_{lya.n_r}x{lya.n_z}x{lya.n_nu}_fp64"
    traffic = load_traffic(dom, workload)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": ALGO_BYTES[dom] * upd_per_step,
                "avg_launch_ms": round(1e3 * avg_s, 4),
                "kernels_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in sorted(stats.items())}}
    step_gbs = STEP_ALGO_BYTES * value / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline(factory, args.cpu_seconds)
    line = {
        "metric": "Lyα M1 cell-freq updates/s + %HBM peak; FP32/batched vs FP64 at M_F err≤1e-4",
        "value": value,
        "unit": "cell-freq updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded developed radiation field, every bin active)",
        "config": {"workload": "static cloud 256x256 r-Z, 128 bins, tau0=1e8, T=1e4 K, cone pi/4, tau_abs=1, "
                               "point source; FP64 plain step",
                   "grid": f"{lya.n_r}x{lya.n_z}", "n_nu": lya.n_nu, "precision": "fp64",
                   "step_hbm_frac_72B": round(step_gbs / HBM_PEAK_GBS, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
