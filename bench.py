#!/usr/bin/env python
"""Benchmark of the cubed-sphere advection step on one MI355X.

Default workload (N=1): the plain FP64 GPU run -- equiangular cubed sphere N=384
(6x384x384 cells), one tracer of two Gaussian hills (b=5) at lon 5pi/6 and 7pi/6 on
the equator, solid-body rotation about an axis tilted pi/4 from the pole, CFL 0.5,
1000 steps.  A "step" is one full advection step of every cell: ghost fill, both
directional passes of the split, update and mass fixer.

Prints ONE JSON line:
  value        tracer cell-updates/s over the timed steps (inputs resident in HBM)
  roofline     the step kernel against the HBM roofline (algorithmic bytes per launch
               over its event-timed mean duration); traffic from the committed
               rocprofv3 counter summary when present
  cpu_baseline the numpy FP64 oracle on this host, one core, on a bounded sample of the
               same workload (a few steps), scaled to cell-updates/s
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "py-cubed-sphere_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published

WORKLOADS = {
    # name: (N, case name, dtype, cfl, default steps)
    "two_hills_n384_f64": (384, "two_hills_tilted", "f64", 0.5, 1000),
    "nl10_n768_f32": (768, "nl10_deformation", "f32", 0.5, 200),
    "polar_n48_f64": (48, "solid_body_polar", "f64", 0.5, 384),
}


def algorithmic_bytes_per_update(dtype, ntr):
    """B_alg = 2 s_Q + 2 s_U / n_t: read Q^n, write Q^{n+1}, the two face winds per cell
    shared by the solve's tracers (halo re-reads, tables and scratch excluded)."""
    s = 8 if dtype == "f64" else 4
    return 2 * s + 2 * s / ntr


def traffic_from_profile(name):
    """HBM bytes per step-kernel launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "k_step_%s_hbm_traffic.txt" % name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(N, case, dt, seconds):
    """Time the oracle (numpy FP64, single thread) on the same workload: 1 warm-up step,
    then as many steps as fit in about ``seconds``."""
    from oracle import timestep as ots
    from oracle.cases import from_case
    grid, sim = ots.make(N, from_case(case), dt)
    ots.init_vars_adv(grid, sim)
    ots.adv_time_step(grid, sim, 1, dt)
    steps, t0 = 0, time.perf_counter()
    while True:
        ots.adv_time_step(grid, sim, steps + 2, (steps + 2) * dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 1000:
            break
    per_step = el / steps
    return {
        "value": 6.0 * N * N * case.ntr / per_step,
        "unit": "tracer cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": "numpy FP64 oracle, N=%d, %d timed steps after 1 warm-up (%.1f s, %.3f s/step)" % (N, steps, el, per_step),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="two_hills_n384_f64", choices=sorted(WORKLOADS))
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--graph", type=int, default=0,
                    help="replay the timed steps from HIP graphs of this many steps (0 = plain launches); "
                         "the step-kernel duration then comes from a separate plain-launch pass")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="stream a buffer of this many MiB through one elementwise kernel before the timed "
                         "steps (a kernel of known traffic for scaling hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every solve of this workload fits one MI355X and "
              "nothing here uses a second GPU (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    from cubedsphere.cases import workload
    from cubedsphere.grid import Grid
    from cubedsphere.solver import Advection, cfl_dt

    N, case_name, dtype, cfl, default_steps = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else default_steps
    W = args.warmup
    case = workload(case_name)
    grid = Grid(N)
    dt = cfl_dt(grid, case, cfl)
    sol = Advection(N, case, dt, dtype=dtype, device="cuda:0", tile=args.tile,
                    max_steps=W + 2 * K + max(args.graph, 0), grid=grid)
    sol.init_vars_adv()
    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)          # one elementwise kernel: reads and writes n * 8 bytes
        torch.cuda.synchronize()
        del src, dst
    sol.step(W)
    if args.graph > 0:
        sol.use_graph(args.graph)
        sol.step(args.graph)                 # capture outside the timed region
    else:
        sol.profile(K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.step(K)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if args.graph > 0:
        sol.use_graph(0)
        sol.profile(K)
        sol.step(K)
    kern_ms, launches = sol.profile_read()
    sol.profile(0)

    updates = 6.0 * N * N * case.ntr * K
    elapsed = t1 - t0
    value = updates / elapsed
    b_alg = algorithmic_bytes_per_update(dtype, case.ntr)
    bytes_per_launch = b_alg * 6.0 * N * N * case.ntr
    kern_s = kern_ms / 1e3 / max(launches, 1)
    achieved = bytes_per_launch / kern_s / 1e9
    traffic = traffic_from_profile(args.workload)

    out = {
        "metric": "tracer cell-updates/s",
        "value": value,
        "unit": "tracer cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (analytic initial condition and wind on the generated grid)",
        "config": {"workload": args.workload, "N": N, "cells": 6 * N * N, "tracers": case.ntr,
                   "case": case_name, "cfl": cfl, "dt": dt, "tile": args.tile,
                   "options": "recon PPM-PL07, dp RK1, split AVLT, ET-DG, MT-0, MF-PR"},
        "hbm_frac_whole_step": value * b_alg / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_step",
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "mean_launch_us": kern_s * 1e6,
            "launches_timed": launches,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, case, dt, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
