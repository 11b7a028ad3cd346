#!/usr/bin/env python
"""Benchmark of the cubed-sphere advection step on one MI355X.

Workloads (--workload):
  two_hills_n384_f64   default; the plain FP64 GPU run: equiangular cubed sphere N=384
                       (6x384x384 cells), one tracer of two Gaussian hills (b=5) at lon
                       5pi/6 and 7pi/6 on the equator, solid-body rotation about an axis
                       tilted pi/4 from the pole, CFL 0.5, 1000 steps
  nl10_n768_f32        precision form: N=768, 4 tracers (Gaussian hills, cosine bells,
                       slotted cylinders, their mix) in the Nair-Lauritzen deformational
                       wind (k=2.4, T=5), FP32 state and fluxes, FP64 tables and mass
  batched_64_f32       batched form: 64 seeded solves, N in {32,48,96,192,384,512}, 1-8
                       tracers, random Gaussian centre / width / rotation-axis tilt, FP32,
                       one launch per kernel per step over all solves, graph-replayed,
                       tile shape from the autotuner
  polar_n48_f64        the reference's CPU-sized case (N=48 polar rotation), for reference
  two_hills_n384_f64_etpl07    the N=384 FP64 case through the general step with schemes 1
  two_hills_n384_f64_rk2_mfaf  and 3 of the reference's convergence study (ET-PL07 with the
                       PL07 split; RK2 with MF-AF); its roofline is over the whole step

A "step" is one full advection step of every cell of every solve: ghost fill, both
directional passes of the split, update and mass fixer (plus the wind update of
time-dependent winds).

Prints ONE JSON line:
  value        tracer cell-updates/s over the timed steps (inputs resident in HBM)
  roofline     the step kernel against the HBM roofline (algorithmic bytes per launch
               over its mean duration: HIP events on the launch stream around one graph
               of 200 back-to-back launches); traffic from the committed rocprofv3
               counter summary of the workload when present
  cpu_baseline the numpy FP64 oracle on this host, one core, on a bounded sample of the
               same workload (a few steps of one solve), scaled to cell-updates/s
  l2_vs_f64    (FP32 workloads) relative L2 difference of the final states from an FP64
               GPU run of the same solves and steps
"""
import argparse
import json
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
    # name: (kind, N, case name, dtype, cfl, default steps)
    "two_hills_n384_f64": ("single", 384, "two_hills_tilted", "f64", 0.5, 1000),
    "nl10_n768_f32": ("single", 768, "nl10_deformation", "f32", 0.5, 300),
    "batched_64_f32": ("batched", None, None, "f32", 0.5, 200),
    "polar_n48_f64": ("single", 48, "solid_body_polar", "f64", 0.5, 384),
    "served_poisson_f32": ("served", None, None, "f32", 0.5, None),
    # the general step (the options the fused step kernel does not cover), on the N=384 case
    "two_hills_n384_f64_etpl07": ("single", 384, "two_hills_tilted", "f64", 0.5, 300),
    "two_hills_n384_f64_rk2_mfaf": ("single", 384, "two_hills_tilted", "f64", 0.5, 300),
}
# par/advection.par option codes of a workload (default: recon 3, dp 1, split 1, et 3, mt 1, mf 3);
# the two general-step workloads are schemes 1 and 3 of src/advection_error.py:61-66
WORKLOAD_OPTS = {
    "two_hills_n384_f64_etpl07": dict(opsplit=3, et=2, mt=2, mf=1),       # PPM-PL07, SP-PL07, ET-PL07, MT-PL07, MF-0
    "two_hills_n384_f64_rk2_mfaf": dict(dp=2, opsplit=1, et=3, mt=1, mf=2),  # RK2, SP-AVLT, ET-DG, MT-0, MF-AF
}
OPTION_NAMES = {
    "recon": ("PPM-0", "PPM-CW84", "PPM-PL07", "PPM-L04"), "dp": ("RK1", "RK2"),
    "opsplit": ("AVLT", "L04", "PL07"), "et": ("ET-S72", "ET-PL07", "ET-DG"), "mt": ("MT-0", "MT-PL07"),
    "mf": ("MF-0", "MF-AF", "MF-PR"),
}


def options_string(opts):
    o = dict(recon=3, dp=1, opsplit=1, et=3, mt=1, mf=3)
    o.update(opts)
    n = {k: OPTION_NAMES[k][o[k] - 1] for k in OPTION_NAMES}
    return "recon %s, dp %s, split %s, %s, %s, %s" % (n["recon"], n["dp"], n["opsplit"], n["et"], n["mt"], n["mf"])


def is_general(opts):
    """The general step runs the options the fused step kernel does not: ET-S72 and ET-PL07."""
    return opts.get("et", 3) != 3
SERVED_REQUESTS, SERVED_RATE = 200, 12.0      # Poisson stream: count, arrivals per second
BATCH_NS = (32, 48, 96, 192, 384, 512)
# tile ids from the autotuner for the single-solve workloads (the batched workload tunes
# itself at start-up over BATCH_TILE_CANDIDATES)
TUNED_TILE = {("f64", 384): 0, ("f32", 768): 3, ("f64", 48): 1}
BATCH_TILE_CANDIDATES = (0, 2, 6, 8, 9, 10)


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


def cpu_baseline(N, case, dt, seconds, opts=None):
    """Time the oracle (numpy FP64, single thread) on one solve of the workload, with the
    workload's options: 1 warm-up step, then as many steps as fit in about ``seconds``."""
    from oracle import timestep as ots
    from oracle.cases import from_case
    grid, sim = ots.make(N, from_case(case), dt, **(opts or {}))
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
        "sample": "numpy FP64 oracle, N=%d, %d tracer(s), %d timed steps after 1 warm-up (%.1f s, %.3f s/step)"
                  % (N, case.ntr, steps, el, per_step),
    }


def batch_specs(seed, steps):
    """The 64 seeded solves of the batched form: (N, case, steps)."""
    from cubedsphere.batch import seeded_batch
    return seeded_batch(seed, steps, BATCH_NS)


def timed_run(stepper, K, W, graph):
    """W warm-up steps, then K timed steps (graph replay when graph > 0, else plain launches
    with event pairs around each step-kernel launch); returns (elapsed s, kernel ms, launches)."""
    import torch
    stepper.step(W)
    if graph > 0:
        stepper.use_graph(graph)
        stepper.step(graph)                  # capture outside the timed region
    else:
        stepper.profile(K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stepper.step(K)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if graph > 0:                            # separate plain-launch pass for the kernel duration
        stepper.use_graph(0)
        stepper.profile(K)
        stepper.step(K)
    kern_ms, launches = stepper.profile_read()
    stepper.profile(0)
    return t1 - t0, kern_ms, launches


def l2_vs_f64(members32, members64):
    """max over solves and tracers of ||q32 - q64|| / ||q64|| (interiors)."""
    import numpy as np
    worst = 0.0
    for a, b in zip(members32, members64):
        qa, qb = a.state(), b.state()
        for n in range(qa.shape[0]):
            worst = max(worst, float(np.linalg.norm(qa[n] - qb[n]) / np.linalg.norm(qb[n])))
    return worst


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="two_hills_n384_f64", choices=sorted(WORKLOADS))
    ap.add_argument("--tile", type=int, default=None, help="tile id (default: tuned for the workload)")
    ap.add_argument("--graph", type=int, default=16,
                    help="replay the timed steps from HIP graphs of this many steps (0 = plain launches); "
                         "the step-kernel duration then comes from a separate plain-launch pass")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-l2", action="store_true", help="skip the FP64 comparison run of FP32 workloads")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="stream a buffer of this many MiB through one elementwise kernel before the timed "
                         "steps (a kernel of known traffic for scaling hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every solve of this workload fits one MI355X and "
              "nothing here uses a second GPU (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    from cubedsphere.autotune import tile_rows, tune_batch
    from cubedsphere.batch import Batch, make_members
    from cubedsphere.cases import workload

    kind, N, case_name, dtype, cfl, default_steps = WORKLOADS[args.workload]
    opts = WORKLOAD_OPTS.get(args.workload, {})
    general = is_general(opts)
    if kind == "served":
        print(json.dumps(run_served(args)))
        return
    K = args.steps if args.steps is not None else default_steps
    W = args.warmup
    nstep_total = W + 2 * K + max(args.graph, 0)
    tuned = None
    if kind == "single":
        specs = [(N, workload(case_name), nstep_total)]
        tile = args.tile if args.tile is not None else TUNED_TILE.get((dtype, N), 0)
    else:
        specs = batch_specs(args.seed, nstep_total)
        if args.tile is not None:
            tile = args.tile
        else:
            tile, timings = tune_batch(specs, dtype=dtype, candidates=BATCH_TILE_CANDIDATES, steps=16, warmup=8,
                                       device_grid=True)
            tuned = {str(t): round(v * 1e6, 1) for t, v in timings.items()}
    tiles = {s[0]: tile for s in specs}
    # many solves: geometry, fields and winds prepared on the device
    device_grid = kind == "batched"
    members = make_members(specs, dtype=dtype, cfl=cfl, tiles=tiles, device_grid=device_grid, **opts)
    stepper = Batch(members, graph_steps=0).init()
    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)          # one elementwise kernel: reads and writes n * 8 bytes
        torch.cuda.synchronize()
        del src, dst

    elapsed, kern_ms, launches = timed_run(stepper, K, W, args.graph)
    per_step = stepper.cell_updates_per_step()
    value = per_step * K / elapsed
    bytes_per_launch = sum(algorithmic_bytes_per_update(dtype, m.case.ntr) * 6.0 * m.grid.N ** 2 * m.case.ntr
                           for m in members)
    b_alg = bytes_per_launch / per_step
    in_step_s = kern_ms / 1e3 / max(launches, 1)       # event pair around each launch inside the step
    burst_reps = 200
    if general:
        # a sequence of simple kernels per step: the roofline is taken over the whole step
        kern_s = elapsed / K
        kernel, timing = "general step (its whole kernel sequence)", "wall time of the timed steps"
    else:
        kern_s = stepper.time_step_kernel(burst_reps) * 1e-6   # one event pair around a graph of launches
        kernel = "k_step"
        timing = ("HIP events on the launch stream around one graph of %d back-to-back step-kernel "
                  "launches (each redoes the same step)" % burst_reps)
    achieved = bytes_per_launch / kern_s / 1e9
    Ns = sorted({m.grid.N for m in members})

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
        "data": "synthetic (analytic initial conditions and winds on the generated grids)",
        "config": {"workload": args.workload, "solves": len(members), "N": Ns if len(Ns) > 1 else Ns[0],
                   "cells_per_step": per_step, "tracers": sum(m.case.ntr for m in members), "cfl": cfl,
                   "tile": tile, "tile_rows": tile_rows(tile), "graph_steps": args.graph,
                   "options": options_string(opts)},
        "hbm_frac_whole_step": value * b_alg / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if general else traffic_from_profile(args.workload),
            "kernel": kernel,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "mean_launch_us": kern_s * 1e6,
            "launches_timed": K if general else burst_reps,
            "timing": timing,
            "mean_launch_us_in_step": None if general else in_step_s * 1e6,
        },
    }
    if tuned is not None:
        out["config"]["autotune_us_per_step"] = tuned
    if dtype == "f32" and not args.no_l2:
        # the same solves in FP64 for as many steps as the FP32 run took
        done = members[0].steps_done
        m64 = make_members([(s[0], s[1], done) for s in specs], dtype="f64", cfl=cfl, tiles=tiles,
                           device_grid=device_grid)
        b64 = Batch(m64, graph_steps=16).init()
        b64.step(done)
        out["l2_vs_f64"] = l2_vs_f64(members, m64)
        out["l2_vs_f64_steps"] = done
        b64.close()
    if not args.no_cpu_baseline:
        m0 = members[0]
        out["cpu_baseline"] = cpu_baseline(m0.grid.N, m0.case, m0.dt, args.cpu_seconds, opts)
    print(json.dumps(out))


def run_served(args):
    """Served form: a seeded Poisson stream of requests through the continuous-batching
    server (N log-uniform in [48, 1024], 1-4 tracers, 50-2000 steps, Gaussian or cap
    fields, tilted solid-body or deformational wind, FP32 with 10% FP64 requests)."""
    from cubedsphere.serve import Server, random_requests
    reqs = random_requests(args.seed, SERVED_REQUESTS, SERVED_RATE)
    Server(chunk=16).run(random_requests(args.seed + 1, 4, 1000.0, steps=(20, 40)))   # warm-up
    srv = Server(chunk=16, budget_bytes=96 << 30, max_slots=64)
    srv.run(reqs)
    st = srv.stats
    return {
        "metric": "tracer cell-updates/s",
        "value": st["cell_updates_per_s"],
        "unit": "tracer cell-updates/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 4,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Poisson request stream of analytic cases)",
        "config": {"workload": args.workload, "requests": len(reqs), "rate_per_s": SERVED_RATE,
                   "chunk_steps": 16, "budget_bytes": 96 << 30, "max_slots": 64,
                   "f64_requests": sum(r.dtype == "f64" for r in reqs)},
        "p50_latency_s": st["p50_latency_s"],
        "p99_latency_s": st["p99_latency_s"],
        "served": st,
        "roofline": None,
        "roofline_note": "served rounds replay graphs; the step kernel's roofline is measured by the "
                         "single-solve and batched workloads",
    }


if __name__ == "__main__":
    main()
