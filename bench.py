#!/usr/bin/env python
"""Benchmark of the cubed-sphere advection step on one MI355X.

With no --workload, prints ONE JSON line for the plain FP64 GPU run (its `value` is the
headline) and, under "forms", the precision, batched and served forms measured in the same
run.  With --workload, prints the line of that workload alone.

Workloads:
  two_hills_n384_f64   the plain FP64 GPU run: equiangular cubed sphere N=384 (6x384x384
                       cells), one tracer of two Gaussian hills (b=5) at lon 5pi/6 and 7pi/6 on
                       the equator, solid-body rotation about an axis tilted pi/4 from the pole,
                       CFL 0.5, 1000 steps
  nl10_n768_f32        precision form: N=768, 4 tracers (Gaussian hills, cosine bells, slotted
                       cylinders, their mix) in the Nair-Lauritzen deformational wind (k=2.4,
                       T=5), FP32 state and fluxes, FP64 tables and mass
  batched_64_f32       batched form: 64 seeded solves, N in {32,48,96,192,384,512}, 1-8 tracers,
                       random Gaussian centre / width / rotation-axis tilt, FP32, one launch per
                       kernel per step over all solves, graph-replayed, tiles from the autotuner
  served_poisson_f32   served form: a seeded Poisson stream of 200 requests at 12/s, N
                       log-uniform in [48, 1024], 1-4 tracers, 50-2000 steps, FP32 (10% FP64)
  polar_n48_f64        the reference's CPU-sized case (N=48 polar rotation)
  two_hills_n384_f64_etpl07, two_hills_n384_f64_rk2_mfaf
                       the N=384 FP64 case with schemes 1 and 3 of the reference's convergence
                       study (ET-PL07 with the PL07 split: the general step; RK2 with MF-AF)

A "step" is one full advection step of every cell of every solve: ghost fill, both
directional passes of the split, update and mass fixer (plus the wind update of
time-dependent winds).

Fields of a line:
  value        tracer cell-updates/s over the timed steps (inputs resident in HBM)
  roofline     the step kernel against the HBM roofline: algorithmic bytes per launch over its
               mean duration -- HIP events on the launch stream around one graph of 200
               back-to-back launches (the state cache-resident, as in the stepping itself,
               whose kernel-trace durations agree); achieved_cold / frac_cold: the same over a
               graph rotating through enough copies of the solve to exceed the 256 MB
               last-level cache (every launch streams from HBM); frac_in_step: from event pairs
               around each launch inside the timed step sequence (they also hold the launch)
  cpu_baseline the numpy FP64 oracle on this host on a bounded sample of the same workload,
               one core (plain form) or a process pool over the host's cores (batched and
               served forms), with the host's CPU model and core counts
  l2_vs_f64    (FP32 workloads) relative L2 difference of the final states from an FP64 GPU
               run of the same solves and steps
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
MALL_BYTES = 256 << 20  # MI355X last-level (Infinity) cache

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
# the two extra N=384 workloads are schemes 1 and 3 of src/advection_error.py:61-66
WORKLOAD_OPTS = {
    "two_hills_n384_f64_etpl07": dict(opsplit=3, et=2, mt=2, mf=1),       # PPM-PL07, SP-PL07, ET-PL07, MT-PL07, MF-0
    "two_hills_n384_f64_rk2_mfaf": dict(dp=2, opsplit=1, et=3, mt=1, mf=2),  # RK2, SP-AVLT, ET-DG, MT-0, MF-AF
}
OPTION_NAMES = {
    "recon": ("PPM-0", "PPM-CW84", "PPM-PL07", "PPM-L04"), "dp": ("RK1", "RK2"),
    "opsplit": ("AVLT", "L04", "PL07"), "et": ("ET-S72", "ET-PL07", "ET-DG"), "mt": ("MT-0", "MT-PL07"),
    "mf": ("MF-0", "MF-AF", "MF-PR"),
}
FORMS = ("nl10_n768_f32", "batched_64_f32", "served_poisson_f32")
SERVED_REQUESTS, SERVED_RATE = 200, 12.0      # Poisson stream: count, arrivals per second
BATCH_NS = (32, 48, 96, 192, 384, 512)
# tile ids from sweeps of every tile shape for the single-solve workloads (the batched workload
# tunes a tile per N at start-up over every tile shape)
TUNED_TILE = {"two_hills_n384_f64": 2, "nl10_n768_f32": 3, "polar_n48_f64": 1,
              "two_hills_n384_f64_etpl07": 0, "two_hills_n384_f64_rk2_mfaf": 2}
POOL_MAX = 16           # the host cores one GPU's share of the box holds


def options_string(opts):
    o = dict(recon=3, dp=1, opsplit=1, et=3, mt=1, mf=3)
    o.update(opts)
    n = {k: OPTION_NAMES[k][o[k] - 1] for k in OPTION_NAMES}
    return "recon %s, dp %s, split %s, %s, %s, %s" % (n["recon"], n["dp"], n["opsplit"], n["et"], n["mt"], n["mf"])


def is_general(opts):
    """The general step runs the options the fused step kernel does not: ET-S72 and ET-PL07."""
    return opts.get("et", 3) != 3


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


# ------------------------------------------------------------------ CPU baselines
def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"host_cpu": model, "host_cpu_count": os.cpu_count(), "host_cores_usable": usable}


def _oracle_rate(job):
    """Worker: time the numpy FP64 oracle on one solve (N, case, dt, opts) for about ``seconds``
    after one warm-up step; returns (cell-updates, seconds, steps)."""
    N, case, dt, opts, seconds = job
    from oracle import timestep as ots
    from oracle.cases import from_case
    grid, sim = ots.make(N, from_case(case), dt, **(opts or {}))
    ots.init_vars_adv(grid, sim)
    ots.adv_time_step(grid, sim, 1, dt)
    ots.update_adv(grid, sim, dt)
    steps, t0 = 0, time.perf_counter()
    while True:
        k = steps + 2
        ots.adv_time_step(grid, sim, k, k * dt)
        ots.update_adv(grid, sim, k * dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 1000:
            break
    return 6.0 * N * N * case.ntr * steps, el, steps


def cpu_baseline(N, case, dt, seconds, opts=None):
    """The oracle (numpy FP64, single thread) on one solve of the workload."""
    upd, el, steps = _oracle_rate((N, case, dt, opts, seconds))
    out = {"value": upd / el, "unit": "tracer cell-updates/s", "cores": 1, "kind": "port",
           "sample": "numpy FP64 oracle, N=%d, %d tracer(s), %d timed steps after 1 warm-up (%.1f s, %.3f s/step)"
                     % (N, case.ntr, steps, el, el / steps)}
    out.update(host_info())
    return out


def cpu_pool_baseline(jobs, what):
    """The oracle on a process pool over the host's usable cores (at most POOL_MAX), one solve
    per process, all running together: the aggregate is the sum of the processes' rates.  Runs
    before anything initialises the GPU (the pool forks)."""
    import multiprocessing as mp
    info = host_info()
    workers = max(1, min(POOL_MAX, info["host_cores_usable"] or 1, len(jobs)))
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_oracle_rate, jobs[:workers])
    rate = sum(u / e for u, e, _ in res)
    out = {"value": rate, "unit": "tracer cell-updates/s", "cores": workers, "kind": "port",
           "sample": "numpy FP64 oracle, %d processes each stepping one solve of %s for about %.0f s "
                     "(N %s, %d tracers in all, %d steps in all)"
                     % (workers, what, jobs[0][4], sorted({j[0] for j in jobs[:workers]}),
                        sum(j[1].ntr for j in jobs[:workers]), sum(s for _, _, s in res))}
    out.update(info)
    return out


def batch_specs(seed, steps):
    """The 64 seeded solves of the batched form: (N, case, steps)."""
    from cubedsphere.batch import seeded_batch
    return seeded_batch(seed, steps, BATCH_NS)


_GRIDS = {}


def _cfl_dt(N, case, cfl=0.5):
    from cubedsphere.grid import Grid
    from cubedsphere.solver import cfl_dt
    if N not in _GRIDS:
        _GRIDS[N] = Grid(N)
    return cfl_dt(_GRIDS[N], case, cfl)


def cpu_jobs_batched(seed, seconds, count=POOL_MAX):
    """The first ``count`` solves of the seeded batch (N as drawn), for the pool baseline."""
    return [(N, case, _cfl_dt(N, case), {}, seconds) for N, case, _ in batch_specs(seed, 1)[:count]]


def cpu_jobs_served(seed, seconds, nmax=256, count=POOL_MAX):
    """The first ``count`` requests of the served stream with N <= nmax (the oracle's set-up at
    N ~ 1024 alone takes minutes), in arrival order."""
    from cubedsphere.serve import random_requests
    reqs = [r for r in random_requests(seed, SERVED_REQUESTS, SERVED_RATE) if r.N <= nmax][:count]
    return [(r.N, r.case, _cfl_dt(r.N, r.case), {}, seconds) for r in reqs]


# ------------------------------------------------------------------ GPU workloads
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


def cold_timing(spec, dtype, cfl, tile, opts, state_bytes):
    """Mean step-kernel microseconds over a graph rotating through enough copies of one solve
    (own state and wind buffers each, shared tables) that their states exceed the last-level
    cache; None for the general step."""
    import torch
    from cubedsphere.batch import make_member
    from cubedsphere.solver import time_step_kernel_rotating
    copies = max(3, int(2 * MALL_BYTES // max(state_bytes, 1)) + 1)
    copies = min(copies, 16)
    solves = []
    try:
        for _ in range(copies):
            m = make_member(spec[0], spec[1], 64, dtype=dtype, cfl=cfl, tile=tile, **opts)
            m.init_vars_adv()
            m.step(1)
            solves.append(m)
        torch.cuda.synchronize()
        us = time_step_kernel_rotating(solves, reps=max(2, 48 // copies))
    finally:
        for m in solves:
            m.close()
    return us, copies


def run_workload(name, K=None, W=10, graph=16, tile=None, seed=2024, cpu_seconds=12.0, cpu=True, l2=True,
                 calibrate_copy_mib=0, cpu_result=None):
    """One workload's JSON object (the GPU part; the CPU baseline from ``cpu_result`` when given,
    else measured here when ``cpu``)."""
    import torch

    from cubedsphere.autotune import tile_ids, tile_rows, tune_batch_per_n
    from cubedsphere.batch import Batch, make_members
    from cubedsphere.cases import workload

    kind, N, case_name, dtype, cfl, default_steps = WORKLOADS[name]
    opts = WORKLOAD_OPTS.get(name, {})
    general = is_general(opts)
    if kind == "served":
        return run_served(seed, cpu_result)
    K = K if K is not None else default_steps
    nstep_total = W + 2 * K + max(graph, 0)
    tuned = None
    if kind == "single":
        specs = [(N, workload(case_name), nstep_total)]
        tile = tile if tile is not None else TUNED_TILE.get(name, 0)
    else:
        specs = batch_specs(seed, nstep_total)
        if tile is None:
            tiles, timings = tune_batch_per_n(specs, dtype=dtype, candidates=tile_ids(), steps=16, warmup=8,
                                              device_grid=True)
            tuned = {"tiles": {str(n): t for n, t in sorted(tiles.items())},
                     "trials_us_per_step": {k: round(v * 1e6, 1) for k, v in timings.items()}}
        else:
            tiles = {s[0]: tile for s in specs}
    if kind == "single":
        tiles = {s[0]: tile for s in specs}
    device_grid = kind == "batched"       # many solves: geometry, fields and winds prepared on the device
    members = make_members(specs, dtype=dtype, cfl=cfl, tiles=tiles, device_grid=device_grid, **opts)
    stepper = Batch(members, graph_steps=0).init()
    if calibrate_copy_mib > 0:
        n = calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)          # one elementwise kernel: reads and writes n * 8 bytes
        torch.cuda.synchronize()
        del src, dst

    elapsed, kern_ms, launches = timed_run(stepper, K, W, graph)
    per_step = stepper.cell_updates_per_step()
    value = per_step * K / elapsed
    bytes_per_launch = sum(algorithmic_bytes_per_update(dtype, m.case.ntr) * 6.0 * m.grid.N ** 2 * m.case.ntr
                           for m in members)
    b_alg = bytes_per_launch / per_step
    in_step_s = kern_ms / 1e3 / max(launches, 1)       # event pair around each launch inside the step
    burst_reps = 200
    cold_us = cold_copies = None
    if general:
        # a sequence of simple kernels per step: the roofline is taken over the whole step
        kern_s = elapsed / K
        kernel, timing = "general step (its whole kernel sequence)", "wall time of the timed steps"
    else:
        kern_s = stepper.time_step_kernel(burst_reps) * 1e-6   # one event pair around a graph of launches
        kernel = "k_step"
        timing = ("HIP events on the launch stream around one graph of %d back-to-back step-kernel "
                  "launches (each redoes the same step; the state cache-resident, as in the stepping)" % burst_reps)
        if kind == "single":
            m = members[0]
            state_bytes = sum(t.numel() * t.element_size() for t in list(m.q) + [m.uavg, m.vavg])
            cold_us, cold_copies = cold_timing(specs[0], dtype, cfl, tile, opts, state_bytes)
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
        "config": {"workload": name, "solves": len(members), "N": Ns if len(Ns) > 1 else Ns[0],
                   "cells_per_step": per_step, "tracers": sum(m.case.ntr for m in members), "cfl": cfl,
                   "tile": tile if kind == "single" else {str(n): t for n, t in sorted(tiles.items())},
                   "tile_rows": tile_rows(tile) if kind == "single" else
                   {str(n): tile_rows(t) for n, t in sorted(tiles.items())}, "graph_steps": graph,
                   "options": options_string(opts)},
        "hbm_frac_whole_step": value * b_alg / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if general else traffic_from_profile(name),
            "kernel": kernel,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "mean_launch_us": kern_s * 1e6,
            "launches_timed": K if general else burst_reps,
            "timing": timing,
            "mean_launch_us_in_step": None if general else in_step_s * 1e6,
            "frac_in_step": None if general else bytes_per_launch / in_step_s / 1e9 / HBM_PEAK_GBS,
        },
    }
    if cold_us is not None:
        out["roofline"].update({
            "achieved_cold": bytes_per_launch / (cold_us * 1e-6) / 1e9,
            "frac_cold": bytes_per_launch / (cold_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "mean_launch_us_cold": cold_us,
            "timing_cold": "one graph rotating through %d copies of the solve (states %.0f MB in all, above the "
                           "256 MB last-level cache)" % (cold_copies, cold_copies * state_bytes / 1e6),
        })
    if tuned is not None:
        out["config"]["autotune_us_per_step"] = tuned
    if dtype == "f32" and l2:
        # the same solves in FP64 for as many steps as the FP32 run took
        done = members[0].steps_done
        m64 = make_members([(s[0], s[1], done) for s in specs], dtype="f64", cfl=cfl, tiles=tiles,
                           device_grid=device_grid, **opts)
        b64 = Batch(m64, graph_steps=16).init()
        b64.step(done)
        out["l2_vs_f64"] = l2_vs_f64(members, m64)
        out["l2_vs_f64_steps"] = done
        b64.close()
    stepper.close()
    if cpu_result is not None:
        out["cpu_baseline"] = cpu_result
    elif cpu:
        m0 = members[0]
        out["cpu_baseline"] = cpu_baseline(m0.grid.N, m0.case, m0.dt, cpu_seconds, opts)
    return out


def run_served(seed, cpu_result=None):
    """Served form: a seeded Poisson stream of requests through the continuous-batching
    server (N log-uniform in [48, 1024], 1-4 tracers, 50-2000 steps, Gaussian or cap
    fields, tilted solid-body or deformational wind, FP32 with 10% FP64 requests)."""
    from cubedsphere.serve import Server, random_requests
    reqs = random_requests(seed, SERVED_REQUESTS, SERVED_RATE)
    Server(chunk=16).run(random_requests(seed + 1, 4, 1000.0, steps=(20, 40)))   # warm-up
    srv = Server(chunk=16, budget_bytes=96 << 30, max_slots=64)
    srv.run(reqs)
    st = srv.stats
    out = {
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
        "config": {"workload": "served_poisson_f32", "requests": len(reqs), "rate_per_s": SERVED_RATE,
                   "chunk_steps": 16, "budget_bytes": 96 << 30, "max_slots": 64,
                   "f64_requests": sum(r.dtype == "f64" for r in reqs)},
        "p50_latency_s": st["p50_latency_s"],
        "p99_latency_s": st["p99_latency_s"],
        "served": st,
        "roofline": None,
        "roofline_note": "served rounds replay graphs of the same step kernels; their roofline is measured by the "
                         "single-solve and batched workloads",
    }
    if cpu_result is not None:
        out["cpu_baseline"] = cpu_result
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=None, choices=sorted(WORKLOADS),
                    help="one workload alone (default: the plain FP64 run with the three forms)")
    ap.add_argument("--tile", type=int, default=None, help="tile id (default: tuned for the workload)")
    ap.add_argument("--graph", type=int, default=16,
                    help="replay the timed steps from HIP graphs of this many steps (0 = plain launches); "
                         "the step-kernel duration then comes from a separate plain-launch pass")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="the plain FP64 line without the forms")
    ap.add_argument("--no-l2", action="store_true", help="skip the FP64 comparison run of FP32 workloads")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="stream a buffer of this many MiB through one elementwise kernel before the timed "
                         "steps (a kernel of known traffic for scaling hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every solve of this workload fits one MI355X and "
              "nothing here uses a second GPU (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    cpu = not args.no_cpu_baseline
    kw = dict(K=args.steps, W=args.warmup, graph=args.graph, tile=args.tile, seed=args.seed,
              cpu_seconds=args.cpu_seconds, l2=not args.no_l2, calibrate_copy_mib=args.calibrate_copy_mib)
    if args.workload is not None:
        cpu_result = None
        if cpu and args.workload in ("batched_64_f32", "served_poisson_f32"):
            jobs = (cpu_jobs_batched(args.seed, args.cpu_seconds) if args.workload == "batched_64_f32"
                    else cpu_jobs_served(args.seed, args.cpu_seconds))
            cpu_result = cpu_pool_baseline(jobs, args.workload)
        print(json.dumps(run_workload(args.workload, cpu=cpu, cpu_result=cpu_result, **kw)))
        return

    # the headline line: the plain FP64 run, with the three forms under "forms".  The CPU pools
    # fork, so they run before anything initialises the GPU.
    pools = {}
    if cpu and not args.no_forms:
        pools["batched_64_f32"] = cpu_pool_baseline(cpu_jobs_batched(args.seed, 10.0), "batched_64_f32")
        pools["served_poisson_f32"] = cpu_pool_baseline(cpu_jobs_served(args.seed, 10.0), "served_poisson_f32")
    out = run_workload("two_hills_n384_f64", cpu=cpu, **kw)
    if not args.no_forms:
        forms = {}
        for name in FORMS:
            fkw = dict(kw, K=None, tile=None)
            cpu_result = pools.get(name)
            if name == "nl10_n768_f32" and cpu:
                # a bounded sample: the first tracer alone (the oracle steps tracers one by one)
                from cubedsphere.cases import Case, workload
                full = workload("nl10_deformation")
                case = Case(full.tracers[:1], full.wind, "nl10_hills")
                cpu_result = cpu_baseline(768, case, _cfl_dt(768, full), 4.0)
            forms[name] = run_workload(name, cpu=False, cpu_result=cpu_result, **fkw)
        out["forms"] = forms
    print(json.dumps(out))


if __name__ == "__main__":
    main()
