"""Autotuner of the step kernel's tile shape.

A tile id selects how many rows one wavefront marches down a column strip
(``csadv_tile_rows``).  Short segments give more wavefronts (parallelism) but recompute
the 3-row halo more often; long segments waste less but may leave the chip idle.  The
best shape depends on N, the tracer count, the dtype and on what else shares the launch,
so it is measured: each candidate runs a few warm-up steps and then a timed block of
steps replayed from HIP graphs, and the fastest wins.

The reference has no counterpart (it has no kernels to tune).
"""
import json
import os
import time

import numpy as np
import torch

from . import _lib
from .batch import Batch, make_members
from .cases import random_solid_body


def tile_ids():
    lib = _lib.load()
    ids = []
    while lib.csadv_tile_rows(len(ids)) > 0:
        ids.append(len(ids))
    return ids


def tile_rows(tile):
    return _lib.load().csadv_tile_rows(int(tile))


def _time_steps(stepper, steps, warmup):
    stepper.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stepper.step(steps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def tune_batch(specs, dtype="f32", candidates=None, steps=32, warmup=16, device="cuda:0", graph_steps=16, **opts):
    """Time one uniform tile for every member of the batch ``specs`` = [(N, case, nsteps)]
    per candidate; returns (best tile, {tile: seconds per step})."""
    cands = list(candidates) if candidates is not None else tile_ids()
    need = warmup + steps
    specs = [(N, case, max(n, need)) for N, case, n in specs]
    timings = {}
    for t in cands:
        members = make_members(specs, dtype=dtype, device=device, tiles={N: t for N, _, _ in specs}, **opts)
        b = Batch(members, graph_steps=graph_steps).init()
        timings[t] = _time_steps(b, steps, warmup)
        b.close()
        del members
    best = min(timings, key=timings.get)
    return best, timings


def tune_batch_per_n(specs, dtype="f32", candidates=None, steps=32, warmup=16, device="cuda:0", graph_steps=16,
                     gain=0.01, **opts):
    """A tile per N for the batch ``specs`` = [(N, case, nsteps)]: the best uniform tile first,
    then a coordinate descent over the sizes (most cells first), each trying every other
    candidate for its solves with the others fixed and keeping a change that is faster by more
    than ``gain``.  One batch is built; tiles change through the descriptors.  Returns
    ({N: tile}, {trial: seconds per step})."""
    cands = list(candidates) if candidates is not None else tile_ids()
    need = warmup + steps
    Ns = sorted({N for N, _, _ in specs})
    members = make_members([(N, case, max(n, need)) for N, case, n in specs], dtype=dtype, device=device,
                           tiles={N: cands[0] for N in Ns}, **opts)
    b = Batch(members, graph_steps=graph_steps).init()
    timings = {}

    def measure(tiles):
        b.set_tiles(tiles)
        sec = _time_steps(b, steps, warmup)
        timings[",".join("%d:%d" % (N, tiles[N]) for N in Ns)] = sec
        return sec

    uniform = {t: measure({N: t for N in Ns}) for t in cands}
    t0 = min(uniform, key=uniform.get)
    best, cur = {N: t0 for N in Ns}, uniform[t0]
    work = {N: sum(6 * N * N * case.ntr for n_, case, _ in specs if n_ == N) for N in Ns}
    for N in sorted(Ns, key=lambda n: -work[n]):
        for t in cands:
            if t == best[N]:
                continue
            trial = dict(best)
            trial[N] = t
            sec = measure(trial)
            if sec < cur * (1.0 - gain):
                best, cur = trial, sec
    b.close()
    del members
    return best, timings


def tune_solve(N, dtype="f32", ntr=1, candidates=None, steps=32, warmup=16, device="cuda:0", seed=0, **opts):
    """Best tile of one solve of size N (a seeded random solid-body case with ntr tracers)."""
    case = random_solid_body(np.random.default_rng(seed), ntr)
    return tune_batch([(N, case, steps + warmup)], dtype=dtype, candidates=candidates, steps=steps, warmup=warmup,
                      device=device, **opts)


class TileTable:
    """Tuned tile per (dtype, N), kept in memory and optionally in a JSON file."""

    def __init__(self, path=None):
        self.path = path
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = {k: int(v) for k, v in json.load(f).items()}

    @staticmethod
    def key(dtype, N):
        return "%s:%d" % (dtype, N)

    def get(self, dtype, N, default=None):
        return self.table.get(self.key(dtype, N), default)

    def tune(self, Ns, dtype="f32", **kw):
        """Tune every N not yet in the table; returns {N: tile}."""
        for N in sorted(set(Ns)):
            if self.get(dtype, N) is None:
                best, _ = tune_solve(N, dtype=dtype, **kw)
                self.table[self.key(dtype, N)] = best
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
        return {N: self.get(dtype, N) for N in Ns}
