"""Served form: a request loop with continuous batching on one GPU.

Requests (a solve: N, tracers, initial fields, wind, number of steps, dtype) arrive over
time.  The server admits them in arrival order while the device memory they need fits the
HBM budget and a batch slot is free, keeps their state resident on the device, and advances
every admitted solve together: each round is one graph-replayed chunk of steps over one
handle per (dtype) group, so requests that arrive while others run join the next round and
their work shares the same kernel launches.  A solve stops at its own step count inside a
chunk (the kernels skip finished solves), its result is read back (folding only its own
pending mass-fixer correction), its memory is released and the freed budget admits the
next request.  A solve's arithmetic does not depend on its companions, so every result is
bitwise the one it would have alone (same tile rule per N).

Latency of a request = completion time (result on the host) - arrival time.  The reference
has no server; each of its runs is one solve in one process (src/advection_sphere.py).
"""
import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .batch import Batch, make_member
from .cases import Case, PI


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    N: int
    case: Case
    nsteps: int
    dtype: str = "f32"


@dataclass
class Completed:
    rid: int
    arrival: float
    admitted: float
    done: float
    cell_updates: float
    result: np.ndarray = field(repr=False)     # [ntr][6][N][N] in the request's precision
    dt: float = 0.0         # time step the solve ran with (CFL 0.5 at t = 0)

    @property
    def latency(self):
        return self.done - self.arrival


def random_case(rng, ntr):
    """Gaussian or cosine-cap tracers at random centres; tilted solid-body rotation or the
    Nair-Lauritzen deformational wind (k = 2.4, T = 5)."""
    tracers = []
    for _ in range(ntr):
        lon0, lat0 = rng.uniform(-PI, PI), math.asin(rng.uniform(-1.0, 1.0))
        if rng.random() < 0.5:
            tracers.append({"field": "gaussians", "centres": [(lon0, lat0)], "b": rng.uniform(5.0, 20.0)})
        else:
            tracers.append({"field": "cosine_bells", "centres": [(lon0, lat0)], "radius": 0.5})
    if rng.random() < 0.5:
        wind = ("solid_body", -rng.uniform(0.0, PI / 2.0), 5.0)
    else:
        wind = ("deformational", 2.4, 5.0)
    return Case(tracers, wind, "served")


def random_requests(seed, count, rate, n_range=(48, 1024), tracers=(1, 4), steps=(50, 2000), f64_fraction=0.1):
    """Seeded Poisson stream: exponential inter-arrival times of mean 1/rate seconds, N
    log-uniform in n_range, tracer count and step count uniform, FP64 on request with
    probability f64_fraction (FP32 otherwise)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        N = int(round(math.exp(rng.uniform(lo, hi))))
        ntr = int(rng.integers(tracers[0], tracers[1] + 1))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        dtype = "f64" if rng.random() < f64_fraction else "f32"
        out.append(Request(rid, t, N, random_case(rng, ntr), nsteps, dtype))
    return out


def default_tile(N):
    """Tile id per N (rows per wavefront segment grow with N: short segments keep small
    solves parallel, long ones waste less halo work on large solves)."""
    if N < 96:
        return 1       # 4 rows
    if N < 256:
        return 0       # 8 rows
    return 2           # 16 rows


class Server:
    """Continuous-batching request loop.

    budget_bytes  device memory the admitted solves' own buffers may hold together
    max_slots     solves admitted at once
    chunk         steps per round (one captured graph)
    tile_of       N -> tile id (fixed per N, so results do not depend on the mix)
    """

    def __init__(self, budget_bytes=64 << 30, max_slots=64, chunk=16, tile_of=default_tile, cfl=0.5,
                 device="cuda:0", clock=time.perf_counter, device_grid=True):
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.chunk = int(chunk)
        self.tile_of = tile_of
        self.cfl = cfl
        self.device = device
        self.clock = clock
        self.device_grid = device_grid  # prepare geometry, fields and winds on the device
        self.groups = {}            # dtype -> Batch
        self.used = 0
        self.stats = {}

    # ------------------------------------------------------------ admission
    def _build(self, req):
        m = make_member(req.N, req.case, req.nsteps, dtype=req.dtype, cfl=self.cfl, device=self.device,
                        tile=self.tile_of(req.N), device_grid=self.device_grid)
        m.stop = req.nsteps
        m.request = req
        return m

    @staticmethod
    def estimate_bytes(req):
        """Device bytes a request's own buffers take (before the wind scratch is released)."""
        s = 8 if req.dtype == "f64" else 4
        M = req.N + 8
        ld, ldu = (M + 3) // 4 * 4, (M + 4) // 4 * 4
        K = len(req.case.basis())
        ntr = req.case.ntr
        cells = 6 * M * ld
        b = 2 * ntr * cells * s + 6 * M * ldu * s + 6 * (M + 1) * ld * s          # state, winds
        b += K * (6 * M * ldu + 6 * (M + 1) * ld) * 8 + K * 2 * cells * 8       # basis, scratch
        b += req.nsteps * 2 * K * 8 + ntr * 6 * req.N * req.N // 16 * 8          # coefficients, partials
        return b

    # ------------------------------------------------------------ loop
    def run(self, requests, on_complete=None):
        """Serve ``requests`` (arrival times relative to the call); returns {rid: Completed}."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        done = {}
        active = {}                 # dtype -> list of members
        fresh = {}                  # dtype -> members admitted since the last round
        t0 = self.clock()
        busy, rounds, peak_slots, peak_bytes = 0.0, 0, 0, 0
        while pending or queue or any(active.values()):
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # admission in arrival order, by budget and slots
            nactive = sum(len(v) for v in active.values())
            while queue and nactive < self.max_slots:
                req = queue[0]
                need = self.estimate_bytes(req)
                if self.used + need > self.budget and nactive > 0:
                    break
                queue.popleft()
                m = self._build(req)
                m.init_buffers()
                m.admitted = self.clock() - t0
                m.bytes = need
                self.used += need
                active.setdefault(req.dtype, []).append(m)
                fresh.setdefault(req.dtype, []).append(m)
                nactive += 1
            peak_slots = max(peak_slots, nactive)
            peak_bytes = max(peak_bytes, self.used)
            if nactive == 0:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.01)))
                continue
            # one round: rebuild handles whose membership changed, then one chunk each
            tb = self.clock()
            for dt, members in active.items():
                if not members:
                    continue
                new = fresh.pop(dt, [])
                b = self.groups.get(dt)
                if b is None:
                    b = self.groups[dt] = _empty_batch(members[0], self.chunk)
                if new or b.members != members:
                    b.rebuild(members, fresh=new)
                b.step(self.chunk)
            torch.cuda.synchronize(self.device)
            rounds += 1
            # completions
            for dt, members in active.items():
                finished = [m for m in members if m.steps_done >= m.stop]
                if not finished:
                    continue
                for m in finished:
                    res = m.state(native=True)
                    t = self.clock() - t0
                    req = m.request
                    c = Completed(req.rid, req.arrival, m.admitted, t, 6.0 * req.N ** 2 * req.case.ntr * req.nsteps, res,
                                  m.dt)
                    done[req.rid] = c
                    if on_complete:
                        on_complete(c)
                    self.used -= m.bytes
                keep = [m for m in members if m.steps_done < m.stop]
                active[dt] = keep
                b = self.groups[dt]
                b.rebuild(keep)
                for m in finished:
                    m.close()
                    m.q = None
            busy += self.clock() - tb
        for b in self.groups.values():
            b.close()
        self.groups = {}
        wall = self.clock() - t0
        lat = np.array([c.latency for c in done.values()])
        self.stats = {
            "requests": len(done),
            "wall_s": wall,
            "busy_s": busy,
            "rounds": rounds,
            "peak_slots": peak_slots,
            "peak_bytes": peak_bytes,
            "cell_updates": float(sum(c.cell_updates for c in done.values())),
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        }
        self.stats["cell_updates_per_s"] = self.stats["cell_updates"] / wall if wall > 0 else None
        return done


def _empty_batch(first, graph_steps):
    """A Batch object with no handle yet (members are attached by rebuild)."""
    b = Batch.__new__(Batch)
    b.members = []
    b.lib = first.lib
    b.device = first.device
    b.handle = None
    b.graph_steps = int(graph_steps)
    return b


def serve_alone(req, **kw):
    """Result of one request served on its own (reference for the alone-vs-mix check)."""
    r = Request(req.rid, 0.0, req.N, req.case, req.nsteps, req.dtype)
    return Server(**kw).run([r])[req.rid].result
