"""Batched form: many independent solves advanced by one handle.

Every kernel of a step is one launch over all solves of the batch (uneven N and tracer
counts share the launch; each solve keeps its own buffers, tables and tile shape), and
groups of steps are replayed from one captured HIP graph.  A solve's arithmetic does not
depend on its companions: its tile is its own, its mass-fixer sums are its own and reading
one result folds only that solve's pending correction, so a solve gives the same bits
alone and in any batch.

The reference runs one solve per process (src/advection_sphere.py:13-61); this module has
no counterpart there beyond running ``adv_time_step`` for each member.
"""
import ctypes

import torch

from . import _lib
from .solver import Advection, cfl_dt, grid_tables, interior_winds


class Batch:
    """One handle over ``members`` (constructed :class:`Advection` solves, not initialised).

    All members share dtype, recon, split and dp (one kernel instantiation per launch)."""

    def __init__(self, members, graph_steps=16):
        if not members:
            raise ValueError("empty batch")
        first = members[0]
        for m in members:
            if (m.dtype, m.opts["recon"], m.opts["dp"]) != (first.dtype, first.opts["recon"], first.opts["dp"]):
                raise ValueError("a batch shares dtype, recon and dp")
            if m.handle is not None:
                raise ValueError("batch members must not be initialised on their own")
        self.members = list(members)
        self.lib = first.lib
        self.device = first.device
        self.handle = None
        self.graph_steps = int(graph_steps)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def init(self):
        """Initial states of every member and one handle over all of them."""
        for m in self.members:
            m.init_buffers()
        self._open(self.members)
        return self

    def _create(self, members):
        descs = (_lib.SolveDesc * len(members))(*[m.desc() for m in members])
        h = ctypes.c_void_p()
        rc = self.lib.csadv_create(descs, len(members), 0, ctypes.byref(h))
        if rc != 0:
            msg = self.lib.csadv_last_error(h).decode()
            self.lib.csadv_destroy(h)
            raise _lib.CsadvError("csadv_create: %s" % msg)
        return h

    def _open(self, fresh):
        """Prepare the winds of the fresh members (on a handle of their own, so members that
        are already running are not touched), then one handle over every member."""
        if fresh:
            h = self._create(fresh)
            rc = self.lib.csadv_wind_init(h, self._stream())
            msg = self.lib.csadv_last_error(h).decode() if rc else ""
            self.lib.csadv_destroy(h)
            if rc:
                raise _lib.CsadvError("csadv_wind_init: %s" % msg)
            for m in fresh:
                m.release_wind_scratch()
        h = self._create(self.members)
        self.handle = h
        for i, m in enumerate(self.members):
            m.handle, m.index, m.owner = h, i, self
        if self.graph_steps > 0:
            _lib.check(self.lib.csadv_set_graph(h, self.graph_steps), h, "csadv_set_graph")

    def rebuild(self, members, fresh=()):
        """Continue with ``members`` (solves already in the batch keep their state; ``fresh``
        ones must have had init_buffers() and get their winds prepared here)."""
        self._close_handle()
        self.members = list(members)
        if self.members:
            self._open(list(fresh))

    def step(self, nsteps=1):
        grew = False
        for m in self.members:
            todo = nsteps if not m.stop else max(0, min(nsteps, m.stop - m.steps_done))
            grew = m.ensure_steps(m.steps_done + todo, refresh=False) or grew
        if grew:
            self.refresh()
        _lib.check(self.lib.csadv_step(self.handle, int(nsteps), self._stream()), self.handle, "csadv_step")
        for m in self.members:
            m.steps_done += nsteps if not m.stop else max(0, min(nsteps, m.stop - m.steps_done))

    def refresh(self):
        """Hand the members' changed buffer pointers (a grown wind table) to the handle."""
        if not self.handle:
            return
        descs = (_lib.SolveDesc * len(self.members))(*[m.desc() for m in self.members])
        _lib.check(self.lib.csadv_update(self.handle, descs, len(self.members), self._stream()), self.handle,
                   "csadv_update")

    def set_tiles(self, tiles):
        """Tile id per N ({N: tile}) from the next step on: the pending mass-fixer sums are folded
        with the tiles they were written with, then the handle takes the new ones."""
        _lib.check(self.lib.csadv_finalize(self.handle, -1, self._stream()), self.handle, "csadv_finalize")
        for m in self.members:
            m.set_tile(tiles.get(m.grid.N, m.tile))
        self.refresh()

    def cell_updates_per_step(self):
        return sum(6 * m.grid.N ** 2 * m.case.ntr for m in self.members)

    def profile(self, capacity):
        _lib.check(self.lib.csadv_profile(self.handle, int(capacity)), self.handle, "csadv_profile")

    def profile_read(self):
        ms, n = ctypes.c_double(), ctypes.c_int32()
        _lib.check(self.lib.csadv_profile_read(self.handle, ctypes.byref(ms), ctypes.byref(n)), self.handle,
                   "csadv_profile_read")
        return ms.value, n.value

    def time_step_kernel(self, reps=200):
        """Mean microseconds of the step kernel alone over ``reps`` graph-replayed launches
        (the repeats redo the current step, so the state is unchanged by the call)."""
        us = ctypes.c_double()
        _lib.check(self.lib.csadv_time_step_kernel(self.handle, int(reps), self._stream(), ctypes.byref(us)), self.handle,
                   "csadv_time_step_kernel")
        return us.value

    def use_graph(self, steps_per_graph):
        _lib.check(self.lib.csadv_set_graph(self.handle, int(steps_per_graph)), self.handle, "csadv_set_graph")

    def _close_handle(self):
        if self.handle:
            for m in self.members:
                m.handle, m.owner = None, None
            self.lib.csadv_destroy(self.handle)
            self.handle = None

    def close(self):
        self._close_handle()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_member(N, case, nsteps, dtype="f32", cfl=0.5, device="cuda:0", tile=0, device_grid=False, **opts):
    """One Advection solve with dt from the CFL number at t=0 (tables shared per N and dtype).
    device_grid: geometry, initial fields and winds evaluated on the device (FP64 torch)."""
    tab = grid_tables(N, dtype, device, device_grid=device_grid)
    winds = interior_winds(tab.grid, case)
    dt = cfl_dt(tab.grid, case, cfl, winds=winds)
    m = Advection(N, case, dt, dtype=dtype, device=device, tile=tile, max_steps=max(1, nsteps), tables=tab, **opts)
    m.basis_winds = winds
    return m


def make_members(specs, dtype="f32", cfl=0.5, device="cuda:0", tiles=None, device_grid=False, **opts):
    """Advection solves of ``specs`` = [(N, case, nsteps)], tile per N from ``tiles``
    ({N: tile}; default 0)."""
    return [make_member(N, case, nsteps, dtype=dtype, cfl=cfl, device=device, tile=(tiles or {}).get(N, 0),
                        device_grid=device_grid, **opts) for N, case, nsteps in specs]


BATCH_NS = (32, 48, 96, 192, 384, 512)


def seeded_batch(seed, steps, sizes=BATCH_NS, count=64):
    """The seeded solves of the batched form: ``count`` specs (N, case, steps), N drawn from
    ``sizes``, 1-8 tracers each, Gaussian centres uniform on the sphere, width b in [5, 20] and
    rotation-axis tilt in [0, pi/2] (cases.random_solid_body)."""
    import numpy as np
    from .cases import random_solid_body
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        N = int(rng.choice(sizes))
        ntr = int(rng.integers(1, 9))
        specs.append((N, random_solid_body(rng, ntr), steps))
    return specs
