"""Device solver: torch-owned buffers plus libcsadv through its C ABI.

``Advection`` is the drop-in for the reference's ``adv_simulation_par`` + ``init_vars_adv``
+ ``adv_time_step`` / ``update_adv`` loop (src/advection_ic.py:21-203,
src/advection_vars.py:19-107, src/advection_timestep.py:19-75).  The host builds the
FP64 tables once (grid, metric, duo-grid weights, wind basis, initial condition), the
device does everything per step.  There is no CPU fallback: without libcsadv.so or a
GPU every call raises.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .grid import Grid
from .topology import ghost_recipes, lagrange_tables, orientation

_TDT = {"f64": torch.float64, "f32": torch.float32}


def _round_up(x, m):
    return (x + m - 1) // m * m


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _pad_last(a, width):
    """Zero-pad the last axis of a numpy array or torch tensor to ``width`` columns."""
    if torch.is_tensor(a):
        out = torch.zeros(a.shape[:-1] + (width,), dtype=a.dtype, device=a.device)
    else:
        a = np.asarray(a)
        out = np.zeros(a.shape[:-1] + (width,), dtype=a.dtype)
    out[..., :a.shape[-1]] = a
    return out


def _host(a):
    return a.cpu().numpy() if torch.is_tensor(a) else np.asarray(a)


def _as_tensor(a):
    return a if torch.is_tensor(a) else torch.from_numpy(np.ascontiguousarray(a))


def _stack(arrs):
    return torch.stack(arrs) if torch.is_tensor(arrs[0]) else np.stack(arrs)


class GridTables:
    """Device copies of the per-N tables (shared by every solve of that N and dtype)."""

    def __init__(self, grid, dtype, device):
        self.grid = grid
        N, M = grid.N, grid.M
        self.ld = M if M % 4 == 0 else _round_up(M, 4)
        self.ldu = _round_up(M + 1, 4)
        td = _TDT[dtype]
        dev = device
        ld, ldu = self.ld, self.ldu

        def up(a, tdt=torch.float64):
            if torch.is_tensor(a):
                return a.to(device=dev, dtype=tdt).contiguous()
            return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=tdt)

        g0 = grid.pc.g
        self.gc64 = up(_pad_last(g0, ld))
        self.gc = up(_pad_last(g0, ld), td)
        self.rgc = up(_pad_last(1.0 / g0, ld), td)
        self.gu = up(_pad_last(grid.pu.g, ldu), td)
        self.gv = up(_pad_last(grid.pv.g, ld), td)
        w, k = lagrange_tables(grid)
        self.lagw = up(w)
        self.lagk = up(k, torch.int32)
        self.orient = up(orientation(N), torch.int32)
        rec = ghost_recipes(grid, ld)
        gflat = _host(_pad_last(g0, ld)).reshape(-1)
        self.n_strip, self.n_corner = len(rec["strip_tgt"]), len(rec["corner_tgt"])
        self.strip_tgt, self.strip_src = up(rec["strip_tgt"], torch.int32), up(rec["strip_src"], torch.int32)
        self.strip_w = up(rec["strip_w"])
        self.corner_tgt, self.corner_src = up(rec["corner_tgt"], torch.int32), up(rec["corner_src"], torch.int32)
        self.corner_w = up(rec["corner_w"])
        self.gfix = up(correction_metric(rec, gflat, M * ld).reshape(M, ld), td)
        pc, pu, pv = grid.pc, grid.pu, grid.pv
        self.conv_pc = up(_pad_last(_stack([pc.ex_lon, pc.ex_lat, pc.ey_lon, pc.ey_lat]), ld))
        self.conv_pu = up(_pad_last(_stack([pu.ex_lon, pu.ex_lat, pu.ey_lon, pu.ey_lat, pu.det]), ldu))
        self.conv_pv = up(_pad_last(_stack([pv.ex_lon, pv.ex_lat, pv.ey_lon, pv.ey_lat, pv.det]), ld))


def correction_metric(rec, gflat, pstride):
    """The metric the mass-fixer correction total C multiplies at every cell a step reads:
    g inside the panel and, at a ghost cell, sum_l w_l g(src_l) over its recipe, so that a
    ghost value interpolated from the stored values S_l takes the correction as
    sum_l w_l (S_l + C g_l) = sum_l w_l S_l + C gfix.  By the cube's symmetry the sum is the
    same at a ghost position of every panel (checked here); panel 0's values are kept."""
    gfix = np.array(gflat, dtype=np.float64, copy=True)
    for kind in ("strip", "corner"):
        gw = np.sum(rec[kind + "_w"] * gflat[rec[kind + "_gsrc"]], axis=1)
        tgt = rec[kind + "_tgt"].astype(np.int64)
        loc = tgt % pstride
        # spread of the values at each position over the panels (sorted groups of equal loc)
        order = np.argsort(loc, kind="stable")
        sl, sg = loc[order], gw[order]
        starts = np.flatnonzero(np.r_[True, sl[1:] != sl[:-1]])
        spread = np.max(np.maximum.reduceat(sg, starts) - np.minimum.reduceat(sg, starts))
        if spread > 1e-12 * np.max(np.abs(gflat)):
            raise RuntimeError("ghost metric weights differ between panels by %.3e" % spread)
        first = tgt < pstride
        gfix[tgt[first]] = gw[first]
    return gfix


def time_step_kernel_rotating(solves, reps=20):
    """Mean microseconds of the step kernel over ``reps`` rounds of one launch per solve (each
    solve its own handle and state buffers): with the solves' states together larger than the
    256 MB last-level cache, every launch streams its state from HBM."""
    lib = _lib.load()
    handles = (ctypes.c_void_p * len(solves))(*[s.handle for s in solves])
    us = ctypes.c_double()
    stream = ctypes.c_void_p(torch.cuda.current_stream(solves[0].device).cuda_stream)
    _lib.check(lib.csadv_time_step_kernel_rotating(handles, len(solves), int(reps), stream, ctypes.byref(us)),
               solves[0].handle, "csadv_time_step_kernel_rotating")
    return us.value


_TABLES = {}


def grid_tables(N, dtype, device, device_grid=False):
    """GridTables of (N, dtype, device), built once per process and shared by every solve.
    device_grid: build the grid geometry on the device (torch FP64) instead of the host."""
    dev = torch.device(device)
    key = (int(N), dtype, str(dev), bool(device_grid))
    if key not in _TABLES:
        _TABLES[key] = GridTables(Grid(N, device=dev if device_grid else None), dtype, dev)
    return _TABLES[key]


def interior_winds(grid, case):
    """Contravariant basis winds at the interior edges (src/advection_vars.py:37-53):
    returns (wu [K][6][M][M+1], wv [K][6][M+1][M]) with zeros outside the interior."""
    i0, iend = grid.i0, grid.iend
    M = grid.M
    basis = case.basis()
    K = len(basis)
    if torch.is_tensor(grid.pu.lon):
        wu = torch.zeros((K, 6, M, M + 1), dtype=torch.float64, device=grid.pu.lon.device)
        wv = torch.zeros((K, 6, M + 1, M), dtype=torch.float64, device=grid.pu.lon.device)
    else:
        wu = np.zeros((K, 6, M, M + 1))
        wv = np.zeros((K, 6, M + 1, M))
    su = (slice(None), slice(i0, iend), slice(i0, iend + 1))
    sv = (slice(None), slice(i0, iend + 1), slice(i0, iend))
    pu, pv = grid.pu, grid.pv
    for b, (fu, fv) in enumerate(basis):
        ul, vl = fu(pu.lon[su], pu.lat[su]), fv(pu.lon[su], pu.lat[su])
        wu[b][su] = (pu.ey_lat[su] * ul - pu.ey_lon[su] * vl) / pu.det[su]      # src/sphgeo.py:120-125
        ul, vl = fu(pv.lon[sv], pv.lat[sv]), fv(pv.lon[sv], pv.lat[sv])
        wv[b][sv] = (-pv.ex_lat[sv] * ul + pv.ex_lon[sv] * vl) / pv.det[sv]
    return wu, wv


def cfl_dt(grid, case, cfl, winds=None):
    """Largest dt with max |contravariant wind| * dt / dx <= cfl over the interior edges at t=0."""
    wu, wv = winds if winds is not None else interior_winds(grid, case)
    s = case.coefficients(0.0)
    if torch.is_tensor(wu):
        st = torch.as_tensor(s, dtype=torch.float64, device=wu.device)
        u, v = torch.tensordot(st, wu, dims=1), torch.tensordot(st, wv, dims=1)
        vmax = max(float(u.abs().max()), float(v.abs().max()))
    else:
        u, v = np.tensordot(s, wu, axes=1), np.tensordot(s, wv, axes=1)
        vmax = max(np.abs(u).max(), np.abs(v).max())
    return cfl * grid.dx / vmax


class Advection:
    """One solve on the device.

    Parameters mirror ``adv_simulation_par`` (src/advection_ic.py:21): ``dt``, option
    codes ``recon, dp, opsplit, et, mt, mf`` from par/advection.par, and a
    :class:`cubedsphere.cases.Case`.  ``dtype`` is "f64" (parity mode) or "f32".
    """

    def __init__(self, N, case, dt, recon=3, dp=1, opsplit=1, et=3, mt=1, mf=3, dtype="f64",
                 device="cuda", tile=0, max_steps=1, tables=None, grid=None):
        if not torch.cuda.is_available():
            raise _lib.CsadvError("libcsadv needs a GPU; no CPU fallback exists")
        self.lib = _lib.load()
        self.grid = grid if grid is not None else (tables.grid if tables is not None else Grid(N))
        self.case = case
        self.dt = float(dt)
        self.opts = dict(recon=recon, dp=dp, split=opsplit, et=et, mt=mt, mf=mf)
        self.dtype = dtype
        self.device = torch.device(device)
        self.tile = tile
        self.tab = tables if tables is not None else GridTables(self.grid, dtype, self.device)
        self.max_steps = max(1, int(max_steps))
        self.handle = None
        self.index = 0              # solve index inside the handle (a batch shares one handle)
        self.owner = None           # the Batch whose handle this solve uses, if any
        self.stop = 0               # > 0: the kernels stop advancing this solve after `stop` steps
        self.steps_done = 0
        self._alloc()

    # ------------------------------------------------------------ buffers
    def _alloc(self):
        g, tab, dev = self.grid, self.tab, self.device
        N, M, ld, ldu = g.N, g.M, tab.ld, tab.ldu
        td = _TDT[self.dtype]
        ntr = self.case.ntr
        K = len(self.case.basis())
        f64 = dict(device=dev, dtype=torch.float64)
        self.q = [torch.zeros((ntr, 6, M, ld), device=dev, dtype=td) for _ in range(2)]
        self.uavg = torch.zeros((6, M, ldu), device=dev, dtype=td)
        self.vavg = torch.zeros((6, M + 1, ld), device=dev, dtype=td)
        # RK2 picks the upwind side from the un-averaged wind (src/averaged_velocity.py:21-27)
        rk2 = self.opts["dp"] == 2
        self.usgn = torch.zeros_like(self.uavg) if rk2 else None
        self.vsgn = torch.zeros_like(self.vavg) if rk2 else None
        self.wbu = torch.zeros((K, 6, M, ldu), **f64)
        self.wbv = torch.zeros((K, 6, M + 1, ld), **f64)
        self.wpc = torch.zeros((K, 2, 6, M, ld), **f64)
        steps = 1 if self.case.steady else self.max_steps
        self.coef = torch.from_numpy(self.case.coefficient_table(steps, self.dt).reshape(-1)).to(**f64)
        self.corr = torch.zeros(2 * ntr, **f64)      # MF-PR correction totals, one slot per step parity
        # FP32 with the fused step: the state is stored from a per-tracer constant (the initial
        # field's minimum), Q = S + C gfix + offset, so that on a flat background S is ~0 and
        # increments far below an FP32 ulp of Q there are kept instead of rounded away
        general = self.opts["et"] != 3          # ET-S72 / ET-PL07: the library's general step
        self.offset = torch.zeros(ntr, **f64) if (self.dtype == "f32" and not general) else None
        # partial sums for the tile of most wavefronts, so that the tile may change (set_tile)
        nparts = 0
        t = 0
        while self.lib.csadv_tile_rows(t) > 0:
            nparts = max(nparts, _lib.check(self.lib.csadv_partials_per_tracer(N, t), None, "csadv_partials_per_tracer"))
            t += 1
        self.partials = torch.zeros(ntr * nparts, **f64)
        self.ctl = torch.zeros(4, device=dev, dtype=torch.int32)
        # ET-S72 and ET-PL07 run the library's general step, which needs a work buffer; the fused
        # step with MF-AF keeps the second pass's cube-edge fluxes there ([ntr][6][4][N])
        self.work = None
        if general:
            nbytes = ctypes.c_int64()
            code = _lib.load().csadv_work_bytes(N, _lib.CSADV_F64 if self.dtype == "f64" else _lib.CSADV_F32,
                                                ctypes.byref(nbytes))
            _lib.check(code, None, "csadv_work_bytes")
            self.work = torch.zeros(int(nbytes.value), device=dev, dtype=torch.uint8)
        elif self.opts["mf"] == 2:
            self.work = torch.zeros(ntr * 24 * N, device=dev, dtype=td)

    def device_bytes(self):
        """Device memory held by this solve's own buffers (tables shared per N excluded)."""
        ts = [self.uavg, self.vavg, self.usgn, self.vsgn, self.wbu, self.wbv, self.wpc, self.coef, self.corr,
              self.partials, self.ctl, self.work, self.offset] + list(self.q)
        return sum(t.numel() * t.element_size() for t in ts if t is not None)

    def release_wind_scratch(self):
        """Free the lat-lon wind scratch once the winds are prepared (only wind init uses it)."""
        self.wpc = None

    def desc(self):
        g, tab = self.grid, self.tab
        d = _lib.SolveDesc()
        d.N, d.ntr = g.N, self.case.ntr
        d.dtype = _lib.CSADV_F64 if self.dtype == "f64" else _lib.CSADV_F32
        d.ld, d.ldu = tab.ld, tab.ldu
        for k, v in self.opts.items():
            setattr(d, k, v)
        d.nbasis = len(self.case.basis())
        d.unsteady = 0 if self.case.steady else 1
        d.max_steps = 1 if self.case.steady else self.max_steps
        d.dt, d.dx, d.dy, d.a2 = self.dt, g.dx, g.dy, g.a2
        d.q0, d.q1 = _ptr(self.q[0]), _ptr(self.q[1])
        d.uavg, d.vavg, d.usgn, d.vsgn = _ptr(self.uavg), _ptr(self.vavg), _ptr(self.usgn), _ptr(self.vsgn)
        d.gc, d.gu, d.gv, d.rgc, d.gc64 = _ptr(tab.gc), _ptr(tab.gu), _ptr(tab.gv), _ptr(tab.rgc), _ptr(tab.gc64)
        d.gfix = _ptr(tab.gfix)
        d.lagw, d.lagk, d.orient = _ptr(tab.lagw), _ptr(tab.lagk), _ptr(tab.orient)
        d.wbu, d.wbv = _ptr(self.wbu), _ptr(self.wbv)
        d.conv_pc, d.conv_pu, d.conv_pv = _ptr(tab.conv_pc), _ptr(tab.conv_pu), _ptr(tab.conv_pv)
        d.wpc = _ptr(self.wpc)      # null once the winds are prepared (scratch released)
        d.n_strip, d.n_corner = tab.n_strip, tab.n_corner
        d.strip_tgt, d.strip_src, d.strip_w = _ptr(tab.strip_tgt), _ptr(tab.strip_src), _ptr(tab.strip_w)
        d.corner_tgt, d.corner_src, d.corner_w = _ptr(tab.corner_tgt), _ptr(tab.corner_src), _ptr(tab.corner_w)
        d.coef, d.corr, d.partials, d.ctl = _ptr(self.coef), _ptr(self.corr), _ptr(self.partials), _ptr(self.ctl)
        d.tile = self.tile
        d.stop = self.stop
        d.work = _ptr(self.work)
        d.offset = _ptr(self.offset)
        return d

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # ------------------------------------------------------------ init (src/advection_vars.py:19-107)
    def init_vars_adv(self):
        """Initial state, wind basis and a handle of this one solve."""
        self.init_buffers()
        h = ctypes.c_void_p()
        d = self.desc()
        rc = self.lib.csadv_create(ctypes.byref(d), 1, self.tile, ctypes.byref(h))
        self.handle = h
        _lib.check(rc, h, "csadv_create")
        _lib.check(self.lib.csadv_wind_init(h, self._stream()), h, "csadv_wind_init")
        return self

    def init_buffers(self):
        """Initial condition at the interior, zero ghosts, wind-basis interiors, cleared
        control words (the device part of init_vars_adv; no handle yet)."""
        g = self.grid
        i0, iend = g.i0, g.iend
        pre = getattr(self, "basis_winds", None)       # interior basis winds computed with dt
        wu, wv = pre if pre is not None else interior_winds(g, self.case)
        self.basis_winds = None
        ldu, ld = self.tab.ldu, self.tab.ld
        self.wbu.copy_(_as_tensor(_pad_last(wu, ldu)))
        self.wbv.copy_(_as_tensor(_pad_last(wv, ld)))
        I = slice(i0, iend)
        q0 = _stack([self.case.q(n, g.pc.lon[:, I, I], g.pc.lat[:, I, I], 0.0) for n in range(self.case.ntr)])
        self.q[0].zero_()
        q0 = _as_tensor(q0).to(self.q[0].device, self.q[0].dtype)
        if self.offset is not None:
            off = q0.amin(dim=(1, 2, 3))                 # exact in the state's precision
            self.offset.copy_(off.double())
            q0 = q0 - off[:, None, None, None]
        self.q[0][:, :, I, I] = q0
        self.q[1].zero_()
        self.ctl.zero_()
        self.corr.zero_()
        torch.cuda.synchronize(self.device)
        self.steps_done = 0
        self._wind_t = 0.0            # time of the last wind update (update_adv)
        self._wind_rows = {}          # coefficient rows written for non-uniform update times
        return self

    # ------------------------------------------------------------ stepping
    def step(self, nsteps=1):
        """nsteps calls of adv_time_step + update_adv (src/advection_timestep.py:19-75)."""
        if self.handle is None:
            raise _lib.CsadvError("init_vars_adv() first")
        if self.owner is not None:
            raise _lib.CsadvError("this solve is stepped by its batch (Batch.step)")
        self.ensure_steps(self.steps_done + nsteps)
        _lib.check(self.lib.csadv_step(self.handle, int(nsteps), self._stream()), self.handle, "csadv_step")
        self.steps_done += nsteps

    # ------------------------------------------------------------ time-dependent winds
    def ensure_steps(self, total, refresh=True):
        """Grow the wind-coefficient table of a time-dependent wind so that it covers ``total``
        steps (rows of t_n = n dt, the times src/advection_sphere.py:45-57 passes to update_adv).
        Steady winds need one row and never grow.  Rows already written (update_adv at other
        times) are kept.  Returns True when the table grew (the handle then needs ``refresh``,
        which this call makes unless ``refresh`` is False)."""
        if self.case.steady or total <= self.max_steps:
            return False
        new = max(int(total), 2 * self.max_steps)
        tab = torch.from_numpy(self.case.coefficient_table(new, self.dt).reshape(-1)).to(self.coef)
        tab[:self.coef.numel()] = self.coef          # keep rows set from update_adv
        self.coef = tab
        self.max_steps = new
        if refresh:
            self.refresh()
        return True

    def set_wind_times(self, t, t_prev):
        """Winds of the next step from the analytic wind at ``t`` (and ``t_prev``, the time of
        the previous wind update, for RK2's extrapolation): what update_adv leaves behind for the
        next adv_time_step (src/advection_timestep.py:48-75).  Writes one coefficient row only
        when it differs from the row already there (at the uniform times t = n dt it never does)."""
        if self.case.steady:
            return
        n = self.steps_done
        self.ensure_steps(n + 1)
        row = np.stack([self.case.coefficients(t), self.case.coefficients(t_prev)]).reshape(-1)
        rows = self._wind_rows
        have = rows.get(n)
        if have is None:     # the row coefficient_table wrote: (s(t_n), s(t_(n-1)))
            have = np.stack([self.case.coefficients(n * self.dt),
                             self.case.coefficients(max(n - 1, 0) * self.dt)]).reshape(-1)
        if not np.array_equal(have, row):
            self.coef[n * row.size:(n + 1) * row.size].copy_(torch.from_numpy(row))
            rows[n] = row

    def set_tile(self, tile):
        """Use tile id ``tile`` from the next step on (the handle needs refresh(); the pending
        mass-fixer partials must be folded first, as Batch.set_tiles does)."""
        self.tile = int(tile)

    def refresh(self):
        """Hand changed buffer pointers to the library (the handle's descriptor copy)."""
        if self.handle is None:
            return
        if self.owner is not None:
            self.owner.refresh()
            return
        d = self.desc()
        _lib.check(self.lib.csadv_update(self.handle, ctypes.byref(d), 1, self._stream()), self.handle,
                   "csadv_update")

    def load_state(self, q):
        """Replace the interior of every tracer with ``q`` ([ntr][6][N][N] or [6][N][N] for one
        tracer, FP64 or the solve's dtype): the pending mass-fixer totals are folded first, so the
        device then holds exactly ``q``."""
        self.finalize()
        g = self.grid
        I = slice(g.i0, g.iend)
        t = torch.as_tensor(np.asarray(q) if not torch.is_tensor(q) else q)
        if t.dim() == 3:
            t = t.unsqueeze(0)
        self.q[self.current()][:, :, I, I] = t.to(self.q[0].device, self.q[0].dtype)

    def use_graph(self, steps_per_graph):
        """Replay groups of ``steps_per_graph`` steps from one captured HIP graph (0 = off)."""
        _lib.check(self.lib.csadv_set_graph(self.handle, int(steps_per_graph)), self.handle, "csadv_set_graph")

    def ghost_fill(self):
        """Duo-grid ghost fill of the current state (src/edges_treatment.py:284-290)."""
        _lib.check(self.lib.csadv_ghost_fill(self.handle, self._stream()), self.handle, "csadv_ghost_fill")

    def winds(self):
        """Time-averaged winds of the next step: (uavg [6][M][M+1], vavg [6][M+1][M]) numpy.
        A handle whose solves are all RK1 with a one-field time-dependent wind forms the winds
        inside the step kernel: its arrays then keep the winds of the first step."""
        M = self.grid.M
        return self.uavg[:, :, :M + 1].double().cpu().numpy(), self.vavg[:, :, :M].double().cpu().numpy()

    def current(self):
        return _lib.check(self.lib.csadv_current(self.handle, self.index), self.handle, "csadv_current")

    def finalize(self):
        _lib.check(self.lib.csadv_finalize(self.handle, self.index, self._stream()), self.handle, "csadv_finalize")

    def state(self, native=False):
        """Interior of every tracer after folding in pending corrections: [ntr][6][N][N] numpy,
        FP64, or with ``native`` in the solve's own precision (the same values; an FP32 solve's
        result then crosses PCIe at half the size)."""
        self.finalize()
        g = self.grid
        I = slice(g.i0, g.iend)
        q = self.q[self.current()][:, :, I, I]
        return (q.contiguous() if native else q.double()).cpu().numpy()

    def full_state(self):
        """Every tracer including ghosts: [ntr][6][M][M] numpy (dtype of the solve)."""
        self.finalize()
        M = self.grid.M
        return self.q[self.current()][:, :, :, :M].cpu().numpy()

    def remap_latlon(self, tracer=0, nlat=720, nlon=1440, field=None, tables=False):
        """Nearest-cell remap onto the plotting lat-lon grid (src/interpolation.py:38-149):
        [nlon][nlat] FP64 numpy of the tracer's state, or of ``field`` (an interior
        [6][N][N] array, i fastest) when given; with ``tables`` also the (i, j, panel) index
        tables of src/interpolation.py:38-139."""
        dev = self.device
        out = torch.empty((nlon, nlat), dtype=torch.float64, device=dev)
        f = None
        if field is not None:
            f = torch.as_tensor(np.ascontiguousarray(field) if not torch.is_tensor(field) else field,
                                dtype=torch.float64).to(dev).contiguous()
        tabs = [torch.empty((nlon, nlat), dtype=torch.int32, device=dev) for _ in range(3)] if tables else [None] * 3
        _lib.check(self.lib.csadv_remap_latlon(self.handle, self.index, int(tracer), _ptr(f), int(nlat), int(nlon),
                                               _ptr(out), _ptr(tabs[0]), _ptr(tabs[1]), _ptr(tabs[2]), self._stream()),
                   self.handle, "csadv_remap_latlon")
        q = out.cpu().numpy()
        return (q, tuple(t.cpu().numpy() for t in tabs)) if tables else q

    def mass(self, tracer=0):
        out = ctypes.c_double()
        _lib.check(self.lib.csadv_mass(self.handle, self.index, tracer, ctypes.byref(out), self._stream()), self.handle,
                   "csadv_mass")
        return out.value

    def diagnostics(self, qexact=None, blowup_limit=0.0, out=None, tracer0=0, ntracers=None):
        """Device diagnostics of the tracers, queued on the stream without a host synchronisation
        (graph-capturable): a float64 device tensor [ntracers][4] of (mass, Linf, L1, L2) per
        tracer (src/diagnostics.py:14-26, src/errors.py:99-113), the norms against ``qexact``
        (device FP64 [ntracers][6][N][N]) or of the field itself.  ``blowup_limit`` > 0 halts the
        solve when a Linf exceeds it (src/output.py:44-47); ``status()`` reports it."""
        n = self.case.ntr - tracer0 if ntracers is None else int(ntracers)
        if out is None:
            out = torch.empty((n, 4), dtype=torch.float64, device=self.device)
        qe = None
        if qexact is not None:
            qe = qexact if torch.is_tensor(qexact) else torch.from_numpy(np.ascontiguousarray(qexact))
            qe = qe.to(self.device, torch.float64).contiguous()
            self._diag_exact = qe            # kept alive until the queued kernels have read it
        _lib.check(self.lib.csadv_diagnostics(self.handle, self.index, int(tracer0), n, _ptr(qe), float(blowup_limit),
                                              _ptr(out), self._stream()), self.handle, "csadv_diagnostics")
        return out

    def status(self):
        """(steps completed, halted) from the device; the host's step count follows it."""
        steps, halted = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self.lib.csadv_status(self.handle, self.index, ctypes.byref(steps), ctypes.byref(halted),
                                         self._stream()), self.handle, "csadv_status")
        self.steps_done = steps.value
        return steps.value, bool(halted.value)

    def exact(self, tracer=0, t=None):
        g = self.grid
        t = self.steps_done * self.dt if t is None else t
        I = slice(g.i0, g.iend)
        return self.case.q(tracer, g.pc.lon[:, I, I], g.pc.lat[:, I, I], t)

    def norms(self, tracer=0, t=None):
        """(Linf, L1, L2) against the exact solution (src/errors.py:99-113)."""
        qe = _as_tensor(self.exact(tracer, t)).to(self.device, torch.float64).contiguous()
        out = (ctypes.c_double * 3)()
        _lib.check(self.lib.csadv_norms(self.handle, self.index, tracer, _ptr(qe), out, self._stream()), self.handle,
                   "csadv_norms")
        return tuple(out)

    def profile(self, capacity):
        """Record HIP events around each step-kernel launch of the next ``capacity`` steps."""
        _lib.check(self.lib.csadv_profile(self.handle, int(capacity)), self.handle, "csadv_profile")

    def profile_read(self):
        """(summed step-kernel milliseconds, launches) since the last read."""
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

    def close(self):
        if self.handle and self.owner is None:
            self.lib.csadv_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
