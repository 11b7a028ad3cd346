"""Advection time stepping -- oracle restatement (test infrastructure).

Restates:
  src/advection_ic.py:21-203          option decoding and state arrays (``Simulation``)
  src/cs_datastruct.py:701-735        velocity containers (``Velocity``)
  src/advection_vars.py:19-107        ``init_vars_adv``
  src/averaged_velocity.py:14-62      ``time_averaged_velocity``
  src/advection_timestep.py:19-75     ``adv_time_step`` / ``update_adv``
  src/advection_sphere.py:13-61       ``adv_sphere`` time loop (norms at the end only)
  src/diagnostics.py:14-26            ``mass_computation``
  src/errors.py:99-113                ``compute_errors``

Multi-tracer runs (an extension) advance each tracer with the same winds; the
reference has one tracer, and one tracer here reproduces it exactly.
"""
import numpy as np

from . import cases
from .edges import vector_edges_rk2_copy
from .grid import CubedSphere, latlon_to_contravariant
from .interp import (ghost_cell_pc_lagrange_interpolation, ghost_cells_adjacent_panels,
                     wind_center2ghostedge_cubic_interpolation, wind_edges2center_cubic_interpolation)
from .lagrange import lagrange_poly_ghostcell_pc
from .operators import cfl_x, cfl_y, divergence
from .recon import Parabola

RECON = {1: "PPM-0", 2: "PPM-CW84", 3: "PPM-PL07", 4: "PPM-L04"}
DP = {1: "RK1", 2: "RK2"}
SPLIT = {1: "SP-AVLT", 2: "SP-L04", 3: "SP-PL07"}
ET = {1: "ET-S72", 2: "ET-PL07", 3: "ET-DG"}
MT = {1: "MT-0", 2: "MT-PL07"}
MF = {1: "MF-0", 2: "MF-AF", 3: "MF-PR"}


class Velocity:
    """src/cs_datastruct.py:701-735"""

    def __init__(self, M, pos):
        shape = {"pu": (M + 1, M, 6), "pv": (M, M + 1, 6), "pc": (M, M, 6)}[pos]
        for name in ("ulon", "vlat", "ucontra", "vcontra", "ucontra_averaged", "vcontra_averaged",
                     "ucontra_old", "vcontra_old"):
            setattr(self, name, np.zeros(shape))
        self.upos = self.uneg = self.vpos = self.vneg = None


class Simulation:
    """Option codes and state arrays, src/advection_ic.py:21-203.

    ``case`` is a reference ``(ic, vf)`` pair turned into a dict, or an extension case dict
    (see :mod:`oracle.cases`).
    """

    def __init__(self, grid, dt, Tf, case, recon=3, dp=1, opsplit=1, et=3, mt=1, mf=3):
        self.dt, self.Tf = dt, Tf
        self.dto2, self.twodt = dt * 0.5, dt * 2.0
        self.case = case
        # the reference object's own option attributes (src/advection_ic.py:24-49, :124-168)
        self.ic, self.vf, self.tc = case.get("ic"), case.get("vf"), 1
        self.recon, self.dp, self.opsplit = recon, dp, opsplit
        self.edge_treatment, self.metric_tensor, self.mass_fixer = et, mt, mf
        self.degree = 3
        self.recon_name, self.dp_name = RECON[recon], DP[dp]
        self.opsplit_name, self.et_name = SPLIT[opsplit], ET[et]
        self.mt_name, self.mf_name = MT[mt], MF[mf]
        valid = (opsplit in (1, 2) and mt == 1) or (opsplit == 3 and mt == 2)
        if not valid:
            raise NameError("invalid (opsplit, mt) pair (src/discrete_operators.py:48-82)")
        if self.et_name == "ET-DG" and grid.projection != "gnomonic_equiangular":
            raise ValueError("ET-DG needs the equiangular grid (src/advection_vars.py:76-77)")
        vf = case["vf"]
        # the reference refreshes winds every step for every vf >= 2 (src/advection_timestep.py:31)
        self.unsteady = (vf >= 2) if isinstance(vf, int) else (vf == "deformational")
        M = grid.N + grid.ng
        self.tracers = cases.case_tracers(case)
        self.ntr = len(self.tracers)
        self.Qs = [np.zeros((M, M, 6)) for _ in range(self.ntr)]
        self.Q = self.Qs[0]
        self.gQ = np.zeros((M, M, 6))
        self.div = np.zeros((M, M, 6))
        self.cx = np.zeros((M + 1, M, 6))
        self.cy = np.zeros((M, M + 1, 6))
        self.U_pu, self.U_pv, self.U_pc = Velocity(M, "pu"), Velocity(M, "pv"), Velocity(M, "pc")
        self.lagrange = None
        self.px = self.py = None
        self.CFL = 0.0


def _interior_winds(grid, sim, t):
    """Analytic lat-lon wind at interior pu / pv edges, src/advection_vars.py:37-40."""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    su = (slice(i0, iend + 1), slice(j0, jend), slice(None))
    sv = (slice(i0, iend), slice(j0, jend + 1), slice(None))
    sim.U_pu.ulon[su], sim.U_pu.vlat[su] = cases.case_wind(sim.case, grid.pu.lon[su], grid.pu.lat[su], t)
    sim.U_pv.ulon[sv], sim.U_pv.vlat[sv] = cases.case_wind(sim.case, grid.pv.lon[sv], grid.pv.lat[sv], t)
    return su, sv


def _conv_pu(grid, U, sl):
    return latlon_to_contravariant(U.ulon[sl], U.vlat[sl], grid.prod_ex_elon_pu[sl], grid.prod_ex_elat_pu[sl],
                                   grid.prod_ey_elon_pu[sl], grid.prod_ey_elat_pu[sl], grid.determinant_ll2contra_pu[sl])


def _conv_pv(grid, U, sl):
    return latlon_to_contravariant(U.ulon[sl], U.vlat[sl], grid.prod_ex_elon_pv[sl], grid.prod_ex_elat_pv[sl],
                                   grid.prod_ey_elon_pv[sl], grid.prod_ey_elat_pv[sl], grid.determinant_ll2contra_pv[sl])


def edges_ghost_cell_treatment_vector(grid, sim):
    """src/edges_treatment.py:296-347"""
    if sim.et_name == "ET-DG":
        wind_edges2center_cubic_interpolation(sim.U_pc, sim.U_pu, sim.U_pv, grid, sim.lagrange)
        wind_center2ghostedge_cubic_interpolation(sim.U_pc, sim.U_pu, sim.U_pv, grid)
    elif sim.dp_name == "RK2":
        vector_edges_rk2_copy(sim.U_pu, sim.U_pv, grid)


def edges_ghost_cell_treatment_scalar(Qx, Qy, grid, sim):
    """src/edges_treatment.py:284-290"""
    if sim.et_name in ("ET-S72", "ET-PL07"):
        ghost_cells_adjacent_panels(Qx, Qy, grid)
    else:
        ghost_cell_pc_lagrange_interpolation(Qx, grid, sim.lagrange)


def time_averaged_velocity(grid, sim):
    """src/averaged_velocity.py:14-62"""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    Upu, Upv = sim.U_pu, sim.U_pv
    Upu.upos = Upu.ucontra[i0:iend + 1, :, :] >= 0
    Upu.uneg = ~Upu.upos
    Upv.vpos = Upv.vcontra[:, j0:jend + 1, :] >= 0
    Upv.vneg = ~Upv.vpos
    if sim.dp_name == "RK1":
        Upu.ucontra_averaged[:] = Upu.ucontra
        Upv.vcontra_averaged[:] = Upv.vcontra
        return
    dto2 = sim.dto2
    # x direction (:38-49)
    ui = 1.5 * Upu.ucontra - 0.5 * Upu.ucontra_old
    a = Upu.ucontra[i0:iend + 1] * dto2 / grid.dx
    u1, u2 = ui[i0 - 1:iend], ui[i0:iend + 1]
    u3, u4 = ui[i0:iend + 1], ui[i0 + 1:iend + 2]
    Upu.ucontra_averaged[i0:iend + 1] = np.where(Upu.upos, (1.0 - a) * u2 + a * u1, -a * u4 + (1.0 + a) * u3)
    # y direction (:51-62)
    vi = 1.5 * Upv.vcontra - 0.5 * Upv.vcontra_old
    a = Upv.vcontra[:, j0:jend + 1] * dto2 / grid.dy
    v1, v2 = vi[:, j0 - 1:jend], vi[:, j0:jend + 1]
    v3, v4 = vi[:, j0:jend + 1], vi[:, j0 + 1:jend + 2]
    Upv.vcontra_averaged[:, j0:jend + 1] = np.where(Upv.vpos, (1.0 - a) * v2 + a * v1, -a * v4 + (1.0 + a) * v3)


def init_vars_adv(grid, sim):
    """src/advection_vars.py:19-107"""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    su, sv = _interior_winds(grid, sim, 0.0)
    sim.U_pu.ucontra[su], sim.U_pu.vcontra[su] = _conv_pu(grid, sim.U_pu, su)
    sim.U_pv.ucontra[sv], sim.U_pv.vcontra[sv] = _conv_pv(grid, sim.U_pv, sv)
    if grid.projection == "gnomonic_equiangular":
        sim.lagrange = lagrange_poly_ghostcell_pc(grid)
    edges_ghost_cell_treatment_vector(grid, sim)
    sim.U_pu.ucontra_old[:] = sim.U_pu.ucontra
    sim.U_pv.vcontra_old[:] = sim.U_pv.vcontra
    time_averaged_velocity(grid, sim)
    sim.cx = cfl_x(sim.U_pu.ucontra, grid, sim)
    sim.cy = cfl_y(sim.U_pv.vcontra, grid, sim)
    sim.CFL = max(abs(np.amax(sim.cx[i0:iend + 1])), abs(np.amax(sim.cy[:, j0:jend + 1])))
    M = grid.N + grid.ng
    sim.px = Parabola(M, sim.recon_name, "x")
    sim.py = Parabola(M, sim.recon_name, "y")
    I, J = slice(i0, iend), slice(j0, jend)
    for n, tr in enumerate(sim.tracers):
        sim.Qs[n][I, J, :] = cases.case_q(sim.case, tr, grid.pc.lon[I, J, :], grid.pc.lat[I, J, :], 0.0)


def adv_time_step(grid, sim, k=None, t=None):
    """src/advection_timestep.py:19-43 (every tracer advanced with the same winds)."""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    if sim.unsteady:
        edges_ghost_cell_treatment_vector(grid, sim)
        time_averaged_velocity(grid, sim)
    for n in range(sim.ntr):
        sim.Q = sim.Qs[n]
        edges_ghost_cell_treatment_scalar(sim.Q, sim.Q, grid, sim)
        divergence(grid, sim)
        sim.Q[i0:iend, j0:jend, :] = sim.Q[i0:iend, j0:jend, :] - sim.dt * sim.div[i0:iend, j0:jend, :]
    sim.Q = sim.Qs[0]


def update_adv(grid, sim, t):
    """src/advection_timestep.py:48-75"""
    if not sim.unsteady:
        return
    _interior_winds(grid, sim, t)
    sim.U_pu.ucontra_old[:] = sim.U_pu.ucontra
    sim.U_pv.vcontra_old[:] = sim.U_pv.vcontra
    full_u = (slice(None),) * 3
    sim.U_pu.ucontra[:], sim.U_pu.vcontra[:] = _conv_pu(grid, sim.U_pu, full_u)
    sim.U_pv.ucontra[:], sim.U_pv.vcontra[:] = _conv_pv(grid, sim.U_pv, full_u)


def mass_computation(Q, grid, total_mass0):
    """src/diagnostics.py:14-26"""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    total_mass = np.sum(Q[i0:iend, j0:jend, :] * grid.metric_tensor_pc[i0:iend, j0:jend, :] * grid.dx * grid.dy)
    if abs(total_mass0) > 10 ** (-10):
        change = abs(total_mass0 - total_mass) / abs(total_mass0)
    else:
        change = abs(total_mass0 - total_mass)
    return total_mass, change


def compute_errors(Q, Qref):
    """src/errors.py:99-113 (unweighted, absolute)."""
    E = abs(Qref - Q)
    n = E.size
    return np.amax(abs(E)), np.sum(E) / n, np.sqrt(np.sum(E * E) / n)


def exact_interior(grid, sim, t, tracer_index=0):
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    I, J = slice(i0, iend), slice(j0, jend)
    return cases.case_q(sim.case, sim.tracers[tracer_index], grid.pc.lon[I, J, :], grid.pc.lat[I, J, :], t)


def run(grid, sim, nsteps, norms=True):
    """The loop of src/advection_sphere.py:13-61 without plotting: init, ``nsteps`` steps,
    returns (linf, l1, l2) of tracer 0 against its exact solution at t = nsteps*dt."""
    init_vars_adv(grid, sim)
    for k in range(1, nsteps + 1):
        t = k * sim.dt
        adv_time_step(grid, sim, k, t)
        update_adv(grid, sim, t)
    if not norms:
        return None
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    return compute_errors(sim.Qs[0][i0:iend, j0:jend, :], exact_interior(grid, sim, nsteps * sim.dt))


def make(N, case, dt, recon=3, dp=1, opsplit=1, et=3, mt=1, mf=3, projection="gnomonic_equiangular"):
    grid = CubedSphere(N, projection)
    sim = Simulation(grid, dt, 0.0, case, recon, dp, opsplit, et, mt, mf)
    return grid, sim
