"""Dimension-split divergence -- oracle restatement (test infrastructure).

Restates src/discrete_operators.py:18-138 (``divergence``, ``F_operator``,
``G_operator``) and src/cfl.py:10-19.
"""
import numpy as np

from .edges import average_flux_cube_edges
from .flux import compute_fluxes
from .interp import ghost_cells_adjacent_panels


def cfl_x(u_edges, grid, sim):
    """src/cfl.py:10-12"""
    return u_edges * sim.dt / grid.dx


def cfl_y(v_edges, grid, sim):
    """src/cfl.py:17-19"""
    return v_edges * sim.dt / grid.dy


def F_operator(grid, sim):
    """src/discrete_operators.py:109-120"""
    i0, iend = grid.i0, grid.iend
    f1 = sim.px.f_upw[i0 + 1:iend + 1]
    f2 = sim.px.f_upw[i0:iend]
    sim.px.dF[i0:iend] = -(f1 - f2)
    sim.px.dF[i0:iend] = sim.px.dF[i0:iend] * sim.dt / grid.dx


def G_operator(grid, sim):
    """src/discrete_operators.py:128-138"""
    j0, jend = grid.j0, grid.jend
    g1 = sim.py.f_upw[:, j0 + 1:jend + 1]
    g2 = sim.py.f_upw[:, j0:jend]
    sim.py.dF[:, j0:jend] = -(g1 - g2)
    sim.py.dF[:, j0:jend] = sim.py.dF[:, j0:jend] * sim.dt / grid.dy


def split_half_steps(grid, sim):
    """src/discrete_operators.py:31-73: the first flux pass on Q and the split's inner
    half-steps (Qx, Qy), divided by the metric under MT-0."""
    N, ng = grid.N, grid.ng
    g = grid.metric_tensor_pc

    sim.gQ[:] = sim.Q * g
    sim.cx[:] = cfl_x(sim.U_pu.ucontra_averaged, grid, sim)
    sim.cy[:] = cfl_y(sim.U_pv.vcontra_averaged, grid, sim)

    compute_fluxes(sim.Q, sim.Q, sim.px, sim.py, sim.U_pu, sim.U_pv, sim.cx, sim.cy, grid, sim)
    F_operator(grid, sim)
    G_operator(grid, sim)
    pxdF, pydF, gQ, Q = sim.px.dF, sim.py.dF, sim.gQ, sim.Q

    if sim.opsplit_name == "SP-AVLT":
        gQx = gQ + 0.5 * pxdF
        gQy = gQ + 0.5 * pydF
    elif sim.opsplit_name in ("SP-L04", "SP-PL07"):
        gpu, gpv = grid.metric_tensor_pu, grid.metric_tensor_pv
        c1x, c2x = gpu[1:] * sim.cx[1:], gpu[:N + ng] * sim.cx[:N + ng]
        c1y, c2y = gpv[:, 1:] * sim.cy[:, 1:], gpv[:, :N + ng] * sim.cy[:, :N + ng]
        if sim.opsplit_name == "SP-L04":
            # L04 eqs. 7-8 (:55-62)
            gQx = gQ + 0.5 * pxdF + 0.5 * (c1x - c2x) * Q
            gQy = gQ + 0.5 * pydF + 0.5 * (c1y - c2y) * Q
        else:
            # PL07 eqs. 17-18 (:64-69)
            Qx = 0.5 * (Q + (Q + pxdF) / (1.0 - (c1x - c2x)))
            Qy = 0.5 * (Q + (Q + pydF) / (1.0 - (c1y - c2y)))
    else:
        raise ValueError("invalid splitting %r" % sim.opsplit_name)

    if sim.mt_name == "MT-0":
        if sim.opsplit_name == "SP-PL07":
            raise NameError("SP-PL07 with MT-0 is not a valid pair (src/discrete_operators.py:72-73)")
        Qx, Qy = gQx / g, gQy / g
    elif sim.opsplit_name != "SP-PL07":
        raise NameError("MT-PL07 needs SP-PL07 (src/discrete_operators.py:72-82)")
    return Qx, Qy


def divergence(grid, sim):
    """src/discrete_operators.py:18-101"""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    dt = sim.dt
    g = grid.metric_tensor_pc
    Qx, Qy = split_half_steps(grid, sim)

    if sim.et_name in ("ET-S72", "ET-PL07"):
        ghost_cells_adjacent_panels(Qx, Qy, grid)

    # second pass with swapped arguments (:81-82)
    compute_fluxes(Qy, Qx, sim.px, sim.py, sim.U_pu, sim.U_pv, sim.cx, sim.cy, grid, sim)
    if sim.mf_name == "MF-AF":
        average_flux_cube_edges(sim.px, sim.py, grid)
    F_operator(grid, sim)
    G_operator(grid, sim)

    pxdF, pydF = sim.px.dF, sim.py.dF
    sim.div[:] = -(pxdF + pydF) / (dt * g)

    if sim.mf_name == "MF-PR":
        # projection mass fixer (:98-101)
        I, J = slice(i0, iend), slice(j0, jend)
        m0 = np.sum(sim.div[I, J, :] * g[I, J, :])
        a2 = np.sum(g[I, J, :] * g[I, J, :])
        sim.div[I, J, :] = sim.div[I, J, :] - g[I, J, :] * m0 / a2
