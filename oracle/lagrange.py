"""Degree-3 Lagrange ghost-cell weights and stencils -- oracle restatement (test infrastructure).

Restates src/lagrange.py:18-163 (``lagrange_basis`` and ``lagrange_poly_ghostcell_pc``):
the ghost-cell centres east of panel 0 are mapped into panel 1's angular
coordinates; each ghost value is interpolated along the neighbour's grid line at
the same depth from 4 support points whose first index ``Kmin`` is clamped as in
the reference (interior targets to the neighbour's interior, ghost targets to the
extended array).  West / north / south tables follow by flip / transpose.
"""
import math

import numpy as np

from .grid import inverse_equiangular_gnomonic_map


def lagrange_basis(x, nodes, N, j):
    """src/lagrange.py:18-23"""
    Lj = 1.0
    for i in range(0, N + 1):
        if i != j:
            Lj = Lj * (x - nodes[i]) / (nodes[j] - nodes[i])
    return Lj


def lagrange_poly_ghostcell_pc(grid, degree=3):
    """src/lagrange.py:28-163.

    Returns ``(weights, stencils)`` with
      weights  = [east (4,M,4), west (4,M,4), north (M,4,4), south (M,4,4)]
      stencils = [Kmin list, Kmax list] in the same east/west/north/south order.
    """
    if grid.projection != "gnomonic_equiangular":
        raise ValueError("ET-DG ghost interpolation needs the equiangular grid (src/lagrange.py:40-45)")
    N, ng, ngl, ngr = grid.N, grid.ng, grid.ngl, grid.ngr
    i0, iend = grid.i0, grid.iend
    dy = grid.dy
    order = degree + 1
    M = N + ng
    p, east = 0, 1

    # ghost centres east of panel 0, seen from panel 1 (src/lagrange.py:59-63)
    _, y_ghost = inverse_equiangular_gnomonic_map(
        grid.pc.X[iend:iend + ngr, :, p], grid.pc.Y[iend:iend + ngr, :, p], grid.pc.Z[iend:iend + ngr, :, p], east)
    # support points: panel 1's first ngr interior columns (src/lagrange.py:66-69)
    _, y = inverse_equiangular_gnomonic_map(
        grid.pc.X[i0:i0 + ngr, :, east], grid.pc.Y[i0:i0 + ngr, :, east], grid.pc.Z[i0:i0 + ngr, :, east], east)

    ghost_pts = np.zeros((ngl, M))
    ghost_pts[:, :] = y_ghost

    # float stencil indices and clamping (src/lagrange.py:77-103)
    K = np.zeros((ngl, M))
    Kmin = np.zeros((ngl, M))
    Kmax = np.zeros((ngl, M))
    half = math.ceil(order / 2)
    for g in range(ngl):
        for i in range(M):
            K[g, i] = (ghost_pts[g, i] - y[g, 0]) / dy
            Kmax[g, i] = K[g, i] + half
            Kmin[g, i] = Kmax[g, i] - order + 1
            if i0 <= i < iend:
                if Kmax[g, i] >= iend:
                    Kmax[g, i] = iend - 1
                    Kmin[g, i] = Kmax[g, i] - order + 1
                elif Kmin[g, i] < i0:
                    Kmin[g, i] = i0
                    Kmax[g, i] = Kmin[g, i] + order - 1
            elif i >= iend:
                if Kmax[g, i] >= M:
                    Kmax[g, i] = M - 1
                    Kmin[g, i] = Kmax[g, i] - order + 1
            else:
                if Kmin[g, i] < 0:
                    Kmin[g, i] = 0
                    Kmax[g, i] = Kmin[g, i] + order - 1
    Kmin = Kmin.astype(int)
    Kmax = Kmax.astype(int)

    # self-checks on interior targets (src/lagrange.py:110-126)
    for g in range(ngl):
        for i in range(i0, iend):
            if Kmax[g, i] - Kmin[g, i] != degree:
                raise RuntimeError("lagrange stencil width")
            if Kmin[g, i] < i0 or Kmax[g, i] > iend:
                raise RuntimeError("lagrange stencil range")
            if order > 1 and not (y[g, Kmin[g, i]] <= y_ghost[g, i] <= y[g, Kmax[g, i]]):
                raise RuntimeError("lagrange stencil bracket")

    # weights (src/lagrange.py:127-142)
    nodes = np.zeros((ngr, M, order))
    weights = np.zeros((ngr, M, order))
    for g in range(ngl):
        for j in range(M):
            nodes[g, j, :] = y[g, Kmin[g, j]:Kmax[g, j] + 1]
    for g in range(ngr):
        for k in range(M):
            for l in range(order):
                weights[g, k, l] = lagrange_basis(ghost_pts[g, k], nodes[g, k, :], degree, l)

    # flips / transposes (src/lagrange.py:144-157)
    w_east = weights
    w_west = np.flip(weights, axis=0)
    w_north = np.transpose(w_east, (1, 0, 2))
    w_south = np.flip(w_north, axis=1)
    kmin_e, kmax_e = Kmin, Kmax
    kmin_w, kmax_w = np.flip(Kmin, axis=0), np.flip(Kmax, axis=0)
    kmin_n, kmax_n = np.transpose(kmin_e, (1, 0)), np.transpose(kmax_e, (1, 0))
    kmin_s, kmax_s = np.flip(kmin_n, axis=1), np.flip(kmax_n, axis=1)
    stencil = [[kmin_e, kmin_w, kmin_n, kmin_s], [kmax_e, kmax_w, kmax_n, kmax_s]]
    return [w_east, w_west, w_north, w_south], stencil
