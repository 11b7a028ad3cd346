"""Cube topology and duo-grid interpolation tables (FP64, host, once per N).

``orientation(N)`` -> int32 [6][4][7]: for panel p and side s (0 E, 1 W, 2 N, 3 S) the
neighbour cell feeding the halo strip at depth d (0 = next to the shared edge) and
along-edge index r is  (q, ib + id*d + ir*r, jb + jd*d + jr*r)  in (panel, i, j).  This
is the connectivity of src/halo_data.py:15-185 expressed as affine index maps.

``lagrange_tables(grid)`` -> (weights [4][M][4] FP64, kmin [4][M] int32): the degree-3
duo-grid interpolation of src/lagrange.py:28-163 for the east side of panel 0; by the
cube's symmetry the same table serves every side once indexed by depth (the reference
derives its west / north / south tables from it by flips and transposes, :144-157).
"""
import math

import numpy as np

E, W, N_, S = 0, 1, 2, 3


def orientation(N):
    """Affine halo maps, derived from src/halo_data.py:33-182."""
    M = N + 8
    i0, iend = 4, N + 4
    j0, jend = 4, N + 4
    F, R = M - 1, None  # F: index of the far end, r reversed as F - r
    # entries: (q, ib, id, ir, jb, jd, jr)
    t = {
        (0, E): (1, i0, 1, 0, 0, 0, 1), (0, W): (3, iend - 1, -1, 0, 0, 0, 1),
        (0, N_): (4, 0, 0, 1, j0, 1, 0), (0, S): (5, 0, 0, 1, jend - 1, -1, 0),
        (1, E): (2, i0, 1, 0, 0, 0, 1), (1, W): (0, iend - 1, -1, 0, 0, 0, 1),
        (1, N_): (4, iend - 1, -1, 0, 0, 0, 1), (1, S): (5, iend - 1, -1, 0, F, 0, -1),
        (2, E): (3, i0, 1, 0, 0, 0, 1), (2, W): (1, iend - 1, -1, 0, 0, 0, 1),
        (2, N_): (4, F, 0, -1, jend - 1, -1, 0), (2, S): (5, F, 0, -1, j0, 1, 0),
        (3, E): (0, i0, 1, 0, 0, 0, 1), (3, W): (2, iend - 1, -1, 0, 0, 0, 1),
        (3, N_): (4, i0, 1, 0, F, 0, -1), (3, S): (5, i0, 1, 0, 0, 0, 1),
        (4, E): (1, 0, 0, 1, jend - 1, -1, 0), (4, W): (3, F, 0, -1, jend - 1, -1, 0),
        (4, N_): (2, F, 0, -1, jend - 1, -1, 0), (4, S): (0, 0, 0, 1, jend - 1, -1, 0),
        (5, E): (1, F, 0, -1, j0, 1, 0), (5, W): (3, 0, 0, 1, j0, 1, 0),
        (5, N_): (0, 0, 0, 1, j0, 1, 0), (5, S): (2, F, 0, -1, j0, 1, 0),
    }
    del R
    out = np.zeros((6, 4, 7), dtype=np.int32)
    for (p, s), row in t.items():
        out[p, s] = row
    return out


def halo_source(orient, p, side, d, r):
    """(q, i, j) of the neighbour cell feeding strip (p, side) at depth d, position r."""
    q, ib, id_, ir, jb, jd, jr = (int(v) for v in orient[p, side])
    return q, ib + id_ * d + ir * r, jb + jd * d + jr * r


def lagrange_tables(grid, degree=3):
    """Duo-grid ghost stencils and weights (src/lagrange.py:28-163), east side of panel 0.

    The arithmetic follows the reference operation for operation so that the integer
    stencils agree exactly: ghost centres of panel 0 are expressed in panel 1's y angle,
    the fractional index is clamped with the same float comparisons and truncated.
    """
    N, M = grid.N, grid.M
    i0, iend = grid.i0, grid.iend
    order = degree + 1
    x = grid.x_pc
    # panel-0 ghost centres east of the edge and panel-1 support centres, [g][k] = (i, j)
    xg, yk = np.meshgrid(x[iend:iend + 4], x, indexing="ij")
    txg, tyg = np.tan(xg), np.tan(yk)
    invD = 1.0 / np.sqrt(1.0 + txg**2 + tyg**2)
    # panel 0 point seen from panel 1: tan y = Z / Y (src/cs_transform.py:105-107)
    y_ghost = np.arctan((invD * tyg) / (invD * txg))
    xs, ys = np.meshgrid(x[i0:i0 + 4], x, indexing="ij")
    txs, tys = np.tan(xs), np.tan(ys)
    invDs = 1.0 / np.sqrt(1.0 + txs**2 + tys**2)
    y_sup = np.arctan((invDs * tys) / invDs)   # panel 1: Y = 1/D, Z = tan y / D

    K = (y_ghost - y_sup[:, 0:1]) / grid.dy
    Kmax = K + math.ceil(order / 2)
    Kmin = Kmax - order + 1
    k = np.arange(M)[None, :]
    inner = (k >= i0) & (k < iend)
    hi = inner & (Kmax >= iend)
    lo = inner & ~hi & (Kmin < i0)
    Kmax = np.where(hi, iend - 1, Kmax)
    Kmin = np.where(hi, Kmax - order + 1, Kmin)
    Kmin = np.where(lo, i0, Kmin)
    Kmax = np.where(lo, Kmin + order - 1, Kmax)
    right = (k >= iend) & (Kmax >= M)
    Kmax = np.where(right, M - 1, Kmax)
    Kmin = np.where(right, Kmax - order + 1, Kmin)
    left = (k < i0) & (Kmin < 0)
    Kmin = np.where(left, 0, Kmin)
    Kmax = np.where(left, Kmin + order - 1, Kmax)
    kmin = Kmin.astype(np.int64)
    kmax = Kmax.astype(np.int64)
    if np.any(kmax - kmin != degree):
        raise RuntimeError("duo-grid stencil width mismatch")
    ii = inner[0]
    if np.any(kmin[:, ii] < i0) or np.any(kmax[:, ii] > iend):
        raise RuntimeError("duo-grid stencil leaves the neighbour's interior")

    g = np.arange(4)[:, None, None]
    nodes = y_sup[g, kmin[:, :, None] + np.arange(order)[None, None, :]]   # [4][M][4]
    xq = y_ghost[:, :, None]
    w = np.empty((4, M, order))
    for j in range(order):
        L = np.ones((4, M))
        for i in range(order):
            if i != j:
                L = L * (xq[..., 0] - nodes[..., i]) / (nodes[..., j] - nodes[..., i])
        w[..., j] = L
    return np.ascontiguousarray(w), np.ascontiguousarray(kmin.astype(np.int32))
