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


def _ghost_target(side, d, k, N):
    iend = N + 4
    return {E: (iend + d, k), W: (3 - d, k), N_: (k, iend + d), S: (k, 3 - d)}[side]


def ghost_recipes(grid, ld, depth=3):
    """Flattened duo-grid recipes of the ghost cells a step reads (depth < 3).

    Every ghost value is a weighted sum of interior cells of other panels:
      strips : 4 sources (src/interpolation.py:191-248)
      corners: up to 16 sources -- a corner interpolates along the east / west halo, whose
               points may be the neighbour's strip ghosts, themselves 4-source sums
               (src/interpolation.py:250-314); the two levels are multiplied out.
    Returns dict of numpy arrays (offsets are element offsets inside one tracer:
    panel * M * ld + j * ld + i):
      strip_tgt [n] int32, strip_src [n][4] int32, strip_w [n][4] f64
      corner_tgt [m] int32, corner_src [m][16] int32, corner_w [m][16] f64
    ``*_gsrc`` hold the offsets inside one panel (j * ld + i) for the metric of each source.
    Corner stencils may read a neighbour strip's 4th layer (depth 3), whose recipe is formed
    where it is read.
    """
    N, M = grid.N, grid.M
    orient = orientation(N)
    w, kmin = lagrange_tables(grid)
    pstride = M * ld
    ks = np.arange(4, N + 4)
    # ---- strips
    # Order within one (panel, side): when the source strip runs along x in the neighbour
    # (ir != 0), the along-edge index is fastest, so consecutive cells read consecutive
    # addresses; when it runs down a column, the depth is fastest, so consecutive cells read
    # (and, on the east and west sides, write) the neighbouring cells of one row.
    tgt, src, wts, gsrc = [], [], [], []
    iend = N + 4

    orient_l = orient.tolist()
    kmin_l = kmin.tolist()
    w_l = w.tolist()

    def strip_of(p, ti, tj):
        """Recipe (source offsets, panel offsets, weights; Python lists) of the strip ghost at
        (ti, tj) of panel p: the side and depth follow from where the cell lies (corners query
        only strip cells with an interior along-edge index, up to depth 3)."""
        if ti >= iend:
            side, d, k = E, ti - iend, tj
        elif ti < 4:
            side, d, k = W, 3 - ti, tj
        elif tj >= iend:
            side, d, k = N_, tj - iend, ti
        else:
            side, d, k = S, 3 - tj, ti
        if not (0 <= d < 4 and 4 <= k < iend):
            raise ValueError("(%d, %d) of panel %d is not a strip ghost cell" % (ti, tj, p))
        q, ib, id_, ir, jb, jd, jr = orient_l[p][side]
        k0 = kmin_l[d][k]
        ss, gg = [], []
        for l in range(4):
            si, sj = ib + id_ * d + ir * (k0 + l), jb + jd * d + jr * (k0 + l)
            ss.append(q * pstride + sj * ld + si)
            gg.append(sj * ld + si)
        return ss, gg, w_l[d][k]

    for p in range(6):
        for side in (E, W, N_, S):
            q, ib, id_, ir, jb, jd, jr = (int(v) for v in orient[p, side])
            parts = []
            for d in range(depth):
                K = kmin[d, ks][:, None] + np.arange(4)[None, :]
                si = ib + id_ * d + ir * K
                sj = jb + jd * d + jr * K
                ti, tj = _ghost_target(side, d, ks, N)
                ti = np.broadcast_to(ti, ks.shape)
                tj = np.broadcast_to(tj, ks.shape)
                parts.append((p * pstride + tj * ld + ti, q * pstride + sj * ld + si, sj * ld + si, w[d, ks, :]))
            if ir != 0:
                for t_, s_, g_, w_ in parts:
                    tgt.append(t_)
                    src.append(s_)
                    gsrc.append(g_)
                    wts.append(w_)
            else:
                tgt.append(np.stack([t_ for t_, _, _, _ in parts], axis=1).reshape(-1))
                src.append(np.stack([s_ for _, s_, _, _ in parts], axis=1).reshape(-1, 4))
                gsrc.append(np.stack([g_ for _, _, g_, _ in parts], axis=1).reshape(-1, 4))
                wts.append(np.stack([w_ for _, _, _, w_ in parts], axis=1).reshape(-1, 4))
    out = {
        "strip_tgt": np.concatenate(tgt).astype(np.int32),
        "strip_src": np.concatenate(src).astype(np.int32),
        "strip_gsrc": np.concatenate(gsrc).astype(np.int32),
        "strip_w": np.ascontiguousarray(np.concatenate(wts)),
    }
    # ---- corners (east / west halos, rows outside the panel)
    rows = [r for r in range(4 - depth, 4)] + [r for r in range(iend, iend + depth)]
    ctgt, csrc, cgsrc, cw = [], [], [], []
    for p in range(6):
        for side in (E, W):
            q, ib, id_, ir, jb, jd, jr = orient_l[p][side]
            for d in range(depth):
                for k in rows:
                    ti, tj = _ghost_target(side, d, k, N)
                    s_, g_, w_ = [], [], []
                    for l in range(4):
                        r = kmin_l[d][k] + l
                        si, sj = ib + id_ * d + ir * r, jb + jd * d + jr * r
                        wl = w_l[d][k][l]
                        if 4 <= si < iend and 4 <= sj < iend:
                            s_.append(q * pstride + sj * ld + si)
                            g_.append(sj * ld + si)
                            w_.append(wl)
                        else:
                            ss, gg, ww = strip_of(q, si, sj)
                            s_.extend(ss)
                            g_.extend(gg)
                            w_.extend(wl * v for v in ww)
                    while len(s_) < 16:          # pad with zero-weight copies of the first source
                        s_.append(s_[0])
                        g_.append(g_[0])
                        w_.append(0.0)
                    ctgt.append(p * pstride + tj * ld + ti)
                    csrc.append(s_)
                    cgsrc.append(g_)
                    cw.append(w_)
    out["corner_tgt"] = np.array(ctgt, dtype=np.int32)
    out["corner_src"] = np.array(csrc, dtype=np.int32)
    out["corner_gsrc"] = np.array(cgsrc, dtype=np.int32)
    out["corner_w"] = np.array(cw, dtype=np.float64)
    return out


_OUT_SIGN = {E: 1, W: -1, N_: 1, S: -1}      # outward normal of a side, in the panel's own axes


def cube_edge_pairs(N):
    """The 24 panel sides and what lies across each cube edge, for the cross-panel steps of
    ET-PL07 (src/edges_treatment.py:31-178) and MF-AF (:231-278).

    Returns int32 [6][4][4]: for panel p and side s, (q, s_q, rev, sign):
      q     the neighbour panel;  s_q its side on the shared edge;
      rev   1 if the along-edge index runs backwards on the neighbour (interior index k of
            ours is N-1-k of the neighbour's), else 0;
      sign  +1 / -1: the neighbour's axis across the edge points the same / opposite way as
            ours (edge values swap q_L <-> q_R, fluxes change sign).
    Derived from the halo orientation maps: the neighbour cell next to our ghost at depth 0
    is the neighbour's boundary cell, and the direction in which depth grows tells its side.
    """
    o = orientation(N)
    out = np.zeros((6, 4, 4), dtype=np.int32)
    for p in range(6):
        for s in (E, W, N_, S):
            q, ib, id_, ir, jb, jd, jr = (int(v) for v in o[p, s])
            if id_ != 0:
                s_q = W if id_ > 0 else E          # depth grows into q from its west / east side
                along = jr
            else:
                s_q = S if jd > 0 else N_
                along = ir
            inward = {W: 1, E: -1, S: 1, N_: -1}[s_q]
            out[p, s] = (q, s_q, 1 if along < 0 else 0, _OUT_SIGN[s] * inward)
    return out
