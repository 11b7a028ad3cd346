"""FP64 geometry of the extended gnomonic cubed sphere, built once per N.

Everything here is init-time work: the step itself runs in libcsadv.  Arrays are produced
directly in the device layout ``[panel][j][i]`` (i = x fastest), as numpy arrays on the
host (the parity path) or, with ``device=``, as FP64 torch tensors built on that device
(the same formulas; used to prepare the many solves of the batched and served forms).

Design: every panel is the panel-0 construction rotated by a signed permutation
matrix (``PANEL_ROT``), so positions, tangent vectors and lat-lon unit vectors come
from one code path.  Positions reproduce src/cs_transform.py:41-94 bit for bit (a
signed permutation is exact); tangent vectors and the metric use the closed forms

    e_x = (1+tx^2)/D^3 (-tx, 1+ty^2, -tx ty),   e_y = (1+ty^2)/D^3 (-ty, -tx ty, 1+tx^2),
    g   = (1+tx^2)(1+ty^2)/D^3,                D = sqrt(1 + tx^2 + ty^2),

which agree with the reference's construction (src/cs_transform.py:245-422,
src/cs_datastruct.py:404-493) to rounding.
"""
import math

import numpy as np
import torch

NG = 4
PIO4 = math.pi / 4.0

# world = R_p @ (c0, c1, c2) with c = (1, tan x, tan y)/D  (src/cs_transform.py:62-91)
PANEL_ROT = np.array([
    [[1, 0, 0], [0, 1, 0], [0, 0, 1]],
    [[0, -1, 0], [1, 0, 0], [0, 0, 1]],
    [[-1, 0, 0], [0, -1, 0], [0, 0, 1]],
    [[0, 1, 0], [-1, 0, 0], [0, 0, 1]],
    [[0, 0, -1], [0, 1, 0], [1, 0, 0]],
    [[0, 0, 1], [0, 1, 0], [-1, 0, 0]],
], dtype=np.int64)


# PANEL_ROT as (source component, sign) per output component, [6][3] each
_ROT_IDX = np.array([[int(np.nonzero(row)[0][0]) for row in R] for R in PANEL_ROT], dtype=np.int64)
_ROT_SGN = np.array([[float(row[np.nonzero(row)[0][0]]) for row in R] for R in PANEL_ROT])


def _rotate(p, c0, c1, c2):
    """Apply the signed permutation of panel p exactly (no multiplications by +-1 rounding)."""
    comps = (c0, c1, c2)
    out = []
    for row in PANEL_ROT[p]:
        k = int(np.nonzero(row)[0][0])
        out.append(comps[k] if row[k] > 0 else -comps[k])
    return out


class PointSet:
    """Geometry at one family of points (centres, x-edges or y-edges), layout [6][ny][nx].

    device None: numpy on the host (the parity path, positions bitwise those of the
    reference); a torch device: the same formulas in FP64 torch on that device (fast
    preparation of many solves; agrees with the host path to rounding)."""

    def __init__(self, x, y, device=None):
        # x varies along the last axis, y along the middle axis
        if device is None:
            xp = np
            yy, xx = np.meshgrid(y, x, indexing="ij")
        else:
            xp = torch
            f64 = dict(dtype=torch.float64, device=device)
            yy, xx = torch.meshgrid(torch.as_tensor(y, **f64), torch.as_tensor(x, **f64), indexing="ij")
        tx, ty = xp.tan(xx), xp.tan(yy)
        D2 = 1.0 + tx**2 + ty**2
        D = xp.sqrt(D2)
        invD = 1.0 / D
        c0, c1, c2 = invD, invD * tx, invD * ty
        D3 = D2 * D
        fx = (1.0 + tx * tx) / D3
        fy = (1.0 + ty * ty) / D3
        ex0 = (-tx * fx, (1.0 + ty * ty) * fx, -tx * ty * fx)
        ey0 = (-ty * fy, -tx * ty * fy, (1.0 + tx * tx) * fy)
        self.g = (1.0 + tx * tx) * (1.0 + ty * ty) / D3
        shape = (6,) + tuple(xx.shape)
        if device is not None:
            self._panels_device(c0, c1, c2, ex0, ey0, f64)
            return

        def empty():
            return np.empty(shape) if device is None else torch.empty(shape, **f64)

        self.X, self.Y, self.Z = empty(), empty(), empty()
        self.lon, self.lat = empty(), empty()
        self.ex_lon, self.ex_lat = empty(), empty()
        self.ey_lon, self.ey_lat = empty(), empty()
        for p in range(6):
            X, Y, Z = _rotate(p, c0, c1, c2)
            ex = _rotate(p, *ex0)
            ey = _rotate(p, *ey0)
            h = xp.hypot(X, Y)
            lon = xp.arctan2(Y, X)
            lat = xp.arctan2(Z, h)
            # unit lat-lon vectors from the position itself (|X,Y,Z| = 1): sin lon = Y/h,
            # cos lon = X/h, sin lat = Z, cos lat = h; no trigonometric call per point.  At a
            # pole (h = 0, a cell centre of odd N) lon = atan2(+-0, +-0) follows the signed
            # zeros, so those few points take sin / cos of lon itself
            pole = h == 0.0
            if device is None:
                with np.errstate(divide="ignore", invalid="ignore"):
                    sl, cl = Y / h, X / h
                if pole.any():
                    sl[pole], cl[pole] = np.sin(lon[pole]), np.cos(lon[pole])
            else:
                sl = torch.where(pole, torch.sin(lon), Y / h)
                cl = torch.where(pole, torch.cos(lon), X / h)
            st, ct = Z, h
            elon = (-sl, cl, 0.0)
            elat = (-st * cl, -st * sl, ct)
            self.X[p], self.Y[p], self.Z[p], self.lon[p], self.lat[p] = X, Y, Z, lon, lat
            self.ex_lon[p] = ex[0] * elon[0] + ex[1] * elon[1]
            self.ex_lat[p] = ex[0] * elat[0] + ex[1] * elat[1] + ex[2] * elat[2]
            self.ey_lon[p] = ey[0] * elon[0] + ey[1] * elon[1]
            self.ey_lat[p] = ey[0] * elat[0] + ey[1] * elat[1] + ey[2] * elat[2]
        self.det = self.ex_lon * self.ey_lat - self.ey_lon * self.ex_lat

    def _panels_device(self, c0, c1, c2, ex0, ey0, f64):
        """The loop above for all six panels at once on the device: the signed permutations as a
        gather and a product by +-1 (exact), then the same expressions elementwise over
        [6][ny][nx], so every value has the bits of the per-panel construction."""
        idx = torch.as_tensor(_ROT_IDX, device=f64["device"])
        sgn = torch.as_tensor(_ROT_SGN, **f64)[:, :, None, None]
        W = torch.stack([c0, c1, c2])[idx] * sgn                 # [6][3][ny][nx]
        EX = torch.stack(list(ex0))[idx] * sgn
        EY = torch.stack(list(ey0))[idx] * sgn
        X, Y, Z = W[:, 0], W[:, 1], W[:, 2]
        h = torch.hypot(X, Y)
        lon = torch.atan2(Y, X)
        lat = torch.atan2(Z, h)
        pole = h == 0.0
        sl = torch.where(pole, torch.sin(lon), Y / h)
        cl = torch.where(pole, torch.cos(lon), X / h)
        st, ct = Z, h
        elon0, elon1 = -sl, cl
        elat0, elat1, elat2 = -st * cl, -st * sl, ct
        self.X, self.Y, self.Z, self.lon, self.lat = X, Y, Z, lon, lat
        self.ex_lon = EX[:, 0] * elon0 + EX[:, 1] * elon1
        self.ex_lat = EX[:, 0] * elat0 + EX[:, 1] * elat1 + EX[:, 2] * elat2
        self.ey_lon = EY[:, 0] * elon0 + EY[:, 1] * elon1
        self.ey_lat = EY[:, 0] * elat0 + EY[:, 1] * elat1 + EY[:, 2] * elat2
        self.det = self.ex_lon * self.ey_lat - self.ey_lon * self.ex_lat


class Grid:
    """Equiangular gnomonic cubed sphere with 4 ghost layers (src/cs_datastruct.py:195-541).

    Attributes: N, M = N+8, dx = dy = (pi/2)/N, i0 = j0 = 4, iend = jend = N+4, point sets
    ``pc`` (centres, [6][M][M]), ``pu`` (x-edges, [6][M][M+1]), ``pv`` (y-edges, [6][M+1][M]).
    """

    projection = "gnomonic_equiangular"

    def __init__(self, N, device=None):
        if N < 8:
            raise ValueError("N must be >= 8 (the 4-cell boundary frames of opposite edges may not overlap)")
        self.N = N
        self.M = N + 2 * NG
        self.ng, self.ngl, self.ngr = 2 * NG, NG, NG
        a = PIO4
        self.a = a
        dx = (a - (-a)) / N
        self.dx = self.dy = dx
        self.i0 = self.j0 = NG
        self.iend = self.jend = N + NG
        # same 1-D coordinates as src/cs_datastruct.py:272-273, 298-300
        self.x_pc = np.linspace(-a + dx / 2.0 - NG * dx, a - dx / 2.0 + NG * dx, N + 2 * NG)
        self.x_pu = np.linspace(-a - NG * dx, a + NG * dx, N + 1 + 2 * NG)
        self.device = device
        self.pc = PointSet(self.x_pc, self.x_pc, device)
        self.pu = PointSet(self.x_pu, self.x_pc, device)
        self.pv = PointSet(self.x_pc, self.x_pu, device)
        I = slice(self.i0, self.iend)
        gi = self.pc.g[I, I]
        self.a2 = 6.0 * float((gi * gi).sum())   # sum of g^2 over the 6 panel interiors

    # helpers for the reference layout Q[i, j, p]
    @staticmethod
    def to_reference(a):
        """[..., 6, ny, nx] -> [nx, ny, 6] (reference numpy layout)."""
        return np.ascontiguousarray(np.moveaxis(np.asarray(a), (-3, -2, -1), (-1, -2, -3)))

    @staticmethod
    def from_reference(a):
        return np.ascontiguousarray(np.moveaxis(np.asarray(a), (-1, -2, -3), (-3, -2, -1)))
