"""FP64 cubed-sphere grid and metric terms -- oracle restatement (test infrastructure).

This module belongs to the CPU oracle. Only tests, the smoke check and the
benchmark's CPU-baseline leg may import it; the product never does.

Layout follows the reference exactly: arrays are indexed ``[i, j, p]`` (numpy C
order, panel index fastest), cell centres have shape ``(N+8, N+8, 6)``, x-edges
``(N+9, N+8, 6)`` and y-edges ``(N+8, N+9, 6)``; 4 ghost layers per side.

Reference files restated here:
  src/sphgeo.py:18-33          sph2cart / cart2sph
  src/sphgeo.py:72-88          tangent_geo_lon / tangent_geo_lat
  src/sphgeo.py:120-133        latlon_to_contravariant / contravariant_to_latlon
  src/cs_transform.py:41-126   equiangular map and its inverse
  src/cs_transform.py:146-238  equidistant map and its inverse
  src/cs_transform.py:245-400  tangent vectors (equidistant, equiangular)
  src/cs_datastruct.py:195-541 grid generation, metric tensor, conversion factors
"""
import math

import numpy as np

NBFACES = 6
PI = math.pi
PIO4 = PI * 0.25


def sph2cart(lon, lat):
    """src/sphgeo.py:18-22"""
    x = np.cos(lat) * np.cos(lon)
    y = np.cos(lat) * np.sin(lon)
    z = np.sin(lat)
    return x, y, z


def cart2sph(X, Y, Z):
    """src/sphgeo.py:28-32"""
    hypotxy = np.hypot(X, Y)
    lat = np.arctan2(Z, hypotxy)
    lon = np.arctan2(Y, X)
    return lon, lat


def tangent_geo_lon(lon):
    """src/sphgeo.py:72-77 (unit eastward vector)."""
    e = np.zeros(lon.shape + (3,))
    e[..., 0] = -np.sin(lon)
    e[..., 1] = np.cos(lon)
    return e


def tangent_geo_lat(lon, lat):
    """src/sphgeo.py:83-88 (unit northward vector)."""
    e = np.zeros(lon.shape + (3,))
    e[..., 0] = -np.sin(lat) * np.cos(lon)
    e[..., 1] = -np.sin(lat) * np.sin(lon)
    e[..., 2] = np.cos(lat)
    return e


def latlon_to_contravariant(u_lon, v_lat, ex_lon, ex_lat, ey_lon, ey_lat, det):
    """src/sphgeo.py:120-125"""
    ucontra = ey_lat * u_lon - ey_lon * v_lat
    vcontra = -ex_lat * u_lon + ex_lon * v_lat
    return ucontra / det, vcontra / det


def contravariant_to_latlon(ucontra, vcontra, ex_lon, ex_lat, ey_lon, ey_lat):
    """src/sphgeo.py:130-133"""
    u_lon = ex_lon * ucontra + ey_lon * vcontra
    v_lat = ex_lat * ucontra + ey_lat * vcontra
    return u_lon, v_lat


def _panels_from_aux(invD, XoD, YoD, shape):
    X = np.zeros(shape + (NBFACES,))
    Y = np.zeros(shape + (NBFACES,))
    Z = np.zeros(shape + (NBFACES,))
    # panel 0..5 orientation, src/cs_transform.py:62-91
    X[..., 0], Y[..., 0], Z[..., 0] = invD, XoD, YoD
    Y[..., 1], X[..., 1], Z[..., 1] = invD, -XoD, YoD
    X[..., 2], Y[..., 2], Z[..., 2] = -invD, -XoD, YoD
    Y[..., 3], X[..., 3], Z[..., 3] = -invD, XoD, YoD
    Z[..., 4], X[..., 4], Y[..., 4] = invD, -YoD, XoD
    Z[..., 5], X[..., 5], Y[..., 5] = -invD, YoD, XoD
    return X, Y, Z


def equiangular_gnomonic_map(x, y):
    """src/cs_transform.py:41-94 (R = 1)."""
    xx, yy = np.meshgrid(x, y, indexing="ij")
    tanx = np.tan(xx)
    tany = np.tan(yy)
    D2 = 1.0 + tanx**2 + tany**2
    D = np.sqrt(D2)
    invD = 1.0 / D
    X, Y, Z = _panels_from_aux(invD, invD * tanx, invD * tany, xx.shape)
    lon, lat = cart2sph(X, Y, Z)
    return X, Y, Z, lon, lat


def equidistant_gnomonic_map(x, y):
    """src/cs_transform.py:146-192 (R = 1)."""
    a = 1.0 / np.sqrt(3.0)
    xx, yy = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(a**2 + xx**2 + yy**2)
    invr = 1.0 / r
    X, Y, Z = _panels_from_aux(invr * a, invr * xx, invr * yy, xx.shape)
    lon, lat = cart2sph(X, Y, Z)
    return X, Y, Z, lon, lat


def inverse_equiangular_gnomonic_map(X, Y, Z, panel):
    """src/cs_transform.py:99-126"""
    if panel == 0:
        tanx, tany = Y / X, Z / X
    elif panel == 1:
        tanx, tany = -X / Y, Z / Y
    elif panel == 2:
        tanx, tany = Y / X, -Z / X
    elif panel == 3:
        tanx, tany = -X / Y, -Z / Y
    elif panel == 4:
        tanx, tany = Y / Z, -X / Z
    elif panel == 5:
        tanx, tany = -Y / Z, -X / Z
    else:
        raise ValueError("invalid panel")
    return np.arctan(tanx), np.arctan(tany)


def equidistant_tg_xdir(x, y):
    """src/cs_transform.py:245-295 (R = 1); x, y are cube-face coordinates."""
    a = 1.0 / np.sqrt(3.0)
    xx, yy = np.meshgrid(x, y, indexing="ij")
    r2 = a**2 + xx**2 + yy**2
    r32 = np.sqrt(r2) ** 3
    invr = 1.0 / r32
    ax, xy, y2, a2 = a * xx, xx * yy, yy * yy, a * a
    s = xx.shape + (NBFACES,)
    X, Y, Z = np.zeros(s), np.zeros(s), np.zeros(s)
    X[..., 0], Y[..., 0], Z[..., 0] = -ax * invr, (a2 + y2) * invr, -xy * invr
    X[..., 1], Y[..., 1], Z[..., 1] = -(a2 + y2) * invr, -ax * invr, -xy * invr
    X[..., 2], Y[..., 2], Z[..., 2] = ax * invr, -(a2 + y2) * invr, -xy * invr
    X[..., 3], Y[..., 3], Z[..., 3] = (a2 + y2) * invr, ax * invr, -xy * invr
    X[..., 4], Y[..., 4], Z[..., 4] = xy * invr, (a2 + y2) * invr, -ax * invr
    X[..., 5], Y[..., 5], Z[..., 5] = -xy * invr, (a2 + y2) * invr, ax * invr
    return X, Y, Z


def equidistant_tg_ydir(x, y):
    """src/cs_transform.py:300-350 (R = 1)."""
    a = 1.0 / np.sqrt(3.0)
    xx, yy = np.meshgrid(x, y, indexing="ij")
    r2 = a**2 + xx**2 + yy**2
    r32 = np.sqrt(r2) ** 3
    invr = 1.0 / r32
    ay, xy, x2, a2 = a * yy, xx * yy, xx * xx, a * a
    s = xx.shape + (NBFACES,)
    X, Y, Z = np.zeros(s), np.zeros(s), np.zeros(s)
    X[..., 0], Y[..., 0], Z[..., 0] = -ay * invr, -xy * invr, (a2 + x2) * invr
    X[..., 1], Y[..., 1], Z[..., 1] = xy * invr, -ay * invr, (a2 + x2) * invr
    X[..., 2], Y[..., 2], Z[..., 2] = ay * invr, xy * invr, (a2 + x2) * invr
    X[..., 3], Y[..., 3], Z[..., 3] = -xy * invr, ay * invr, (a2 + x2) * invr
    X[..., 4], Y[..., 4], Z[..., 4] = -(a2 + x2) * invr, -xy * invr, -ay * invr
    X[..., 5], Y[..., 5], Z[..., 5] = (a2 + x2) * invr, -xy * invr, ay * invr
    return X, Y, Z


def equiangular_tg_xdir(x, y):
    """src/cs_transform.py:365-379: equidistant tangent at a*tan(x), scaled by a/cos^2 x."""
    a = 1.0 / np.sqrt(3.0)
    X, Y, Z = equidistant_tg_xdir(a * np.tan(x), a * np.tan(y))
    xx, _ = np.meshgrid(x, y, indexing="ij")
    cos2x = np.cos(xx) * np.cos(xx)
    for p in range(NBFACES):
        X[..., p], Y[..., p], Z[..., p] = a * X[..., p] / cos2x, a * Y[..., p] / cos2x, a * Z[..., p] / cos2x
    return X, Y, Z


def equiangular_tg_ydir(x, y):
    """src/cs_transform.py:384-398"""
    a = 1.0 / np.sqrt(3.0)
    X, Y, Z = equidistant_tg_ydir(a * np.tan(x), a * np.tan(y))
    _, yy = np.meshgrid(x, y, indexing="ij")
    cos2y = np.cos(yy) * np.cos(yy)
    for p in range(NBFACES):
        X[..., p], Y[..., p], Z[..., p] = a * X[..., p] / cos2y, a * Y[..., p] / cos2y, a * Z[..., p] / cos2y
    return X, Y, Z


class Points:
    """Cartesian and spherical coordinates of one family of grid points (src/sphgeo.py:point)."""

    def __init__(self, X, Y, Z, lon, lat):
        self.X, self.Y, self.Z, self.lon, self.lat = X, Y, Z, lon, lat


class Vec3:
    def __init__(self, X, Y, Z):
        self.X, self.Y, self.Z = X, Y, Z


def _metric(ex, ey):
    """src/cs_datastruct.py:407-416: sqrt(|ex|^2 |ey|^2 - (ex.ey)^2) on panel 0, copied to all."""
    g0 = -(ex.X[..., 0] * ey.X[..., 0] + ex.Y[..., 0] * ey.Y[..., 0] + ex.Z[..., 0] * ey.Z[..., 0]) ** 2 + (
        ex.X[..., 0] ** 2 + ex.Y[..., 0] ** 2 + ex.Z[..., 0] ** 2
    ) * (ey.X[..., 0] ** 2 + ey.Y[..., 0] ** 2 + ey.Z[..., 0] ** 2)
    g0 = np.sqrt(g0)
    g = np.zeros(g0.shape + (NBFACES,))
    for p in range(NBFACES):
        g[..., p] = g0
    return g


def _conversion(ex, ey, pts):
    """src/cs_datastruct.py:456-493: dot products of the map tangents with e_lon / e_lat."""
    elon = tangent_geo_lon(pts.lon)
    elat = tangent_geo_lat(pts.lon, pts.lat)
    ex_lon = ex.X * elon[..., 0] + ex.Y * elon[..., 1] + ex.Z * elon[..., 2]
    ex_lat = ex.X * elat[..., 0] + ex.Y * elat[..., 1] + ex.Z * elat[..., 2]
    ey_lon = ey.X * elon[..., 0] + ey.Y * elon[..., 1] + ey.Z * elon[..., 2]
    ey_lat = ey.X * elat[..., 0] + ey.Y * elat[..., 1] + ey.Z * elat[..., 2]
    det = ex_lon * ey_lat - ey_lon * ex_lat
    return ex_lon, ex_lat, ey_lon, ey_lat, det


class CubedSphere:
    """Extended (ghosted) gnomonic cubed sphere, src/cs_datastruct.py:24-541 (gridload=False path).

    Attributes use the reference's names (``pc``, ``pu``, ``pv``, ``metric_tensor_pc`` ...,
    ``prod_ex_elon_pu`` ..., ``determinant_ll2contra_pv``) so the oracle reads like the
    reference.
    """

    def __init__(self, N, projection="gnomonic_equiangular"):
        self.R = 1.0
        self.N = N
        self.projection = projection
        if projection == "gnomonic_equiangular":
            a = PIO4
            gmap, tgx, tgy = equiangular_gnomonic_map, equiangular_tg_xdir, equiangular_tg_ydir
        elif projection == "gnomonic_equidistant":
            a = self.R / np.sqrt(3.0)
            gmap, tgx, tgy = equidistant_gnomonic_map, equidistant_tg_xdir, equidistant_tg_ydir
        else:
            raise ValueError("invalid grid transformation: %r" % projection)
        self.a = a
        x_min, x_max, y_min, y_max = -a, a, -a, a
        dx = (x_max - x_min) / N
        dy = (y_max - y_min) / N
        self.dx, self.dy = dx, dy
        ngl = ngr = 4
        ng = ngl + ngr
        self.ngl, self.ngr, self.ng = ngl, ngr, ng
        self.i0, self.iend = ngl, ngl + N
        self.j0, self.jend = ngl, ngl + N

        # src/cs_datastruct.py:272-273, 298-300, 330-332
        x_pc = np.linspace(x_min + dx / 2.0 - ngl * dx, x_max - dx / 2.0 + ngr * dx, N + ng)
        y_pc = np.linspace(y_min + dy / 2.0 - ngl * dy, y_max - dy / 2.0 + ngr * dy, N + ng)
        x_pu = np.linspace(x_min - ngl * dx, x_max + ngr * dx, N + 1 + ng)
        y_pv = np.linspace(y_min - ngl * dx, y_max + ngr * dy, N + 1 + ng)
        self.x_pc, self.y_pc, self.x_pu, self.y_pv = x_pc, y_pc, x_pu, y_pv

        self.pc = Points(*gmap(x_pc, y_pc))
        self.pu = Points(*gmap(x_pu, y_pc))
        self.pv = Points(*gmap(x_pc, y_pv))
        self.ex_pc, self.ey_pc = Vec3(*tgx(x_pc, y_pc)), Vec3(*tgy(x_pc, y_pc))
        self.ex_pu, self.ey_pu = Vec3(*tgx(x_pu, y_pc)), Vec3(*tgy(x_pu, y_pc))
        self.ex_pv, self.ey_pv = Vec3(*tgx(x_pc, y_pv)), Vec3(*tgy(x_pc, y_pv))

        self.metric_tensor_pc = _metric(self.ex_pc, self.ey_pc)
        self.metric_tensor_pu = _metric(self.ex_pu, self.ey_pu)
        self.metric_tensor_pv = _metric(self.ex_pv, self.ey_pv)

        (self.prod_ex_elon_pc, self.prod_ex_elat_pc, self.prod_ey_elon_pc, self.prod_ey_elat_pc,
         self.determinant_ll2contra_pc) = _conversion(self.ex_pc, self.ey_pc, self.pc)
        (self.prod_ex_elon_pu, self.prod_ex_elat_pu, self.prod_ey_elon_pu, self.prod_ey_elat_pu,
         self.determinant_ll2contra_pu) = _conversion(self.ex_pu, self.ey_pu, self.pu)
        (self.prod_ex_elon_pv, self.prod_ex_elat_pv, self.prod_ey_elon_pv, self.prod_ey_elat_pv,
         self.determinant_ll2contra_pv) = _conversion(self.ex_pv, self.ey_pv, self.pv)
