"""Lat-lon nearest-cell remap for output -- oracle restatement (test infrastructure).

This module belongs to the CPU oracle. Only tests, the smoke check and the
benchmark's CPU-baseline leg may import it; the product never does.

Reference files restated here:
  src/cs_datastruct.py:569-582   latlon_grid: lat = linspace(-pi/2, pi/2, Nlat),
                                 lon = linspace(-pi, pi, Nlon), meshgrid (lon-major)
  src/interpolation.py:38-139    ll2cs: the panel of each point is the face of its largest
                                 Cartesian component (masks assigned in panel order 0..5, so a
                                 later panel wins a tie), the cell is floor((x + a)/dx),
                                 floor((y + a)/dy) of the inverse equiangular map
  src/interpolation.py:144-149   nearest_neighbour: Q_ll[lon, lat] = Q[ix, jy, panel]
  src/constants.py:36-37         Nlat = 720, Nlon = 2 Nlat
"""
import numpy as np

from .grid import NBFACES, PIO4, inverse_equiangular_gnomonic_map, sph2cart

NLAT, NLON = 720, 1440


class LatLonGrid:
    """src/cs_datastruct.py:569-582"""

    def __init__(self, nlat=NLAT, nlon=NLON):
        self.Nlat, self.Nlon = nlat, nlon
        lat = np.linspace(-np.pi * 0.5, np.pi * 0.5, nlat)
        lon = np.linspace(-np.pi, np.pi, nlon)
        self.lat, self.lon = np.meshgrid(lat, lon)          # [Nlon][Nlat]
        self.X, self.Y, self.Z = sph2cart(self.lon, self.lat)


def ll2cs(N, ll):
    """src/interpolation.py:38-139 for the equiangular grid: (i, j, panel), each [Nlon][Nlat]."""
    X, Y, Z = ll.X, ll.Y, ll.Z
    i = np.zeros(X.shape, dtype=np.uint32)
    j = np.zeros(X.shape, dtype=np.uint32)
    panel = np.zeros(X.shape, dtype=np.uint32)
    a = PIO4
    dx = 2 * a / N
    dy = 2 * a / N
    PM = np.stack([abs(X), abs(Y), abs(Z)], axis=-1).max(axis=-1)     # axis 2 of the [Nlon][Nlat] grids
    panel[np.logical_and(PM == abs(X), X > 0)] = 0
    panel[np.logical_and(PM == abs(Y), Y > 0)] = 1
    panel[np.logical_and(PM == abs(X), X < 0)] = 2
    panel[np.logical_and(PM == abs(Y), Y < 0)] = 3
    panel[np.logical_and(PM == abs(Z), Z > 0)] = 4
    panel[np.logical_and(PM == abs(Z), Z <= 0)] = 5
    for p in range(NBFACES):
        m = panel == p
        x, y = inverse_equiangular_gnomonic_map(X[m], Y[m], Z[m], p)
        i[m] = np.array(np.floor((x - (-a)) / dx), dtype=np.uint32)
        j[m] = np.array(np.floor((y - (-a)) / dy), dtype=np.uint32)
    return i, j, panel


def edge_distance(N, ll):
    """Distance, in cells, of each point's equiangular coordinates from the nearest cell edge
    (min over x and y): where it is below a few ulps, libm differences in atan may pick
    either neighbour."""
    X, Y, Z = ll.X, ll.Y, ll.Z
    a, dx = PIO4, 2 * PIO4 / N
    _, _, panel = ll2cs(N, ll)
    out = np.empty(X.shape)
    for p in range(NBFACES):
        m = panel == p
        x, y = inverse_equiangular_gnomonic_map(X[m], Y[m], Z[m], p)
        fx, fy = (x + a) / dx, (y + a) / dx
        out[m] = np.minimum(abs(fx - np.round(fx)), abs(fy - np.round(fy)))
    return out


def nearest_neighbour(Qint, i, j, panel):
    """src/interpolation.py:144-149; Qint is the interior [N][N][6] (reference layout i, j, p)."""
    Q_ll = np.zeros(i.shape)
    for p in range(NBFACES):
        m = panel == p
        Q_ll[m] = Qint[i[m], j[m], p]
    return Q_ll
