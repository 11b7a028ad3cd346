"""Grid files in the reference's layout: ``grid/<projection>_cs_<N>.nc``.

The reference writes its grid with ``save_grid_netcdf4`` (src/plot.py:520-669) and reads it back
in ``cubed_sphere.__init__`` with ``gridload`` (src/cs_datastruct.py:44-190).  Layout (reference
arrays are ``[i, j, panel]``, n = N + 8):

    scalars  dx, dy, a (f8); ng, ngl, ngr, N, i0, iend, j0, jend (i4); attribute title
    po     (ix,  jy,  panel, coorddim)   cell vertices           ix = jy = n + 1
    pc     (ix2, jy2, panel, coorddim)   cell centres            ix2 = jy2 = n
    pu     (ix,  jy2, panel, coorddim)   x-edge midpoints
    pv     (ix2, jy,  panel, coorddim)   y-edge midpoints
    ex_*, ey_*  the tangent vectors dP/dx, dP/dy at the same points
    coorddim = (X, Y, Z, lon, lat); the tangent vectors' lon, lat are 0, as the reference's
    ``point`` objects hold them (src/sphgeo.py:138-147).

The reference writes NETCDF4_CLASSIC, an HDF5 file; this image has neither netCDF4 nor an HDF5
reader, so ``save`` writes netCDF-3 (64-bit offsets, scipy.io) with the same dimensions,
variables, types and title.  ``netCDF4.Dataset`` opens netCDF-3 files, so the reference's loader
reads them; files the reference wrote itself cannot be read here.

``load`` reads such a file, rebuilds ``Grid(N)`` and checks the file's geometry against it
(positions to 1e-13, tangent vectors to 1e-12 relative): the step uses the closed-form metric
of the generated grid, so a file is accepted only when it describes that same grid.
"""
import os

import numpy as np
from scipy.io import netcdf_file

from .grid import Grid, _rotate

PROJECTION = "gnomonic_equiangular"


def grid_name(N):
    """The reference's grid name (src/cs_datastruct.py:40)."""
    return "%s_cs_%d" % (PROJECTION, N)


def _points(x, y):
    """Positions (X, Y, Z, lon, lat) and tangent vectors (X, Y, Z) at the points (x[i], y[j]),
    each [nx][ny][6] (the reference's [i, j, panel])."""
    xx, yy = np.meshgrid(x, y, indexing="ij")
    tx, ty = np.tan(xx), np.tan(yy)
    D2 = 1.0 + tx**2 + ty**2
    D = np.sqrt(D2)
    invD = 1.0 / D
    c = (invD, invD * tx, invD * ty)
    D3 = D2 * D
    fx, fy = (1.0 + tx * tx) / D3, (1.0 + ty * ty) / D3
    ex0 = (-tx * fx, (1.0 + ty * ty) * fx, -tx * ty * fx)
    ey0 = (-ty * fy, -tx * ty * fy, (1.0 + tx * tx) * fy)
    shape = xx.shape + (6, 5)
    pos, ex, ey = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    for p in range(6):
        X, Y, Z = _rotate(p, *c)
        pos[:, :, p, 0], pos[:, :, p, 1], pos[:, :, p, 2] = X, Y, Z
        pos[:, :, p, 3] = np.arctan2(Y, X)
        pos[:, :, p, 4] = np.arctan2(Z, np.hypot(X, Y))
        for k, v in enumerate(_rotate(p, *ex0)):
            ex[:, :, p, k] = v
        for k, v in enumerate(_rotate(p, *ey0)):
            ey[:, :, p, k] = v
    return pos, ex, ey


def _families(grid):
    """The four point families of the reference's file: name -> (x, y, x dimension, y dimension)."""
    return {"po": (grid.x_pu, grid.x_pu, "ix", "jy"), "pc": (grid.x_pc, grid.x_pc, "ix2", "jy2"),
            "pu": (grid.x_pu, grid.x_pc, "ix", "jy2"), "pv": (grid.x_pc, grid.x_pu, "ix2", "jy")}


def save(grid, path):
    """Write ``grid`` (a ``Grid``) to ``path`` in the reference's layout (netCDF-3)."""
    n = grid.N + grid.ng
    with netcdf_file(path, "w", version=2) as f:
        f.title = grid_name(grid.N)
        # (scalars through [()]: scipy's assignValue indexes a 0-d array with [:])
        for name, val in (("dx", grid.dx), ("dy", grid.dy), ("a", grid.a)):
            f.createVariable(name, "d", ())[()] = val
        for name in ("ng", "ngl", "ngr", "N", "i0", "iend", "j0", "jend"):
            f.createVariable(name, "i", ())[()] = int(getattr(grid, name))
        f.createDimension("panel", 6)
        f.createDimension("ix", n + 1)
        f.createDimension("jy", n + 1)
        f.createDimension("ix2", n)
        f.createDimension("jy2", n)
        f.createDimension("coorddim", 5)
        f.createDimension("r3dim", 3)
        f.createDimension("ed", 4)
        for fam, (x, y, dx_, dy_) in _families(grid).items():
            pos, ex, ey = _points(x, y)
            dims = (dx_, dy_, "panel", "coorddim")
            f.createVariable(fam, "d", dims)[:] = pos
            f.createVariable("ex_" + fam, "d", dims)[:] = ex
            f.createVariable("ey_" + fam, "d", dims)[:] = ey


def load(path, device=None):
    """Read a grid file in the reference's layout and return the ``Grid`` it describes."""
    with netcdf_file(path, "r", mmap=False) as f:
        v = f.variables
        N = int(v["N"].getValue())
        grid = Grid(N, device=device)
        for name in ("ng", "ngl", "ngr", "i0", "iend", "j0", "jend"):
            if int(v[name].getValue()) != int(getattr(grid, name)):
                raise ValueError("%s: %s = %d, the equiangular grid with 4 ghost layers has %d"
                                 % (path, name, int(v[name].getValue()), int(getattr(grid, name))))
        for name in ("dx", "dy", "a"):
            if abs(float(v[name].getValue()) - getattr(grid, name)) > 1e-15:
                raise ValueError("%s: %s differs from the equiangular grid of N=%d" % (path, name, N))
        for fam, (x, y, _, _) in _families(grid).items():
            pos, ex, ey = _points(x, y)
            got = np.asarray(v[fam][:])
            if got.shape != pos.shape:
                raise ValueError("%s: %s has shape %s, expected %s" % (path, fam, got.shape, pos.shape))
            # positions (X, Y, Z) and lat; lon compared on the circle (+-pi are one longitude)
            dxyz = np.abs(got[..., :3] - pos[..., :3]).max()
            dlon = np.abs(np.angle(np.exp(1j * (got[..., 3] - pos[..., 3])))).max()
            dlat = np.abs(got[..., 4] - pos[..., 4]).max()
            if max(dxyz, dlon, dlat) > 1e-13:
                raise ValueError("%s: %s is not the equiangular grid of N=%d (max difference %.3e)"
                                 % (path, fam, N, max(dxyz, dlon, dlat)))
            for tname, t in (("ex_" + fam, ex), ("ey_" + fam, ey)):
                tg = np.asarray(v[tname][:])
                if np.abs(tg[..., :3] - t[..., :3]).max() > 1e-12 * np.abs(t[..., :3]).max():
                    raise ValueError("%s: %s is not the equiangular grid's tangent field" % (path, tname))
    return grid


def cubed_sphere(N, gridload=False, griddir="grid/", device=None, verbose=False):
    """The grid of the reference's ``cubed_sphere(N, "gnomonic_equiangular", showonscreen,
    gridload)`` (src/cs_datastruct.py:25-60): with ``gridload``, ``griddir/<name>.nc`` is read
    when it exists; when it does not, the grid is generated, as the reference does for the
    gnomonic grids."""
    path = os.path.join(griddir, grid_name(N) + ".nc")
    if gridload:
        if os.path.isfile(path):
            if verbose:
                print("Loading grid file " + path)
            return load(path, device=device)
        print("Couldnt find the grid file to be read: ", path)
    return Grid(N, device=device)
