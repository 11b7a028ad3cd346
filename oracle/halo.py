"""Panel connectivity: depth-4 neighbour strips -- oracle restatement (test infrastructure).

Restates src/halo_data.py:15-400.  For every panel ``p`` the four nearest rows or
columns of its east / west / north / south neighbour are copied with the
reference's transposes and flips.  Output shapes follow the reference:
east / west ``(4, M, 6)``, north / south ``(M, 4, 6)``.

``src`` selects which field each strip is read from; the plain halo
(src/halo_data.py:15-185) reads everything from one field, while the
``_NS`` / ``_WE`` variants (src/halo_data.py:191-400) read some panels from ``Qx``
and others from ``Qy``.
"""
import numpy as np

NG = 4


def _strips(Qsel, i0, iend, j0, jend):
    """Common strip table.  ``Qsel(p, side)`` returns the source array for panel ``p``.

    Each entry restates the matching block of src/halo_data.py:33-182.
    """
    east, west, north, south = {}, {}, {}, {}
    # panel 0 (src/halo_data.py:33-52)
    east[0] = Qsel(0, "E")[i0:i0 + NG, :, 1]
    west[0] = Qsel(0, "W")[iend - NG:iend, :, 3]
    north[0] = Qsel(0, "N")[:, j0:j0 + NG, 4]
    south[0] = Qsel(0, "S")[:, jend - NG:jend, 5]
    # panel 1 (src/halo_data.py:54-78)
    east[1] = Qsel(1, "E")[i0:i0 + NG, :, 2]
    west[1] = Qsel(1, "W")[iend - NG:iend, :, 0]
    north[1] = np.flip(np.transpose(Qsel(1, "N")[iend - NG:iend, :, 4]), axis=1)
    south[1] = np.flip(np.transpose(Qsel(1, "S")[iend - NG:iend, :, 5]), axis=0)
    # panel 2 (src/halo_data.py:80-103)
    east[2] = Qsel(2, "E")[i0:i0 + NG, :, 3]
    west[2] = Qsel(2, "W")[iend - NG:iend, :, 1]
    north[2] = np.flip(np.flip(Qsel(2, "N")[:, jend - NG:jend, 4], axis=0), axis=1)
    south[2] = np.flip(np.flip(Qsel(2, "S")[:, j0:j0 + NG, 5], axis=1), axis=0)
    # panel 3 (src/halo_data.py:105-129)
    east[3] = Qsel(3, "E")[i0:i0 + NG, :, 0]
    west[3] = Qsel(3, "W")[iend - NG:iend, :, 2]
    north[3] = np.flip(np.transpose(Qsel(3, "N")[i0:i0 + NG, :, 4]), axis=0)
    south[3] = np.flip(np.transpose(Qsel(3, "S")[i0:i0 + NG, :, 5]), axis=1)
    # panel 4 (src/halo_data.py:131-155)
    east[4] = np.transpose(np.flip(Qsel(4, "E")[:, jend - NG:jend, 1], axis=1))
    west[4] = np.flip(np.transpose(Qsel(4, "W")[:, jend - NG:jend, 3]), axis=1)
    north[4] = np.flip(np.flip(Qsel(4, "N")[:, jend - NG:jend, 2], axis=0), axis=1)
    south[4] = Qsel(4, "S")[:, jend - NG:jend, 0]
    # panel 5 (src/halo_data.py:157-182)
    east[5] = np.flip(np.transpose(Qsel(5, "E")[:, j0:j0 + NG, 1]), axis=1)
    west[5] = np.flip(np.transpose(Qsel(5, "W")[:, j0:j0 + NG, 3]), axis=0)
    north[5] = Qsel(5, "N")[:, j0:j0 + NG, 0]
    south[5] = np.flip(np.flip(Qsel(5, "S")[:, j0:j0 + NG, 2], axis=0), axis=1)
    return east, west, north, south


def _stack(d):
    return np.stack([np.array(d[p]) for p in range(6)], axis=2)


def get_halo_data_interpolation(Q, grid):
    """src/halo_data.py:15-185 -> (east, west, north, south)."""
    e, w, n, s = _strips(lambda p, side: Q, grid.i0, grid.iend, grid.j0, grid.jend)
    return _stack(e), _stack(w), _stack(n), _stack(s)


# which field feeds each strip in the _NS / _WE variants
_NS_SRC = {0: ("x", "x"), 1: ("y", "y"), 2: ("x", "x"), 3: ("y", "y"), 4: ("x", "x"), 5: ("x", "x")}
_WE_SRC = {0: ("y", "y"), 1: ("y", "y"), 2: ("y", "y"), 3: ("y", "y"), 4: ("x", "x"), 5: ("x", "x")}


def get_halo_data_interpolation_NS(Qx, Qy, grid):
    """src/halo_data.py:191-296 -> (north, south)."""
    def sel(p, side):
        return Qx if _NS_SRC[p][0 if side == "N" else 1] == "x" else Qy
    _, _, n, s = _strips(sel, grid.i0, grid.iend, grid.j0, grid.jend)
    return _stack(n), _stack(s)


def get_halo_data_interpolation_WE(Qx, Qy, grid):
    """src/halo_data.py:302-400 -> (east, west)."""
    def sel(p, side):
        return Qx if _WE_SRC[p][0 if side == "E" else 1] == "x" else Qy
    e, w, _, _ = _strips(sel, grid.i0, grid.iend, grid.j0, grid.jend)
    return _stack(e), _stack(w)
