"""Ghost-cell fills (scalar duo-grid, adjacent-panel copy, vector wind) -- oracle (test infrastructure).

Restates src/interpolation.py:154-532:
  ghost_cell_pc_lagrange_interpolation   :154-315  (ET-DG scalar halo, edges then corners)
  ghost_cells_adjacent_panels            :320-340  (ET-S72 / ET-PL07 copy halo)
  wind_edges2center_cubic_interpolation  :347-430
  wind_center2ghostedge_cubic_interpolation :436-532

Pins: the edge strips through the reference's own Lagrange tables and halo strips (golden
vectors), exactness for cubics along the neighbour lines and 4th-order convergence on every
side; the vector fill by 3rd-order convergence to the analytic wind on every ghost edge.  The
composition of the corners (pass 2 of the scalar fill) and of the vector fill is fixed by
those only up to truncation order: their bitwise agreement with the reference is parity
unpinned.  src/interpolation.py cannot be imported here to write vectors (it imports netCDF4,
and cartopy through src/plot.py, neither of which this image has).
"""
import numpy as np

from .grid import contravariant_to_latlon, latlon_to_contravariant
from .halo import (get_halo_data_interpolation, get_halo_data_interpolation_NS,
                   get_halo_data_interpolation_WE)

_L = np.arange(4)


def _interp_x(support, kmin, w):
    """sum_l w[g,k,l] * support[g, kmin[g,k]+l]  (src/interpolation.py:203-209)."""
    g = np.arange(support.shape[0])[:, None, None]
    data = support[g, kmin[:, :, None] + _L[None, None, :]]
    return np.sum(data * w, axis=2)


def _interp_y(support, kmin, w):
    """sum_l w[k,g,l] * support[kmin[k,g]+l, g]  (src/interpolation.py:229-233)."""
    g = np.arange(support.shape[1])[None, :, None]
    data = support[kmin[:, :, None] + _L[None, None, :], g]
    return np.sum(data * w, axis=2)


def ghost_cell_pc_lagrange_interpolation(Q, grid, lagrange):
    """ET-DG halo fill in place, src/interpolation.py:154-315.

    ``lagrange`` is ``(weights, stencils)`` from :func:`oracle.lagrange.lagrange_poly_ghostcell_pc`.
    """
    (w_e, w_w, w_n, w_s), (kmin, _kmax) = lagrange
    k_e, k_w, k_n, k_s = kmin
    i0, iend, j0, jend, ngr = grid.i0, grid.iend, grid.j0, grid.jend, grid.ngr

    # pass 1: edge strips from the neighbours' interiors (:191-248)
    he, hw, hn, hs = get_halo_data_interpolation(Q, grid)
    for p in range(6):
        Q[iend:iend + ngr, j0:jend, p] = _interp_x(he[:, :, p], k_e, w_e)[:, j0:jend]
        Q[0:i0, j0:jend, p] = _interp_x(hw[:, :, p], k_w, w_w)[:, j0:jend]
        Q[i0:iend, jend:jend + ngr, p] = _interp_y(hn[:, :, p], k_n, w_n)[i0:iend, :]
        Q[i0:iend, 0:j0, p] = _interp_y(hs[:, :, p], k_s, w_s)[i0:iend, :]

    # pass 2: corners from the east / west halo, which now holds the strips (:250-314)
    he, hw, _, _ = get_halo_data_interpolation(Q, grid)
    for p in range(6):
        Q[iend:iend + ngr, jend:jend + ngr, p] = _interp_x(he[:, :, p], k_e[:, jend:jend + ngr], w_e[:, jend:jend + ngr, :])
        Q[iend:iend + ngr, 0:j0, p] = _interp_x(he[:, :, p], k_e[:, 0:j0], w_e[:, 0:j0, :])
        Q[0:i0, jend:jend + ngr, p] = _interp_x(hw[:, :, p], k_w[:, jend:jend + ngr], w_w[:, jend:jend + ngr, :])
        Q[0:i0, 0:j0, p] = _interp_x(hw[:, :, p], k_w[:, 0:j0], w_w[:, 0:j0, :])


def ghost_cells_adjacent_panels(Qx, Qy, grid):
    """ET-S72 / ET-PL07 halo: plain copies of the adjacent panels, src/interpolation.py:320-340.

    All strips are read before any is written; when ``Qx is Qy`` the y-ghost writes
    land last, as in the reference.
    """
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    he, hw = get_halo_data_interpolation_WE(Qx, Qy, grid)
    hn, hs = get_halo_data_interpolation_NS(Qx, Qy, grid)
    Qy[0:i0, :, :] = hw
    Qy[iend:, :, :] = he
    Qx[:, 0:j0, :] = hs
    Qx[:, jend:, :] = hn


def wind_edges2center_cubic_interpolation(U_pc, U_pu, U_pv, grid, lagrange):
    """src/interpolation.py:347-430: edge winds -> centres in the 4-deep boundary frame,
    contravariant -> lat-lon, then duo-grid ghost fills of u_lon and v_lat."""
    i0, iend, j0, jend, ngl = grid.i0, grid.iend, grid.j0, grid.jend, grid.ngl
    a1, a2, a3, a4 = 5.0 / 16.0, 15.0 / 16.0, -5.0 / 16.0, 1.0 / 16.0
    b1, b2, b3, b4 = -1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0
    uc, vc = U_pc.ucontra, U_pc.vcontra
    u, v = U_pu.ucontra, U_pv.vcontra

    # west boundary (:358-373)
    uc[i0, j0:jend, :] = a1 * u[i0, j0:jend, :] + a2 * u[i0 + 1, j0:jend, :] + a3 * u[i0 + 2, j0:jend, :] + a4 * u[i0 + 3, j0:jend, :]
    uc[i0 + 1:i0 + ngl, j0:jend, :] = (b1 * u[i0:i0 - 1 + ngl, j0:jend, :] + b2 * u[i0 + 1:i0 + ngl, j0:jend, :]
                                       + b3 * u[i0 + 2:i0 + 1 + ngl, j0:jend, :] + b4 * u[i0 + 3:i0 + 2 + ngl, j0:jend, :])
    vc[i0:i0 + ngl, j0 + ngl:jend - ngl, :] = (b1 * v[i0:j0 + ngl, i0 + ngl - 1:iend - ngl - 1, :] + b2 * v[i0:i0 + ngl, j0 + ngl:jend - ngl, :]
                                               + b3 * v[i0:i0 + ngl, j0 + ngl + 1:jend - ngl + 1, :] + b4 * v[i0:i0 + ngl, j0 + ngl + 2:jend - ngl + 2, :])
    # east boundary (:375-389)
    uc[iend - 1, j0:jend, :] = a4 * u[iend - 3, j0:jend, :] + a3 * u[iend - 2, j0:jend, :] + a2 * u[iend - 1, j0:jend, :] + a1 * u[iend, j0:jend, :]
    uc[iend - ngl:iend - 1, j0:jend, :] = (b4 * u[iend - ngl - 1:iend - 2, j0:jend, :] + b3 * u[iend - ngl + 0:iend - 1, j0:jend, :]
                                           + b2 * u[iend - ngl + 1:iend, j0:jend, :] + b1 * u[iend - ngl + 2:iend + 1, j0:jend, :])
    vc[iend - ngl:iend, j0 + ngl:jend - ngl, :] = (b1 * v[iend - ngl:iend, i0 + ngl - 1:iend - ngl - 1, :] + b2 * v[iend - ngl:iend, j0 + ngl:jend - ngl, :]
                                                   + b3 * v[iend - ngl:iend, j0 + ngl + 1:jend - ngl + 1, :] + b4 * v[iend - ngl:iend, j0 + ngl + 2:jend - ngl + 2, :])
    # south boundary (:391-404)
    vc[i0:iend, j0, :] = a1 * v[i0:iend, j0, :] + a2 * v[i0:iend, j0 + 1, :] + a3 * v[i0:iend, j0 + 2, :] + a4 * v[i0:iend, j0 + 3, :]
    vc[i0:iend, j0 + 1:j0 + ngl, :] = (b1 * v[i0:iend, j0:j0 - 1 + ngl, :] + b2 * v[i0:iend, j0 + 1:j0 + ngl, :]
                                       + b3 * v[i0:iend, j0 + 2:j0 + 1 + ngl, :] + b4 * v[i0:iend, j0 + 3:j0 + 2 + ngl, :])
    uc[i0 + ngl:iend - ngl, j0:j0 + ngl, :] = (b1 * u[i0 + ngl - 1:iend - ngl - 1, j0:j0 + ngl, :] + b2 * u[i0 + ngl:iend - ngl, j0:j0 + ngl, :]
                                               + b3 * u[i0 + ngl + 1:iend - ngl + 1, j0:j0 + ngl, :] + b4 * u[i0 + ngl + 2:iend - ngl + 2, j0:j0 + ngl, :])
    # north boundary (:406-418)
    vc[i0:iend, jend - 1, :] = a4 * v[i0:iend, jend - 3, :] + a3 * v[i0:iend, jend - 2, :] + a2 * v[i0:iend, jend - 1, :] + a1 * v[i0:iend, jend, :]
    vc[i0:iend, jend - ngl:jend - 1, :] = (b4 * v[i0:iend, jend - ngl - 1:jend - 2, :] + b3 * v[i0:iend, jend - ngl + 0:iend - 1, :]
                                           + b2 * v[i0:iend, jend - ngl + 1:iend, :] + b1 * v[i0:iend, jend - ngl + 2:iend + 1, :])
    uc[i0 + ngl:iend - ngl, jend - ngl:jend, :] = (b1 * u[i0 + ngl - 1:iend - ngl - 1, jend - ngl:jend, :] + b2 * u[i0 + ngl:iend - ngl, jend - ngl:jend, :]
                                                   + b3 * u[i0 + ngl + 1:iend - ngl + 1, jend - ngl:jend, :] + b4 * u[i0 + ngl + 2:iend - ngl + 2, jend - ngl:jend, :])

    # contravariant -> lat-lon on the interior centres (:421-425)
    U_pc.ulon[i0:iend, j0:jend, :], U_pc.vlat[i0:iend, j0:jend, :] = contravariant_to_latlon(
        uc[i0:iend, j0:jend, :], vc[i0:iend, j0:jend, :],
        grid.prod_ex_elon_pc[i0:iend, j0:jend, :], grid.prod_ex_elat_pc[i0:iend, j0:jend, :],
        grid.prod_ey_elon_pc[i0:iend, j0:jend, :], grid.prod_ey_elat_pc[i0:iend, j0:jend, :])

    if grid.projection == "gnomonic_equiangular":
        ghost_cell_pc_lagrange_interpolation(U_pc.ulon, grid, lagrange)
        ghost_cell_pc_lagrange_interpolation(U_pc.vlat, grid, lagrange)


def _to_contra_pu(U_pu, grid, sl):
    U_pu.ucontra[sl], U_pu.vcontra[sl] = latlon_to_contravariant(
        U_pu.ulon[sl], U_pu.vlat[sl], grid.prod_ex_elon_pu[sl], grid.prod_ex_elat_pu[sl],
        grid.prod_ey_elon_pu[sl], grid.prod_ey_elat_pu[sl], grid.determinant_ll2contra_pu[sl])


def _to_contra_pv(U_pv, grid, sl):
    U_pv.ucontra[sl], U_pv.vcontra[sl] = latlon_to_contravariant(
        U_pv.ulon[sl], U_pv.vlat[sl], grid.prod_ex_elon_pv[sl], grid.prod_ex_elat_pv[sl],
        grid.prod_ey_elon_pv[sl], grid.prod_ey_elat_pv[sl], grid.determinant_ll2contra_pv[sl])


def wind_center2ghostedge_cubic_interpolation(U_pc, U_pu, U_pv, grid):
    """src/interpolation.py:436-532: lat-lon centre winds -> ghost edges, then contravariant."""
    i0, iend, j0, jend = grid.i0, grid.iend, grid.j0, grid.jend
    a1, a2 = 9.0 / 16.0, -1.0 / 16.0
    ul, vl = U_pc.ulon, U_pc.vlat

    # x-flux ghost rows (:446-468)
    for sl_j in (slice(None, j0), slice(jend, None)):
        sl = (slice(i0, iend + 1), sl_j, slice(None))
        U_pu.ulon[sl] = a1 * (ul[i0:iend + 1, sl_j, :] + ul[i0 - 1:iend, sl_j, :]) + a2 * (ul[i0 + 1:iend + 2, sl_j, :] + ul[i0 - 2:iend - 1, sl_j, :])
        U_pu.vlat[sl] = a1 * (vl[i0:iend + 1, sl_j, :] + vl[i0 - 1:iend, sl_j, :]) + a2 * (vl[i0 + 1:iend + 2, sl_j, :] + vl[i0 - 2:iend - 1, sl_j, :])
        _to_contra_pu(U_pu, grid, sl)

    # y-flux ghost columns (:474-496)
    for sl_i in (slice(None, i0), slice(iend, None)):
        sl = (sl_i, slice(j0, jend + 1), slice(None))
        U_pv.ulon[sl] = a1 * (ul[sl_i, j0:jend + 1, :] + ul[sl_i, j0 - 1:jend, :]) + a2 * (ul[sl_i, j0 + 1:jend + 2, :] + ul[sl_i, j0 - 2:jend - 1, :])
        U_pv.vlat[sl] = a1 * (vl[sl_i, j0:jend + 1, :] + vl[sl_i, j0 - 1:jend, :]) + a2 * (vl[sl_i, j0 + 1:jend + 2, :] + vl[sl_i, j0 - 2:jend - 1, :])
        _to_contra_pv(U_pv, grid, sl)

    # RK2 departure-point edges, full length (:502-532)
    sl = (i0 - 1, slice(None), slice(None))
    U_pu.ulon[sl] = a1 * (ul[i0 - 2] + ul[i0 - 1]) + a2 * (ul[i0] + ul[i0 - 3])
    U_pu.vlat[sl] = a1 * (vl[i0 - 2] + vl[i0 - 1]) + a2 * (vl[i0] + vl[i0 - 3])
    _to_contra_pu(U_pu, grid, sl)
    sl = (iend + 1, slice(None), slice(None))
    U_pu.ulon[sl] = a1 * (ul[iend] + ul[iend + 1]) + a2 * (ul[iend - 1] + ul[iend + 2])
    U_pu.vlat[sl] = a1 * (vl[iend] + vl[iend + 1]) + a2 * (vl[iend - 1] + vl[iend + 2])
    _to_contra_pu(U_pu, grid, sl)
    sl = (slice(None), j0 - 1, slice(None))
    U_pv.ulon[sl] = a1 * (ul[:, j0 - 2] + ul[:, j0 - 1]) + a2 * (ul[:, j0] + ul[:, j0 - 3])
    U_pv.vlat[sl] = a1 * (vl[:, j0 - 2] + vl[:, j0 - 1]) + a2 * (vl[:, j0] + vl[:, j0 - 3])
    _to_contra_pv(U_pv, grid, sl)
    sl = (slice(None), jend + 1, slice(None))
    U_pv.ulon[sl] = a1 * (ul[:, jend] + ul[:, jend + 1]) + a2 * (ul[:, jend - 1] + ul[:, jend + 2])
    U_pv.vlat[sl] = a1 * (vl[:, jend] + vl[:, jend + 1]) + a2 * (vl[:, jend - 1] + vl[:, iend + 2])
    _to_contra_pv(U_pv, grid, sl)
