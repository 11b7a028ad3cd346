"""Cube-edge treatments inside the flux path -- oracle restatement (test infrastructure).

Restates src/edges_treatment.py:31-347:
  average_parabola_cube_edges :31-76    (ET-PL07 parabola averaging, 12 cube edges)
  edges_extrapolation         :82-178   (ET-PL07, PL07 appendix C eqs. 47 and 49)
  ghost_parabolas             :119-178  (its ghost-cell copies of the averaged parabolas)
  average_flux_cube_edges     :231-278  (MF-AF)
  vector_edges_rk2_copy       :305-347  (ET-S72 / ET-PL07 wind halo for RK2)
The statement order is kept exactly: later statements read values written by
earlier ones.
"""
import numpy as np

fl = np.flip


def average_parabola_cube_edges(px, py, g):
    """src/edges_treatment.py:31-76"""
    i0, j0, iend, jend = g.i0, g.j0, g.iend, g.jend
    J, I = slice(j0, jend), slice(i0, iend)
    px.q_R[iend - 1, J, 0:3] = (px.q_R[iend - 1, J, 0:3] + px.q_L[i0, J, 1:4]) * 0.5
    px.q_L[i0, J, 1:4] = px.q_R[iend - 1, J, 0:3]
    px.q_R[iend - 1, J, 3] = (px.q_R[iend - 1, J, 3] + px.q_L[i0, J, 0]) * 0.5
    px.q_L[i0, J, 0] = px.q_R[iend - 1, J, 3]
    py.q_L[I, j0, 4] = (py.q_L[I, j0, 4] + py.q_R[I, jend - 1, 0]) * 0.5
    py.q_R[I, jend - 1, 0] = py.q_L[I, j0, 4]
    px.q_R[iend - 1, J, 4] = (px.q_R[iend - 1, J, 4] + py.q_R[I, jend - 1, 1]) * 0.5
    py.q_R[I, jend - 1, 1] = px.q_R[iend - 1, J, 4]
    py.q_R[I, jend - 1, 4] = (py.q_R[I, jend - 1, 4] + fl(py.q_R[I, jend - 1, 2])) * 0.5
    py.q_R[I, jend - 1, 2] = fl(py.q_R[I, jend - 1, 4])
    px.q_L[i0, J, 4] = (fl(py.q_R[I, jend - 1, 3]) + px.q_L[i0, J, 4]) * 0.5
    py.q_R[I, jend - 1, 3] = fl(px.q_L[i0, J, 4])
    py.q_R[I, jend - 1, 5] = (py.q_R[I, jend - 1, 5] + py.q_L[I, j0, 0]) * 0.5
    py.q_L[I, j0, 0] = py.q_R[I, jend - 1, 5]
    py.q_L[I, j0, 1] = (py.q_L[I, j0, 1] + fl(px.q_R[iend - 1, J, 5])) * 0.5
    px.q_R[iend - 1, J, 5] = fl(py.q_L[I, j0, 1])
    py.q_L[I, j0, 2] = (py.q_L[I, j0, 2] + fl(py.q_L[I, j0, 5])) * 0.5
    py.q_L[I, j0, 5] = fl(py.q_L[I, j0, 2])
    py.q_L[I, j0, 3] = (px.q_L[i0, J, 5] + py.q_L[I, j0, 3]) * 0.5
    px.q_L[i0, J, 5] = py.q_L[I, j0, 3]


def edges_extrapolation(Qx, Qy, px, py, g, recon_name):
    """src/edges_treatment.py:82-178 (non-overlapped grids)."""
    i0, j0, iend, jend = g.i0, g.j0, g.iend, g.jend
    J, I = slice(j0, jend), slice(i0, iend)
    # PL07 eq. 47 (:88-94)
    px.q_L[i0, J, :] = 1.5 * Qx[i0, J, :] - 0.5 * Qx[i0 + 1, J, :]
    px.q_R[iend - 1, J, :] = 1.5 * Qx[iend - 1, J, :] - 0.5 * Qx[iend - 2, J, :]
    py.q_L[I, j0, :] = 1.5 * Qy[I, j0, :] - 0.5 * Qy[I, j0 + 1, :]
    py.q_R[I, jend - 1, :] = 1.5 * Qy[I, jend - 1, :] - 0.5 * Qy[I, jend - 2, :]
    # PL07 eq. 49 (:96-108)
    px.q_R[i0, J, :] = (3.0 * Qx[i0, J, :] + 11.0 * Qx[i0 + 1, J, :] - 2.0 * (Qx[i0 + 2, J, :] - Qx[i0, J, :])) / 14.0
    px.q_L[i0 + 1, J, :] = px.q_R[i0, J, :]
    px.q_L[iend - 1, J, :] = (3.0 * Qx[iend - 1, J, :] + 11.0 * Qx[iend - 2, J, :] - 2.0 * (Qx[iend - 3, J, :] - Qx[iend - 1, J, :])) / 14.0
    px.q_R[iend - 2, J, :] = px.q_L[iend - 1, J, :]
    py.q_R[I, j0, :] = (3.0 * Qy[I, j0, :] + 11.0 * Qy[I, j0 + 1, :] - 2.0 * (Qy[I, j0 + 2, :] - Qy[I, j0, :])) / 14.0
    py.q_L[I, j0 + 1, :] = py.q_R[I, j0, :]
    py.q_L[I, jend - 1, :] = (3.0 * Qy[I, jend - 1, :] + 11.0 * Qy[I, jend - 2, :] - 2.0 * (Qy[I, jend - 3, :] - Qy[I, jend - 1, :])) / 14.0
    py.q_R[I, jend - 2, :] = py.q_L[I, jend - 1, :]
    if recon_name == "PPM-PL07":
        # (:110-114)
        px.q_R[i0 + 1, J, :] = px.q_L[i0 + 2, J, :]
        px.q_L[iend - 2, J, :] = px.q_R[iend - 3, J, :]
        py.q_R[I, j0 + 1, :] = py.q_L[I, j0 + 2, :]
        py.q_L[I, jend - 2, :] = py.q_R[I, jend - 3, :]

    average_parabola_cube_edges(px, py, g)
    ghost_parabolas(px, py, g)


def ghost_parabolas(px, py, g):
    """src/edges_treatment.py:119-178: ghost-cell copies of the averaged parabolas."""
    i0, j0, iend, jend = g.i0, g.j0, g.iend, g.jend
    J, I = slice(j0, jend), slice(i0, iend)
    px.q_L[iend, J, 0:3] = px.q_L[i0, J, 1:4]
    px.q_R[iend, J, 0:3] = px.q_R[i0, J, 1:4]
    px.q_L[i0 - 1, J, 1:4] = px.q_L[iend - 1, J, 0:3]
    px.q_R[i0 - 1, J, 1:4] = px.q_R[iend - 1, J, 0:3]
    px.q_L[iend, J, 3] = px.q_L[i0, J, 0]
    px.q_R[iend, J, 3] = px.q_R[i0, J, 0]
    px.q_L[i0 - 1, J, 0] = px.q_L[iend - 1, J, 3]
    px.q_R[i0 - 1, J, 0] = px.q_R[iend - 1, J, 3]
    py.q_L[I, jend, 0] = py.q_L[I, j0, 4]
    py.q_R[I, jend, 0] = py.q_R[I, j0, 4]
    py.q_L[I, j0 - 1, 4] = py.q_L[I, jend - 1, 0]
    py.q_R[I, j0 - 1, 4] = py.q_R[I, jend - 1, 0]
    py.q_L[I, jend, 1] = px.q_R[iend - 1, J, 4]
    py.q_R[I, jend, 1] = px.q_L[iend - 1, J, 4]
    px.q_R[iend, J, 4] = py.q_L[I, jend - 1, 1]
    px.q_L[iend, J, 4] = py.q_R[I, jend - 1, 1]
    py.q_L[I, jend, 2] = fl(py.q_R[I, jend - 1, 4])
    py.q_R[I, jend, 2] = fl(py.q_L[I, jend - 1, 4])
    py.q_R[I, jend, 4] = fl(py.q_L[I, jend - 1, 2])
    py.q_L[I, jend, 4] = fl(py.q_R[I, jend - 1, 2])
    py.q_L[I, jend, 3] = fl(px.q_L[i0, J, 4])
    py.q_R[I, jend, 3] = fl(px.q_R[i0, J, 4])
    px.q_L[i0 - 1, J, 4] = fl(py.q_L[I, jend - 1, 3])
    px.q_R[i0 - 1, J, 4] = fl(py.q_R[I, jend - 1, 3])
    py.q_L[I, jend, 5] = py.q_L[I, j0, 0]
    py.q_R[I, jend, 5] = py.q_R[I, j0, 0]
    py.q_L[I, j0 - 1, 0] = py.q_L[I, jend - 1, 5]
    py.q_R[I, j0 - 1, 0] = py.q_R[I, jend - 1, 5]
    px.q_L[iend, J, 5] = fl(py.q_L[I, j0, 1])
    px.q_R[iend, J, 5] = fl(py.q_R[I, j0, 1])
    py.q_L[I, j0 - 1, 1] = fl(px.q_L[iend - 1, J, 5])
    py.q_R[I, j0 - 1, 1] = fl(px.q_R[iend - 1, J, 5])
    py.q_L[I, j0 - 1, 5] = fl(py.q_R[I, j0, 2])
    py.q_R[I, j0 - 1, 5] = fl(py.q_L[I, j0, 2])
    py.q_L[I, j0 - 1, 2] = fl(py.q_R[I, j0, 5])
    py.q_R[I, j0 - 1, 2] = fl(py.q_L[I, j0, 5])
    px.q_L[i0 - 1, J, 5] = py.q_R[I, j0, 3]
    px.q_R[i0 - 1, J, 5] = py.q_L[I, j0, 3]
    py.q_R[I, j0 - 1, 3] = px.q_L[i0, J, 5]
    py.q_L[I, j0 - 1, 3] = px.q_R[i0, J, 5]


def average_flux_cube_edges(px, py, g):
    """MF-AF, src/edges_treatment.py:231-278"""
    i0, j0, iend, jend = g.i0, g.j0, g.iend, g.jend
    J, I = slice(j0, jend), slice(i0, iend)
    a = 0.5
    b = 1.0 - a
    fx, fy = px.f_upw, py.f_upw
    fx[iend, J, 0:3] = a * fx[iend, J, 0:3] + b * fx[i0, J, 1:4]
    fx[i0, J, 1:4] = fx[iend, J, 0:3]
    fx[iend, J, 3] = a * fx[iend, J, 3] + b * fx[i0, J, 0]
    fx[i0, J, 0] = fx[iend, J, 3]
    fy[I, j0, 4] = a * fy[I, j0, 4] + b * fy[I, jend, 0]
    fy[I, jend, 0] = fy[I, j0, 4]
    fx[iend, J, 4] = a * fx[iend, J, 4] - b * fy[I, jend, 1]
    fy[I, jend, 1] = -fx[iend, J, 4]
    fy[I, jend, 4] = a * fy[I, jend, 4] - b * fl(fy[I, jend, 2])
    fy[I, jend, 2] = -fl(fy[I, jend, 4])
    fx[i0, J, 4] = a * fx[i0, J, 4] + b * fl(fy[I, jend, 3])
    fy[I, jend, 3] = fl(fx[i0, J, 4])
    fy[I, jend, 5] = a * fy[I, jend, 5] + b * fy[I, j0, 0]
    fy[I, j0, 0] = fy[I, jend, 5]
    fy[I, j0, 1] = a * fy[I, j0, 1] + b * fl(fx[iend, J, 5])
    fx[iend, J, 5] = fl(fy[I, j0, 1])
    fy[I, j0, 2] = a * fy[I, j0, 2] - b * fl(fy[I, j0, 5])
    fy[I, j0, 5] = -fl(fy[I, j0, 2])
    fy[I, j0, 3] = -a * fx[i0, J, 5] + b * fy[I, j0, 3]
    fx[i0, J, 5] = -fy[I, j0, 3]


def vector_edges_rk2_copy(U_pu, U_pv, g):
    """ET-S72 / ET-PL07 wind halo for RK2, src/edges_treatment.py:305-347"""
    i0, j0, iend, jend = g.i0, g.j0, g.iend, g.jend
    J, I = slice(j0, jend), slice(i0, iend)
    u, v = U_pu.ucontra, U_pv.vcontra
    u[iend + 1, J, 0:3] = u[i0 + 1, J, 1:4]
    u[iend + 1, J, 3] = u[i0 + 1, J, 0]
    u[i0 - 1, J, 1:4] = u[iend - 1, J, 0:3]
    u[i0 - 1, J, 0] = u[iend - 1, J, 3]
    v[I, jend + 1, 0] = v[I, j0 + 1, 4]
    v[I, j0 - 1, 4] = v[I, jend - 1, 0]
    u[iend + 1, J, 4] = -v[I, jend - 1, 1]
    v[I, jend + 1, 1] = -u[iend - 1, J, 4]
    v[I, jend + 1, 4] = -fl(v[I, jend - 1, 2])
    v[I, jend + 1, 2] = -fl(v[I, jend - 1, 4])
    u[i0 - 1, J, 4] = fl(v[I, jend - 1, 3])
    v[I, jend + 1, 3] = fl(u[i0 + 1, J, 4])
    v[I, jend + 1, 5] = v[I, j0 + 1, 0]
    v[I, j0 - 1, 0] = v[I, jend - 1, 5]
    v[I, j0 - 1, 1] = fl(u[iend - 1, J, 5])
    u[iend + 1, J, 5] = fl(v[I, j0 + 1, 1])
    v[I, j0 - 1, 2] = -fl(v[I, j0 + 1, 5])
    v[I, j0 - 1, 5] = -fl(v[I, j0 + 1, 2])
    v[I, j0 - 1, 3] = -u[i0 + 1, J, 5]
    u[i0 - 1, J, 5] = -v[I, j0 + 1, 3]
