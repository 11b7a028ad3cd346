"""PPM upwind fluxes -- oracle restatement (test infrastructure).

Restates src/flux.py:9-128 (``compute_fluxes``, ``numerical_flux_ppm_x/_y``):
metric treatment of the edge values, parabola coefficients, the CW84 eq. 1.12
flux selected by the sign of the contravariant wind, times the time-averaged
wind.  Boolean-mask gathers become ``np.where`` on identical formulas.
"""
import numpy as np

from .recon import ppm_reconstruction


def _flux_axis0(Qx, p, upos, c, uavg, g_pc, g_pu, mt_name, i0, iend):
    """src/flux.py:20-70 on axis 0.  ``c``/``uavg``/``g_pu`` are edge arrays (M+1 along axis 0)."""
    s = slice(i0 - 1, iend + 1)
    if mt_name == "MT-0":
        p.q_L[s] = p.q_L[s] * g_pu[i0 - 1:iend + 1]
        p.q_R[s] = p.q_R[s] * g_pu[i0:iend + 2]
        q = Qx[s] * g_pc[s]
    elif mt_name == "MT-PL07":
        q = Qx[s]
    else:
        raise ValueError("invalid mt %r" % mt_name)
    q_L, q_R = p.q_L[s], p.q_R[s]
    p.dq[s] = q_R - q_L
    p.q6[s] = 6 * q - 3 * (q_R + q_L)

    e = slice(i0, iend + 1)
    cc = c[e]
    # flux from the left cell (:48-54) and from the right cell (:56-62)
    qR, dq, q6 = p.q_R[i0 - 1:iend], p.dq[i0 - 1:iend], p.q6[i0 - 1:iend]
    fL = qR + cc * 0.5 * (q6 - dq) - q6 * cc * cc / 3.0
    qL, dq, q6 = p.q_L[i0:iend + 1], p.dq[i0:iend + 1], p.q6[i0:iend + 1]
    fR = qL - cc * 0.5 * (q6 + dq) - q6 * cc * cc / 3.0
    p.f_L[e] = np.where(upos, fL, p.f_L[e])
    p.f_R[e] = np.where(upos, p.f_R[e], fR)
    # CW84 eq. 1.13 (:64-70)
    p.f_upw[e] = np.where(upos, p.f_L[e], p.f_R[e])
    p.f_upw[e] = p.f_upw[e] * uavg[e]
    if mt_name == "MT-PL07":
        p.f_upw[e] = p.f_upw[e] * g_pu[e]


def numerical_flux_ppm_x(Qx, px, U_pu, cx, grid, mt_name):
    """src/flux.py:20-70"""
    _flux_axis0(Qx, px, U_pu.upos, cx, U_pu.ucontra_averaged, grid.metric_tensor_pc,
                grid.metric_tensor_pu, mt_name, grid.i0, grid.iend)


def numerical_flux_ppm_y(Qy, py, U_pv, cy, grid, mt_name):
    """src/flux.py:75-128 (axis 1)."""
    t = (1, 0, 2)
    tp = _Transposed(py)
    _flux_axis0(np.transpose(Qy, t), tp, np.transpose(U_pv.vpos, t), np.transpose(cy, t),
                np.transpose(U_pv.vcontra_averaged, t), np.transpose(grid.metric_tensor_pc, t),
                np.transpose(grid.metric_tensor_pv, t), mt_name, grid.j0, grid.jend)


class _Transposed:
    """Axis-swapped views of a y parabola so the x code can run on it."""

    def __init__(self, p):
        t = (1, 0, 2)
        for name in ("q_L", "q_R", "dq", "q6", "f_L", "f_R", "f_upw"):
            setattr(self, name, np.transpose(getattr(p, name), t))


def compute_fluxes(Qx, Qy, px, py, U_pu, U_pv, cx, cy, grid, sim):
    """src/flux.py:9-15"""
    ppm_reconstruction(Qx, Qy, px, py, grid, sim.et_name)
    numerical_flux_ppm_x(Qx, px, U_pu, cx, grid, sim.mt_name)
    numerical_flux_ppm_y(Qy, py, U_pv, cy, grid, sim.mt_name)
