"""PPM edge reconstruction -- oracle restatement (test infrastructure).

Restates src/reconstruction_1d.py:26-394.  Masked numpy assignments of the
reference become ``np.where`` on the same elementwise formulas.  Only the x
direction is written out; the y direction runs the same code on the transposed
view, which is exactly how src/reconstruction_1d.py:199-384 mirrors :26-193.

``px`` / ``py`` are :class:`Parabola` objects holding ``q_L``, ``q_R``, ``dq``, ``q6``.
"""
import numpy as np

from .edges import edges_extrapolation


class Parabola:
    """src/cs_datastruct.py:635-684 (only the arrays the hot path reads)."""

    def __init__(self, M, recon_name, direction):
        self.recon_name = recon_name
        self.q_L = np.zeros((M, M, 6))
        self.q_R = np.zeros((M, M, 6))
        self.dq = np.zeros((M, M, 6))
        self.q6 = np.zeros((M, M, 6))
        if direction == "x":
            self.f_L = np.zeros((M + 1, M, 6))
            self.f_R = np.zeros((M + 1, M, 6))
            self.f_upw = np.zeros((M + 1, M, 6))
        else:
            self.f_L = np.zeros((M, M + 1, 6))
            self.f_R = np.zeros((M, M + 1, 6))
            self.f_upw = np.zeros((M, M + 1, 6))
        self.dF = np.zeros((M, M, 6))


def _recon_axis0(Q, qL, qR, dqa, q6a, recon, i0, iend):
    """Edge values along axis 0 for cells i0-1 .. iend (src/reconstruction_1d.py:26-193).

    ``qL``, ``qR``, ``dqa``, ``q6a`` are written in place (views allowed).
    """
    c = slice(i0 - 1, iend + 1)
    if recon == "PPM-0":
        # CW84 eq. 1.9 (:33-44)
        Qe = np.zeros((Q.shape[0] + 1,) + Q.shape[1:])
        Q1, Q2 = Q[i0 - 1:iend + 2], Q[i0 - 2:iend + 1]
        Q3, Q4 = Q[i0:iend + 3], Q[i0 - 3:iend]
        Qe[i0 - 1:iend + 2] = (7.0 / 12.0) * (Q1 + Q2) - (Q3 + Q4) / 12.0
        qR[c] = Qe[i0:iend + 2]
        qL[c] = Qe[i0 - 1:iend + 1]
    elif recon == "PPM-PL07":
        # PL07 eqs. 41-42 (:46-62)
        a1, a2, a3, a4, a5 = 2.0 / 60.0, -13.0 / 60.0, 47.0 / 60.0, 27.0 / 60.0, -3.0 / 60.0
        q1, q2, q3 = Q[i0 - 3:iend - 1], Q[i0 - 2:iend], Q[i0 - 1:iend + 1]
        q4, q5 = Q[i0:iend + 2], Q[i0 + 1:iend + 3]
        qR[c] = a1 * q1 + a2 * q2 + a3 * q3 + a4 * q4 + a5 * q5
        qL[c] = a5 * q1 + a4 * q2 + a3 * q3 + a2 * q4 + a1 * q5
    elif recon == "PPM-CW84":
        # CW84 eqs. 1.6-1.10 with the overshoot fixes (:64-153)
        dQ = np.zeros_like(Q)
        s = slice(i0 - 2, iend + 2)
        Q1, Q2, Q3 = Q[i0 - 3:iend + 1], Q[i0 - 2:iend + 2], Q[i0 - 1:iend + 3]
        dQ0 = 0.5 * (Q3 - Q1)
        dQ1 = 2.0 * (Q3 - Q2)
        dQ2 = 2.0 * (Q2 - Q1)
        d = np.minimum(np.abs(dQ0), np.abs(dQ1))
        d = np.minimum(d, np.abs(dQ2)) * np.sign(dQ0)
        mask = (Q3 - Q2) * (Q2 - Q1) > 0.0
        dQ[s] = np.where(mask, d, 0.0)
        Qe = np.zeros((Q.shape[0] + 1,) + Q.shape[1:])
        Qe[i0 - 1:iend + 2] = 0.5 * (Q[i0 - 1:iend + 2] + Q[i0 - 2:iend + 1]) - (dQ[i0 - 1:iend + 2] - dQ[i0 - 2:iend + 1]) / 6.0
        qR[c] = Qe[i0:iend + 2]
        qL[c] = Qe[i0 - 1:iend + 1]
        q = Q[c]
        dq_old = qR[c] - qL[c]
        q6_old = 6 * q - 3 * (qR[c] + qL[c])
        dqa[c] = dq_old
        q6a[c] = q6_old
        local_max = (qR[c] - q) * (q - qL[c]) <= 0
        qR[c] = np.where(local_max, q, qR[c])
        qL[c] = np.where(local_max, q, qL[c])
        overshoot = np.abs(dq_old) < np.abs(q6_old)
        r, l = qR[c], qL[c]
        move_left = (r - l) * (q - 0.5 * (r + l)) > ((r - l) ** 2) / 6.0
        move_right = -((r - l) ** 2) / 6.0 > (r - l) * (q - 0.5 * (r + l))
        qL[c] = np.where(overshoot & move_left, 3.0 * q - 2.0 * qR[c], qL[c])
        qR[c] = np.where(overshoot & move_right, 3.0 * q - 2.0 * qL[c], qR[c])
    elif recon == "PPM-L04":
        # L04 appendix B, eqs. B1-B4 (:155-192)
        dQ = np.zeros_like(Q)
        dmono = np.zeros_like(Q)
        s = slice(i0 - 3, iend + 3)
        Qm, Q0, Qp = Q[i0 - 4:iend + 2], Q[i0 - 3:iend + 3], Q[i0 - 2:iend + 4]
        dQ[s] = 0.25 * (Qp - Qm)
        Q_max = np.maximum(np.maximum(Qm, Q0), Qp)
        Q_min = np.minimum(np.minimum(Qm, Q0), Qp)
        dQ_min = Q_max - Q0
        dQ_max = Q0 - Q_min
        dmono[s] = np.minimum(np.minimum(np.abs(dQ[s]), dQ_min), dQ_max) * np.sign(dQ[s])
        Qe = np.zeros((Q.shape[0] + 1,) + Q.shape[1:])
        Qe[i0 - 1:iend + 2] = 0.5 * (Q[i0 - 1:iend + 2] + Q[i0 - 2:iend + 1]) - (dmono[i0 - 1:iend + 2] - dmono[i0 - 2:iend + 1]) / 3.0
        qR[c] = Qe[i0:iend + 2]
        qL[c] = Qe[i0 - 1:iend + 1]
        q0 = Q[c]
        dm = dmono[c]
        qmin = np.minimum(2.0 * np.abs(dm), np.abs(qL[c] - q0)) * np.sign(2.0 * dm)
        qL[c] = q0 - qmin
        qmin = np.minimum(2.0 * np.abs(dm), np.abs(qR[c] - q0)) * np.sign(2.0 * dm)
        qR[c] = q0 + qmin
    else:
        raise ValueError("invalid reconstruction %r" % recon)


def ppm_reconstruction_x(Q, px, grid):
    """src/reconstruction_1d.py:26-193"""
    _recon_axis0(Q, px.q_L, px.q_R, px.dq, px.q6, px.recon_name, grid.i0, grid.iend)


def ppm_reconstruction_y(Q, py, grid):
    """src/reconstruction_1d.py:199-384 (axis 1 = y)."""
    t = (1, 0, 2)
    _recon_axis0(np.transpose(Q, t), np.transpose(py.q_L, t), np.transpose(py.q_R, t),
                 np.transpose(py.dq, t), np.transpose(py.q6, t), py.recon_name, grid.j0, grid.jend)


def ppm_reconstruction(Qx, Qy, px, py, grid, et_name):
    """src/reconstruction_1d.py:387-394"""
    ppm_reconstruction_x(Qx, px, grid)
    ppm_reconstruction_y(Qy, py, grid)
    if et_name == "ET-PL07":
        edges_extrapolation(Qx, Qy, px, py, grid, px.recon_name)
