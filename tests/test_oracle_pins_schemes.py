"""Pins for the oracle's split, winds, limiters and cube-edge rules (CPU only).

Each test checks one oracle function against what the mathematics fixes -- an exactness
property, a convergence order, a closed form, a geometric consistency on all twelve cube
edges -- chosen so that a plausible slip (a dropped or sign-flipped term, a wrong factor, a
swapped or flipped neighbour) fails it:

  split_half_steps        SP-L04 / SP-PL07 inner steps are advective forms (src/discrete_operators.py:55-69)
  velocity_adv, div_exact divergence of the reference winds vf 1-4 (src/advection_ic.py:287-329)
  _recon_axis0            CW84 / L04 limiter branches (src/reconstruction_1d.py:64-192)
  vector_edges_rk2_copy   ET-S72 / ET-PL07 RK2 wind halo (src/edges_treatment.py:305-347)
  average_flux_cube_edges MF-AF pairing and signs (src/edges_treatment.py:231-278)
  average_parabola_cube_edges, edges_extrapolation, ghost_parabolas
                          ET-PL07 (src/edges_treatment.py:31-178)
  ghost_cells_adjacent_panels  ET-S72 / ET-PL07 copy halo (src/interpolation.py:320-340)
"""
import math

import numpy as np
import pytest

from oracle import cases as ocases
from oracle import edges, recon
from oracle import timestep as ots
from oracle.grid import latlon_to_contravariant
from oracle.interp import ghost_cells_adjacent_panels
from oracle.operators import split_half_steps

PI = math.pi


def _dt(N, cfl=0.5):
    """dt of Courant number ``cfl`` for the reference's vf=1 wind (max contravariant speed ~1.42 u0)."""
    return cfl * (PI / 2 / N) / (2 * PI / 5 * 1.42)


# ------------------------------------------------------------------ operator splitting
def _half_step_error(N, opsplit, mt):
    g, s = ots.make(N, {"ic": 1, "vf": 1}, _dt(N), opsplit=opsplit, mt=mt, mf=1)
    ots.init_vars_adv(g, s)                           # ic = 1: q = 1
    ots.edges_ghost_cell_treatment_scalar(s.Q, s.Q, g, s)
    Qx, Qy = split_half_steps(g, s)
    I = slice(g.i0, g.iend)
    return max(np.abs(Qx[I, I, :] - 1.0).max(), np.abs(Qy[I, I, :] - 1.0).max())


def test_split_inner_steps_are_advective_forms():
    """q = 1 under the non-divergent vf=1 wind at a fixed Courant number.  AVLT's inner step is
    the flux form gQ + dF/2: it leaves 1 by O(dt) = O(h).  L04's adds (c1 - c2) Q / 2, which
    cancels the one-directional divergence of g u to O(h^2) (L04 eqs. 7-8).  PL07's
    (Q + (Q + dF) / (1 - (c1 - c2))) / 2 with MT-PL07 keeps q = 1 exactly (PL07 eqs. 17-18):
    there dF = -(c1 - c2) for q = 1.  A dropped term or a flipped sign in either leaves O(h)."""
    avlt = [_half_step_error(N, 1, 1) for N in (16, 32, 64)]
    l04 = [_half_step_error(N, 2, 1) for N in (16, 32, 64)]
    pl07 = [_half_step_error(N, 3, 2) for N in (16, 32)]
    assert all(1.8 < avlt[i] / avlt[i + 1] < 2.2 for i in range(2)), avlt
    assert all(l04[i] / l04[i + 1] > 3.5 for i in range(2)), l04
    assert l04[0] < 0.05 * avlt[0]
    assert max(pl07) < 1e-15, pl07


# ------------------------------------------------------------------ reference winds
def _spherical_divergence(vf, lon, lat, t, h):
    """(1/cos lat) (du/dlon + d(v cos lat)/dlat), fourth-order central differences of step h."""
    def u(lo, la):
        return ocases.velocity_adv(lo, la, t, vf)[0]

    def vc(lo, la):
        return ocases.velocity_adv(lo, la, t, vf)[1] * np.cos(la)

    d = lambda f, a, b, da, db: (-f(a + 2 * da, b + 2 * db) + 8 * f(a + da, b + db) - 8 * f(a - da, b - db)
                                 + f(a - 2 * da, b - 2 * db)) / (12 * h)
    return (d(u, lon, lat, h, 0.0) + d(vc, lon, lat, 0.0, h)) / np.cos(lat)


@pytest.mark.parametrize("vf", [1, 2, 3, 4])
def test_reference_winds_have_their_closed_form_divergence(vf):
    """vf 1 (solid body) and 2 (NL10 case 4 plus the translation 2 pi cos(lat) / T) are
    non-divergent; vf 3 (NL10 case 3) has div = 3 sin(lon) sin(lat) cos^2(lat) cos(pi t / T);
    vf 4 has the reference's div_exact.  A wrong factor or sign between the u and v formulas
    leaves an O(1) divergence.  The reference's div_exact returns the vf = 4 formula for vf = 3
    as well (src/advection_ic.py:318-329), which the oracle keeps."""
    rng = np.random.default_rng(vf)
    lon = rng.uniform(-PI, PI, 400)
    lat = rng.uniform(-1.3, 1.3, 400)
    t = 0.7
    if vf <= 2:
        exact = np.zeros_like(lon)
    elif vf == 3:
        exact = 3.0 * np.sin(lon) * np.sin(lat) * np.cos(lat) ** 2 * np.cos(PI * t / 5.0)
    else:
        exact = ocases.div_exact(lon, lat, 4)
    errs = [np.abs(_spherical_divergence(vf, lon, lat, t, h) - exact).max() for h in (4e-3, 2e-3)]
    assert errs[1] < 1e-8, errs
    if vf >= 3:
        assert errs[0] / errs[1] > 12.0 or errs[0] < 1e-11, errs      # fourth order
        assert np.abs(exact).max() > 0.5
    if vf == 3:
        assert np.array_equal(ocases.div_exact(lon, lat, 3), ocases.div_exact(lon, lat, 4))


@pytest.mark.parametrize("vf", [2, 3])
def test_reversing_winds_return_the_gaussian_after_one_period(vf):
    """ic=2 in the time-reversing winds vf 2 and 3: the exact solution at T = 5 is the initial
    Gaussian (qexact_adv, src/advection_ic.py:248-254); the whole run's L2 error converges
    (the convergence study's dt of the vf, halved with N; src/advection_error.py:86-153).  A
    wrong translation speed or time dependence leaves an O(1) error."""
    errs = []
    for N, dt in ((16, {2: 0.0125, 3: 0.00625}[vf]), (32, {2: 0.00625, 3: 0.003125}[vf])):
        g, s = ots.make(N, {"ic": 2, "vf": vf}, dt)
        errs.append(ots.run(g, s, int(round(5.0 / dt)))[2])
    # vf 2's deformation stretches the hill into filaments: at N = 16, 32 its error is still
    # spatial and converges slowly (measured ratio 1.53); vf 3's converges at ~3.2
    bound, ratio = {2: (0.1, 1.4), 3: (0.04, 2.5)}[vf]
    assert errs[0] < bound and errs[0] / errs[1] > ratio, errs


# ------------------------------------------------------------------ CW84 and L04 limiters
def _edges(Q1d, name):
    n = len(Q1d)
    i0, iend = 4, n - 4
    Q = np.repeat(np.asarray(Q1d, float)[:, None, None], 6, axis=2)
    qL, qR, dq, q6 = (np.zeros_like(Q) for _ in range(4))
    recon._recon_axis0(Q, qL, qR, dq, q6, name, i0, iend)
    c = np.arange(i0 - 1, iend + 1)
    return Q[c, 0, 0], qL[c, 0, 0], qR[c, 0, 0]


def _averages(f, n):
    x = np.arange(n + 1, dtype=float)
    F = f(x)                                          # antiderivative at the cell edges
    return F[1:] - F[:-1]


@pytest.mark.parametrize("name", ["PPM-CW84", "PPM-L04"])
def test_limiters_inactive_on_smooth_monotone_data_give_ppm0(name):
    """Where the slope limiters and the overshoot / B3-B4 constraints are inactive (smooth,
    strictly monotone data whose successive differences stay within a factor 3), CW84's edge
    (Q_i + Q_(i-1))/2 - (dQ_i - dQ_(i-1))/6 (CW84 eq. 1.6) with dQ = (Q_(i+1) - Q_(i-1))/2, and
    L04's (Q_i + Q_(i-1))/2 - (dm_i - dm_(i-1))/3 (L04 eq. B2) with dm = (Q_(i+1) - Q_(i-1))/4,
    are both CW84 eq. 1.9, the unlimited PPM-0 edge.  Non-linear data: the /6 and /3 terms
    matter (a wrong divisor fails)."""
    f = lambda x: 2.0 * np.sin(0.08 * x) + 0.02 * x ** 2    # antiderivative of a smooth increasing field
    Q = _averages(f, 32)
    assert np.all(np.diff(Q) > 0)
    _, l0, r0 = _edges(Q, "PPM-0")
    _, l1, r1 = _edges(Q, name)
    assert np.allclose(l1, l0, rtol=0, atol=1e-14) and np.allclose(r1, r0, rtol=0, atol=1e-14)
    # the slope-difference term is far above rounding here (the data are not linear)
    c = np.arange(3, len(Q) - 3)                            # cells i0-1 .. iend
    assert np.abs(r0 - 0.5 * (Q[c] + Q[c + 1])).max() > 1e-5


@pytest.mark.parametrize("name", ["PPM-CW84", "PPM-L04"])
def test_limiters_flatten_local_extrema(name):
    """At a cell that is a strict local extremum both limiters give the constant parabola
    q_L = q_R = Q (CW84 eq. 1.10 first case; L04's dm = 0 there, so B3-B4 collapse)."""
    x = np.arange(30, dtype=float)
    for Q in (1.0 - 0.004 * (x - 14.6) ** 2, np.cos(0.5 * x), np.where(x == 15, 2.0, 1.0) + 0.01 * x):
        q, qL, qR = _edges(Q, name)
        ext = [k for k in range(1, len(q) - 1) if (q[k] - q[k - 1]) * (q[k + 1] - q[k]) < 0]
        assert ext
        for k in ext:
            assert qL[k] == q[k] and qR[k] == q[k], (k, q[k], qL[k], qR[k])


def test_cw84_overshoot_fix_makes_every_parabola_monotone():
    """CW84 eq. 1.10: where the parabola of a monotone cell would overshoot (|dq| < |q6|), one
    edge moves to 3q - 2q_other, so that the parabola's derivative vanishes at that edge and
    keeps one sign inside the cell.  On random monotone data the fix fires and every final
    parabola is monotone; swapping q_L and q_R in either branch leaves overshooting cells."""
    rng = np.random.default_rng(11)
    fired = 0
    for _ in range(40):
        Q = np.cumsum(rng.exponential(1.0, 40) ** 3)
        q, qL, qR = _edges(Q, "PPM-CW84")
        dq, q6 = qR - qL, 6 * q - 3 * (qR + qL)
        d0, d1 = dq + q6, dq - q6                         # derivative of the parabola at 0 and 1
        scale = np.abs(dq).max()
        assert np.all(d0 * d1 >= -1e-12 * scale ** 2)
        fired += int(np.sum(np.isclose(qL, 3 * q - 2 * qR, rtol=0, atol=1e-12 * scale) & (dq != 0)))
        fired += int(np.sum(np.isclose(qR, 3 * q - 2 * qL, rtol=0, atol=1e-12 * scale) & (dq != 0)))
    assert fired > 20


def test_l04_edges_stay_within_twice_the_limited_slope_and_keep_monotone_order():
    """L04 eqs. B3-B4: |q_L - Q| and |q_R - Q| are at most 2 |dm| with dm the monotone slope
    (B1), q_L <= Q <= q_R on increasing data, and a step stays a step (dm = 0 beside it, so the
    cells there are constant: first-order upwind)."""
    rng = np.random.default_rng(12)
    for _ in range(40):
        Q = np.cumsum(rng.exponential(1.0, 40) ** 2)
        q, qL, qR = _edges(Q, "PPM-L04")
        assert np.all(qL <= q + 1e-12) and np.all(q <= qR + 1e-12)
        Qf = np.asarray(Q)
        ii = np.arange(3, 37)
        dm = np.minimum(np.minimum(0.25 * np.abs(Qf[ii + 1] - Qf[ii - 1]), Qf[ii + 1] - Qf[ii]), Qf[ii] - Qf[ii - 1])
        assert np.all(np.abs(qR - q) <= 2 * dm + 1e-12) and np.all(np.abs(q - qL) <= 2 * dm + 1e-12)
    step = np.where(np.arange(30) < 15, 0.0, 1.0)
    q, qL, qR = _edges(step, "PPM-L04")
    assert np.array_equal(qL, q) and np.array_equal(qR, q)


# ------------------------------------------------------------------ cube-edge rules
def _contra_winds(g, wind, t=0.0):
    """Analytic contravariant winds (u at every x-edge, v at every y-edge, ghosts included)."""
    ul, vl = wind(g.pu.lon, g.pu.lat, t)
    u, _ = latlon_to_contravariant(ul, vl, g.prod_ex_elon_pu, g.prod_ex_elat_pu, g.prod_ey_elon_pu,
                                   g.prod_ey_elat_pu, g.determinant_ll2contra_pu)
    ul, vl = wind(g.pv.lon, g.pv.lat, t)
    _, v = latlon_to_contravariant(ul, vl, g.prod_ex_elon_pv, g.prod_ex_elat_pv, g.prod_ey_elon_pv,
                                   g.prod_ey_elat_pv, g.determinant_ll2contra_pv)
    return u, v


def _generic_wind(lon, lat, t):
    """A smooth wind with no symmetry of the cube: solid body about a tilted axis plus vf 4."""
    u1, v1 = ocases.velocity_adv(lon, lat, t, 1)
    u4, v4 = ocases.velocity_adv(lon, lat, t, 4)
    return u1 + 0.3 * u4 + 0.2 * np.cos(lat), v1 + 0.3 * v4


def test_flux_average_leaves_a_consistent_flux_field_unchanged_on_all_cube_edges():
    """MF-AF averages each of the 12 cube edges' fluxes from its two panels, with flips and
    signs for edges between differently oriented panels.  The physical normal mass flux g u^n
    of one wind field through a cube edge is the same from both sides (the edge and its
    coordinate spacing are shared), so a consistent field passes unchanged to rounding; a
    wrong pairing, flip or sign changes it at O(1)."""
    g = ots.make(16, {"ic": 2, "vf": 1}, 0.01, mf=2)[0]
    M = g.N + g.ng
    px, py = recon.Parabola(M, "PPM-PL07", "x"), recon.Parabola(M, "PPM-PL07", "y")
    u, v = _contra_winds(g, _generic_wind)
    fx = g.metric_tensor_pu * u
    fy = g.metric_tensor_pv * v
    px.f_upw[:] = fx
    py.f_upw[:] = fy
    edges.average_flux_cube_edges(px, py, g)
    scale = max(np.abs(fx).max(), np.abs(fy).max())
    assert np.abs(px.f_upw - fx).max() <= 1e-14 * scale
    assert np.abs(py.f_upw - fy).max() <= 1e-14 * scale
    # the flux does vary along every cube edge (the check is not vacuous)
    J = slice(g.j0, g.jend)
    assert np.abs(np.diff(fx[g.iend, J, :], axis=0)).min() > 0


def _edge_values(g, f):
    """Parabola edge values from a field given at points: q_L / q_R of cell i at x-edges i, i+1
    (px) and at y-edges j, j+1 (py)."""
    M = g.N + g.ng
    fu = f(g.pu.X, g.pu.Y, g.pu.Z)
    fv = f(g.pv.X, g.pv.Y, g.pv.Z)
    px = recon.Parabola(M, "PPM-PL07", "x")
    py = recon.Parabola(M, "PPM-PL07", "y")
    px.q_L[:], px.q_R[:] = fu[:M], fu[1:M + 1]
    py.q_L[:], py.q_R[:] = fv[:, :M], fv[:, 1:M + 1]
    return px, py, fu, fv


def _smooth(X, Y, Z):
    return np.exp(-2 * ((X - 0.4) ** 2 + (Y + 0.3) ** 2 + (Z - 0.5) ** 2)) + 0.3 * X * Z - 0.2 * Y


def test_parabola_average_leaves_consistent_edge_values_unchanged_on_all_cube_edges():
    """ET-PL07's averaging of the 12 cube edges (src/edges_treatment.py:31-76) pairs edge values
    that sit at the same point of the sphere: values of one field there pass unchanged."""
    g = ots.make(16, {"ic": 2, "vf": 1}, 0.01)[0]
    px, py, _, _ = _edge_values(g, _smooth)
    before = [a.copy() for a in (px.q_L, px.q_R, py.q_L, py.q_R)]
    edges.average_parabola_cube_edges(px, py, g)
    for a, b in zip((px.q_L, px.q_R, py.q_L, py.q_R), before):
        assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()


def test_ghost_parabolas_follow_the_geometry():
    """The ghost-cell parabolas ET-PL07 copies from the neighbours (src/edges_treatment.py:119-178):
    the ghost's edge on the cube edge gets the neighbour's value at that very point (equal to
    rounding); its far edge gets the neighbour's value one cell inside, which converges to the
    field at the ghost's own far edge at first order.  A swapped q_L / q_R or a missing flip
    leaves O(1) errors."""
    far = []
    for N in (16, 32, 64):
        g = ots.make(N, {"ic": 2, "vf": 1}, 0.01)[0]
        px, py, fu, fv = _edge_values(g, _smooth)
        M = N + g.ng
        exp = [a.copy() for a in (px.q_L, px.q_R, py.q_L, py.q_R)]
        edges.ghost_parabolas(px, py, g)
        i0, iend = g.i0, g.iend
        J = slice(i0, iend)
        near = max(np.abs(px.q_L[iend, J] - fu[iend, J]).max(), np.abs(px.q_R[i0 - 1, J] - fu[i0, J]).max(),
                   np.abs(py.q_L[J, iend] - fv[J, iend]).max(), np.abs(py.q_R[J, i0 - 1] - fv[J, i0]).max())
        assert near <= 1e-14, near
        far.append(max(np.abs(px.q_R[iend, J] - fu[iend + 1, J]).max(), np.abs(px.q_L[i0 - 1, J] - fu[i0 - 1, J]).max(),
                       np.abs(py.q_R[J, iend] - fv[J, iend + 1]).max(), np.abs(py.q_L[J, i0 - 1] - fv[J, i0 - 1]).max()))
        assert M == N + 8 and len(exp) == 4
    assert far[0] < 0.15 and all(far[i] / far[i + 1] > 1.8 for i in range(2)), far


def _linear_near_sides(n_cells, i0, a, b):
    """Cell values a + b |x - centre| (x = cell centre index): linear over the three cells next
    to each panel side, and the same value a + b N / 2 at both sides, so the cube-edge
    averaging keeps it."""
    x = np.arange(n_cells) + 0.5
    return a + b * np.abs(x - (i0 + (n_cells - 2 * i0) / 2.0))


def test_pl07_edge_extrapolation_exact_for_linear_data():
    """PL07 eq. 47 (1.5 Q_0 - 0.5 Q_1 at the side) and eq. 49 ((3 Q_0 + 11 Q_1 - 2 (Q_2 - Q_0)) / 14
    one edge inside) are exact for linear data, on every side of every panel."""
    N = 16
    g, s = ots.make(N, {"ic": 2, "vf": 1}, 0.01, et=2, opsplit=3, mt=2, mf=1)
    ots.init_vars_adv(g, s)
    M, i0, iend = N + 8, g.i0, g.iend
    line = _linear_near_sides(M, i0, 0.7, 0.05)
    Qx = np.repeat(np.repeat(line[:, None], M, axis=1)[:, :, None], 6, axis=2)
    Qy = np.ascontiguousarray(np.transpose(Qx, (1, 0, 2)))
    edge = lambda k: 0.7 + 0.05 * abs(k - (i0 + N / 2.0))         # the linear field at edge index k
    for name in ("PPM-PL07", "PPM-0"):
        s.px.recon_name = s.py.recon_name = name
        recon.ppm_reconstruction(Qx, Qy, s.px, s.py, g, "ET-PL07")
        J = slice(i0, iend)
        for arr, k in ((s.px.q_L[i0, J], i0), (s.px.q_R[iend - 1, J], iend), (s.px.q_R[i0, J], i0 + 1),
                       (s.px.q_L[i0 + 1, J], i0 + 1), (s.px.q_L[iend - 1, J], iend - 1), (s.px.q_R[iend - 2, J], iend - 1),
                       (s.py.q_L[J, i0], i0), (s.py.q_R[J, iend - 1], iend), (s.py.q_R[J, i0], i0 + 1),
                       (s.py.q_L[J, iend - 1], iend - 1)):
            assert np.allclose(arr, edge(k), rtol=0, atol=1e-14), (name, k)


def test_rk2_wind_copy_converges_to_the_analytic_wind():
    """ET-S72 / ET-PL07 with RK2 fill the ghost edge one outside each panel side with the
    neighbour's normal wind one edge inside, with the orientation signs and flips of
    src/edges_treatment.py:305-347.  That is the wind at the ghost edge up to O(h) (the two
    panels' coordinate lines part at first order away from the cube edge); a wrong sign or
    flip leaves an O(1) error."""
    errs = []
    for N in (16, 32, 64):
        g, s = ots.make(N, {"ic": 2, "vf": 1}, 0.01, et=1, dp=2)
        u, v = _contra_winds(g, _generic_wind)
        i0, iend = g.i0, g.iend
        I = slice(i0, iend)
        inner_u = (slice(i0, iend + 1), I)
        inner_v = (I, slice(i0, iend + 1))
        s.U_pu.ucontra[:] = 0.0
        s.U_pv.vcontra[:] = 0.0
        s.U_pu.ucontra[inner_u] = u[inner_u]
        s.U_pv.vcontra[inner_v] = v[inner_v]
        edges.vector_edges_rk2_copy(s.U_pu, s.U_pv, g)
        U, V = s.U_pu.ucontra, s.U_pv.vcontra
        errs.append(max(np.abs(U[[i0 - 1, iend + 1]][:, I] - u[[i0 - 1, iend + 1]][:, I]).max(),
                        np.abs(V[:, [i0 - 1, iend + 1]][I] - v[:, [i0 - 1, iend + 1]][I]).max()))
    scale = np.abs(u).max()
    assert errs[0] < 0.1 * scale and all(errs[i] / errs[i + 1] > 1.8 for i in range(2)), errs


@pytest.mark.parametrize("separate", [False, True], ids=["one-field", "between-passes"])
def test_copy_halo_converges_to_the_field_at_the_ghost_centres(separate):
    """ET-S72 / ET-PL07 copy halo (src/interpolation.py:320-340 with src/halo_data.py): every
    ghost cell of the four side strips takes a neighbour cell that converges to the ghost's own
    position at first order; a wrong neighbour, flip or transpose leaves O(1).  With separate
    arrays (the copy between the two passes, which reads Qx or Qy by panel) both hold the same
    field, so only the geometry is checked here."""
    errs = []
    for N in (16, 32, 64):
        g = ots.make(N, {"ic": 2, "vf": 1}, 0.01, et=1)[0]
        exact = _smooth(g.pc.X, g.pc.Y, g.pc.Z)
        i0, iend = g.i0, g.iend
        I = slice(i0, iend)
        Qx = np.zeros_like(exact)
        Qx[I, I, :] = exact[I, I, :]
        Qy = Qx.copy() if separate else Qx
        ghost_cells_adjacent_panels(Qx, Qy, g)
        err = max(np.abs(Qy[:i0, I] - exact[:i0, I]).max(), np.abs(Qy[iend:, I] - exact[iend:, I]).max(),
                  np.abs(Qx[I, :i0] - exact[I, :i0]).max(), np.abs(Qx[I, iend:] - exact[I, iend:]).max())
        errs.append(err)
    assert errs[0] < 0.3 and all(errs[i] / errs[i + 1] > 1.8 for i in range(2)), errs
