"""Pins for the CPU oracle: golden vectors from the reference, closed forms, invariants.

Each test checks one oracle function against something that is not the oracle's own
formula: the reference's golden vectors (tests/golden), a closed form, an exactness
property of the numerical method, or a conservation / symmetry invariant.
"""
import math

import numpy as np
import pytest

from oracle import cases as ocases
from oracle import edges, grid as ogrid, halo, interp, lagrange, recon
from oracle import timestep as ots
from oracle.flux import _flux_axis0
from oracle.operators import divergence

GOLD = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_lagrange_halo.npz"))


# ------------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("N", [8, 16, 48])
def test_lagrange_tables_match_reference_golden(N):
    g = ogrid.CubedSphere(N)
    if N <= 16:
        assert np.array_equal(g.pc.X, GOLD["N%d_pcX" % N])
        assert np.array_equal(g.pc.Z, GOLD["N%d_pcZ" % N])
    w, (kmin, kmax) = lagrange.lagrange_poly_ghostcell_pc(g)
    for s, side in enumerate(("east", "west", "north", "south")):
        assert np.array_equal(w[s], GOLD["N%d_w_%s" % (N, side)])
        assert np.array_equal(kmin[s], GOLD["N%d_kmin_%s" % (N, side)])
        assert np.array_equal(kmax[s], GOLD["N%d_kmax_%s" % (N, side)])


@pytest.mark.parametrize("N", [8, 16])
def test_halo_strips_match_reference_golden(N):
    g = ogrid.CubedSphere(N)
    Q, Qx, Qy = GOLD["N%d_Q" % N], GOLD["N%d_Qx" % N], GOLD["N%d_Qy" % N]
    for got, key in zip(halo.get_halo_data_interpolation(Q, g), ("east", "west", "north", "south")):
        assert np.array_equal(got, GOLD["N%d_halo_%s" % (N, key)])
    n, s = halo.get_halo_data_interpolation_NS(Qx, Qy, g)
    e, w = halo.get_halo_data_interpolation_WE(Qx, Qy, g)
    assert np.array_equal(n, GOLD["N%d_haloNS_north" % N]) and np.array_equal(s, GOLD["N%d_haloNS_south" % N])
    assert np.array_equal(e, GOLD["N%d_haloWE_east" % N]) and np.array_equal(w, GOLD["N%d_haloWE_west" % N])


# ------------------------------------------------------------------ grid closed forms
@pytest.mark.parametrize("N", [8, 32])
def test_grid_metric_matches_closed_form(N):
    """g = 1/(r^3 cos^2 x cos^2 y), r = sqrt(1+tan^2 x+tan^2 y) (src/cs_transform.py:414-418)."""
    g = ogrid.CubedSphere(N)
    for x, y, G in ((g.x_pc, g.y_pc, g.metric_tensor_pc), (g.x_pu, g.y_pc, g.metric_tensor_pu),
                    (g.x_pc, g.y_pv, g.metric_tensor_pv)):
        xx, yy = np.meshgrid(x, y, indexing="ij")
        r = np.sqrt(1 + np.tan(xx) ** 2 + np.tan(yy) ** 2)
        closed = 1.0 / (r ** 3 * np.cos(xx) ** 2 * np.cos(yy) ** 2)
        for p in range(6):
            assert np.allclose(G[:, :, p], closed, rtol=1e-14, atol=0)
    # total area of the sphere from the interior cells
    I = slice(g.i0, g.iend)
    area = np.sum(g.metric_tensor_pc[I, I, :]) * g.dx * g.dy
    assert abs(area - 4 * math.pi) < (math.pi / 2 / N) ** 2  # midpoint quadrature, second order


def test_grid_points_unit_and_conversion_roundtrip():
    g = ogrid.CubedSphere(12)
    for P in (g.pc, g.pu, g.pv):
        assert np.allclose(P.X ** 2 + P.Y ** 2 + P.Z ** 2, 1.0, atol=1e-14)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(g.pu.lon.shape)
    v = rng.standard_normal(g.pu.lon.shape)
    uc, vc = ogrid.latlon_to_contravariant(u, v, g.prod_ex_elon_pu, g.prod_ex_elat_pu, g.prod_ey_elon_pu,
                                           g.prod_ey_elat_pu, g.determinant_ll2contra_pu)
    u2, v2 = ogrid.contravariant_to_latlon(uc, vc, g.prod_ex_elon_pu, g.prod_ex_elat_pu, g.prod_ey_elon_pu,
                                           g.prod_ey_elat_pu)
    assert np.allclose(u2, u, atol=1e-12) and np.allclose(v2, v, atol=1e-12)


def test_inverse_map_recovers_panel_coordinates():
    g = ogrid.CubedSphere(10)
    for p in range(6):
        x, y = ogrid.inverse_equiangular_gnomonic_map(g.pc.X[:, :, p], g.pc.Y[:, :, p], g.pc.Z[:, :, p], p)
        xx, yy = np.meshgrid(g.x_pc, g.y_pc, indexing="ij")
        assert np.allclose(x, xx, atol=1e-14) and np.allclose(y, yy, atol=1e-14)


# ------------------------------------------------------------------ duo-grid ghost fill
def test_ghost_fill_exact_for_cubics_along_neighbour_lines():
    """Degree-3 interpolation reproduces a cubic in the neighbour's along-edge angle exactly."""
    N = 16
    g = ogrid.CubedSphere(N)
    lag = lagrange.lagrange_poly_ghostcell_pc(g)
    P = lambda y: 0.3 - 1.2 * y + 2.5 * y ** 2 - 4.0 * y ** 3
    Q = np.zeros((N + 8, N + 8, 6))
    I = slice(g.i0, g.iend)
    # panel 1 depends on its y angle only: panel 0's east strip interpolates along panel 1's lines
    Q[I, I, 1] = P(g.y_pc[I])[None, :]
    interp.ghost_cell_pc_lagrange_interpolation(Q, g, lag)
    xg, yg = ogrid.inverse_equiangular_gnomonic_map(g.pc.X[g.iend:g.iend + 4, I, 0], g.pc.Y[g.iend:g.iend + 4, I, 0],
                                                    g.pc.Z[g.iend:g.iend + 4, I, 0], 1)
    assert np.allclose(Q[g.iend:g.iend + 4, I, 0], P(yg), atol=1e-13)


def test_ghost_fill_fourth_order_and_geometric_orientation():
    """Ghost values of a smooth global field converge at ~4th order to the field at the ghost
    centres on every side and corner; a wrong flip or neighbour would give O(1) errors."""
    errs = []
    for N in (16, 32):
        g = ogrid.CubedSphere(N)
        lag = lagrange.lagrange_poly_ghostcell_pc(g)
        f = lambda X, Y, Z: np.exp(-3 * ((X - 0.3) ** 2 + (Y + 0.5) ** 2 + (Z - 0.6) ** 2)) + 0.5 * X * Y
        exact = f(g.pc.X, g.pc.Y, g.pc.Z)
        Q = np.zeros_like(exact)
        I = slice(g.i0, g.iend)
        Q[I, I, :] = exact[I, I, :]
        interp.ghost_cell_pc_lagrange_interpolation(Q, g, lag)
        mask = np.ones(Q.shape[:2], bool)
        mask[I, I] = False
        errs.append(np.abs(Q - exact)[mask].max())
    assert errs[0] < 1e-3
    assert errs[0] / errs[1] > 8.0, errs


# ------------------------------------------------------------------ PPM reconstruction
def _cell_averages(poly, edges):
    P = np.polynomial.Polynomial(poly).integ()
    return (P(edges[1:]) - P(edges[:-1])) / (edges[1:] - edges[:-1])


@pytest.mark.parametrize("name,degree", [("PPM-PL07", 4), ("PPM-0", 3)])
def test_ppm_edges_exact_for_polynomials(name, degree):
    """PL07 eqs. 41-42 are exact for degree <= 4, CW84 eq. 1.9 for degree <= 3."""
    n = 24
    h = 0.1
    e = np.arange(n + 1) * h
    rng = np.random.default_rng(1)
    poly = rng.standard_normal(degree + 1)
    Qc = _cell_averages(poly, e)
    Q = np.repeat(Qc[:, None, None], 6, axis=2).repeat(3, axis=1)
    qL, qR = np.zeros_like(Q), np.zeros_like(Q)
    dq, q6 = np.zeros_like(Q), np.zeros_like(Q)
    i0, iend = 4, n - 4
    recon._recon_axis0(Q, qL, qR, dq, q6, name, i0, iend)
    p = np.polynomial.Polynomial(poly)
    c = np.arange(i0 - 1, iend + 1)
    assert np.allclose(qR[c, 0, 0], p(e[c + 1]), atol=1e-12)
    assert np.allclose(qL[c, 0, 0], p(e[c]), atol=1e-12)
    if name == "PPM-PL07":
        bad = _cell_averages(np.r_[poly, 1.0], e)     # degree 5 is not reproduced
        Q[:] = bad[:, None, None]
        recon._recon_axis0(Q, qL, qR, dq, q6, name, i0, iend)
        assert not np.allclose(qR[c, 0, 0], np.polynomial.Polynomial(np.r_[poly, 1.0])(e[c + 1]), atol=1e-9)


@pytest.mark.parametrize("name", ["PPM-CW84", "PPM-L04"])
def test_monotone_ppm_linear_exact_and_bounded(name):
    n = 30
    i0, iend = 4, n - 4
    c = np.arange(i0 - 1, iend + 1)
    # linear data: edge values exact
    Q = np.repeat((0.5 + 0.2 * np.arange(n))[:, None, None], 6, axis=2)
    qL, qR, dq, q6 = (np.zeros_like(Q) for _ in range(4))
    recon._recon_axis0(Q, qL, qR, dq, q6, name, i0, iend)
    assert np.allclose(qR[c, 0, 0], 0.5 + 0.2 * (c + 0.5), atol=1e-13)
    # a step: edge values stay within the neighbouring cell means (no new extrema)
    Q = np.repeat(np.where(np.arange(n) < n // 2, 0.0, 1.0)[:, None, None], 6, axis=2)
    recon._recon_axis0(Q, qL, qR, dq, q6, name, i0, iend)
    assert qL[c].min() >= -1e-14 and qR[c].max() <= 1 + 1e-14


# ------------------------------------------------------------------ upwind flux
def test_flux_is_parabola_average_over_departure_interval():
    """CW84 eq. 1.12: f = (1/c) * integral of the upwind parabola over the swept fraction c."""
    rng = np.random.default_rng(3)
    n, i0, iend = 12, 4, 8
    Q = rng.uniform(0.5, 1.5, (n, 1, 1))
    p = recon.Parabola(n, "PPM-PL07", "x")
    p.q_L, p.q_R = np.zeros((n, 1, 1)), np.zeros((n, 1, 1))
    p.dq, p.q6 = np.zeros((n, 1, 1)), np.zeros((n, 1, 1))
    p.f_L, p.f_R, p.f_upw = np.zeros((n + 1, 1, 1)), np.zeros((n + 1, 1, 1)), np.zeros((n + 1, 1, 1))
    recon._recon_axis0(Q, p.q_L, p.q_R, p.dq, p.q6, "PPM-PL07", i0, iend)
    c = rng.uniform(-0.9, 0.9, (n + 1, 1, 1))
    u = np.sign(c) * 1.0
    upos = u[i0:iend + 1] >= 0
    ones = np.ones((n + 1, 1, 1))
    _flux_axis0(Q, p, upos, c, u, np.ones((n, 1, 1)), ones, "MT-PL07", i0, iend)
    z = np.linspace(0, 1, 20001)
    for e in range(i0, iend + 1):
        ce = c[e, 0, 0]
        cell = e - 1 if ce >= 0 else e
        qL, qR, Qc = p.q_L[cell, 0, 0], p.q_R[cell, 0, 0], Q[cell, 0, 0]
        dqv, q6v = qR - qL, 6 * Qc - 3 * (qR + qL)
        par = lambda s: qL + s * (dqv + q6v * (1 - s))
        zz = (1 - ce) + ce * z if ce >= 0 else -ce * z
        avg = np.trapezoid(par(zz), z)
        assert abs(p.f_upw[e, 0, 0] / u[e, 0, 0] - avg) < 1e-8


# ------------------------------------------------------------------ divergence, fixers, symmetry
def _state(N, dt, **opts):
    g, s = ots.make(N, {"ic": 1, "vf": 1}, dt, **opts)
    ots.init_vars_adv(g, s)
    return g, s


def test_divergence_of_constant_field_converges():
    """One-step divergence of q = 1 under a non-divergent wind (the operator_accuracy pattern,
    src/operator_accuracy.py:29-147).  At a fixed Courant number the time-averaged PPM flux
    leaves an O(dt) = O(h) term, so the error halves when N doubles."""
    vals = []
    for N in (16, 32, 64):
        g, s = _state(N, 0.2 / N, mf=1)
        ots.edges_ghost_cell_treatment_scalar(s.Q, s.Q, g, s)
        divergence(g, s)
        I = slice(g.i0, g.iend)
        vals.append(np.abs(s.div[I, I, :]).max())
    assert vals[0] < 0.05
    assert vals[0] / vals[1] > 1.8 and vals[1] / vals[2] > 1.8, vals


def test_mass_fixer_projection_zero_mean_divergence():
    g, s = ots.make(16, {"ic": 2, "vf": 1}, 0.02)
    ots.init_vars_adv(g, s)
    ots.edges_ghost_cell_treatment_scalar(s.Q, s.Q, g, s)
    divergence(g, s)
    I = slice(g.i0, g.iend)
    m0 = np.sum(s.div[I, I, :] * g.metric_tensor_pc[I, I, :])
    assert abs(m0) < 1e-13 * np.sum(np.abs(s.div[I, I, :] * g.metric_tensor_pc[I, I, :]))


@pytest.mark.parametrize("opts", [dict(mf=3), dict(mf=2, dp=2), dict(et=2, opsplit=3, mt=2, mf=1)])
def test_conservative_options_conserve_mass(opts):
    g, s = ots.make(16, {"ic": 2, "vf": 2}, 0.01, **opts)
    ots.init_vars_adv(g, s)
    m0, _ = ots.mass_computation(s.Q, g, 1.0)
    for k in range(1, 6):
        ots.adv_time_step(g, s, k, k * s.dt)
        ots.update_adv(g, s, k * s.dt)
    m1, change = ots.mass_computation(s.Q, g, m0)
    assert change < 1e-14


def test_zonal_symmetry_equatorial_panels():
    """Polar solid-body rotation of a zonally symmetric field: a 90-degree rotation about the
    axis maps panel p to p+1, so panels 0-3 stay identical to round-off."""
    case = {"vf": "solid_body", "alpha": 0.0, "period": 5.0,
            "tracers": [{"field": "gaussians", "centres": [(0.0, 0.6)], "b": 0.0}]}
    g, s = ots.make(16, case, 0.01)
    ots.init_vars_adv(g, s)
    I = slice(g.i0, g.iend)
    lat = g.pc.lat[I, I, :]
    s.Q[I, I, :] = np.exp(-4 * (lat - 0.3) ** 2)
    for k in range(1, 6):
        ots.adv_time_step(g, s, k, k * s.dt)
    for p in range(1, 4):
        assert np.allclose(s.Q[I, I, p], s.Q[I, I, 0], atol=1e-13)
    # panels 4 and 5 are invariant under 90-degree rotation of their own coordinates
    q4 = s.Q[I, I, 4]
    assert np.allclose(q4, np.rot90(q4), atol=1e-13)


def test_rk2_departure_velocity_closed_form():
    """src/averaged_velocity.py:34-62 on a linear, steady wind: u_avg = u(x - u dt/2)."""
    g, s = ots.make(16, {"ic": 2, "vf": 1}, 0.01, dp=2)
    M = g.N + 8
    x = np.arange(M + 1)[:, None, None] * np.ones((1, M, 6))
    s.U_pu.ucontra[:] = 0.5 + 0.01 * x
    s.U_pu.ucontra_old[:] = s.U_pu.ucontra
    s.U_pv.vcontra[:] = -0.3
    s.U_pv.vcontra_old[:] = -0.3
    ots.time_averaged_velocity(g, s)
    e = slice(g.i0, g.iend + 1)
    u = s.U_pu.ucontra[e]
    a = u * s.dto2 / g.dx
    expect = u - a * 0.01   # linear interpolation of a linear field
    assert np.allclose(s.U_pu.ucontra_averaged[e], expect, atol=1e-14)
    assert np.allclose(s.U_pv.vcontra_averaged[:, g.j0:g.jend + 1], -0.3, atol=1e-15)


def test_edges_extrapolation_and_flux_average_keep_cube_edges_consistent():
    """ET-PL07 makes both panels' parabola edge values equal across a cube edge; MF-AF makes the
    two panels' fluxes through a cube edge equal (with the orientation signs)."""
    g, s = ots.make(12, {"ic": 2, "vf": 1}, 0.01, et=2, opsplit=3, mt=2, mf=2)
    ots.init_vars_adv(g, s)
    rng = np.random.default_rng(5)
    Q = rng.uniform(0.5, 1.5, s.Q.shape)
    recon.ppm_reconstruction(Q, Q, s.px, s.py, g, "ET-PL07")
    J = slice(g.j0, g.jend)
    assert np.array_equal(s.px.q_R[g.iend - 1, J, 0], s.px.q_L[g.i0, J, 1])
    s.px.f_upw[:] = rng.standard_normal(s.px.f_upw.shape)
    s.py.f_upw[:] = rng.standard_normal(s.py.f_upw.shape)
    edges.average_flux_cube_edges(s.px, s.py, g)
    assert np.array_equal(s.px.f_upw[g.iend, J, 0], s.px.f_upw[g.i0, J, 1])
    I = slice(g.i0, g.iend)
    assert np.array_equal(s.py.f_upw[I, g.jend, 1], -s.px.f_upw[g.iend, J, 4])


def test_reference_gaussian_exact_solution_is_the_rodrigues_rotation():
    """The reference's hand-expanded rotation (src/advection_ic.py:222-246) equals the
    Rodrigues rotation about (-sin a, 0, cos a) used for the extension cases."""
    rng = np.random.default_rng(7)
    lon = rng.uniform(-math.pi, math.pi, 200)
    lat = rng.uniform(-1.4, 1.4, 200)
    t = 1.3
    ref = ocases.qexact_ref(lon, lat, t, 2, 1)
    Xr, Yr, Zr = ocases.rotate_back(lon, lat, -45.0 * ocases.DEG2RAD, 2 * math.pi / 5.0 * t)
    ext = ocases.gaussians_xyz(Xr, Yr, Zr, [(math.pi / 4, math.pi / 6)], 10.0)
    assert np.allclose(ref, ext, atol=1e-13)


def test_whole_step_converges():
    """L2 error of the rotated Gaussian after t = 0.5 (the src/advection_error.py procedure with
    dt halved with N): the error drops by more than 2.5x from N=16 to N=32."""
    errs = []
    for N in (16, 32):
        dt = 0.4 / N
        g, s = ots.make(N, {"ic": 2, "vf": 1}, dt)
        nst = int(round(0.5 / dt))
        errs.append(ots.run(g, s, nst)[2])
    assert errs[0] < 2e-3
    assert errs[0] / errs[1] > 2.5, errs


# ------------------------------------------------------------------ vector ghost fill
def test_vector_ghost_fill_converges_to_analytic_wind():
    """src/interpolation.py:347-532: the contravariant winds the vector fill writes outside the
    interior (x-edge ghost rows, y-edge ghost columns and the RK2 departure edges i0-1, iend+1)
    converge to the analytic wind at those points; a wrong orientation, sign or stencil in any
    strip leaves an O(1) error there instead."""
    from oracle import cases as ocases
    from oracle.grid import latlon_to_contravariant
    errs = []
    for N in (16, 32, 64):
        grid, sim = ots.make(N, {"ic": 2, "vf": 1}, 0.01)
        ots.init_vars_adv(grid, sim)
        i0, iend = grid.i0, grid.iend
        ul, vl = ocases.case_wind(sim.case, grid.pu.lon, grid.pu.lat, 0.0)
        uex, _ = latlon_to_contravariant(ul, vl, grid.prod_ex_elon_pu, grid.prod_ex_elat_pu, grid.prod_ey_elon_pu,
                                         grid.prod_ey_elat_pu, grid.determinant_ll2contra_pu)
        ul, vl = ocases.case_wind(sim.case, grid.pv.lon, grid.pv.lat, 0.0)
        _, vex = latlon_to_contravariant(ul, vl, grid.prod_ex_elon_pv, grid.prod_ex_elat_pv, grid.prod_ey_elon_pv,
                                         grid.prod_ey_elat_pv, grid.determinant_ll2contra_pv)
        J, E = slice(i0, iend), slice(i0, iend + 1)
        ghost = list(range(0, i0)) + list(range(iend, grid.N + 2 * grid.ngl))
        rk2 = [i0 - 1, iend + 1]
        U, V = sim.U_pu.ucontra, sim.U_pv.vcontra
        errs.append(max(np.abs(U[E][:, ghost] - uex[E][:, ghost]).max(),
                        np.abs(V[:, E][ghost] - vex[:, E][ghost]).max(),
                        np.abs(U[rk2][:, J] - uex[rk2][:, J]).max(),
                        np.abs(V[:, rk2][J] - vex[:, rk2][J]).max()))
    errs = np.array(errs)
    assert errs[-1] < 1e-5
    assert np.all(errs[:-1] / errs[1:] > 7.0), errs          # third order or better


# ------------------------------------------------------------------ lat-lon output remap
class _Pts:
    def __init__(self, X, Y, Z):
        self.X, self.Y, self.Z = X, Y, Z


def test_remap_sends_every_cell_centre_to_its_own_cell():
    """A cell centre lies inside its cell, so ll2cs must return that cell and panel."""
    from oracle import remap
    N = 16
    g = ogrid.CubedSphere(N)
    I = slice(g.i0, g.iend)
    i, j, p = remap.ll2cs(N, _Pts(g.pc.X[I, I, :], g.pc.Y[I, I, :], g.pc.Z[I, I, :]))
    ii, jj, pp = np.meshgrid(np.arange(N), np.arange(N), np.arange(6), indexing="ij")
    assert np.array_equal(i, ii) and np.array_equal(j, jj) and np.array_equal(p, pp)


def test_remap_panel_is_the_face_the_point_projects_into_and_ties_go_to_the_later_face():
    from oracle import remap
    rng = np.random.default_rng(3)
    v = rng.normal(size=(3, 4000))
    v /= np.linalg.norm(v, axis=0)
    pts = _Pts(v[0], v[1], v[2])
    _, _, p = remap.ll2cs(8, pts)
    for q in range(6):
        m = p == q
        x, y = ogrid.inverse_equiangular_gnomonic_map(v[0][m], v[1][m], v[2][m], q)
        assert np.all(abs(x) <= ogrid.PIO4 * (1 + 1e-12)) and np.all(abs(y) <= ogrid.PIO4 * (1 + 1e-12))
    # ties: |X| = |Y| with X, Y > 0 goes to panel 1; the cube corner X = Y = Z > 0 to panel 4
    s = 1.0 / math.sqrt(2.0)
    c = 1.0 / math.sqrt(3.0)
    _, _, p = remap.ll2cs(8, _Pts(np.array([s, c, -s]), np.array([s, c, -s]), np.array([0.0, c, 0.0])))
    assert list(p) == [1, 4, 3]


def test_remap_of_a_smooth_field_converges_first_order():
    """Nearest-cell sampling of q = sin(lat) errs by at most about half a cell diagonal."""
    from oracle import remap
    ll = remap.LatLonGrid(90, 180)
    errs = []
    for N in (16, 32):
        g = ogrid.CubedSphere(N)
        I = slice(g.i0, g.iend)
        i, j, p = remap.ll2cs(N, ll)
        q_ll = remap.nearest_neighbour(np.sin(g.pc.lat[I, I, :]), i, j, p)
        errs.append(np.abs(q_ll - np.sin(ll.lat)).max())
    assert errs[0] < 1.2 * (math.pi / 2 / 16) and 1.6 < errs[0] / errs[1] < 2.5, errs
