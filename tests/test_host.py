"""Host-side product code on the CPU: FP64 tables, topology, problem definitions, par
readers and the C ABI's exported symbols (no device calls)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from cubedsphere import config
from cubedsphere.cases import Case, reference_case, workload, random_solid_body
from cubedsphere.grid import Grid
from cubedsphere.topology import halo_source, lagrange_tables, orientation
from oracle import cases as ocases
from oracle import grid as ogrid
from oracle import lagrange as olagrange

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "ref_lagrange_halo.npz"))
R = Grid.to_reference


def _b6(a):
    return np.broadcast_to(a, (6,) + a.shape)


@pytest.mark.parametrize("N", [8, 16, 17, 40])
def test_grid_tables_match_oracle(N):
    g, o = Grid(N), ogrid.CubedSphere(N)
    assert np.array_equal(R(g.pc.X), o.pc.X) and np.array_equal(R(g.pu.Z), o.pu.Z)
    assert np.array_equal(R(g.pc.lon), o.pc.lon) and np.array_equal(R(g.pv.lat), o.pv.lat)
    for mine, ref in ((g.pc.g, o.metric_tensor_pc), (g.pu.g, o.metric_tensor_pu), (g.pv.g, o.metric_tensor_pv)):
        # the outermost ghost edge of N=8 sits at x = pi/2, where g -> 0: compare on the array scale
        assert np.abs(R(_b6(mine)) - ref).max() <= 3e-15 * np.abs(ref).max()
    for P, sfx in ((g.pc, "pc"), (g.pu, "pu"), (g.pv, "pv")):
        for mine, name in ((P.ex_lon, "prod_ex_elon_"), (P.ex_lat, "prod_ex_elat_"), (P.ey_lon, "prod_ey_elon_"),
                           (P.ey_lat, "prod_ey_elat_"), (P.det, "determinant_ll2contra_")):
            assert np.allclose(R(mine), getattr(o, name + sfx), rtol=0, atol=3e-15)
    I = slice(o.i0, o.iend)
    assert abs(g.a2 - np.sum(o.metric_tensor_pc[I, I, :] ** 2)) < 1e-12 * g.a2


@pytest.mark.parametrize("N", [16, 33])
def test_correction_metric_is_the_ghost_fill_of_the_metric(N):
    """gfix, the metric the step multiplies the mass-fixer total by, is g inside every panel and,
    at each of the 3 ghost layers a step reads (corners included), the oracle's duo-grid ghost
    fill of the metric field itself (src/interpolation.py:154-315) on every one of the six
    panels: a ghost value interpolated from S_l + C g_l is then the stored sum plus C gfix."""
    import torch
    from oracle import interp as ointerp
    from cubedsphere.solver import GridTables
    o = ogrid.CubedSphere(N)
    tab = GridTables(Grid(N), "f64", torch.device("cpu"))
    M = N + 8
    gfix = tab.gfix.numpy()[:, :M].T                       # reference layout [i, j]
    gc = tab.gc.numpy()[:, :M].T
    I = slice(o.i0, o.iend)
    assert np.array_equal(gfix[I, I], gc[I, I])
    G = np.zeros((M, M, 6))
    G[I, I, :] = o.metric_tensor_pc[I, I, :]
    ointerp.ghost_cell_pc_lagrange_interpolation(G, o, olagrange.lagrange_poly_ghostcell_pc(o))
    reach = np.arange(1, o.iend + 3)                       # columns / rows a step reads
    ring = np.zeros((M, M), bool)
    ring[np.ix_(reach, reach)] = True
    ring[I, I] = False
    assert ring.sum() == 12 * N + 36
    for p in range(6):
        assert np.abs(G[:, :, p][ring] - gfix[ring]).max() <= 1e-14 * np.abs(gc).max(), p
    # a plausible mistake, the cell's own metric at the ghost positions, is far off
    assert np.abs(gc[ring] - gfix[ring]).max() > 1e-3


@pytest.mark.parametrize("N", [8, 16, 48])
def test_lagrange_tables_bitwise_equal_reference(N):
    w, k = lagrange_tables(Grid(N))
    assert np.array_equal(w, GOLD["N%d_w_east" % N])
    assert np.array_equal(k, GOLD["N%d_kmin_east" % N])


@pytest.mark.parametrize("N", [24, 96, 384])
def test_lagrange_tables_equal_oracle(N):
    w, k = lagrange_tables(Grid(N))
    wo, (ko, _) = olagrange.lagrange_poly_ghostcell_pc(ogrid.CubedSphere(N))
    assert np.array_equal(w, wo[0]) and np.array_equal(k, ko[0])


@pytest.mark.parametrize("N", [8, 16])
def test_orientation_reproduces_reference_halo(N):
    orient = orientation(N)
    M = N + 8
    Q = GOLD["N%d_Q" % N]
    for side, key in ((0, "east"), (1, "west"), (2, "north"), (3, "south")):
        H = GOLD["N%d_halo_%s" % (N, key)]
        for p in range(6):
            for d in range(4):
                g = d if side in (0, 2) else 3 - d   # reference strips store W/S farthest-first
                for r in range(M):
                    q, i, j = halo_source(orient, p, side, d, r)
                    ref = H[g, r, p] if side in (0, 1) else H[r, g, p]
                    assert Q[i, j, q] == ref


@pytest.mark.parametrize("N", [8, 17, 384, 1024])
def test_orientation_and_stencils_stay_in_bounds(N):
    """Every halo source index lies inside [0, M) for d in 0..3 and r in 0..M-1."""
    o, M = orientation(N), N + 8
    d = np.arange(4)[:, None]
    r = np.arange(M)[None, :]
    for p in range(6):
        for s in range(4):
            q, ib, id_, ir, jb, jd, jr = o[p, s]
            i, j = ib + id_ * d + ir * r, jb + jd * d + jr * r
            assert 0 <= q < 6 and q != p
            assert i.min() >= 0 and i.max() < M and j.min() >= 0 and j.max() < M
    if N <= 384:
        _, k = lagrange_tables(Grid(N))
        assert k.min() >= 0 and (k + 3).max() < M
        assert k[:, 4:N + 4].min() >= 4 and (k[:, 4:N + 4] + 3).max() < N + 4


@pytest.mark.parametrize("vf", [1, 2, 3, 4])
def test_wind_basis_recombines_to_reference_wind(vf):
    """sum_m s_m(t) b_m(lon, lat) equals the reference wind (src/advection_ic.py:287-313)."""
    rng = np.random.default_rng(vf)
    lon, lat = rng.uniform(-math.pi, math.pi, 500), rng.uniform(-1.5, 1.5, 500)
    case = reference_case(2, vf)
    for t in (0.0, 0.37, 2.9):
        s = case.coefficients(t)
        u = sum(sm * fu(lon, lat) for sm, (fu, _) in zip(s, case.basis()))
        v = sum(sm * fv(lon, lat) for sm, (_, fv) in zip(s, case.basis()))
        ur, vr = ocases.velocity_adv(lon, lat, t, vf)
        assert np.allclose(u, ur, atol=1e-13) and np.allclose(v, vr, atol=1e-13)


def test_extension_winds_and_fields_match_oracle():
    rng = np.random.default_rng(11)
    lon, lat = rng.uniform(-math.pi, math.pi, 400), rng.uniform(-1.5, 1.5, 400)
    for name in ("solid_body_polar", "two_hills_tilted", "nl10_deformation"):
        case = workload(name)
        ocase = ocases.from_case(case)
        for t in (0.0, 1.1):
            s = case.coefficients(t)
            u = sum(sm * fu(lon, lat) for sm, (fu, _) in zip(s, case.basis()))
            ur, _ = ocases.case_wind(ocase, lon, lat, t)
            assert np.allclose(u, ur, atol=1e-13), name
            for n in range(case.ntr):
                a = case.q(n, lon, lat, t)
                b = ocases.case_q(ocase, ocases.case_tracers(ocase)[n], lon, lat, t)
                assert np.allclose(a, b, atol=1e-13), (name, n)


def test_coefficient_table_rows_follow_update_order():
    case = reference_case(2, 3)
    tab = case.coefficient_table(4, 0.1)
    assert tab.shape == (4, 2, 1)
    assert tab[0, 0, 0] == tab[0, 1, 0] == 1.0                       # step 1: winds at t=0, old = new
    assert abs(tab[2, 0, 0] - math.cos(math.pi * 0.2 / 5.0)) < 1e-15  # step 3: t_2
    assert abs(tab[2, 1, 0] - math.cos(math.pi * 0.1 / 5.0)) < 1e-15  # old: t_1


def test_random_batched_case_is_seeded():
    a = random_solid_body(np.random.default_rng(5), 3)
    b = random_solid_body(np.random.default_rng(5), 3)
    assert a.wind == b.wind and a.tracers == b.tracers
    assert -math.pi / 2 <= a.wind[1] <= 0.0 and all(5.0 <= t["b"] <= 20.0 for t in a.tracers)


# ------------------------------------------------------------------ par files
def test_par_files_positional(tmp_path):
    (tmp_path / "configuration.par").write_text("a\nb\n48\nc\n2\nd\n0\ne\n1\nf\n5\ng\n2\n")
    (tmp_path / "advection.par").write_text(
        "\n".join(["x", "y", "5", "#", "0.01", "#", "3", "#", "2", "#", "1", "#", "4", "#", "2", "#", "2", "#",
                   "3", "#", "1", "#", "3"]) + "\n")
    gp = config.get_parameters(str(tmp_path))
    assert gp.N == 48 and gp.transformation == "gnomonic_equiangular" and gp.test_case == 5
    ap = config.get_advection_parameters(str(tmp_path))
    assert (ap.Tf, ap.dt, ap.ic, ap.vf, ap.recon, ap.dp, ap.opsplit, ap.et, ap.mt, ap.mf) == \
        (5.0, 0.01, 3, 2, 4, 2, 2, 3, 1, 3)
    config.check_options(ap)


def test_repo_par_files_are_reference_defaults():
    ap = config.get_advection_parameters(os.path.join(ROOT, "par"))
    assert (ap.Tf, ap.dt, ap.ic, ap.vf, ap.tc, ap.recon, ap.dp, ap.opsplit, ap.et, ap.mt, ap.mf) == \
        (5.0, 0.0025, 2, 1, 2, 3, 1, 1, 3, 1, 3)


@pytest.mark.parametrize("bad", [dict(opsplit=3, mt=1), dict(opsplit=1, mt=2), dict(recon=5), dict(et=0)])
def test_invalid_options_raise(bad):
    p = config.AdvectionParams(0.01, 5.0, 1, 2, 1, 3, 1, 1, 3, 1, 3)
    for k, v in bad.items():
        setattr(p, k, v)
    with pytest.raises(ValueError):
        config.check_options(p)


def test_missing_par_file_raises(tmp_path):
    with pytest.raises(FileNotFoundError):
        config.get_advection_parameters(str(tmp_path))


# ------------------------------------------------------------------ C ABI
def test_library_loads_and_exports_every_declared_symbol():
    from cubedsphere import _lib
    header = open(os.path.join(ROOT, "include", "csadv.h")).read()
    declared = sorted(set(re.findall(r"\b(csadv_[a-z_]+)\s*\(", header)))
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.SYMBOLS)
    loaded = _lib.load()
    assert loaded.csadv_abi_version() == _lib.ABI_VERSION == 7
    assert loaded.csadv_partials_per_tracer(384, 0) > 0
    assert loaded.csadv_partials_per_tracer(384, 99) < 0
    rows = []
    while loaded.csadv_tile_rows(len(rows)) > 0:
        rows.append(loaded.csadv_tile_rows(len(rows)))
    assert len(rows) >= 6 and len(set(rows)) == len(rows)


def test_build_flags_of_the_device_units():
    """Every device unit is built for gfx950 without packed FP32 arithmetic; the FP32 fused-step
    units fuse multiply-adds only within one source expression (their one-solve and multi-solve
    kernels, a solve alone and in a batch, must round alike whatever their code layout), and no
    other unit sets the contraction mode."""
    from cubedsphere import build
    cmds = {os.path.basename(c[-1]): c for c in build.compile_commands()}
    hip = {n: c for n, c in cmds.items() if n.endswith(".hip")}
    assert any(n.startswith("step_f32_") for n in hip) and any(n.startswith("step_f64_") for n in hip)
    for name, cmd in hip.items():
        assert "--offload-arch=gfx950" in cmd, name
        k = cmd.index("-packed-fp32-ops")
        assert cmd[k - 3:k] == ["-Xclang", "-target-feature", "-Xclang"], name
        contract = [f for f in cmd if f.startswith("-ffp-contract")]
        assert contract == (["-ffp-contract=on"] if name.startswith("step_f32_") else []), name


def test_descriptor_struct_matches_header_field_order():
    from cubedsphere import _lib
    header = open(os.path.join(ROOT, "include", "csadv.h")).read()
    body = header[header.index("typedef struct csadv_solve_desc"):header.index("} csadv_solve_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\*?\s*([a-zA-Z_0-9]+)\s*(?=[,;])", body.split("{", 1)[1])
    assert names == [f[0] for f in _lib.SolveDesc._fields_]


def test_create_rejects_invalid_descriptor_without_gpu_work():
    from cubedsphere import _lib
    lib = _lib.load()
    d = _lib.SolveDesc()
    d.N, d.ntr, d.ld, d.ldu = 16, 1, 24, 28
    d.recon, d.dp, d.split, d.et, d.mt, d.mf, d.nbasis = 3, 1, 3, 3, 1, 3, 1   # split 3 needs mt 2
    h = ctypes.c_void_p()
    rc = lib.csadv_create(ctypes.byref(d), 1, 0, ctypes.byref(h))
    assert rc == -1
    assert b"opsplit" in lib.csadv_last_error(h)
    lib.csadv_destroy(h)


def test_create_rejects_a_general_step_past_its_addressing_limit():
    """The general step addresses a tracer's Qx and Qy (two [6][M][ld] arrays) from one base
    with 32-bit byte offsets: in FP64 N = 6680 passes that check and N = 6681 is rejected with
    CSADV_E_RANGE before any GPU work; FP32 fits up to the library's N = 8192."""
    from cubedsphere import _lib
    lib = _lib.load()
    for dtype, N, want in ((0, 6680, None), (0, 6681, -4), (1, 8192, None)):
        d = _lib.SolveDesc()
        d.N, d.ntr, d.dtype = N, 1, dtype
        d.ld, d.ldu = (N + 11) // 4 * 4, (N + 12) // 4 * 4
        d.recon, d.dp, d.split, d.et, d.mt, d.mf, d.nbasis = 3, 1, 3, 2, 2, 1, 1
        d.work = 1                                  # (never dereferenced: validation only)
        h = ctypes.c_void_p()
        rc = lib.csadv_create(ctypes.byref(d), 1, 0, ctypes.byref(h))
        msg = lib.csadv_last_error(h)
        lib.csadv_destroy(h)
        if want is None:
            assert b"32-bit" not in msg, msg        # fails later, on the null device pointers
        else:
            assert rc == want and b"32-bit" in msg, (rc, msg)


# ------------------------------------------------------------------ cube-edge pairings
def _side_index(side, N, i0):
    """(array axis, boundary index, ghost index) of a side in reference layout [i, j, p]."""
    iend = i0 + N
    return {0: (0, iend - 1, iend), 1: (0, i0, i0 - 1), 2: (1, iend - 1, iend), 3: (1, i0, i0 - 1)}[side]


def _along(a, axis, idx, p, i0, N):
    """The N interior values along the edge at index idx of axis (other axis interior)."""
    I = slice(i0, i0 + N)
    return a[idx, I, p] if axis == 0 else a[I, idx, p]


def _set_along(a, axis, idx, p, i0, N, v):
    I = slice(i0, i0 + N)
    if axis == 0:
        a[idx, I, p] = v
    else:
        a[I, idx, p] = v


def _generic_parabola_average_and_ghosts(pxL, pxR, pyL, pyR, N, i0, tab):
    """What the device does for ET-PL07 from the pair table: average every cube-edge value
    with the one across the edge, then copy each neighbour boundary parabola into the ghost."""
    arrs = {0: (pxL, pxR), 1: (pxL, pxR), 2: (pyL, pyR), 3: (pyL, pyR)}
    def bval(p, s):                          # boundary edge value facing the cube edge
        L, R = arrs[s]
        axis, b, _ = _side_index(s, N, i0)
        return _along(R if s in (0, 2) else L, axis, b, p, i0, N)
    new = {}
    for p in range(6):
        for s in range(4):
            q, sq, rev, sg = (int(v) for v in tab[p, s])
            nb = bval(q, sq)
            nb = nb[::-1] if rev else nb
            new[(p, s)] = (bval(p, s) + nb) * 0.5
    for (p, s), v in new.items():
        L, R = arrs[s]
        axis, b, _ = _side_index(s, N, i0)
        _set_along(R if s in (0, 2) else L, axis, b, p, i0, N, v)
    ghosts = {}
    for p in range(6):
        for s in range(4):
            q, sq, rev, sg = (int(v) for v in tab[p, s])
            Lq, Rq = arrs[sq]
            axis, b, _ = _side_index(sq, N, i0)
            l, r = _along(Lq, axis, b, q, i0, N), _along(Rq, axis, b, q, i0, N)
            if rev:
                l, r = l[::-1], r[::-1]
            ghosts[(p, s)] = (l, r) if sg > 0 else (r, l)
    for (p, s), (l, r) in ghosts.items():
        L, R = arrs[s]
        axis, _, gidx = _side_index(s, N, i0)
        _set_along(L, axis, gidx, p, i0, N, l)
        _set_along(R, axis, gidx, p, i0, N, r)


def test_cube_edge_pairs_reproduce_et_pl07_averaging_and_ghost_copies():
    from cubedsphere.topology import cube_edge_pairs
    from oracle import edges as oedges
    from oracle.recon import Parabola
    from types import SimpleNamespace
    N = 12
    M, i0 = N + 8, 4
    rng = np.random.default_rng(3)
    g = SimpleNamespace(i0=i0, j0=i0, iend=i0 + N, jend=i0 + N)
    px, py = Parabola(M, "PPM-PL07", "x"), Parabola(M, "PPM-PL07", "y")
    for a in (px.q_L, px.q_R, py.q_L, py.q_R):
        a[:] = rng.standard_normal(a.shape)
    ref = [a.copy() for a in (px.q_L, px.q_R, py.q_L, py.q_R)]
    oedges.average_parabola_cube_edges(px, py, g)
    # the ghost copies of edges_extrapolation follow the averaging; run them through the
    # whole routine on arrays whose boundary cells eq. 47 / 49 leave as they are
    mine = [a.copy() for a in ref]
    _generic_parabola_average_and_ghosts(*mine, N, i0, cube_edge_pairs(N))
    I = slice(i0, i0 + N)
    for got, want in zip(mine, (px.q_L, px.q_R, py.q_L, py.q_R)):
        # boundary rows / columns after averaging
        assert np.array_equal(got[i0, I], want[i0, I]) and np.array_equal(got[i0 + N - 1, I], want[i0 + N - 1, I])
        assert np.array_equal(got[I, i0], want[I, i0]) and np.array_equal(got[I, i0 + N - 1], want[I, i0 + N - 1])


def test_cube_edge_pairs_reproduce_et_pl07_ghost_parabolas():
    from cubedsphere.topology import cube_edge_pairs
    from oracle import edges as oedges
    from oracle.recon import Parabola
    from types import SimpleNamespace
    N = 12
    M, i0 = N + 8, 4
    rng = np.random.default_rng(4)
    g = SimpleNamespace(i0=i0, j0=i0, iend=i0 + N, jend=i0 + N)
    Qx, Qy = rng.standard_normal((M, M, 6)), rng.standard_normal((M, M, 6))
    px, py = Parabola(M, "PPM-PL07", "x"), Parabola(M, "PPM-PL07", "y")
    oedges.edges_extrapolation(Qx, Qy, px, py, g, "PPM-0")      # eq. 47 / 49, averaging, ghosts
    # the same from the pair table: eq. 47 / 49 by the oracle's formulas, then the table
    px2, py2 = Parabola(M, "PPM-PL07", "x"), Parabola(M, "PPM-PL07", "y")
    J = slice(i0, i0 + N)
    iend = i0 + N
    px2.q_L[i0, J] = 1.5 * Qx[i0, J] - 0.5 * Qx[i0 + 1, J]
    px2.q_R[iend - 1, J] = 1.5 * Qx[iend - 1, J] - 0.5 * Qx[iend - 2, J]
    py2.q_L[J, i0] = 1.5 * Qy[J, i0] - 0.5 * Qy[J, i0 + 1]
    py2.q_R[J, iend - 1] = 1.5 * Qy[J, iend - 1] - 0.5 * Qy[J, iend - 2]
    px2.q_R[i0, J] = (3.0 * Qx[i0, J] + 11.0 * Qx[i0 + 1, J] - 2.0 * (Qx[i0 + 2, J] - Qx[i0, J])) / 14.0
    px2.q_L[iend - 1, J] = (3.0 * Qx[iend - 1, J] + 11.0 * Qx[iend - 2, J] - 2.0 * (Qx[iend - 3, J] - Qx[iend - 1, J])) / 14.0
    py2.q_R[J, i0] = (3.0 * Qy[J, i0] + 11.0 * Qy[J, i0 + 1] - 2.0 * (Qy[J, i0 + 2] - Qy[J, i0])) / 14.0
    py2.q_L[J, iend - 1] = (3.0 * Qy[J, iend - 1] + 11.0 * Qy[J, iend - 2] - 2.0 * (Qy[J, iend - 3] - Qy[J, iend - 1])) / 14.0
    _generic_parabola_average_and_ghosts(px2.q_L, px2.q_R, py2.q_L, py2.q_R, N, i0, cube_edge_pairs(N))
    for a, b in ((px.q_L, px2.q_L), (px.q_R, px2.q_R)):
        for idx in (i0 - 1, i0, iend - 1, iend):
            assert np.array_equal(a[idx, J], b[idx, J]), idx
    for a, b in ((py.q_L, py2.q_L), (py.q_R, py2.q_R)):
        for idx in (i0 - 1, i0, iend - 1, iend):
            assert np.array_equal(a[J, idx], b[J, idx]), idx


def test_cube_edge_pairs_reproduce_mf_af():
    from cubedsphere.topology import cube_edge_pairs
    from oracle import edges as oedges
    from oracle.recon import Parabola
    from types import SimpleNamespace
    N = 10
    M, i0 = N + 8, 4
    iend = i0 + N
    rng = np.random.default_rng(5)
    g = SimpleNamespace(i0=i0, j0=i0, iend=iend, jend=iend)
    px, py = Parabola(M, "PPM-PL07", "x"), Parabola(M, "PPM-PL07", "y")
    px.f_upw[:] = rng.standard_normal(px.f_upw.shape)
    py.f_upw[:] = rng.standard_normal(py.f_upw.shape)
    fx, fy = px.f_upw.copy(), py.f_upw.copy()
    oedges.average_flux_cube_edges(px, py, g)
    tab = cube_edge_pairs(N)
    J = slice(i0, iend)
    edge = {0: (fx, 0, iend), 1: (fx, 0, i0), 2: (fy, 1, iend), 3: (fy, 1, i0)}
    def val(p, s):
        a, axis, e = edge[s]
        return a[e, J, p] if axis == 0 else a[J, e, p]
    new = {}
    for p in range(6):
        for s in range(4):
            q, sq, rev, sg = (int(v) for v in tab[p, s])
            nb = val(q, sq)[::-1] if rev else val(q, sq)
            new[(p, s)] = 0.5 * val(p, s) + 0.5 * (sg * nb)
    for (p, s), v in new.items():
        a, axis, e = edge[s]
        if axis == 0:
            a[e, J, p] = v
        else:
            a[J, e, p] = v
    assert np.array_equal(fx, px.f_upw) and np.array_equal(fy, py.f_upw)


def test_orientation_reproduces_adjacent_panel_copy_halo():
    """ET-S72 / ET-PL07 halo (src/interpolation.py:320-340) from the orientation maps: Qy's
    x-ghost columns and Qx's y-ghost rows are the neighbours' cells, read from Qx or Qy as
    src/halo_data.py:191-400 selects, all sources read before any ghost is written."""
    from oracle.interp import ghost_cells_adjacent_panels
    from types import SimpleNamespace
    N = 10
    M, i0 = N + 8, 4
    iend = i0 + N
    rng = np.random.default_rng(6)
    g = SimpleNamespace(i0=i0, j0=i0, iend=iend, jend=iend)
    o = orientation(N)
    for same in (True, False):
        Qx = rng.standard_normal((M, M, 6))
        Qy = Qx if same else rng.standard_normal((M, M, 6))
        ref_x, ref_y = Qx.copy(), (Qx.copy() if same else Qy.copy())
        ghost_cells_adjacent_panels(ref_x, ref_x if same else ref_y, g)
        src_x, src_y = Qx.copy(), Qy.copy()
        out_x = Qx.copy()
        out_y = out_x if same else Qy.copy()
        # sides E, W fill Qy's columns, N, S fill Qx's rows; panels 4, 5 read E/W from Qx,
        # panels 1, 3 read N/S from Qy (the _WE / _NS selections)
        writes = []
        for p in range(6):
            for side in (0, 1, 2, 3):
                use_x = (p in (4, 5)) if side in (0, 1) else (p not in (1, 3))
                src = src_x if use_x else src_y
                for d in range(4):
                    for r in range(M):
                        q, i, j = halo_source(o, p, side, d, r)
                        ti, tj = {0: (iend + d, r), 1: (i0 - 1 - d, r), 2: (r, iend + d), 3: (r, i0 - 1 - d)}[side]
                        writes.append((side, p, ti, tj, src[i, j, q]))
        for side in (0, 1, 2, 3):                 # x-ghost columns first, then y-ghost rows
            for s_, p, ti, tj, v in writes:
                if s_ == side:
                    (out_y if side in (0, 1) else out_x)[ti, tj, p] = v
        assert np.array_equal(out_x, ref_x)
        if not same:
            assert np.array_equal(out_y, ref_y)


def test_study_rates_and_printout_follow_the_reference(capsys):
    """Rates are |log E_i - log E_(i-1)| / log 2 (src/advection_error.py:163-164): halving the
    error is rate 1 and growing it fourfold is rate 2, not -2.  The run printout has the layout
    of src/errors.py:84-94."""
    from cubedsphere.advection import STUDY_SCHEMES, convergence_rates, print_errors_simul
    err = np.array([[1.0, 2.0, 4.0], [0.5, 0.5, 16.0]])
    assert np.allclose(convergence_rates(err), [[1.0, 2.0, 2.0]])
    assert len(STUDY_SCHEMES) == 4
    print_errors_simul(err[:, 0], err[:, 1], err[:, 2], 1)
    out = capsys.readouterr().out.splitlines()
    assert out == ["Norms          (Linf,    L1,       L2) ",
                   "Error E_1    : 5.00e-01 5.00e-01 1.60e+01",
                   "Ratio E_1/E_0: 2.00e+00 4.00e+00 2.50e-01"]
