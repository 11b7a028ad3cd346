"""Parity of the HIP path (libcsadv via its C ABI) with the FP64 CPU oracle.

Tolerance: FP64 results agree with the oracle within 1e-12 relative (max |a-b| / max |b|),
the bar the reference's FP64 GPU mode is held to; FP32 within 1e-5 relative L2 of FP64.
"""
import math

import numpy as np
import pytest
import torch

from helpers import interior_dev, oracle_case, ref_to_dev, rel_max, run_oracle
from oracle import timestep as ots

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _solver(N, case, dt, **kw):
    from cubedsphere.solver import Advection
    return Advection(N, case, dt, device="cuda:0", **kw).init_vars_adv()


def _ref(ic, vf):
    from cubedsphere.cases import reference_case
    return reference_case(ic, vf)


@pytest.mark.parametrize("N", [16, 48])
def test_ghost_fill_matches_oracle(N):
    from oracle.interp import ghost_cell_pc_lagrange_interpolation
    sol = _solver(N, _ref(2, 1), 0.01)
    rng = np.random.default_rng(N)
    g = sol.grid
    I = slice(g.i0, g.iend)
    interior = rng.standard_normal((6, N, N))
    sol.q[0].zero_()
    sol.q[0][0, :, I, I] = torch.from_numpy(interior).to(sol.q[0])
    sol.ghost_fill()
    got = sol.full_state()[0]
    grid, sim = ots.make(N, {"ic": 2, "vf": 1}, 0.01)
    ots.init_vars_adv(grid, sim)
    Q = np.zeros((g.M, g.M, 6))
    Q[I, I, :] = np.transpose(interior, (2, 1, 0))
    ghost_cell_pc_lagrange_interpolation(Q, grid, sim.lagrange)
    assert rel_max(got, ref_to_dev(Q)) < 1e-13


@pytest.mark.parametrize("vf,dp", [(1, 1), (1, 2), (2, 1), (2, 2), (3, 1), (4, 2)])
def test_wind_preparation_matches_oracle(vf, dp):
    N = 24
    sol = _solver(N, _ref(2, vf), 0.005, dp=dp, max_steps=4)
    ua, va = sol.winds()
    grid, sim = ots.make(N, {"ic": 2, "vf": vf}, 0.005, dp=dp)
    ots.init_vars_adv(grid, sim)
    i0, iend = grid.i0, grid.iend
    ru = np.transpose(sim.U_pu.ucontra_averaged, (2, 1, 0))   # [p][j][e]
    rv = np.transpose(sim.U_pv.vcontra_averaged, (2, 1, 0))   # [p][e][i]
    assert rel_max(ua[:, :, i0:iend + 1], ru[:, :, i0:iend + 1]) < 1e-13
    assert rel_max(va[:, i0:iend + 1, :], rv[:, i0:iend + 1, :]) < 1e-13


OPTION_SETS = [
    dict(),                                            # par/advection.par defaults
    dict(recon=1), dict(recon=2), dict(recon=4),
    dict(opsplit=2),
    dict(opsplit=3, mt=2),
    dict(dp=2),
    dict(mf=1),
    dict(recon=4, opsplit=2, dp=2, mf=1),
    dict(recon=3, opsplit=3, mt=2, dp=2, mf=3),
]


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda o: "-".join("%s%d" % kv for kv in sorted(o.items())) or "default")
def test_step_parity_options(opts):
    N, dt, nsteps = 20, 0.01, 8
    case = _ref(2, 1)
    sol = _solver(N, case, dt, **opts)
    sol.step(nsteps)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


@pytest.mark.parametrize("ic,vf", [(1, 1), (3, 1), (4, 1), (2, 2), (3, 2), (2, 3), (3, 4)])
@pytest.mark.parametrize("dp", [1, 2])
def test_step_parity_winds(ic, vf, dp):
    N, dt, nsteps = 24, 0.004, 10
    case = _ref(ic, vf)
    sol = _solver(N, case, dt, dp=dp, max_steps=nsteps)
    sol.step(nsteps)
    grid, sim = run_oracle(N, case, dt, nsteps, dp=dp)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


def test_multitracer_deformational():
    from cubedsphere.cases import workload
    N, dt, nsteps = 32, 0.004, 12
    case = workload("nl10_deformation")
    sol = _solver(N, case, dt, max_steps=nsteps)
    sol.step(nsteps)
    grid, sim = run_oracle(N, case, dt, nsteps)
    got = sol.state()
    for n in range(case.ntr):
        assert rel_max(got[n], interior_dev(grid, sim.Qs[n])) < TOL64, n


def test_two_hills_tilted_n384():
    """The N=384 FP64 configuration: a few steps against the oracle on identical inputs."""
    from cubedsphere.cases import workload
    from cubedsphere.grid import Grid
    from cubedsphere.solver import cfl_dt
    N, nsteps = 384, 4
    case = workload("two_hills_tilted")
    dt = cfl_dt(Grid(N), case, 0.5)
    sol = _solver(N, case, dt)
    sol.step(nsteps)
    grid, sim = run_oracle(N, case, dt, nsteps)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


def test_two_hills_tilted_n384_after_1_10_and_1000_steps():
    """The N=384 FP64 configuration at the checkpoints of its parity specification: the device
    state within 1e-12 (relative max) of the oracle after 1, 10 and 1000 steps (about 23% of a
    revolution; the oracle takes about a minute on one core)."""
    from cubedsphere.cases import workload
    from cubedsphere.grid import Grid
    from cubedsphere.solver import cfl_dt
    N = 384
    case = workload("two_hills_tilted")
    dt = cfl_dt(Grid(N), case, 0.5)
    sol = _solver(N, case, dt)
    grid, sim = ots.make(N, oracle_case(case), dt)
    ots.init_vars_adv(grid, sim)
    k = 0
    for target in (1, 10, 1000):
        sol.step(target - k)
        while k < target:
            k += 1
            ots.adv_time_step(grid, sim, k, k * dt)
            ots.update_adv(grid, sim, k * dt)
        assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64, target


def test_mass_and_norms_match_oracle():
    N, dt, nsteps = 32, 0.01, 20
    case = _ref(2, 1)
    sol = _solver(N, case, dt)
    m0 = sol.mass()
    sol.step(nsteps)
    m1 = sol.mass()
    grid, sim = run_oracle(N, case, dt, nsteps)
    ref_m, _ = ots.mass_computation(sim.Q, grid, 1.0)
    assert abs(m1 - ref_m) <= 1e-13 * abs(ref_m)
    assert abs(m1 - m0) <= 1e-13 * abs(m0)          # MF-PR conserves mass to round-off
    I = slice(grid.i0, grid.iend)
    ref_err = ots.compute_errors(sim.Q[I, I, :], ots.exact_interior(grid, sim, nsteps * dt))
    got_err = sol.norms()
    for a, b in zip(got_err, ref_err):
        assert abs(a - b) <= 1e-10 * abs(b)


def test_fp32_precision_form_close_to_fp64():
    from cubedsphere.cases import workload
    N, dt, nsteps = 64, 0.004, 20
    case = workload("nl10_deformation")
    s64 = _solver(N, case, dt, max_steps=nsteps, dtype="f64")
    s32 = _solver(N, case, dt, max_steps=nsteps, dtype="f32")
    s64.step(nsteps)
    s32.step(nsteps)
    a, b = s32.state(), s64.state()
    for n in range(case.ntr):
        l2 = np.linalg.norm(a[n] - b[n]) / np.linalg.norm(b[n])
        assert l2 <= 1e-5, (n, l2)


def test_graph_replay_bitwise_equal_plain_launches():
    """Steps replayed from a captured HIP graph give the same bits as plain launches, over
    several replays with eager device work and synchronisation between them."""
    N, dt = 24, 0.004
    case = _ref(3, 2)                      # unsteady wind: the per-step wind kernel is captured too
    plain = _solver(N, case, dt, max_steps=32)
    graph = _solver(N, case, dt, max_steps=32)
    graph.use_graph(4)
    plain.step(9)
    graph.step(9)                          # 2 replays + 1 plain step
    scratch = torch.randn(1 << 16, device="cuda:0").sum()
    torch.cuda.synchronize()
    plain.step(14)
    graph.step(14)                         # 3 more replays + 2 plain steps
    del scratch
    assert np.array_equal(plain.full_state(), graph.full_state())
    grid, sim = run_oracle(N, case, dt, 23)
    assert rel_max(graph.state()[0], interior_dev(grid, sim.Q)) < TOL64


@pytest.mark.parametrize("tile", range(12))
def test_every_tile_shape_matches_oracle(tile):
    N, dt, nsteps = 40, 0.008, 6
    case = _ref(2, 1)
    sol = _solver(N, case, dt, tile=tile)
    sol.step(nsteps)
    grid, sim = run_oracle(N, case, dt, nsteps)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


def test_drop_in_driver_matches_oracle_run(tmp_path):
    """adv_sphere from the par files (test case 5, tc = 1) against the oracle's run: same
    final error norms and mass conservation (src/advection_sphere.py:13-61)."""
    from cubedsphere import advection
    (tmp_path / "configuration.par").write_text("a\nb\n16\nc\n2\nd\n0\ne\n0\nf\n5\ng\n1\n")
    (tmp_path / "advection.par").write_text(
        # Tf, dt, ic, vf, tc, recon, dp, opsplit, et, mt, mf (one value every other line)
        "\n".join(["x", "y", "0.25", "#", "0.025", "#", "2", "#", "1", "#", "1", "#", "3", "#", "1", "#", "1", "#",
                   "3", "#", "1", "#", "3"]) + "\n")
    linf, l1, l2 = advection.main(str(tmp_path))
    grid, sim = ots.make(16, {"ic": 2, "vf": 1}, 0.025)
    ref = ots.run(grid, sim, 10)
    for a, b in zip((linf, l1, l2), ref):
        assert abs(a - b) <= 1e-10 * abs(b)


@pytest.mark.parametrize("N", [17, 33])
@pytest.mark.parametrize("opts", [dict(), dict(recon=2, opsplit=2, dp=2, mf=1)], ids=["default", "cw84-l04-rk2"])
def test_odd_n_matches_oracle(N, opts):
    """Odd N puts a cell centre on each pole (panels 4 and 5), where lon = atan2(+-0, +-0)."""
    dt, nsteps = 0.5 * (math.pi / 2 / N) / (2 * math.pi / 5 * 1.42), 8
    case = _ref(2, 1)
    sol = _solver(N, case, dt, **opts)
    sol.step(nsteps)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64
