"""Device diagnostics (csadv_diagnostics) and the blow-up halt, through the C ABI.

Bars: the device mass and norms agree with the oracle's mass_computation / compute_errors
within 1e-13 (mass) and 1e-12 (norms) relative, and are bitwise those of csadv_mass /
csadv_norms (the same kernels); a call captured in a HIP graph gives the same bits on every
replay as an eager call; a solve halted by an exceeded limit is left untouched by later steps.
"""
import numpy as np
import pytest
import torch

from helpers import run_oracle
from oracle import timestep as ots

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _solver(N, case, dt, **kw):
    from cubedsphere.solver import Advection
    return Advection(N, case, dt, device="cuda:0", **kw).init_vars_adv()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_diagnostics_match_oracle_and_the_host_entry_points(dtype):
    from cubedsphere.cases import workload
    N, dt, nsteps = 32, 0.004, 12
    case = workload("nl10_deformation")
    sol = _solver(N, case, dt, max_steps=nsteps, dtype=dtype)
    sol.step(nsteps)
    qe = np.stack([sol.exact(n) for n in range(case.ntr)])
    out = sol.diagnostics(qexact=qe).cpu().numpy()
    grid, sim = run_oracle(N, case, dt, nsteps)
    I = slice(grid.i0, grid.iend)
    for n in range(case.ntr):
        # against the host entry points: the same kernels, the same bits
        assert out[n, 0] == sol.mass(n)
        assert tuple(out[n, 1:]) == sol.norms(n)
        if dtype == "f64":
            ref_m, _ = ots.mass_computation(sim.Qs[n], grid, 1.0)
            ref = ots.compute_errors(sim.Qs[n][I, I, :], ots.exact_interior(grid, sim, nsteps * dt, n))
            assert abs(out[n, 0] - ref_m) <= 1e-13 * abs(ref_m)
            for a, b in zip(out[n, 1:], ref):
                assert abs(a - b) <= 1e-12 * abs(b)
    # without an exact field: the norms of the field itself
    own = sol.diagnostics().cpu().numpy()
    q = sol.state()
    for n in range(case.ntr):
        assert abs(own[n, 1] - np.abs(q[n]).max()) <= 1e-15 * np.abs(q[n]).max() + (1e-7 if dtype == "f32" else 0)
        assert abs(own[n, 2] - np.abs(q[n]).mean()) <= 1e-12 * np.abs(q[n]).mean() + (1e-7 if dtype == "f32" else 0)


def test_diagnostics_inside_a_captured_graph_replay_like_eager_calls():
    """Steps plus a diagnostics call captured in one HIP graph: every replay's diagnostics equal
    those of eager calls at the same step (eager work and synchronisation between replays)."""
    from cubedsphere.cases import reference_case
    N, dt = 24, 0.01
    eager = _solver(N, reference_case(2, 1), dt)
    graphed = _solver(N, reference_case(2, 1), dt)
    ring = torch.zeros((3, 1, 4), dtype=torch.float64, device="cuda:0")
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        graphed.step(2)
        graphed.diagnostics(out=ring[0])
        graphed.step(2)
        graphed.diagnostics(out=ring[1])
    graphed.steps_done -= 4                     # the capture ran nothing
    for rep in range(3):
        g.replay()
        scratch = torch.randn(1 << 16, device="cuda:0").sum()
        torch.cuda.synchronize()
        eager.step(2)
        e1 = eager.diagnostics().clone()
        eager.step(2)
        e2 = eager.diagnostics().clone()
        torch.cuda.synchronize()
        assert torch.equal(ring[0], e1) and torch.equal(ring[1], e2), rep
        del scratch
    steps, halted = graphed.status()
    assert steps == 12 and not halted
    assert np.array_equal(graphed.full_state(), eager.full_state())


@pytest.mark.parametrize("opts", [dict(), dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1)], ids=["fused", "general"])
def test_blowup_limit_halts_the_solve(opts):
    """A limit below the field's own Linf (forced from the host) halts the solve at once: later
    steps leave the state and the step count untouched, as the reference's run stops there."""
    from cubedsphere.cases import reference_case
    N, dt = 20, 0.01
    sol = _solver(N, reference_case(2, 1), dt, **opts)
    sol.step(3)
    sol.diagnostics(blowup_limit=1e3)           # not exceeded
    sol.step(2)
    assert sol.status() == (5, False)
    before = sol.full_state()
    sol.diagnostics(blowup_limit=0.5)           # max |q| ~ 1 > 0.5: halt
    sol.step(4)
    steps, halted = sol.status()
    assert halted and steps == 5
    assert np.array_equal(sol.full_state(), before)


def test_adv_sphere_stops_a_run_that_blows_up():
    """A run whose field grows without bound (dt far beyond the CFL limit) is halted on the
    device at the first check past the limit and stops as the reference's does."""
    from cubedsphere import advection
    from cubedsphere._lib import CsadvError
    from cubedsphere.grid import Grid
    g = Grid(16)
    sim = advection.adv_simulation_par(g, 0.4, 40.0, 2, 1, 1, 3, 1, 1, 3, 1, 3)   # CFL ~ 8
    with pytest.raises(CsadvError, match="blew up"):
        advection.adv_sphere(g, sim)
    steps, halted = sim.status()
    assert halted and steps < 100
