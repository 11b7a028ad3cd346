"""The general step (ET-S72, ET-PL07) and MF-AF against the oracle.

ET-S72 and ET-PL07 run the library's general step (two marched passes per tracer of
src/discrete_operators.py:18-101, the copy halo read from the adjacent panels) instead of the
fused step kernel; ET-DG with MF-AF runs the fused step with its cube-edge fix-up.  Bar: FP64
within 1e-12 relative of the oracle on identical inputs, as for the fused path.
"""
import numpy as np
import pytest
import torch

from helpers import interior_dev, rel_max, run_oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _solve(N, case, dt, nsteps, **opts):
    from cubedsphere.solver import Advection
    sol = Advection(N, case, dt, device="cuda:0", max_steps=nsteps, **opts).init_vars_adv()
    sol.step(nsteps)
    return sol


# the four schemes of the reference's convergence study (src/advection_error.py:61-66) and
# the remaining combinations of the edge treatments and mass fixers
GENERAL_OPTS = [
    dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),        # study scheme 1 (ET-PL07)
    dict(recon=3, opsplit=1, et=3, mt=1, dp=2, mf=2),        # study scheme 3 (MF-AF)
    dict(recon=3, opsplit=1, et=1, mt=1, dp=1, mf=3),        # ET-S72 with MF-PR
    dict(recon=3, opsplit=1, et=1, mt=1, dp=2, mf=1),        # ET-S72, RK2 wind copy
    dict(recon=1, opsplit=2, et=2, mt=1, dp=2, mf=2),        # ET-PL07 + MF-AF, PPM-0, L04 split
    dict(recon=2, opsplit=1, et=2, mt=1, dp=1, mf=3),        # CW84 with ET-PL07
    dict(recon=4, opsplit=3, et=1, mt=2, dp=1, mf=2),        # L04 recon, ET-S72, MF-AF
    dict(recon=2, opsplit=3, et=3, mt=2, dp=1, mf=2),        # ET-DG + MF-AF, CW84, PL07 split
    dict(recon=1, opsplit=2, et=3, mt=1, dp=2, mf=2),        # ET-DG + MF-AF, PPM-0, L04 split, RK2
]


@pytest.mark.parametrize("opts", GENERAL_OPTS, ids=lambda o: "-".join("%s%d" % kv for kv in sorted(o.items())))
def test_general_step_matches_oracle(opts):
    from cubedsphere.cases import reference_case
    N, dt, nsteps = 20, 0.01, 6
    case = reference_case(2, 1)
    sol = _solve(N, case, dt, nsteps, **opts)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


@pytest.mark.parametrize("vf", [2, 3])
def test_general_step_time_dependent_winds(vf):
    from cubedsphere.cases import reference_case
    N, dt, nsteps = 24, 0.004, 8
    case = reference_case(3, vf)
    for opts in (dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1), dict(recon=3, opsplit=1, et=3, mt=1, dp=2, mf=2)):
        sol = _solve(N, case, dt, nsteps, **opts)
        grid, sim = run_oracle(N, case, dt, nsteps, **opts)
        assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64, opts


def test_general_step_multitracer_and_conservation():
    """MF-AF makes the fluxes across each cube edge identical, so mass is conserved to round-off."""
    from cubedsphere.cases import workload
    N, dt, nsteps = 24, 0.004, 10
    case = workload("nl10_deformation")
    opts = dict(recon=3, opsplit=1, et=3, mt=1, dp=2, mf=2)
    sol = _solve(N, case, dt, nsteps, **opts)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    got = sol.state()
    for n in range(case.ntr):
        assert rel_max(got[n], interior_dev(grid, sim.Qs[n])) < TOL64, n
    from cubedsphere.solver import Advection
    fresh = Advection(N, case, dt, device="cuda:0", max_steps=nsteps, **opts).init_vars_adv()
    m0 = fresh.mass()
    assert abs(sol.mass() - m0) <= 1e-13 * abs(m0)


@pytest.mark.parametrize("opts", [dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),
                                  dict(recon=2, opsplit=1, et=1, mt=1, dp=2, mf=3),
                                  dict(recon=4, opsplit=2, et=2, mt=1, dp=1, mf=2)],
                         ids=["etpl07", "ets72-rk2-mfpr", "etpl07-mfaf"])
def test_general_step_four_tracers(opts):
    """Four tracers through the general step (passes run tracer by tracer and the last kernel of
    the last tracer ends the step): every tracer against the oracle."""
    from cubedsphere.cases import workload
    N, dt, nsteps = 24, 0.004, 7
    case = workload("nl10_deformation")
    sol = _solve(N, case, dt, nsteps, **opts)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    got = sol.state()
    for n in range(case.ntr):
        assert rel_max(got[n], interior_dev(grid, sim.Qs[n])) < TOL64, n
    assert sol.status() == (nsteps, False)


def test_general_step_graph_replay_bitwise():
    from cubedsphere.cases import reference_case
    from cubedsphere.solver import Advection
    N, dt = 20, 0.01
    opts = dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1)
    a = Advection(N, reference_case(2, 1), dt, device="cuda:0", **opts).init_vars_adv()
    b = Advection(N, reference_case(2, 1), dt, device="cuda:0", **opts).init_vars_adv()
    b.use_graph(4)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.full_state(), b.full_state())


# the smallest N the library takes (8: the ET-PL07 cells beside opposite panel sides, i0-1 ..
# i0+1 and iend-2 .. iend, come closest), odd N (a cell centre on each pole) and N = 9, whose
# last row segment has one row
SMALL_N_OPTS = [
    dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),        # ET-PL07, PL07 recon
    dict(recon=2, opsplit=1, et=2, mt=1, dp=1, mf=3),        # ET-PL07, CW84 recon, MF-PR
    dict(recon=3, opsplit=1, et=3, mt=1, dp=2, mf=2),        # MF-AF
    dict(recon=4, opsplit=2, et=1, mt=1, dp=2, mf=2),        # ET-S72 + MF-AF, RK2 wind copy
]


@pytest.mark.parametrize("N", [8, 9, 17])
@pytest.mark.parametrize("opts", SMALL_N_OPTS, ids=lambda o: "-".join("%s%d" % kv for kv in sorted(o.items())))
def test_general_step_small_and_odd_n(N, opts):
    from cubedsphere.cases import reference_case
    dt, nsteps = 0.4 * (np.pi / 2 / N) / (2 * np.pi / 5 * 1.42), 6
    case = reference_case(2, 1)
    sol = _solve(N, case, dt, nsteps, **opts)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


@pytest.mark.parametrize("N", [61, 128])
@pytest.mark.parametrize("opts", [dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),
                                  dict(recon=4, opsplit=2, et=1, mt=1, dp=2, mf=2),
                                  dict(recon=2, opsplit=1, et=2, mt=1, dp=1, mf=3)],
                         ids=["etpl07", "ets72-rk2-mfaf", "etpl07-cw84-mfpr"])
def test_general_step_several_column_strips(N, opts):
    """N = 61 and 128 give two and three column strips per panel (at most 58 columns each) and
    several row segments: strips away from the panel sides, which take no ET-PL07 x-edge
    treatment, and segment boundaries inside the panel, against the oracle."""
    from cubedsphere.cases import reference_case
    dt, nsteps = 0.45 * (np.pi / 2 / N) / (2 * np.pi / 5 * 1.42), 8
    case = reference_case(2, 1)
    sol = _solve(N, case, dt, nsteps, **opts)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    assert rel_max(sol.state()[0], interior_dev(grid, sim.Q)) < TOL64


@pytest.mark.parametrize("opts", [dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),
                                  dict(recon=1, opsplit=1, et=1, mt=1, dp=2, mf=3)],
                         ids=["etpl07", "ets72-rk2-mfpr"])
def test_general_step_every_tile(opts):
    """Every tile shape (4 to 32 rows per segment: one-row last segments, segments that start at
    the ET-PL07 rows near the panel sides) against one oracle run, N = 61."""
    from cubedsphere.cases import reference_case
    from cubedsphere.solver import Advection
    from cubedsphere import _lib
    N = 61
    dt, nsteps = 0.45 * (np.pi / 2 / N) / (2 * np.pi / 5 * 1.42), 5
    case = reference_case(2, 1)
    grid, sim = run_oracle(N, case, dt, nsteps, **opts)
    ref = interior_dev(grid, sim.Q)
    lib = _lib.load()
    tile = 0
    while lib.csadv_tile_rows(tile) > 0:
        sol = Advection(N, case, dt, device="cuda:0", max_steps=nsteps, tile=tile, **opts).init_vars_adv()
        sol.step(nsteps)
        assert rel_max(sol.state()[0], ref) < TOL64, tile
        sol.close()
        tile += 1
    assert tile == 12


@pytest.mark.parametrize("opts", [dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),
                                  dict(recon=3, opsplit=1, et=1, mt=1, dp=1, mf=3),
                                  dict(recon=3, opsplit=1, et=3, mt=1, dp=2, mf=2)],
                         ids=["etpl07", "ets72-mfpr", "rk2-mfaf"])
def test_general_step_fp32_close_to_fp64(opts):
    """The precision form through the general step: FP32 within 1e-5 relative L2 of FP64 after
    200 steps, and the mass of the conservative options kept to FP32 round-off."""
    from cubedsphere.cases import Case
    N, nsteps = 48, 200
    case = Case([{"field": "gaussians", "centres": [(0.7, 0.3)], "b": 20.0}], ("solid_body", -0.6, 5.0), "narrow")
    dt = 0.5 * (np.pi / 2 / N) / (2 * np.pi / 5 * 1.42)
    out, drift = {}, {}
    from cubedsphere.solver import Advection
    for dtype in ("f32", "f64"):
        sol = Advection(N, case, dt, device="cuda:0", dtype=dtype, **opts).init_vars_adv()
        m0 = sol.mass()
        sol.step(nsteps)
        out[dtype] = sol.state()[0]
        drift[dtype] = abs(sol.mass() - m0) / abs(m0)
    l2 = np.linalg.norm(out["f32"] - out["f64"]) / np.linalg.norm(out["f64"])
    print("general step FP32 %s: L2 vs FP64 %.3e, mass drift FP32 %.2e FP64 %.2e" % (opts, l2, drift["f32"], drift["f64"]))
    assert l2 <= 1e-5, l2
    if opts["mf"] != 1:
        assert drift["f64"] <= 1e-13 and drift["f32"] <= 2e-6, drift
