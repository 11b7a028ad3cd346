"""The reference's loop with only its step swapped: the seam of src/advection_sphere.py:13-61.

The reference's ``adv_sphere`` imports ``init_vars_adv``, ``adv_time_step`` and ``update_adv``
and reads ``simulation.Q`` between steps (``mass_computation`` at the start, ``output_adv``
after every step).  Here that loop runs twice on the oracle's objects, whose classes mirror
the reference's ``cubed_sphere`` and ``adv_simulation_par`` (attributes, array layout
``Q[i, j, panel]``): once with the oracle's own three functions, once with
``cubedsphere.advection``'s, which bind the object to the device step.  Bar: the states and
every mass read agree within 1e-12 relative (FP64).
"""
import math

import numpy as np
import pytest
import torch

from helpers import rel_max
from oracle import timestep as ots

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _loop(mod, grid, sim, nsteps, times, read_every=3, q_init=None):
    """src/advection_sphere.py:33-57 with output_adv's reads of Q (mass each read)."""
    mod.init_vars_adv(grid, sim)
    if q_init is not None:
        I, J = slice(grid.i0, grid.iend), slice(grid.j0, grid.jend)
        if mod is ots:                   # the oracle's own state: written in place
            sim.Q[I, J, :] = q_init
        else:                            # a bound object: the assignment loads the device state
            q = sim.Q.copy()
            q[I, J, :] = q_init
            sim.Q = q
    m0, _ = ots.mass_computation(sim.Q, grid, 1.0)
    masses = [m0]
    for k in range(1, nsteps + 1):
        t = times[k]
        mod.adv_time_step(grid, sim, k, t)
        if k % read_every == 0 or k == nsteps:
            masses.append(ots.mass_computation(sim.Q, grid, m0)[0])
        mod.update_adv(grid, sim, t)
    I, J = slice(grid.i0, grid.iend), slice(grid.j0, grid.jend)
    return sim.Q[I, J, :].copy(), np.array(masses)


def _compare(N, case, dt, nsteps, opts, times=None, q_init=None):
    from cubedsphere import advection
    times = times if times is not None else [k * dt for k in range(nsteps + 1)]
    g_ref, s_ref = ots.make(N, case, dt, **opts)
    q_ref, m_ref = _loop(ots, g_ref, s_ref, nsteps, times, q_init=q_init)
    g_dev, s_dev = ots.make(N, case, dt, **opts)
    q_dev, m_dev = _loop(advection, g_dev, s_dev, nsteps, times, q_init=q_init)
    assert type(s_dev).__name__ == "Simulation" and isinstance(s_dev, ots.Simulation)
    assert rel_max(q_dev, q_ref) < TOL64
    assert np.max(np.abs(m_dev - m_ref)) <= 1e-13 * np.max(np.abs(m_ref))
    assert abs(s_dev.CFL - s_ref.CFL) <= 1e-13 * abs(s_ref.CFL)
    return s_dev


@pytest.mark.parametrize("opts", [dict(), dict(recon=3, opsplit=3, et=2, mt=2, dp=1, mf=1),
                                  dict(recon=4, opsplit=2, et=3, mt=1, dp=2, mf=2)],
                         ids=["default", "etpl07", "rk2-mfaf"])
def test_reference_loop_with_device_step(opts):
    _compare(24, {"ic": 2, "vf": 1}, 0.01, 20, opts)


def test_reference_loop_time_dependent_wind_grows_and_follows_update_times():
    """vf=2 with RK2: the wind table grows past Tf/dt (the object says Tf = 0), and update_adv
    at irregular times sets the next step's winds exactly as the reference's does."""
    dt, nsteps = 0.004, 14
    times = [0.0] + [k * dt + 0.3 * dt * math.sin(k) for k in range(1, nsteps + 1)]
    s = _compare(20, {"ic": 3, "vf": 2}, dt, nsteps, dict(dp=2), times=times)
    from cubedsphere import seam
    assert seam.device_solve(s).max_steps >= nsteps


def test_reference_loop_assigned_initial_state():
    """simulation.Q assigned after init_vars_adv reaches the device state."""
    N = 16
    rng = np.random.default_rng(5)
    q0 = 1.0 + 0.1 * rng.standard_normal((N, N, 6))
    _compare(N, {"ic": 2, "vf": 1}, 0.01, 6, dict(), q_init=q0)
