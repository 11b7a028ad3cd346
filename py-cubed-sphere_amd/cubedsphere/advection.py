"""Drop-in entry points of the advection path, driven by the same par files.

Mirrors the reference's call sequence (src/advection_sphere.py:13-61,
src/advection_error.py:21-188, main.py test case 5) with the step on the GPU:

  sim = adv_simulation_par(cs_grid, dt, Tf, ic, vf, tc, recon, dp, opsplit, et, mt, mf)
  init_vars_adv(cs_grid, sim)            # src/advection_vars.py:19
  adv_time_step(cs_grid, sim, k, t)      # src/advection_timestep.py:19  (one step on the device)
  update_adv(cs_grid, sim, t)            # src/advection_timestep.py:48  (no-op: see below)
  adv_sphere(cs_grid, sim) -> (linf, l1, l2)
  error_analysis_adv(sim) -> errors table of the convergence study (tc = 2)
  main(pardir)                           # test case 5 from par/configuration.par + advection.par

Differences of contract, all deliberate: ``cs_grid`` is a :class:`cubedsphere.grid.Grid`;
``sim`` is a :class:`cubedsphere.solver.Advection` whose state lives on the GPU (read it
with ``sim.state()``); the time-dependent winds of ``update_adv`` are recombined on the
device at the start of the next step, so ``update_adv`` has nothing left to do; plotting
and file output of ``output_adv`` are not part of this path.  Errors are those of
src/errors.py:99-113 at the final step.
"""
import numpy as np

from . import config
from ._lib import CsadvError
from .cases import reference_case
from .grid import Grid
from .solver import Advection

# the four schemes of the convergence study (src/advection_error.py:61-66):
# (recon, opsplit, et, mt, dp, mf)
STUDY_SCHEMES = ((3, 3, 2, 2, 1, 1), (3, 3, 3, 2, 1, 3), (3, 1, 3, 1, 2, 2), (3, 1, 3, 1, 2, 3))
STUDY_DT0 = {1: 0.025, 2: 0.0125, 3: 0.00625, 4: 0.0125}     # dt at N = 16 per vector field


def adv_simulation_par(cs_grid, dt, Tf, ic, vf, tc, recon, dp, opsplit, et, mt, mf, dtype="f64", tile=0,
                       device="cuda:0"):
    """Simulation object of src/advection_ic.py:21-203: options validated, buffers on the device."""
    config.check_options(config.AdvectionParams(dt, Tf, tc, ic, vf, recon, dp, opsplit, et, mt, mf))
    nsteps = int(Tf / dt)
    sim = Advection(cs_grid.N, reference_case(ic, vf), dt, recon=recon, dp=dp, opsplit=opsplit, et=et, mt=mt,
                    mf=mf, dtype=dtype, device=device, tile=tile, max_steps=max(nsteps, 1), grid=cs_grid)
    sim.Tf, sim.tc, sim.ic, sim.vf = Tf, tc, ic, vf
    return sim


def init_vars_adv(cs_grid, simulation):
    """src/advection_vars.py:19-107: initial condition, winds, tables, the device handle."""
    simulation.init_vars_adv()
    simulation.total_mass0 = simulation.mass()


def adv_time_step(cs_grid, simulation, k, t):
    """src/advection_timestep.py:19-43: one step (ghost fill, winds, split fluxes, update,
    mass fixer).  ``k`` and ``t`` must follow the run (the device tracks the step count)."""
    if k != simulation.steps_done + 1:
        raise CsadvError("adv_time_step: step %d requested after %d steps" % (k, simulation.steps_done))
    simulation.step(1)


def update_adv(cs_grid, simulation, t):
    """src/advection_timestep.py:48-75.  The device recombines the time-dependent winds at the
    start of every step (from precomputed basis fields), so nothing is left to update here."""
    return None


def adv_sphere(cs_grid, simulation, divtest_flag=False):
    """src/advection_sphere.py:13-61 without plotting: run Tf/dt steps, return the final
    (L_inf, L_1, L_2) errors against the exact solution."""
    nsteps = 1 if divtest_flag else int(simulation.Tf / simulation.dt)
    init_vars_adv(cs_grid, simulation)
    for k in range(1, nsteps + 1):
        adv_time_step(cs_grid, simulation, k, k * simulation.dt)
        update_adv(cs_grid, simulation, k * simulation.dt)
    simulation.mass_change = _mass_change(simulation.total_mass0, simulation.mass())
    return simulation.norms(t=nsteps * simulation.dt)


def _mass_change(m0, m):
    """src/output.py: relative change, absolute when the initial mass vanishes."""
    return abs(m0 - m) / abs(m0) if abs(m0) > 1e-10 else abs(m0 - m)


def error_analysis_adv(simulation, ntest=6, n0=16, device="cuda:0"):
    """src/advection_error.py:21-188 without plots: the 4-scheme x ntest-resolution study of the
    simulation's (ic, vf); returns (Ns, errors [ntest][4][3] as (linf, l1, l2), rates [ntest-1][4][3])."""
    vf, ic = simulation.vf, simulation.ic
    Ns = [n0 * 2 ** i for i in range(ntest)]
    dts = [STUDY_DT0[vf] * 0.5 ** i for i in range(ntest)]
    err = np.zeros((ntest, len(STUDY_SCHEMES), 3))
    for d, (recon, opsplit, et, mt, dp, mf) in enumerate(STUDY_SCHEMES):
        for i, (N, dt) in enumerate(zip(Ns, dts)):
            g = Grid(N)
            sim = adv_simulation_par(g, dt, 5.0, ic, vf, 2, recon, dp, opsplit, et, mt, mf, device=device)
            err[i, d] = adv_sphere(g, sim)
            sim.close()
    rates = np.log2(err[:-1] / err[1:])
    return Ns, err, rates


def main(pardir="par"):
    """main.py test case 5 from the par files: tc = 1 runs one advection, tc = 2 the study."""
    gp = config.get_parameters(pardir)
    ap = config.get_advection_parameters(pardir)
    if gp.test_case != 5:
        raise CsadvError("only test case 5 (advection) is on this path; par file asks for %d" % gp.test_case)
    g = Grid(gp.N)
    sim = adv_simulation_par(g, ap.dt, ap.Tf, ap.ic, ap.vf, ap.tc, ap.recon, ap.dp, ap.opsplit, ap.et, ap.mt, ap.mf)
    if ap.tc == 1:
        return adv_sphere(g, sim)
    if ap.tc == 2:
        return error_analysis_adv(sim)
    raise CsadvError("invalid advection test case %r" % ap.tc)
