"""Drop-in entry points of the advection path, driven by the same par files.

Mirrors the reference's call sequence (src/advection_sphere.py:13-61,
src/advection_error.py:21-188, main.py test case 5) with the step on the GPU:

  sim = adv_simulation_par(cs_grid, dt, Tf, ic, vf, tc, recon, dp, opsplit, et, mt, mf)
  init_vars_adv(cs_grid, sim)            # src/advection_vars.py:19
  adv_time_step(cs_grid, sim, k, t)      # src/advection_timestep.py:19  (one step on the device)
  update_adv(cs_grid, sim, t)            # src/advection_timestep.py:48  (winds of the next step)
  adv_sphere(cs_grid, sim) -> (linf, l1, l2)
  error_analysis_adv(sim) -> errors table of the convergence study (tc = 2), rates as the reference's
  output_adv(cs_grid, sim, k, nsteps)    # src/output.py: errors, mass, lat-lon remaps, data files
  main(pardir)                           # test case 5 from par/configuration.par + advection.par

``init_vars_adv``, ``adv_time_step`` and ``update_adv`` take either this module's objects (a
:class:`cubedsphere.grid.Grid` and the :class:`cubedsphere.solver.Advection` that
``adv_simulation_par`` returns, state read with ``sim.state()``) or the reference's own
``cubed_sphere`` and ``adv_simulation_par`` objects, whose ``Q`` then reads the device state
(:mod:`cubedsphere.seam`): the reference's own ``adv_sphere`` loop runs unchanged with these
three imported from here.  Time-dependent winds are recombined on the device at the start of
every step from basis fields and one coefficient row per step; ``update_adv`` sets the row of
the next step.  The output side (``output_adv``) computes the errors, mass, the device lat-lon
remaps and the reference's data files; it draws no plots (cartopy is not available).  Errors
are those of src/errors.py:99-113 at the final step.
"""
import os

import numpy as np
import torch

from . import config, gridfile, seam
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
    sim.recon, sim.dp, sim.opsplit, sim.et, sim.mt, sim.mf = recon, dp, opsplit, et, mt, mf
    return sim


def init_vars_adv(cs_grid, simulation):
    """src/advection_vars.py:19-107: initial condition, winds, tables, the device handle, and
    the CFL number of :88-98 (the largest signed Courant number of the interior x-edges of all
    rows and the interior y-edges of all columns, in absolute value).

    ``simulation`` is either an :class:`Advection` (from :func:`adv_simulation_par` here) or the
    reference's own ``adv_simulation_par`` object with its ``cubed_sphere`` grid: the latter is
    bound to a device solve (:mod:`cubedsphere.seam`) and its ``Q`` then reads the device state."""
    if not isinstance(simulation, Advection):
        seam.bind(cs_grid, simulation)
        return
    simulation.init_vars_adv()
    simulation.total_mass0 = simulation.mass()
    M, i0, iend = cs_grid.M, cs_grid.i0, cs_grid.iend
    U = simulation.usgn if simulation.usgn is not None else simulation.uavg
    V = simulation.vsgn if simulation.vsgn is not None else simulation.vavg
    cx = float((U[:, :M, i0:iend + 1].double() * simulation.dt / cs_grid.dx).max())
    cy = float((V[:, i0:iend + 1, :M].double() * simulation.dt / cs_grid.dx).max())
    simulation.CFL = max(abs(cx), abs(cy))


def adv_time_step(cs_grid, simulation, k, t):
    """src/advection_timestep.py:19-43: one step (ghost fill, winds, split fluxes, update,
    mass fixer).  ``k`` and ``t`` must follow the run (the device tracks the step count).  For
    the reference's own objects the step ignores ``k`` and ``t``, as the reference's does."""
    if not isinstance(simulation, Advection):
        seam.step(cs_grid, simulation, k, t)
        return
    if k != simulation.steps_done + 1:
        raise CsadvError("adv_time_step: step %d requested after %d steps" % (k, simulation.steps_done))
    simulation.step(1)


def update_adv(cs_grid, simulation, t):
    """src/advection_timestep.py:48-75: the next step's winds become the analytic wind at ``t``.
    The device recombines the time-dependent winds from precomputed basis fields at the start
    of every step, with one coefficient row per step; this writes the row for time ``t`` when it
    differs from the one of the uniform times ``t = k dt`` (which it never does inside
    :func:`adv_sphere`) and grows the table past its end."""
    if not isinstance(simulation, Advection):
        seam.update(cs_grid, simulation, t)
        return
    if simulation.case.steady:
        return
    prev = getattr(simulation, "_wind_t", 0.0)
    simulation.set_wind_times(float(t), prev)
    simulation._wind_t = float(t)


BLOWUP_LINF = 100.0     # src/output.py:44-47: a run stops once its L_inf error exceeds this
CHECK_EVERY = 16        # steps between the device's blow-up checks in adv_sphere


def adv_sphere(cs_grid, simulation, divtest_flag=False, datadir=None, verbose=False):
    """src/advection_sphere.py:13-61 without plotting: run Tf/dt steps, return the final
    (L_inf, L_1, L_2) errors against the exact solution.  Nothing is read between steps
    (there is no per-step output here), so the steps go to the device in chunks of 16 (one
    replayed graph each) with no host synchronisation, which is the same as Nsteps calls of
    adv_time_step + update_adv.  After every chunk the device checks the field itself
    (csadv_diagnostics): a max |Q| above 100 + max |Q(0)| means an L_inf error above 100, and
    halts the solve there; the run then stops as the reference's does (src/output.py:44-47),
    raising CsadvError, and so does a final L_inf error above 100."""
    nsteps = 1 if divtest_flag else int(simulation.Tf / simulation.dt)
    init_vars_adv(cs_grid, simulation)
    if nsteps >= 64:
        simulation.use_graph(CHECK_EVERY)    # replay graphs: no per-launch host cost
    limit = BLOWUP_LINF + float(simulation.diagnostics()[:, 1].max())
    check = torch.empty((simulation.case.ntr, 4), dtype=torch.float64, device=simulation.device)
    done = 0
    while done < nsteps:
        k = min(CHECK_EVERY, nsteps - done)
        simulation.step(k)
        simulation.diagnostics(blowup_limit=limit, out=check)
        done += k
    steps, halted = simulation.status()
    linf = simulation.norms(t=steps * simulation.dt)[0]
    if halted or linf > BLOWUP_LINF:
        print("Stopping due to large errors.")
        print("The CFL number is:", getattr(simulation, "CFL", None))
        raise CsadvError("advection blew up (L_inf error above %g) by step %d of %d" % (BLOWUP_LINF, steps, nsteps))
    if datadir is not None or verbose:       # src/output.py at the final step
        return output_adv(cs_grid, simulation, nsteps, nsteps, datadir=datadir, verbose=verbose)[0]
    simulation.mass_change = _mass_change(simulation.total_mass0, simulation.mass())
    return simulation.norms(t=nsteps * simulation.dt)


# plotting grid of src/constants.py:36-37
NLAT = 720
NLON = 2 * NLAT

RECON_NAMES = {1: "PPM-0", 2: "PPM-CW84", 3: "PPM-PL07", 4: "PPM-L04"}          # src/advection_ic.py:85-158
DP_NAMES = {1: "RK1", 2: "RK2"}
SPLIT_NAMES = {1: "SP-AVLT", 2: "SP-L04", 3: "SP-PL07"}
ET_NAMES = {1: "ET-S72", 2: "ET-PL07", 3: "ET-DG"}
MT_NAMES = {1: "MT-0", 2: "MT-PL07"}
MF_NAMES = {1: "MF-0", 2: "MF-AF", 3: "MF-PR"}


def ll2cs(cs_grid, simulation, nlat=NLAT, nlon=NLON):
    """src/interpolation.py:38-139 on the device: the (i, j, panel) tables of the plotting grid,
    each [nlon][nlat]."""
    _, tabs = simulation.remap_latlon(nlat=nlat, nlon=nlon, tables=True)
    return tabs


def nearest_neighbour(simulation, field=None, tracer=0, nlat=NLAT, nlon=NLON):
    """src/interpolation.py:144-149 on the device: the tracer's state (or an interior field
    [6][N][N]) on the plotting grid, [nlon][nlat] FP64 numpy."""
    return simulation.remap_latlon(tracer=tracer, nlat=nlat, nlon=nlon, field=field)


def print_diagnostics_adv(error_linf, error_l1, error_l2, mass_change, t, Nsteps):
    """src/output.py:11-14"""
    print("\nStep", t, "from", Nsteps)
    print("Error (Linf, L1, L2) :", "{:.2e}".format(error_linf), "{:.2e}".format(error_l1), "{:.2e}".format(error_l2))
    print("Total mass variation:", "{:.2e}".format(mass_change))


def output_name(simulation, error=True):
    """File name stem of src/output.py:121-123 (error field) and :91-92 (state)."""
    opts = (SPLIT_NAMES[simulation.opsplit], RECON_NAMES[simulation.recon], DP_NAMES[simulation.dp],
            ET_NAMES[simulation.et], MT_NAMES[simulation.mt])
    if error:
        return ("adv_Q_error_ic%d_vf%d_%s_%s_%s_%s_%s_%s_interp3" % ((simulation.ic, simulation.vf) + opts +
                                                                    (MF_NAMES[simulation.mf],)))
    return "adv_Q_ic%d_vf%d_%s_%s_%s_%s_%s_interp3" % ((simulation.ic, simulation.vf) + opts)


def output_adv(cs_grid, simulation, k, nsteps, datadir=None, nlat=NLAT, nlon=NLON, verbose=True):
    """src/output.py:16-150 at step k, without plotting: errors against the exact solution,
    total mass variation, the printout of step k > 0, and the lat-lon fields Q_ll and the
    error Q_exact_ll - Q_ll (both remapped on the device).  With ``datadir`` at the final
    step, writes the reference's files: <grid>_<name>.npy (the error field) and
    <grid>_<name> (text: Linf, L1, L2, CFL, mass change, dt, Tf)."""
    t = k * simulation.dt
    linf, l1, l2 = simulation.norms(t=t)
    simulation.mass_change = _mass_change(simulation.total_mass0, simulation.mass())
    if verbose and k > 0:
        print_diagnostics_adv(linf, l1, l2, simulation.mass_change, k, nsteps)
    q_ll = nearest_neighbour(simulation, nlat=nlat, nlon=nlon)
    qe_ll = nearest_neighbour(simulation, field=simulation.exact(0, t), nlat=nlat, nlon=nlon)
    err_ll = qe_ll - q_ll
    if datadir is not None and k == nsteps:
        import os
        stem = os.path.join(datadir, "%s_cs_%d_%s" % (cs_grid.projection, cs_grid.N, output_name(simulation)))
        np.save(stem, err_ll, allow_pickle=False)
        np.savetxt(stem, np.array([linf, l1, l2, getattr(simulation, "CFL", 0.0), simulation.mass_change,
                                   simulation.dt, simulation.Tf]))
    return (linf, l1, l2), q_ll, err_ll


def _mass_change(m0, m):
    """src/output.py: relative change, absolute when the initial mass vanishes."""
    return abs(m0 - m) / abs(m0) if abs(m0) > 1e-10 else abs(m0 - m)


STUDY_NAMES = ("PL07-RK1", "PL07-RK1-DG-PR", "AVLT-RK2-DG-AF", "ALVT-RK2-DG-PR")   # src/advection_error.py:170


def print_errors_simul(error_linf, error_l1, error_l2, i):
    """src/errors.py:84-94: the errors of run i and, from the second run on, the ratio to run i-1."""
    if i > 0:
        print("Norms          (Linf,    L1,       L2) ")
        print("Error E_" + str(i) + "    :", "{:.2e}".format(error_linf[i]), "{:.2e}".format(error_l1[i]),
              "{:.2e}".format(error_l2[i]))
        print("Ratio E_" + str(i) + "/E_" + str(i - 1) + ":", "{:.2e}".format(error_linf[i - 1] / error_linf[i]),
              "{:.2e}".format(error_l1[i - 1] / error_l1[i]), "{:.2e}".format(error_l2[i - 1] / error_l2[i]))
    else:
        print("Error(Linf, L1, L2) :", "{:.2e}".format(error_linf[i]), "{:.2e}".format(error_l1[i]),
              "{:.2e}".format(error_l2[i]))


def convergence_rates(err):
    """|log E_i - log E_(i-1)| / log 2 along the resolution axis (src/advection_error.py:163-164)."""
    return np.abs(np.log(err[1:]) - np.log(err[:-1])) / np.log(2.0)


def error_analysis_adv(simulation, ntest=6, n0=16, device="cuda:0", verbose=False, dtype="f64"):
    """src/advection_error.py:21-188 without plots: the 4-scheme x ntest-resolution study of the
    simulation's (ic, vf), Tf = 5, N = n0 2^i, dt halving from the vf's first step.  Returns
    (Ns, errors [ntest][4][3] as (linf, l1, l2), rates [ntest-1][4][3])."""
    vf, ic = simulation.vf, simulation.ic
    Ns = [n0 * 2 ** i for i in range(ntest)]
    dts = [STUDY_DT0[vf] * 0.5 ** i for i in range(ntest)]
    err = np.zeros((ntest, len(STUDY_SCHEMES), 3))
    for d, (recon, opsplit, et, mt, dp, mf) in enumerate(STUDY_SCHEMES):
        for i, (N, dt) in enumerate(zip(Ns, dts)):
            g = Grid(N)
            sim = adv_simulation_par(g, dt, 5.0, ic, vf, 2, recon, dp, opsplit, et, mt, mf, dtype=dtype, device=device)
            if verbose:
                print("\nParameters: N=" + str(N) + ", dt=" + str(dt), ", scheme=", STUDY_NAMES[d])
            err[i, d] = adv_sphere(g, sim)
            sim.close()
            if verbose:
                print_errors_simul(err[:, d, 0], err[:, d, 1], err[:, d, 2], i)
    return Ns, err, convergence_rates(err)


def main(pardir="par"):
    """main.py test case 5 from the par files: tc = 1 runs one advection, tc = 2 the study."""
    gp = config.get_parameters(pardir)
    ap = config.get_advection_parameters(pardir)
    if gp.test_case != 5:
        raise CsadvError("only test case 5 (advection) is on this path; par file asks for %d" % gp.test_case)
    g = gridfile.cubed_sphere(gp.N, gp.gridload, griddir=os.path.join(os.path.dirname(os.path.abspath(pardir)), "grid"),
                              verbose=gp.showonscreen)
    sim = adv_simulation_par(g, ap.dt, ap.Tf, ap.ic, ap.vf, ap.tc, ap.recon, ap.dp, ap.opsplit, ap.et, ap.mt, ap.mf)
    if ap.tc == 1:
        return adv_sphere(g, sim, verbose=True)
    if ap.tc == 2:
        sim.close()
        return error_analysis_adv(sim, verbose=True)
    raise CsadvError("invalid advection test case %r" % ap.tc)


def _cli(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m cubedsphere.advection",
                                 description="Test case 5 from the par files, or the convergence study of one (ic, vf).")
    ap.add_argument("pardir", nargs="?", default="par", help="directory holding configuration.par and advection.par")
    ap.add_argument("--study", nargs=2, type=int, metavar=("IC", "VF"),
                    help="run the 4-scheme convergence study of (ic, vf) without par files")
    ap.add_argument("--ntest", type=int, default=6, help="resolutions of the study: N = 16 * 2^i, i < ntest")
    ap.add_argument("--dtype", default="f64", choices=("f64", "f32"))
    args = ap.parse_args(argv)
    if args.study:
        class _P:
            ic, vf = args.study
        Ns, err, rates = error_analysis_adv(_P, ntest=args.ntest, verbose=True, dtype=args.dtype)
        print("\nic=%d vf=%d, Tf=5, %s" % (args.study[0], args.study[1], args.dtype))
        print("%-16s %6s %10s %10s %10s %6s %6s %6s" % ("scheme", "N", "Linf", "L1", "L2", "r_inf", "r_1", "r_2"))
        for d, name in enumerate(STUDY_NAMES):
            for i, N in enumerate(Ns):
                r = rates[i - 1, d] if i else (np.nan, np.nan, np.nan)
                print("%-16s %6d %10.3e %10.3e %10.3e %6.2f %6.2f %6.2f" % ((name, N) + tuple(err[i, d]) + tuple(r)))
        return 0
    print(main(args.pardir))
    return 0


if __name__ == "__main__":
    raise SystemExit(_cli())
