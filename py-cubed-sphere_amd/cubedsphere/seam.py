"""The reference's own objects on the device step: the seam of src/advection_sphere.py.

The reference's time loop (src/advection_sphere.py:13-61) works on two objects it builds itself:
``cs_grid`` (``cubed_sphere``, src/cs_datastruct.py:24-541) and ``simulation``
(``adv_simulation_par``, src/advection_ic.py:21-203).  Between steps it reads
``simulation.Q`` -- a numpy array ``[i, j, panel]`` with the 4-cell halo -- in
``mass_computation`` (src/diagnostics.py:14-26) and ``output_adv`` (src/output.py:22-150).

``bind(cs_grid, simulation)`` (called by ``cubedsphere.advection.init_vars_adv`` for such
objects) attaches a device solve to ``simulation`` and turns its ``Q`` into a property:

* reading ``simulation.Q`` returns the host array with the interior of the current device state
  (pending mass-fixer totals folded), refreshed only when a step ran since the last read, so a
  loop that reads it once per run (plot off) moves no state over PCIe between steps;
* assigning ``simulation.Q = array`` loads the array's interior into the device state.

Ghost cells of the host array are not maintained: the step fills its own halo on the device,
as the reference refills it at the start of every step (src/advection_timestep.py:28), and
nothing after the step reads them.  Writing into the returned array in place does not reach
the device (assign the array instead).

Option codes come from the object's ``recon``, ``dp``, ``opsplit``, ``edge_treatment``,
``metric_tensor`` and ``mass_fixer`` attributes (src/advection_ic.py:38-168), or from their
``*_name`` strings; the case from ``ic`` and ``vf``.  The grid must be the equiangular
gnomonic cubed sphere with the reference's 4-cell halo; its metric is checked against the one
the device tables are built from.
"""
import numpy as np

from ._lib import CsadvError
from .cases import reference_case
from .solver import Advection, grid_tables

_NAME_CODES = {
    "recon": {"PPM-0": 1, "PPM-CW84": 2, "PPM-PL07": 3, "PPM-L04": 4},
    "dp": {"RK1": 1, "RK2": 2},
    "opsplit": {"SP-AVLT": 1, "SP-L04": 2, "SP-PL07": 3},
    "et": {"ET-S72": 1, "ET-PL07": 2, "ET-DG": 3},
    "mt": {"MT-0": 1, "MT-PL07": 2},
    "mf": {"MF-0": 1, "MF-AF": 2, "MF-PR": 3},
}
# option -> (integer attribute, name attribute) of adv_simulation_par
_ATTRS = {"recon": ("recon", "recon_name"), "dp": ("dp", "dp_name"), "opsplit": ("opsplit", "opsplit_name"),
          "et": ("edge_treatment", "et_name"), "mt": ("metric_tensor", "mt_name"), "mf": ("mass_fixer", "mf_name")}


def options_of(simulation):
    """(recon, dp, opsplit, et, mt, mf) codes of a reference-shaped simulation object."""
    out = {}
    for opt, (code_attr, name_attr) in _ATTRS.items():
        code = getattr(simulation, code_attr, None)
        if not isinstance(code, (int, np.integer)):
            name = getattr(simulation, name_attr, None)
            if name not in _NAME_CODES[opt]:
                raise CsadvError("simulation object has neither %s nor a known %s" % (code_attr, name_attr))
            code = _NAME_CODES[opt][name]
        out[opt] = int(code)
    return out


def is_bound(simulation):
    return "_csadv" in getattr(simulation, "__dict__", {})


def device_solve(simulation):
    """The Advection solve bound to a reference-shaped simulation object."""
    if not is_bound(simulation):
        raise CsadvError("init_vars_adv(cs_grid, simulation) first")
    return simulation.__dict__["_csadv"]


def _interior(cs_grid):
    return slice(cs_grid.i0, cs_grid.iend), slice(cs_grid.j0, cs_grid.jend)


def _get_Q(self):
    d = self.__dict__
    dev = d["_csadv"]
    host = d["_csadv_Q"]
    if d["_csadv_Q_step"] != dev.steps_done:
        I, J = d["_csadv_IJ"]
        host[I, J, :] = np.transpose(dev.state()[0], (2, 1, 0))
        d["_csadv_Q_step"] = dev.steps_done
    return host


def _set_Q(self, value):
    d = self.__dict__
    value = np.asarray(value, dtype=np.float64)
    if "_csadv" not in d:           # before binding: a plain attribute
        d["_csadv_Q"] = value
        return
    host = d["_csadv_Q"]
    if value.shape != host.shape:
        raise CsadvError("simulation.Q must keep its shape %s" % (host.shape,))
    host[...] = value
    I, J = d["_csadv_IJ"]
    d["_csadv"].load_state(np.transpose(host[I, J, :], (2, 1, 0)))
    d["_csadv_Q_step"] = d["_csadv"].steps_done


_BOUND_CLASSES = {}


def _bound_class(cls):
    """Subclass of the simulation's own class whose ``Q`` is the device-synchronised property
    (a data descriptor, so it takes precedence over the instance attribute)."""
    if cls not in _BOUND_CLASSES:
        _BOUND_CLASSES[cls] = type(cls.__name__, (cls,), {"Q": property(_get_Q, _set_Q, doc=__doc__),
                                                          "__module__": cls.__module__})
    return _BOUND_CLASSES[cls]


def _check_grid(cs_grid, grid):
    if getattr(cs_grid, "projection", "gnomonic_equiangular") != "gnomonic_equiangular":
        raise CsadvError("the device step runs on the equiangular gnomonic grid only (got %r)" % cs_grid.projection)
    if (cs_grid.i0, cs_grid.j0, cs_grid.iend, cs_grid.jend) != (grid.i0, grid.j0, grid.iend, grid.jend):
        raise CsadvError("the device step expects the reference's 4-cell halo (i0 = 4, iend = N + 4)")
    if abs(cs_grid.dx - grid.dx) > 1e-15 * grid.dx or abs(cs_grid.dy - grid.dy) > 1e-15 * grid.dy:
        raise CsadvError("grid spacing differs from the equiangular grid of N = %d" % grid.N)
    g_ref = getattr(cs_grid, "metric_tensor_pc", None)
    if g_ref is not None:
        ours = np.repeat(np.asarray(grid.pc.g).T[:, :, None], 6, axis=2)      # [i, j, panel]
        if np.max(np.abs(np.asarray(g_ref) - ours)) > 1e-13 * np.max(np.abs(ours)):
            raise CsadvError("cs_grid.metric_tensor_pc differs from the equiangular metric of N = %d" % grid.N)


def bind(cs_grid, simulation, dtype="f64", device="cuda:0", tile=0):
    """init_vars_adv (src/advection_vars.py:19-107) for the reference's own objects: a device
    solve with the object's options and case, its initial condition loaded, and the host-side
    results of the reference's init written back onto the object (``Q``, ``CFL``, ``cx``, ``cy``
    and the normal winds ``U_pu.ucontra`` / ``U_pv.vcontra`` when the object has them)."""
    opts = options_of(simulation)
    ic, vf = int(simulation.ic), int(simulation.vf)
    dt = float(simulation.dt)
    Tf = float(getattr(simulation, "Tf", 0.0) or 0.0)
    N = int(cs_grid.N)
    tab = grid_tables(N, dtype, device)
    _check_grid(cs_grid, tab.grid)
    old = simulation.__dict__.pop("_csadv", None)
    if old is not None:
        old.close()
    dev = Advection(N, reference_case(ic, vf), dt, recon=opts["recon"], dp=opts["dp"], opsplit=opts["opsplit"],
                    et=opts["et"], mt=opts["mt"], mf=opts["mf"], dtype=dtype, device=device, tile=tile,
                    max_steps=max(1, int(Tf / dt)) if Tf > 0 else 1, tables=tab)
    dev.init_vars_adv()
    M = N + 8
    d = simulation.__dict__
    host = d.pop("Q", d.get("_csadv_Q"))
    if not (isinstance(host, np.ndarray) and host.shape == (M, M, 6) and host.dtype == np.float64):
        host = np.zeros((M, M, 6))
    host[...] = 0.0
    d["_csadv_Q"] = host
    d["_csadv_Q_step"] = -1
    d["_csadv_IJ"] = _interior(cs_grid)
    d["_csadv_t"] = 0.0                  # time of the last wind update (RK2 extrapolates from it)
    d["_csadv"] = dev
    if type(simulation) not in _BOUND_CLASSES.values():
        simulation.__class__ = _bound_class(type(simulation))
    _host_winds(cs_grid, simulation, dev)
    return dev


def _host_winds(cs_grid, simulation, dev):
    """CFL number and Courant fields of src/advection_vars.py:88-98, and the normal winds of
    the first step, from the device's prepared winds."""
    M = cs_grid.N + 8
    uu = dev.usgn if dev.usgn is not None else dev.uavg
    vv = dev.vsgn if dev.vsgn is not None else dev.vavg
    u = np.transpose(uu[:, :, :M + 1].double().cpu().numpy(), (2, 1, 0))     # [e, j, p]
    v = np.transpose(vv[:, :, :M].double().cpu().numpy(), (2, 1, 0))         # [i, e, p]
    cx, cy = u * dev.dt / cs_grid.dx, v * dev.dt / cs_grid.dy
    i0, iend, j0, jend = cs_grid.i0, cs_grid.iend, cs_grid.j0, cs_grid.jend
    simulation.cx, simulation.cy = cx, cy
    simulation.CFL = max(abs(np.amax(cx[i0:iend + 1, :, :])), abs(np.amax(cy[:, j0:jend + 1, :])))
    for attr, comp, arr in (("U_pu", "ucontra", u), ("U_pv", "vcontra", v)):
        vel = getattr(simulation, attr, None)
        target = getattr(vel, comp, None) if vel is not None else None
        if isinstance(target, np.ndarray) and target.shape == arr.shape:
            target[...] = arr


def step(cs_grid, simulation, k, t):
    """adv_time_step (src/advection_timestep.py:19-43): one step of the bound device solve with
    the winds the last update_adv left (the reference's step reads neither k nor t)."""
    device_solve(simulation).step(1)


def update(cs_grid, simulation, t):
    """update_adv (src/advection_timestep.py:48-75): for a time-dependent wind (vf >= 2) the
    next step's winds become those at ``t``, and the previous ones those of the last update."""
    dev = device_solve(simulation)
    if int(simulation.vf) < 2 or dev.case.steady:
        return
    d = simulation.__dict__
    dev.set_wind_times(float(t), d["_csadv_t"])
    d["_csadv_t"] = float(t)
