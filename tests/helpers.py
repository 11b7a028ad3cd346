"""Shared test helpers: oracle runs and layout conversion (tests only)."""
import math

import numpy as np

from oracle import timestep as ots
from oracle.cases import from_case

PI = math.pi


def ref_to_dev(a):
    """Reference layout [i, j, p] -> device layout [p, j, i]."""
    return np.ascontiguousarray(np.transpose(a, (2, 1, 0)))


def oracle_case(case):
    """Translate a product Case into the oracle's case dict."""
    return from_case(case)


def run_oracle(N, case, dt, nsteps, **opts):
    grid, sim = ots.make(N, oracle_case(case), dt, **opts)
    ots.init_vars_adv(grid, sim)
    for k in range(1, nsteps + 1):
        t = k * dt
        ots.adv_time_step(grid, sim, k, t)
        ots.update_adv(grid, sim, t)
    return grid, sim


def interior_dev(grid, Q):
    I = slice(grid.i0, grid.iend)
    return ref_to_dev(Q[I, I, :])


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
