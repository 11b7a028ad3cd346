"""Readers for the reference's positional parameter files and the extension file.

``par/configuration.par`` and ``par/advection.par`` are read exactly as
src/configuration.py:13-158 reads them: by line position, every other line being a
comment that is skipped by count (not by '#').  Invalid values raise ``ValueError``
instead of printing and calling ``exit()``.

Knobs that the reference files cannot express (extension initial conditions and
winds, CFL-based time steps, precision, batching) live in a separate optional
``key = value`` file, ``par/extension.par``.
"""
import os
from dataclasses import dataclass, field

TRANSFORMATIONS = {1: "gnomonic_equidistant", 2: "gnomonic_equiangular", 3: "overlapped"}
MAPS = {1: "mercator", 2: "sphere"}


@dataclass
class GridParams:
    N: int
    transformation: str
    showonscreen: bool
    gridload: bool
    test_case: int
    map_projection: str


@dataclass
class AdvectionParams:
    dt: float
    Tf: float
    tc: int
    ic: int
    vf: int
    recon: int
    dp: int
    opsplit: int
    et: int
    mt: int
    mf: int


def _values(path, count):
    """Values on lines 2, 4, ... after the two-line header (src/configuration.py:22-46)."""
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    with open(path) as f:
        lines = f.readlines()
    vals = []
    pos = 2
    for _ in range(count):
        if pos >= len(lines):
            raise ValueError("%s: expected %d values" % (path, count))
        vals.append(lines[pos].strip())
        pos += 2
    return vals


def get_parameters(pardir="par"):
    """src/configuration.py:13-84 -> GridParams."""
    v = _values(os.path.join(pardir, "configuration.par"), 6)
    N, transf, gridload, show, tcase, mp = int(v[0]), int(v[1]), bool(int(v[2])), bool(int(v[3])), int(v[4]), int(v[5])
    if mp not in MAPS:
        raise ValueError("invalid map projection %d" % mp)
    if transf not in TRANSFORMATIONS:
        raise ValueError("invalid transformation %d" % transf)
    if transf == 3:
        gridload = True
    return GridParams(N, TRANSFORMATIONS[transf], show, gridload, tcase, MAPS[mp])


def get_advection_parameters(pardir="par"):
    """src/configuration.py:90-158 -> AdvectionParams (same field order as the reference tuple)."""
    v = _values(os.path.join(pardir, "advection.par"), 11)
    Tf, dt = float(v[0]), float(v[1])
    ic, vf, tc, recon, dp, opsplit, et, mt, mf = (int(x) for x in v[2:])
    return AdvectionParams(dt, Tf, tc, ic, vf, recon, dp, opsplit, et, mt, mf)


def check_options(p):
    """Option ranges as decoded by src/advection_ic.py:53-149 and the (opsplit, mt) pairs
    accepted by src/discrete_operators.py:48-82."""
    rng = {"ic": (1, 4), "vf": (1, 4), "tc": (1, 2), "recon": (1, 4), "dp": (1, 2), "opsplit": (1, 3),
           "et": (1, 3), "mt": (1, 2), "mf": (1, 3)}
    for k, (lo, hi) in rng.items():
        val = getattr(p, k)
        if not lo <= val <= hi:
            raise ValueError("invalid %s = %r" % (k, val))
    if not ((p.opsplit in (1, 2) and p.mt == 1) or (p.opsplit == 3 and p.mt == 2)):
        raise ValueError("invalid (opsplit, mt) pair (%d, %d)" % (p.opsplit, p.mt))


@dataclass
class Extension:
    """Optional knobs (par/extension.par).  ``workload`` names a predefined case."""
    workload: str = ""
    dtype: str = "f64"
    cfl: float = 0.0           # > 0: dt from the CFL number instead of advection.par
    steps: int = 0             # > 0: number of steps instead of int(Tf/dt)
    tile: int = 0
    extra: dict = field(default_factory=dict)


def get_extension(pardir="par"):
    path = os.path.join(pardir, "extension.par")
    ext = Extension()
    if not os.path.exists(path):
        return ext
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if "=" not in line:
                raise ValueError("%s: expected key = value, got %r" % (path, line))
            k, v = (s.strip() for s in line.split("=", 1))
            if k in ("workload", "dtype"):
                setattr(ext, k, v)
            elif k in ("cfl",):
                setattr(ext, k, float(v))
            elif k in ("steps", "tile"):
                setattr(ext, k, int(v))
            else:
                ext.extra[k] = v
    if ext.dtype not in ("f64", "f32"):
        raise ValueError("dtype must be f64 or f32")
    return ext
