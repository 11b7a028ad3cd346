"""Problem definitions: initial conditions, exact solutions and winds (host, FP64).

Reference cases are the ``(ic, vf)`` codes of src/advection_ic.py:208-313.  Winds are
represented as a short *basis*: ``u(lon, lat, t) = sum_m s_m(t) * b_m(lon, lat)``.  The
contravariant conversion and the duo-grid vector ghost fill are linear, so the device
builds each basis field once and recombines them every step with the coefficients
``s_m(t)``:

  vf 1, vf 4, solid body : one field, s = 1 (steady)
  vf 3, NL10 case 1      : one field, s = cos(pi t / T)
  vf 2 (NL10 case 4)     : four fields (the translation lon - 2 pi t / T expanded with
                           sin^2(a-b) = (1 - cos 2a cos 2b - sin 2a sin 2b)/2)

Extension cases (parametrised solid-body rotation, NL10 case 1 and the NL10 tracer
fields) follow Williamson et al. (1992) test 1 and Nair & Lauritzen (2010) sec. 2-3.
"""
import math

import numpy as np
import torch

PI = math.pi
DEG2RAD = 1.0 / (180.0 / PI)


# Every field and wind below takes numpy arrays (host, the parity path) or torch tensors
# (device preparation of the batched and served forms) and computes with the same library.
def _xp(a):
    return torch if torch.is_tensor(a) else np


def sin(a):
    return _xp(a).sin(a)


def cos(a):
    return _xp(a).cos(a)


def sph2cart(lon, lat):
    return cos(lat) * cos(lon), cos(lat) * sin(lon), sin(lat)


# ---------------------------------------------------------------- scalar fields
def gaussian_sum(X, Y, Z, centres, b, hmax=1.0):
    xp = _xp(X)
    q = xp.zeros_like(X)
    for lon0, lat0 in centres:
        X0, Y0, Z0 = sph2cart(lon0, lat0)
        q = q + hmax * xp.exp(-b * ((X - X0) ** 2 + (Y - Y0) ** 2 + (Z - Z0) ** 2))
    return q


def _gc_dist(lon, lat, lon0, lat0):
    xp = _xp(lon)
    c = math.sin(lat0) * sin(lat) + math.cos(lat0) * cos(lat) * cos(lon - lon0)
    return xp.arccos(xp.clip(c, -1.0, 1.0))


def cosine_bells(lon, lat, centres, radius=0.5, base=0.1, scale=0.9):
    xp = _xp(lon)
    q = xp.zeros_like(lon)
    for lon0, lat0 in centres:
        r = _gc_dist(lon, lat, lon0, lat0)
        q = q + xp.where(r < radius, 0.5 * (1.0 + cos(PI * r / radius)), 0.0)
    return base + scale * q


def slotted_cylinders(lon, lat, centres, radius=0.5, high=1.0, low=0.1):
    xp = _xp(lon)
    q = xp.full_like(lon, low)
    for n, (lon0, lat0) in enumerate(centres):
        r = _gc_dist(lon, lat, lon0, lat0)
        d = lon - lon0
        dlon = xp.abs(xp.arctan2(sin(d), cos(d)))        # |lon - lon0| wrapped to [0, pi]
        slot = (lat - lat0 < -5.0 * radius / 12.0) if n % 2 == 0 else (lat - lat0 > 5.0 * radius / 12.0)
        keep = (dlon >= radius / 6.0) | ((dlon < radius / 6.0) & slot)
        q = xp.where((r <= radius) & keep, high, q)
    return q


def named_field(spec, X, Y, Z):
    xp = _xp(X)
    lon = xp.arctan2(Y, X)
    lat = xp.arctan2(Z, xp.hypot(X, Y))
    kind = spec["field"]
    if kind == "gaussians":
        return gaussian_sum(X, Y, Z, spec["centres"], spec.get("b", 5.0), spec.get("hmax", 1.0))
    if kind == "cosine_bells":
        return cosine_bells(lon, lat, spec["centres"], spec.get("radius", 0.5))
    if kind == "slotted_cylinders":
        return slotted_cylinders(lon, lat, spec["centres"], spec.get("radius", 0.5))
    if kind == "nl10_mix":
        return 0.5 * gaussian_sum(X, Y, Z, spec["centres"], spec.get("b", 5.0), 0.95) + \
            0.5 * cosine_bells(lon, lat, spec["centres"], spec.get("radius", 0.5))
    raise ValueError("unknown field %r" % kind)


def reference_q(ic, vf, lon, lat, t):
    """Exact solutions of src/advection_ic.py:215-281."""
    if ic == 1:
        return _xp(lon).ones_like(lon)
    X, Y, Z = sph2cart(lon, lat)
    if ic == 2:
        if vf == 1:
            alpha, u0 = -45.0 * DEG2RAD, 2.0 * PI / 5.0
            wt = -u0 * t
            ca, sa, cw, sw = math.cos(alpha), math.sin(alpha), math.cos(wt), math.sin(wt)
            rX = (cw * ca * ca + sa * sa) * X - sw * ca * Y + (cw * ca * sa - ca * sa) * Z
            rY = sw * ca * X + cw * Y + sa * sw * Z
            rZ = (cw * sa * ca - sa * ca) * X - sw * sa * Y + (cw * sa * sa + ca * ca) * Z
            return gaussian_sum(rX, rY, rZ, [(PI / 4.0, PI / 6.0)], 10.0)
        return gaussian_sum(X, Y, Z, [(0.0, 0.0)], 10.0)
    if ic == 3:
        cs = [(0, PI / 3.0), (0, -PI / 3.0)] if vf == 1 else [(-PI / 6.0, 0), (PI / 6.0, 0)]
        return gaussian_sum(X, Y, Z, cs, 5.0)
    if ic == 4:
        alpha = -45.0 * DEG2RAD
        f = -cos(lon) * cos(lat) * math.sin(alpha) + sin(lat) * math.cos(alpha)
        return 1.0 - f * f
    raise ValueError("invalid initial condition %r" % ic)


def _rotated_back(lon, lat, alpha, angle):
    """Points rotated by -angle about (-sin alpha, 0, cos alpha) (Rodrigues)."""
    X, Y, Z = sph2cart(lon, lat)
    kx, kz = -math.sin(alpha), math.cos(alpha)
    c, s = math.cos(-angle), math.sin(-angle)
    kd = kx * X + kz * Z
    Xr = X * c + (-kz * Y) * s + kx * kd * (1.0 - c)
    Yr = Y * c + (kz * X - kx * Z) * s
    Zr = Z * c + (kx * Y) * s + kz * kd * (1.0 - c)
    return Xr, Yr, Zr


# ---------------------------------------------------------------- cases
class Case:
    """A problem: tracers, wind basis, exact solutions.

    ``tracers``: list of reference ``ic`` codes or extension field specs (dicts).
    ``wind``: ("ref", vf) | ("solid_body", alpha, period) | ("deformational", k, period).
    """

    def __init__(self, tracers, wind, name=""):
        self.tracers = list(tracers)
        self.wind = wind
        self.name = name

    @property
    def ntr(self):
        return len(self.tracers)

    @property
    def ref_vf(self):
        return self.wind[1] if self.wind[0] == "ref" else 0

    @property
    def steady(self):
        kind = self.wind[0]
        return kind == "solid_body" or (kind == "ref" and self.wind[1] in (1, 4))

    def basis(self):
        """[(ulon_fn, vlat_fn)] of the wind basis fields (functions of lon, lat)."""
        kind = self.wind[0]
        if kind == "ref":
            vf = self.wind[1]
            if vf == 1:
                alpha, u0 = -45.0 * DEG2RAD, 2.0 * PI / 5.0
                return [(lambda lo, la: u0 * (cos(la) * cos(alpha) + sin(la) * cos(lo) * sin(alpha)),
                         lambda lo, la: -u0 * sin(lo) * sin(alpha))]
            if vf == 2:
                T, k = 5.0, 2.0
                return [(lambda lo, la: 2.0 * PI * cos(la) / T, lambda lo, la: 0.0 * lo),
                        (lambda lo, la: 0.5 * k * sin(2.0 * la), lambda lo, la: 0.0 * lo),
                        (lambda lo, la: -0.5 * k * sin(2.0 * la) * cos(2.0 * lo),
                         lambda lo, la: k * cos(la) * sin(2.0 * lo)),
                        (lambda lo, la: -0.5 * k * sin(2.0 * la) * sin(2.0 * lo),
                         lambda lo, la: -k * cos(la) * cos(2.0 * lo))]
            if vf == 3:
                k = 1.0
                return [(lambda lo, la: -k * sin((lo + PI) / 2.0) ** 2 * sin(2.0 * la) * cos(la) ** 2,
                         lambda lo, la: (k / 2.0) * sin(lo + PI) * cos(la) ** 3)]
            if vf == 4:
                return [(lambda lo, la: -(sin(lo) * sin(lo) * cos(la) ** 3),
                         lambda lo, la: -4.0 * (cos(la) ** 3) * sin(la) * cos(lo) * sin(lo))]
            raise ValueError("invalid vector field %r" % vf)
        if kind == "solid_body":
            alpha, period = self.wind[1], self.wind[2]
            u0 = 2.0 * PI / period
            return [(lambda lo, la: u0 * (cos(la) * cos(alpha) + sin(la) * cos(lo) * sin(alpha)),
                     lambda lo, la: -u0 * sin(lo) * sin(alpha))]
        if kind == "deformational":
            k = self.wind[1]
            return [(lambda lo, la: k * sin(lo / 2.0) ** 2 * sin(2.0 * la),
                     lambda lo, la: (k / 2.0) * sin(lo) * cos(la))]
        raise ValueError("unknown wind %r" % (kind,))

    def coefficients(self, t):
        """s_m(t) for the basis above."""
        kind = self.wind[0]
        if self.steady:
            return np.ones(1)
        if kind == "ref" and self.wind[1] == 2:
            T = 5.0
            c, w2 = math.cos(PI * t / T), 2.0 * (2.0 * PI / T) * t
            return np.array([1.0, c, c * math.cos(w2), c * math.sin(w2)])
        period = 5.0 if kind == "ref" else self.wind[2]
        return np.array([math.cos(PI * t / period)])

    def coefficient_table(self, nsteps, dt):
        """Rows n = 0..nsteps-1: (s(t_n), s(t_{n-1})) with t_{-1} := t_0, the winds step n+1 uses
        (src/advection_timestep.py:31-37 with update_adv at :48-75)."""
        K = len(self.basis())
        n = max(nsteps, 1)
        cur = np.array([self.coefficients(k * dt) for k in range(n)], dtype=np.float64).reshape(n, K)
        tab = np.empty((n, 2, K))
        tab[:, 0] = cur
        tab[0, 1] = cur[0]              # s(t_{n-1}) is row n-1's s(t): the same value
        tab[1:, 1] = cur[:-1]
        return tab

    def q(self, n, lon, lat, t):
        spec = self.tracers[n]
        if isinstance(spec, int):
            return reference_q(spec, self.ref_vf, lon, lat, t)
        if self.wind[0] == "solid_body":
            X, Y, Z = _rotated_back(lon, lat, self.wind[1], 2.0 * PI / self.wind[2] * t)
        else:
            X, Y, Z = sph2cart(lon, lat)   # time-reversing flows: exact at t = period
        return named_field(spec, X, Y, Z)


def reference_case(ic, vf):
    return Case([ic], ("ref", vf), name="ic%d_vf%d" % (ic, vf))


NL10_CENTRES = [(5.0 * PI / 6.0, 0.0), (7.0 * PI / 6.0, 0.0)]


def workload(name, seed=0):
    """Predefined cases of the benchmark configurations."""
    if name == "solid_body_polar":      # N=48 revolution, 12-day period, polar axis
        return Case([{"field": "gaussians", "centres": [(3.0 * PI / 2.0, 0.0)], "b": 10.0}],
                    ("solid_body", 0.0, 12.0 * 86400.0), name)
    if name == "two_hills_tilted":      # N=384 FP64 parity run
        return Case([{"field": "gaussians", "centres": NL10_CENTRES, "b": 5.0}],
                    ("solid_body", -PI / 4.0, 5.0), name)
    if name == "nl10_deformation":      # N=768 precision form, 4 tracers
        return Case([{"field": "gaussians", "centres": NL10_CENTRES, "b": 5.0, "hmax": 0.95},
                     {"field": "cosine_bells", "centres": NL10_CENTRES, "radius": 0.5},
                     {"field": "slotted_cylinders", "centres": NL10_CENTRES, "radius": 0.5},
                     {"field": "nl10_mix", "centres": NL10_CENTRES}],
                    ("deformational", 2.4, 5.0), name)
    raise ValueError("unknown workload %r" % name)


def random_solid_body(rng, ntr):
    """A seeded random solve of the batched form: Gaussian centre uniform on the sphere,
    width b in [5, 20], rotation-axis tilt in [0, pi/2]."""
    tracers = []
    for _ in range(ntr):
        z = rng.uniform(-1.0, 1.0)
        lon0 = rng.uniform(-PI, PI)
        tracers.append({"field": "gaussians", "centres": [(lon0, math.asin(z))], "b": rng.uniform(5.0, 20.0)})
    return Case(tracers, ("solid_body", -rng.uniform(0.0, PI / 2.0), 5.0), "random_solid_body")
