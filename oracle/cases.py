"""Initial conditions, winds and exact solutions -- oracle (test infrastructure).

Reference cases restate src/advection_ic.py:208-329 (``qexact_adv``, ``velocity_adv``,
``div_exact``).  Extension cases (parametrised solid-body rotation, NL10 case-1
deformational wind, multi-tracer NL10 fields) are the analytic definitions the
benchmark workloads need; they follow Williamson et al. (1992) test 1 and
Nair & Lauritzen (2010) sections 2-3 and are written from those formulas.

A *case* is described by a plain dict (the same dict the product reads):
  {"ic": int | "gaussians" | "cosine_bells" | "slotted_cylinders" | "nl10_mix",
   "vf": int | "solid_body" | "deformational", ... parameters ...}
"""
import math

import numpy as np

from .grid import sph2cart

PI = math.pi
DEG2RAD = 1.0 / (180.0 / PI)


# ----------------------------------------------------------------------------- reference winds
def velocity_adv(lon, lat, t, vf):
    """src/advection_ic.py:287-313"""
    if vf == 1:
        alpha = -45.0 * DEG2RAD
        u0 = 2.0 * PI / 5.0
        ulon = u0 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha))
        vlat = -u0 * np.sin(lon) * np.sin(alpha)
    elif vf == 2:
        T, k = 5.0, 2.0
        lonp = lon - 2 * PI * t / T
        ulon = k * (np.sin((lonp + PI)) ** 2) * (np.sin(2.0 * lat)) * (np.cos(PI * t / T)) + 2.0 * PI * np.cos(lat) / T
        vlat = k * (np.sin(2 * (lonp + PI))) * (np.cos(lat)) * (np.cos(PI * t / T))
    elif vf == 3:
        T, k = 5.0, 1.0
        ulon = -k * (np.sin((lon + PI) / 2.0) ** 2) * (np.sin(2.0 * lat)) * (np.cos(lat) ** 2) * (np.cos(PI * t / T))
        vlat = (k / 2.0) * (np.sin((lon + PI))) * (np.cos(lat) ** 3) * (np.cos(PI * t / T))
    elif vf == 4:
        m = n = 1
        ulon = -m * (np.sin(lon) * np.sin(m * lon) * np.cos(n * lat) ** 3)
        vlat = -4 * n * (np.cos(n * lat) ** 3) * np.sin(n * lat) * np.cos(m * lon) * np.sin(lon)
    else:
        raise ValueError("invalid vector field %r" % vf)
    return ulon, vlat


def qexact_ref(lon, lat, t, ic, vf):
    """src/advection_ic.py:215-281"""
    if ic == 1:
        return np.ones(np.shape(lon))
    if ic == 2:
        if vf == 1:
            alpha = -45.0 * DEG2RAD
            u0 = 2.0 * PI / 5.0
            wt = -u0 * t
            cosa, sina = np.cos(alpha), np.sin(alpha)
            cos2a, sin2a = cosa * cosa, sina * sina
            coswt, sinwt = np.cos(wt), np.sin(wt)
            X, Y, Z = sph2cart(lon, lat)
            rotX = (coswt * cos2a + sin2a) * X - sinwt * cosa * Y + (coswt * cosa * sina - cosa * sina) * Z
            rotY = sinwt * cosa * X + coswt * Y + sina * sinwt * Z
            rotZ = (coswt * sina * cosa - sina * cosa) * X - sinwt * sina * Y + (coswt * sin2a + cos2a) * Z
            X0, Y0, Z0 = sph2cart(PI / 4.0, PI / 6.0)
            b0 = 10.0
            return np.exp(-b0 * ((rotX - X0) ** 2 + (rotY - Y0) ** 2 + (rotZ - Z0) ** 2))
        X, Y, Z = sph2cart(lon, lat)
        X0, Y0, Z0 = sph2cart(0.0, 0.0)
        return np.exp(-10.0 * ((X - X0) ** 2 + (Y - Y0) ** 2 + (Z - Z0) ** 2))
    if ic == 3:
        X, Y, Z = sph2cart(lon, lat)
        if vf == 1:
            (lon1, lat1), (lon2, lat2) = (0, PI / 3.0), (0, -PI / 3.0)
        else:
            (lon1, lat1), (lon2, lat2) = (-PI / 6.0, 0), (PI / 6.0, 0)
        X1, Y1, Z1 = sph2cart(lon1, lat1)
        X2, Y2, Z2 = sph2cart(lon2, lat2)
        b0 = 5.0
        return (np.exp(-b0 * ((X - X1) ** 2 + (Y - Y1) ** 2 + (Z - Z1) ** 2))
                + np.exp(-b0 * ((X - X2) ** 2 + (Y - Y2) ** 2 + (Z - Z2) ** 2)))
    if ic == 4:
        alpha = -45.0 * DEG2RAD
        f = (-np.cos(lon) * np.cos(lat) * np.sin(alpha) + np.sin(lat) * np.cos(alpha))
        return 1.0 - f * f
    raise ValueError("invalid initial condition %r" % ic)


def div_exact(lon, lat, vf):
    """src/advection_ic.py:318-329 (vf=3 returns the vf=4 formula, as in the reference)."""
    if vf <= 2:
        return np.zeros(np.shape(lon))
    m = n = 1
    return (-np.cos(lon) * np.sin(m * lon) * m * np.cos(n * lat) ** 4 / np.cos(lat)
            - np.sin(lon) * np.cos(m * lon) * m ** 2 * np.cos(n * lat) ** 4 / np.cos(lat)
            + 12.0 * np.sin(lon) * np.cos(m * lon) * np.cos(n * lat) ** 2 * np.sin(n * lat) ** 2 * n ** 2 * np.cos(lat)
            - 4.0 * np.sin(lon) * np.cos(m * lon) * np.cos(n * lat) ** 4 * n ** 2 * np.cos(lat)
            + 4.0 * np.sin(lon) * np.cos(m * lon) * np.cos(n * lat) ** 3 * np.sin(n * lat) * n * np.sin(lat)) / np.cos(lat)


# ----------------------------------------------------------------------------- extension winds
def solid_body_wind(lon, lat, alpha, period):
    """Williamson et al. (1992) test 1 wind with tilt ``alpha`` and rotation ``period``.

    Same formula as src/advection_ic.py:288-292 with the angle and period as parameters;
    the rotation axis is (-sin(alpha), 0, cos(alpha)).
    """
    u0 = 2.0 * PI / period
    ulon = u0 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha))
    vlat = -u0 * np.sin(lon) * np.sin(alpha)
    return ulon, vlat


def deformational_wind(lon, lat, t, k, period):
    """Nair & Lauritzen (2010) case 1: non-divergent, time-reversing deformation."""
    s = np.cos(PI * t / period)
    ulon = k * np.sin(lon / 2.0) ** 2 * np.sin(2.0 * lat) * s
    vlat = (k / 2.0) * np.sin(lon) * np.cos(lat) * s
    return ulon, vlat


def rotate_back(lon, lat, alpha, omega_t):
    """Cartesian point rotated by ``-omega_t`` about (-sin a, 0, cos a) (Rodrigues' formula)."""
    X, Y, Z = sph2cart(lon, lat)
    kx, ky, kz = -math.sin(alpha), 0.0, math.cos(alpha)
    c, s = math.cos(-omega_t), math.sin(-omega_t)
    kdotx = kx * X + ky * Y + kz * Z
    cx = ky * Z - kz * Y
    cy = kz * X - kx * Z
    cz = kx * Y - ky * X
    Xr = X * c + cx * s + kx * kdotx * (1.0 - c)
    Yr = Y * c + cy * s + ky * kdotx * (1.0 - c)
    Zr = Z * c + cz * s + kz * kdotx * (1.0 - c)
    return Xr, Yr, Zr


# ----------------------------------------------------------------------------- extension fields
def _great_circle(lon, lat, lon0, lat0):
    return np.arccos(np.clip(np.sin(lat0) * np.sin(lat) + np.cos(lat0) * np.cos(lat) * np.cos(lon - lon0), -1.0, 1.0))


def gaussians_xyz(X, Y, Z, centres, b, hmax=1.0):
    q = np.zeros(np.shape(X))
    for (lon0, lat0) in centres:
        X0, Y0, Z0 = sph2cart(lon0, lat0)
        q = q + hmax * np.exp(-b * ((X - X0) ** 2 + (Y - Y0) ** 2 + (Z - Z0) ** 2))
    return q


def cosine_bells(lon, lat, centres, radius=0.5, hmax=1.0, base=0.1, scale=0.9):
    """NL10 eq. (12): base + scale * sum_i (hmax/2)(1 + cos(pi r_i / radius)) inside the radius."""
    q = np.zeros(np.shape(lon))
    for (lon0, lat0) in centres:
        r = _great_circle(lon, lat, lon0, lat0)
        q = q + np.where(r < radius, 0.5 * hmax * (1.0 + np.cos(PI * r / radius)), 0.0)
    return base + scale * q


def slotted_cylinders(lon, lat, centres, radius=0.5, high=1.0, low=0.1):
    """NL10 eq. (13): two cylinders, the first slotted from below, the second from above."""
    q = np.full(np.shape(lon), low)
    for n, (lon0, lat0) in enumerate(centres):
        r = _great_circle(lon, lat, lon0, lat0)
        dlon = np.abs(np.angle(np.exp(1j * (lon - lon0))))
        inside = r <= radius
        outside_slot = dlon >= radius / 6.0
        if n % 2 == 0:
            rim = (dlon < radius / 6.0) & (lat - lat0 < -5.0 * radius / 12.0)
        else:
            rim = (dlon < radius / 6.0) & (lat - lat0 > 5.0 * radius / 12.0)
        q = np.where(inside & (outside_slot | rim), high, q)
    return q


def field(name, lon, lat, params):
    """Extension scalar fields evaluated at points (lon, lat)."""
    X, Y, Z = sph2cart(lon, lat)
    return field_xyz(name, X, Y, Z, params)


def field_xyz(name, X, Y, Z, params):
    lon = np.arctan2(Y, X)
    lat = np.arctan2(Z, np.hypot(X, Y))
    if name == "gaussians":
        return gaussians_xyz(X, Y, Z, params["centres"], params.get("b", 5.0), params.get("hmax", 1.0))
    if name == "cosine_bells":
        return cosine_bells(lon, lat, params["centres"], params.get("radius", 0.5))
    if name == "slotted_cylinders":
        return slotted_cylinders(lon, lat, params["centres"], params.get("radius", 0.5))
    if name == "nl10_mix":
        # linear combination of the Gaussian-hill and cosine-bell fields
        a = gaussians_xyz(X, Y, Z, params["centres"], params.get("b", 5.0), 0.95)
        c = cosine_bells(lon, lat, params["centres"], params.get("radius", 0.5))
        return 0.5 * a + 0.5 * c
    raise ValueError("unknown field %r" % name)


# ----------------------------------------------------------------------------- case dispatch
def case_wind(case, lon, lat, t):
    vf = case["vf"]
    if isinstance(vf, int):
        return velocity_adv(lon, lat, t, vf)
    if vf == "solid_body":
        return solid_body_wind(lon, lat, case["alpha"], case["period"])
    if vf == "deformational":
        return deformational_wind(lon, lat, t, case.get("k", 2.4), case.get("period", 5.0))
    raise ValueError("unknown wind %r" % vf)


def case_is_steady(case):
    vf = case["vf"]
    return vf in (1, 4, "solid_body")


def case_tracers(case):
    """List of tracer descriptors (one per tracer)."""
    if "tracers" in case:
        return case["tracers"]
    return [case["ic"]]


def case_q(case, tracer, lon, lat, t):
    """Exact solution of one tracer at time t (IC at t=0)."""
    vf = case["vf"]
    if isinstance(tracer, int):
        return qexact_ref(lon, lat, t, tracer, vf if isinstance(vf, int) else 0)
    name, params = tracer["field"], tracer
    if vf == "solid_body":
        Xr, Yr, Zr = rotate_back(lon, lat, case["alpha"], 2.0 * PI / case["period"] * t)
        return field_xyz(name, Xr, Yr, Zr, params)
    # deformational and other time-reversing flows return to the IC at t = period
    return field(name, lon, lat, params)


def from_case(case):
    """Case dict for an object with the product's ``Case`` attributes (``wind``, ``tracers``)."""
    kind = case.wind[0]
    if kind == "ref":
        if case.ntr == 1 and isinstance(case.tracers[0], int):
            return {"ic": case.tracers[0], "vf": case.wind[1]}
        return {"vf": case.wind[1], "tracers": list(case.tracers)}
    if kind == "solid_body":
        return {"vf": "solid_body", "alpha": case.wind[1], "period": case.wind[2], "tracers": list(case.tracers)}
    return {"vf": "deformational", "k": case.wind[1], "period": case.wind[2], "tracers": list(case.tracers)}
