"""Deterministic initial conditions for the force-evaluation workloads.

Star-cluster models in Henon units (G = M = 1) as used by taichi_collisional's
drivers (taichi_with_jerks/taichi.cxx:88-130 read 7 columns m x y z vx vy vz).
All draws come from ``numpy.random.default_rng(seed)`` (PCG64) in a fixed
order, so a (model, N, seed) triple names one array bit for bit.

Plummer (a = 3*pi/16):
  * mass fraction u ~ U(0,1), redrawn while u > 0.999 (radial truncation at
    99.9% of the mass, r_max ~ 38.7 a);  r = a / sqrt(u^(-2/3) - 1)
  * isotropic direction: cos(theta) ~ U(-1,1), phi ~ U(0, 2 pi)
  * speed by the Aarseth-Henon-Wielen (1974) rejection on
    g(q) = q^2 (1-q^2)^(7/2) (max 0.1), v = q * sqrt(2) * (r^2 + a^2)^(-1/4),
    isotropic velocity direction drawn like the position direction
  * equal masses 1/N; no centre-of-mass correction.
Draw order: all u, then position directions, then (q, acceptance) rounds,
then velocity directions.

Dehnen gamma model (a = 1): r = a / (u^(-1/(3-gamma)) - 1) with the same
truncation and directions; velocities from an isotropic Gaussian with the
local Jeans dispersion (computed numerically) -- a stated simplification.

Kroupa IMF with binaries (kroupa_binaries): the Plummer model above for the
systems' centres of mass; stellar masses from the two-segment Kroupa IMF
dN/dm ~ m^-1.3 on [0.08, 0.5] and m^-2.3 on [0.5, 100] (continuous at 0.5),
drawn by inverting each segment's cumulative after choosing the segment by
its number fraction; all masses scaled so the total is 1.  "10% of the stars
in binaries" is taken as 10% of the N bodies being binary members, i.e.
N/20 pairs.  Each pair: semi-major axis log-uniform in [1e-5, 1e-3], thermal
eccentricity f(e) = 2e (e = sqrt(u)), uniform mean anomaly (Kepler's equation
solved by Newton iteration), isotropic orbit orientation; the pair sits at
its system's position and velocity.  Draw order: the Plummer draws for the
N - N/20 systems, then the N masses, then per pair a, e, mean anomaly and
the three orientation angles.
"""
import math

import numpy as np

HENON_A = 3.0 * math.pi / 16.0


def _directions(rng, n):
    c = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    return np.stack([s * np.cos(ph), s * np.sin(ph), c], axis=1)


def _mass_fractions(rng, n, umax=0.999):
    u = rng.uniform(0.0, 1.0, n)
    bad = (u > umax) | (u <= 0.0)
    while bad.any():
        u[bad] = rng.uniform(0.0, 1.0, int(bad.sum()))
        bad = (u > umax) | (u <= 0.0)
    return u


def plummer(n, seed=1, a=HENON_A):
    """Plummer sphere: returns (x[n,3], v[n,3], m[n]) as float64 arrays."""
    rng = np.random.default_rng(seed)
    u = _mass_fractions(rng, n)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    x = r[:, None] * _directions(rng, n)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0.0, 1.0, todo.size)
        yy = rng.uniform(0.0, 0.1, todo.size)
        ok = yy < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = math.sqrt(2.0) * (r * r + a * a) ** (-0.25)
    v = (q * vesc)[:, None] * _directions(rng, n)
    m = np.full(n, 1.0 / n)
    return np.ascontiguousarray(x), np.ascontiguousarray(v), m


def dehnen(n, gamma=1.5, seed=1, a=1.0):
    """Dehnen gamma model with Jeans-dispersion Gaussian velocities."""
    rng = np.random.default_rng(seed)
    u = _mass_fractions(rng, n)
    r = a / (u ** (-1.0 / (3.0 - gamma)) - 1.0)
    x = r[:, None] * _directions(rng, n)
    # isotropic Jeans dispersion sigma^2(r) = (1/rho) int_r^inf rho GM(<r')/r'^2 dr'
    g = gamma
    rr = np.geomspace(1e-6, 1e6, 4000)
    rho = (3 - g) / (4 * math.pi) * a / (rr ** g * (rr + a) ** (4 - g))
    mass = (rr / (rr + a)) ** (3 - g)
    integrand = rho * mass / rr ** 2
    # cumulative integral from the outside in (trapezoid on the log grid)
    seg = 0.5 * (integrand[1:] + integrand[:-1]) * np.diff(rr)
    tail = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    sig2 = np.interp(r, rr, tail / rho)
    v = rng.standard_normal((n, 3)) * np.sqrt(np.maximum(sig2, 0.0))[:, None]
    m = np.full(n, 1.0 / n)
    return np.ascontiguousarray(x), np.ascontiguousarray(v), m


def _kroupa_masses(rng, n):
    # segment number fractions from the integrals of m^-1.3 and m^-2.3 (scaled
    # to be continuous at 0.5)
    a1, a2, m0, mb, m1 = 1.3, 2.3, 0.08, 0.5, 100.0
    n1 = (mb ** (1 - a1) - m0 ** (1 - a1)) / (1 - a1)
    n2 = mb ** (a2 - a1) * (m1 ** (1 - a2) - mb ** (1 - a2)) / (1 - a2)
    low = rng.uniform(0.0, 1.0, n) < n1 / (n1 + n2)
    u = rng.uniform(0.0, 1.0, n)
    lo = (m0 ** (1 - a1) + u * (mb ** (1 - a1) - m0 ** (1 - a1))) ** (1 / (1 - a1))
    hi = (mb ** (1 - a2) + u * (m1 ** (1 - a2) - mb ** (1 - a2))) ** (1 / (1 - a2))
    return np.where(low, lo, hi)


def kroupa_binaries(n, seed=1, a=HENON_A, binary_fraction=0.1):
    """Plummer systems, Kroupa IMF masses, binary_fraction of the n bodies in
    binaries (see the module docstring).  Returns (x, v, m), total mass 1."""
    npair = int(round(n * binary_fraction / 2))
    nsys = n - npair
    xs, vs, _ = plummer(nsys, seed=seed, a=a)
    rng = np.random.default_rng([seed, 2])
    m = _kroupa_masses(rng, n)
    m /= m.sum()
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    # systems 0..nsys-npair-1 are single stars, the last npair are binaries
    ns = nsys - npair
    x[:ns], v[:ns] = xs[:ns], vs[:ns]
    sma = 10.0 ** rng.uniform(-5.0, -3.0, npair)
    ecc = np.sqrt(rng.uniform(0.0, 1.0, npair))
    mean = rng.uniform(0.0, 2.0 * math.pi, npair)
    E = mean.copy()
    for _ in range(60):
        E -= (E - ecc * np.sin(E) - mean) / (1.0 - ecc * np.cos(E))
    m1, m2 = m[ns:ns + 2 * npair:2], m[ns + 1:ns + 2 * npair:2]
    mt = m1 + m2
    nmot = np.sqrt(mt / sma ** 3)
    den = 1.0 - ecc * np.cos(E)
    rel_x = np.stack([sma * (np.cos(E) - ecc), sma * np.sqrt(1 - ecc ** 2) * np.sin(E),
                      np.zeros(npair)], axis=1)
    rel_v = np.stack([-sma * nmot * np.sin(E) / den,
                      sma * nmot * np.sqrt(1 - ecc ** 2) * np.cos(E) / den, np.zeros(npair)], axis=1)
    # isotropic orientation: z-x-z Euler angles with cos(inclination) uniform
    w, om = rng.uniform(0.0, 2.0 * math.pi, (2, npair))
    ci = rng.uniform(-1.0, 1.0, npair)
    si = np.sqrt(1 - ci ** 2)
    cw, sw, co, so = np.cos(w), np.sin(w), np.cos(om), np.sin(om)
    R = np.empty((npair, 3, 3))
    R[:, 0, 0] = co * cw - so * sw * ci
    R[:, 0, 1] = -co * sw - so * cw * ci
    R[:, 0, 2] = so * si
    R[:, 1, 0] = so * cw + co * sw * ci
    R[:, 1, 1] = -so * sw + co * cw * ci
    R[:, 1, 2] = -co * si
    R[:, 2, 0] = sw * si
    R[:, 2, 1] = cw * si
    R[:, 2, 2] = ci
    rx = np.einsum("kij,kj->ki", R, rel_x)
    rv = np.einsum("kij,kj->ki", R, rel_v)
    f1, f2 = (m2 / mt)[:, None], (m1 / mt)[:, None]
    cx, cv = xs[ns:], vs[ns:]
    x[ns:ns + 2 * npair:2] = cx - f1 * rx
    x[ns + 1:ns + 2 * npair:2] = cx + f2 * rx
    v[ns:ns + 2 * npair:2] = cv - f1 * rv
    v[ns + 1:ns + 2 * npair:2] = cv + f2 * rv
    return np.ascontiguousarray(x), np.ascontiguousarray(v), np.ascontiguousarray(m)


def write_ascii(path, x, v, m):
    """Reference IC format: 7 columns m x y z vx vy vz."""
    with open(path, "w") as f:
        for i in range(m.size):
            f.write("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                m[i], x[i, 0], x[i, 1], x[i, 2], v[i, 0], v[i, 1], v[i, 2]))


def read_ascii(path):
    d = np.loadtxt(path, ndmin=2)
    return (np.ascontiguousarray(d[:, 1:4]), np.ascontiguousarray(d[:, 4:7]),
            np.ascontiguousarray(d[:, 0]))
