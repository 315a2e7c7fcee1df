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
