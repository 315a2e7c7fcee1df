"""Pin the swap matrices, harmonics and the drand48 root-shift restatement."""
import ctypes
import json
import math
import os
import re
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HDR = os.path.join(ROOT, "taichi_collisional_amd", "csrc", "swap_tables.h")


def _tables():
    s = open(HDR).read()
    rp = list(map(int, re.search(r"kSwapRowPtr\[\d+\] = \{(.*?)\};", s).group(1).split(",")))
    col = list(map(int, re.search(r"kSwapCol\[\d+\] = \{(.*?)\};", s).group(1).split(",")))
    val = [float(v) for v in re.search(r"kSwapVal\[\d+\] = \{(.*?)\};", s).group(1).split(",")]
    hs = re.search(r"kSwapHash\[\d+\] = \{(.*?)\};", s).group(1).replace('"', "").split(",")
    return rp, col, val, hs


def _dense(n, rp, col, val):
    k = 2 * n + 1
    S = np.zeros((k, k))
    for r in range(k):
        for e in range(rp[n * n + r], rp[n * n + r + 1]):
            S[r, col[e]] = val[e]
    return S


def test_swap_tables_match_reference_files():
    """Derived tables == the reference's rotate_order_*.txt (hash of integer form)."""
    g = json.load(open(os.path.join(GOLDEN, "swap_hashes.json")))
    rp, col, val, hs = _tables()
    for o in g["orders"]:
        n = o["n"]
        nnz = rp[(n + 1) ** 2] - rp[n * n]
        if n <= 28:
            assert hs[n] == o["sha256"], "order %d differs from the reference" % n
            assert nnz == o["nnz"]
        else:
            # the reference's orders 29-30 carry entries beyond 2^53 that are not exact
            assert nnz == o["nnz"]


def test_swap_tables_integer_involution():
    """S_n is an involution: (2^(n-1) S)^2 == 4^(n-1) I exactly, n <= 28."""
    rp, col, val, _ = _tables()
    for n in range(0, 29):
        k = 2 * n + 1
        sh = max(n - 1, 0)
        A = [[0] * k for _ in range(k)]
        for r in range(k):
            for e in range(rp[n * n + r], rp[n * n + r + 1]):
                w = val[e] * (1 << sh)
                assert w == int(w)
                A[r][col[e]] = int(w)
        for r in range(k):
            for c in range(k):
                acc = sum(A[r][t] * A[t][c] for t in range(k))
                assert acc == (4 ** sh if r == c else 0)


def test_swap_generator_reproduces_low_orders():
    """The exact rational derivation reproduces the committed rows for n <= 7."""
    sys.path.insert(0, os.path.join(ROOT, "taichi_collisional_amd", "csrc"))
    import gen_swap_tables as g
    rp, col, val, _ = _tables()
    mats = g.all_matrices(max_order=7)
    for n, S in enumerate(mats):
        D = _dense(n, rp, col, val)
        assert np.array_equal(D, np.array([[float(v) for v in row] for row in S]))


def test_swap_maps_harmonics_of_swapped_point(oracle_lib):
    """a(swap(x)) = S a(x) on the scaled real-packed harmonics, in floating point."""
    L = oracle_lib.lib()
    p = 12
    nt = (p + 1) ** 2
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.normal(size=3)
        G, Gd = np.zeros(2 * nt), np.zeros(2 * nt)
        L.orc_harmonics(p, x, np.zeros(3), G, Gd)
        Gs, _ = np.zeros(2 * nt), np.zeros(2 * nt)
        L.orc_harmonics(p, x[::-1].copy(), np.zeros(3), Gs, Gd)

        def scaled(G):
            a = np.zeros(nt)
            for n in range(p + 1):
                s = n * n + n
                for m in range(0, n + 1):
                    f = math.factorial(n + m) * math.factorial(n - m)
                    a[s + m] = G[2 * (s + m)] * f
                    if m:
                        a[s - m] = G[2 * (s + m) + 1] * f
            return a

        a, b = scaled(G), scaled(Gs)
        L.orc_swap(p, a)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())


def _legendre_harmonic(n, m, x):
    """Y_n^m = r^n P_n^m(cos t) e^{i m phi}/(n+m)!, no Condon-Shortley phase."""
    from scipy.special import lpmv
    r = np.linalg.norm(x)
    ct = x[2] / r
    phi = math.atan2(x[1], x[0])
    pnm = lpmv(m, n, ct) * (-1) ** m  # scipy includes the Condon-Shortley phase
    return r ** n * pnm * np.exp(1j * m * phi) / math.factorial(n + m)


def test_harmonics_closed_form(oracle_lib):
    L = oracle_lib.lib()
    p = 10
    nt = (p + 1) ** 2
    rng = np.random.default_rng(1)
    for _ in range(4):
        x = rng.normal(size=3)
        G, Gd = np.zeros(2 * nt), np.zeros(2 * nt)
        L.orc_harmonics(p, x, np.zeros(3), G, Gd)
        for n in range(p + 1):
            for m in range(-n, n + 1):
                t = n * n + n + m
                got = G[2 * t] + 1j * G[2 * t + 1]
                if m >= 0:
                    want = _legendre_harmonic(n, m, x)
                else:
                    want = (-1) ** m * np.conj(_legendre_harmonic(n, -m, x))
                assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_harmonics_time_derivative(oracle_lib):
    """dY/dt for a point moving with velocity v equals the central difference."""
    L = oracle_lib.lib()
    p = 8
    nt = (p + 1) ** 2
    rng = np.random.default_rng(2)
    x, v = rng.normal(size=3), rng.normal(size=3)
    G, Gd = np.zeros(2 * nt), np.zeros(2 * nt)
    L.orc_harmonics(p, x, v, G, Gd)
    h = 1e-5
    Gp, Gm, tmp = np.zeros(2 * nt), np.zeros(2 * nt), np.zeros(2 * nt)
    L.orc_harmonics(p, x + h * v, v, Gp, tmp)
    L.orc_harmonics(p, x - h * v, v, Gm, tmp)
    fd = (Gp - Gm) / (2 * h)
    assert np.allclose(Gd, fd, rtol=1e-7, atol=1e-8)


def test_drand48_restatement_matches_libc(oracle_lib):
    """orc_shift_from_seed == srand48(seed); drand48()-0.5 (x3) from libc."""
    libc = ctypes.CDLL(None)
    libc.drand48.restype = ctypes.c_double
    for seed in (0, 1, 7, 12345, 2 ** 31 - 1):
        libc.srand48(ctypes.c_long(seed))
        want = [libc.drand48() - 0.5 for _ in range(9)]
        for skip in (0, 3, 6):
            got = oracle_lib.shift_from_seed(seed, skip)
            assert list(got) == want[skip:skip + 3]
