#!/usr/bin/env python3
"""Generate the x<->z swap matrices of the real-packed solid-harmonic basis.

The fast multipole M2L used by taichi_collisional rotates expansions so the
translation runs along z.  Part of every rotation is a fixed linear map S_n per
order n that exchanges the x and z axes.  It acts on the "scaled" real-packed
vector of order n:

    a[n+m] = (n+m)!(n-m)! Re Y_n^m(x),   a[n-m] = (n+m)!(n-m)! Im Y_n^m(x)  (m>=1)
    a[n]   = (n!)^2 Re Y_n^0(x)

where Y_n^m is the regular solid harmonic r^n P_n^m(cos t) e^{i m phi}/(n+m)!
without the Condon-Shortley phase (the recursions of make_Gnm,
taichi_with_jerks/kernel_fast_lean.h:237-268).  S_n is defined by

    a(swap(x)) = S_n a(x),   swap(x, y, z) = (z, y, x).

The reference ships S_n as generated C snippets
(multipole_rotation_matrix/rotate_order_{3..30}.txt, included by
taichi_with_jerks/kernel_rotate.h:25-52).  This script derives the same maps
from the definition above in exact rational arithmetic: the harmonic
recursions are rational in (x, y, z), so evaluating them at 2n+1 rational
points and solving the linear system gives S_n exactly.  Every entry is an
integer times 2^-(n-1) (n>=1), which the script asserts.

Output: ``swap_tables.h`` (next to this file) with, per order n <= 30,
the sparse rows of S_n and of D_n^-1 S_n D_n, where D_n = diag((n+|m|)!(n-|m|)!).
The second form lets the M2L kernels apply the swap to unscaled multipoles
without the scale/unscale passes of taichi_with_jerks/kernel_fast_lean.h:1175-1251.

Run:  python gen_swap_tables.py [--check]   (--check compares with the
committed header instead of writing it).
"""
import hashlib
import os
import random
import sys
from fractions import Fraction as Q

MAX_ORDER = 30
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "swap_tables.h")


def cmul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def harmonics(x, y, z, order):
    """Regular solid harmonics Y_n^m, m>=0, as exact complex rationals."""
    r2 = x * x + y * y + z * z
    Y = {(0, 0): (Q(1), Q(0))}
    for n in range(1, order + 1):
        prev = Y[(n - 1, n - 1)]
        Y[(n, n)] = cmul(prev, (x / (2 * n), y / (2 * n)))
        Y[(n, n - 1)] = (z * prev[0], z * prev[1])
        for m in range(0, n - 1):
            a = Y[(n - 1, m)]
            b = Y[(n - 2, m)]
            d = n * n - m * m
            Y[(n, m)] = (((2 * n - 1) * z * a[0] - r2 * b[0]) / d,
                         ((2 * n - 1) * z * a[1] - r2 * b[1]) / d)
    return Y


def fact(k):
    f = 1
    for i in range(2, k + 1):
        f *= i
    return f


def dscale(n, m):
    m = abs(m)
    return fact(n + m) * fact(n - m)


def scaled_vector(Y, n):
    """Real-packed scaled vector of order n, local index i = n+m (m=-n..n)."""
    v = [Q(0)] * (2 * n + 1)
    for m in range(0, n + 1):
        re, im = Y[(n, m)]
        s = dscale(n, m)
        v[n + m] = re * s
        if m > 0:
            v[n - m] = im * s
    return v


def solve(A, B):
    """Return X with X A = B for square A (rows = components, cols = samples)."""
    k = len(A)
    # X A = B  <=>  A^T X^T = B^T ; Gauss-Jordan on [A^T | B^T]
    At = [[A[r][c] for r in range(k)] for c in range(k)]
    Bt = [[B[r][c] for r in range(k)] for c in range(k)]
    for col in range(k):
        piv = next(r for r in range(col, k) if At[r][col] != 0)
        At[col], At[piv] = At[piv], At[col]
        Bt[col], Bt[piv] = Bt[piv], Bt[col]
        inv = 1 / At[col][col]
        At[col] = [v * inv for v in At[col]]
        Bt[col] = [v * inv for v in Bt[col]]
        for r in range(k):
            if r != col and At[r][col] != 0:
                f = At[r][col]
                At[r] = [a - f * b for a, b in zip(At[r], At[col])]
                Bt[r] = [a - f * b for a, b in zip(Bt[r], Bt[col])]
    # Bt now holds X^T
    return [[Bt[c][r] for c in range(k)] for r in range(k)]


def swap_matrix(n, rng):
    """Exact S_n as a (2n+1)x(2n+1) list of Fractions."""
    k = 2 * n + 1
    while True:
        pts = [tuple(Q(rng.randint(-9, 9), rng.randint(1, 7)) for _ in range(3))
               for _ in range(k)]
        A = [[None] * k for _ in range(k)]
        B = [[None] * k for _ in range(k)]
        for c, (x, y, z) in enumerate(pts):
            a = scaled_vector(harmonics(x, y, z, n), n)
            b = scaled_vector(harmonics(z, y, x, n), n)
            for r in range(k):
                A[r][c] = a[r]
                B[r][c] = b[r]
        try:
            return solve(A, B)
        except StopIteration:  # singular sample set, draw again
            continue


def all_matrices(max_order=MAX_ORDER, seed=12345):
    rng = random.Random(seed)
    return [swap_matrix(n, rng) for n in range(max_order + 1)]


def integer_form(S, n):
    """Entries as integers times 2^-(n-1); returns {(row, col): int}."""
    shift = max(n - 1, 0)
    out = {}
    for r, row in enumerate(S):
        for c, v in enumerate(row):
            if v != 0:
                w = v * (1 << shift)
                assert w.denominator == 1, (n, r, c, v)
                out[(r, c)] = int(w)
    return out


def canonical_hash(intform, n):
    """SHA-256 over 'row col value' lines in row-major order (global indices)."""
    base = n * n
    lines = ["%d %d %d" % (base + r, base + c, v) for (r, c), v in sorted(intform.items())]
    return hashlib.sha256("\n".join(lines).encode()).hexdigest()


def emit(mats):
    rows_ptr = [0]
    cols, vals, svals = [], [], []
    order_row0 = []
    for n, S in enumerate(mats):
        order_row0.append(len(rows_ptr) - 1)
        k = 2 * n + 1
        for r in range(k):
            for c in range(k):
                v = S[r][c]
                if v != 0:
                    cols.append(c)
                    vals.append(v)
                    mr, mc = r - n, c - n
                    svals.append(v * dscale(n, mc) / dscale(n, mr))
            rows_ptr.append(len(cols))
    order_row0.append(len(rows_ptr) - 1)
    hashes = [canonical_hash(integer_form(S, n), n) for n, S in enumerate(mats)]
    out = []
    out.append("// Generated by gen_swap_tables.py -- do not edit.")
    out.append("// x<->z swap maps S_n of the scaled real-packed solid-harmonic basis,")
    out.append("// derived in exact rational arithmetic (see the generator's docstring).")
    out.append("// Per order n: local rows/cols 0..2n (index n+m).  kSwapVal holds S_n,")
    out.append("// kSwapValScaled holds D_n^-1 S_n D_n with D_n = diag((n+|m|)!(n-|m|)!).")
    out.append("#pragma once")
    out.append("namespace tcfmm {")
    out.append("constexpr int kSwapMaxOrder = %d;" % (len(mats) - 1))
    out.append("constexpr int kSwapNnz = %d;" % len(cols))
    out.append("// first global row of order n (= n*n); rows of order n are n*n .. n*n+2n")
    out.append("constexpr int kSwapRowPtr[%d] = {%s};" % (len(rows_ptr), ",".join(map(str, rows_ptr))))
    out.append("constexpr unsigned char kSwapCol[%d] = {%s};" % (len(cols), ",".join(map(str, cols))))
    out.append("constexpr double kSwapVal[%d] = {%s};" % (
        len(vals), ",".join(repr(float(v)) for v in vals)))
    out.append("constexpr double kSwapValScaled[%d] = {%s};" % (
        len(svals), ",".join(repr(float(v)) for v in svals)))
    out.append("// SHA-256 of the integer form of S_n ('row col int' lines, entries x 2^(n-1))")
    out.append("constexpr const char* kSwapHash[%d] = {%s};" % (
        len(hashes), ",".join('"%s"' % h for h in hashes)))
    out.append("}  // namespace tcfmm")
    return "\n".join(out) + "\n"


def main():
    mats = all_matrices()
    for n, S in enumerate(mats):
        k = 2 * n + 1
        # involution check: S S = I
        for r in range(k):
            for c in range(k):
                acc = sum(S[r][t] * S[t][c] for t in range(k))
                assert acc == (1 if r == c else 0), ("not an involution", n, r, c)
        integer_form(S, n)
    text = emit(mats)
    if "--check" in sys.argv:
        with open(OUT) as f:
            if f.read() != text:
                print("swap_tables.h is stale")
                sys.exit(1)
        print("swap_tables.h up to date")
        return
    with open(OUT, "w") as f:
        f.write(text)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
