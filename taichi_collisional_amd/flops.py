"""Algorithmic flop counts of the FMM force+jerk evaluation.

The unit of work is the reference algorithm's arithmetic, counted from its
loop bounds (one add, multiply, divide or square root = 1 flop; a complex
multiply-add = 8), so throughput figures do not depend on how this build
arranges the work:

    FLOP = 44 N_pp + c_M2L(p) N_M2L + c_P2M(p) N + c_L2P(p) N
           + c_M2M(p) E + c_L2L(p) E

N_pp: ordered body pairs of the P2P leaf pairs (self pairs included),
N_M2L: accepted (target, source) cell pairs, E: parent-child edges.
Each constant covers both the moments and their time derivatives (M and
Mdot, L and Ldot), as the reference computes them together.

Sources of the loop bounds (taichi_with_jerks/kernel_fast_lean.h):
  P2P        :566-607  (Vec8d branch, per pair)
  M2L_rotate :1130-1371
  P2M        :701-824 (harmonics :237-304)
  M2M        :826-999
  L2L        :1001-1058
  L2P        :1060-1128
"""
import json
import os

_SWAP_NNZ = None


def swap_nnz():
    """Non-zeros of the x<->z swap per order (from the generated tables)."""
    global _SWAP_NNZ
    if _SWAP_NNZ is None:
        import re
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "swap_tables.h")
        s = open(path).read()
        rp = list(map(int, re.search(r"kSwapRowPtr\[\d+\] = \{(.*?)\};", s).group(1).split(",")))
        _SWAP_NNZ = [rp[(n + 1) ** 2] - rp[n * n] for n in range(0, 31)]
    return _SWAP_NNZ


def c_p2p():
    # dx(3) dv(3) r2(5) vdotr(6) sqrt(1) 1/r(1) r^-2(1) factor(1) pot(1)
    # factor*=r^-2 (1) a+=(6) j+=dv f (6) factor*=-r^-2 vdotr (2) j+=dx f (6)
    return 44


def c_m2l(p):
    nt = (p + 1) ** 2
    tri = p * (p + 1) // 2          # (n, m>=1) pairs
    f = 0
    f += 20                         # dX, r_xy, r, 1/r, first powers
    f += 12 * (p - 1)               # power recurrences (two complex products per m)
    f += 2 * nt                     # scale M, Mdot by (n+m)!(n-m)!
    f += 4 * 2 * 6 * tri            # four z-rotations of two arrays, 6 flops per (n,m)
    f += 4 * 2 * 2 * sum(swap_nnz()[:p + 1])   # four swaps of two arrays
    f += 2 * nt                     # unscale
    for n in range(p + 1):          # translation along z (:1253-1295)
        f += (p - n + 1) * 6
        for m in range(1, min(n, p - n) + 1):
            f += 1 + (p - n - m + 1) * 11 + 1
    f += 2 * nt                     # accumulate L, Ldot
    return f


def c_harmonics(p):
    # make_Gnm: per n: diag complex product (2 div + 6), m < n-1 terms (~12 each), m = n-1 (2)
    f = 1 + 5
    for n in range(1, p + 1):
        f += 8 + 12 * max(n - 1, 0) + 2
    return f


def c_harmonics_dot(p):
    f = 6
    for n in range(p + 1):
        f += 8 * (max(n - 1, 0) + n + (n + 1))
    return f


def c_p2m(p):
    nt = (p + 1) ** 2
    return 3 + c_harmonics(p) + c_harmonics_dot(p) + 2 * 2 * nt + 2


def c_l2p(p):
    terms = sum((2 * k + 1) for k in range(p + 1)) + 2 * sum((2 * k + 1) for k in range(p))
    return 6 + c_harmonics(p) + c_harmonics_dot(p) + 8 * terms + 8 * 2 * 2 * sum(
        (2 * k + 1) for k in range(p)) + 7


def _m2m_terms(p):
    t = 0
    for n in range(p + 1):
        for m in range(n + 1):
            for k in range(n + 1):
                lo, hi = max(-k, m - n + k), min(k, m + n - k)
                t += max(hi - lo + 1, 0)
    return t


def _l2l_terms(p):
    return sum((n + 1) * sum(2 * k + 1 for k in range(p - n + 1)) for n in range(p + 1))


def c_m2m(p):
    return 3 + c_harmonics(p) + 2 * 8 * _m2m_terms(p) + 2 * (p + 1) ** 2


def c_l2l(p):
    return 3 + c_harmonics(p) + 2 * 8 * _l2l_terms(p) + 2 * (p + 1) ** 2


def constants(p):
    return {"p": p, "c_P2P": c_p2p(), "c_M2L": c_m2l(p), "c_P2M": c_p2m(p), "c_L2P": c_l2p(p),
            "c_M2M": c_m2m(p), "c_L2L": c_l2l(p)}


def eval_flops(p, n, n_pp, n_m2l, n_edges):
    c = constants(p)
    parts = {"P2P": c["c_P2P"] * n_pp, "M2L": c["c_M2L"] * n_m2l, "P2M": c["c_P2M"] * n,
             "L2P": c["c_L2P"] * n, "M2M": c["c_M2M"] * n_edges, "L2L": c["c_L2L"] * n_edges}
    parts["total"] = sum(parts.values())
    return parts


if __name__ == "__main__":
    print(json.dumps([constants(p) for p in range(2, 25)], indent=1))
