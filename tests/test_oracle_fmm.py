"""End-to-end pinning of the oracle against the compiled reference's outputs
(golden fixtures) and the FP64 direct sum, plus tree invariants."""
import hashlib

import numpy as np
import pytest

from conftest import golden, relerr_scalar, relerr_vec
from taichi_collisional_amd import ics


def _sha(*arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _inputs(g):
    if "x" in g:
        return g["x"], g["v"], g["m"]
    x, v, m = ics.plummer(int(g["n"]), seed=int(g["ic_seed"]))
    assert _sha(x, v, m) == str(g["input_sha"]), "IC generator changed"
    return x, v, m


@pytest.mark.parametrize("name", ["e2e_dat10.npz", "e2e_c1.npz", "e2e_c2s.npz", "e2e_p5.npz",
                                  "e2e_p16.npz", "e2e_p20.npz", "e2e_p24.npz"])
def test_oracle_matches_reference_theta_mac(oracle_lib, name):
    g = golden(name)
    x, v, m = _inputs(g)
    p, th, seed = int(g["p"]), float(g["theta"]), int(g["seed"])
    u = oracle_lib.shift_from_seed(seed)
    assert np.array_equal(u, g["shift"])
    o = oracle_lib.Oracle(p, ncrit=int(g["ncrit"]), theta=th, shift_frac=u)
    pot, acc, jerk, _ = o.eval(x, v, m)
    idx = g["idx"]
    assert o.stats["n_m2l"] == g["counts"][0]
    assert o.stats["n_p2p_cells"] == g["counts"][1]
    assert o.stats["n_p2p_pairs"] == g["counts"][2]
    assert relerr_scalar(pot[idx], g["pot"])[1] < 1e-13
    assert relerr_vec(acc[idx], g["acc"])[1] < 1e-12
    assert relerr_vec(jerk[idx], g["jerk"])[1] < 1e-12


def test_oracle_reference_default_path(oracle_lib):
    """The reference default (p=2 pre-pass + error-controlled MAC)."""
    g = golden("e2e_default.npz")
    x, v, m = _inputs(g)
    u = oracle_lib.shift_from_seed(int(g["seed"]))
    o = oracle_lib.Oracle(int(g["p"]), ncrit=64, theta=0.5, mac=1, force_accuracy=2e-8,
                          shift_frac=u, direct_below=8000)
    pot, acc, jerk, acco = o.eval(x, v, m)
    idx = g["idx"]
    # acc_old comes from the p=2 pre-pass, whose P2P runs in FP32: the walk's
    # tie decisions on the grid-aligned cells are the reference's (its FMA
    # contraction is reproduced), so only the FP32 summation order differs
    e_ao = relerr_scalar(acco[idx], g["acc_old"])
    assert e_ao[0] < 1e-7 and e_ao[1] < 2e-6
    # every error-MAC decision is the reference's: all bodies at round-off
    ea = np.sqrt(np.sum((acc[idx] - g["acc"]) ** 2, 1) / np.sum(g["acc"] ** 2, 1))
    ej = np.sqrt(np.sum((jerk[idx] - g["jerk"]) ** 2, 1) / np.sum(g["jerk"] ** 2, 1))
    assert ea.max() < 1e-12 and ej.max() < 1e-11
    assert relerr_scalar(pot[idx], g["pot"])[1] < 1e-12


def test_oracle_error_vs_direct(oracle_lib):
    """FMM error against the FP64 direct sum is the truncation error (theta^(p+1) scale)."""
    x, v, m = ics.plummer(20000, seed=4)
    rng = np.random.default_rng(0)
    tg = np.sort(rng.choice(20000, 1024, replace=False))
    dp, da, dj = oracle_lib.direct(x, v, m, tg)
    errs = {}
    for p in (4, 8, 12):
        o = oracle_lib.Oracle(p, ncrit=64, theta=0.5, shift_frac=(0.1, -0.2, 0.3))
        pot, acc, jerk, _ = o.eval(x, v, m)
        errs[p] = relerr_vec(acc[tg], da)[0]
    assert errs[4] > errs[8] > errs[12]
    assert errs[12] < 1e-5 and errs[4] < 1e-2


def test_tree_invariants(oracle_lib):
    x, v, m = ics.plummer(30000, seed=8)
    ncrit = 64
    o = oracle_lib.Oracle(8, ncrit=ncrit, theta=0.5, shift_frac=(0.2, 0.1, -0.3))
    o.eval(x, v, m)
    c = o.cells()
    nc = c["nchild"].size
    leaves = c["nchild"] == 0
    assert np.all(c["nbody"][leaves] <= ncrit)
    assert np.all(c["nbody"][~leaves] > ncrit)
    # children partition the parent's body range, contiguous and in order
    for i in np.nonzero(~leaves)[0]:
        ch = np.arange(c["child"][i], c["child"][i] + c["nchild"][i])
        assert c["body"][ch[0]] == c["body"][i]
        assert np.all(c["body"][ch[1:]] == c["body"][ch[:-1]] + c["nbody"][ch[:-1]])
        assert c["body"][ch[-1]] + c["nbody"][ch[-1]] == c["body"][i] + c["nbody"][i]
        assert np.all(c["level"][ch] == c["level"][i] + 1)
    # leaves cover every body once
    cover = np.zeros(m.size, int)
    for i in np.nonzero(leaves)[0]:
        cover[c["body"][i]:c["body"][i] + c["nbody"][i]] += 1
    assert np.all(cover == 1)
    assert sorted(c["order"]) == list(range(m.size))
    # bounding balls enclose their bodies
    xs = x[c["order"]]
    for i in np.nonzero(leaves)[0][:500]:
        pts = xs[c["body"][i]:c["body"][i] + c["nbody"][i]]
        d = np.linalg.norm(pts - c["X"][i], axis=1)
        assert d.max() <= c["R"][i]
    assert nc == o.stats["ncells"]


def test_minball_is_minimal(oracle_lib):
    """The ball encloses all points and no smaller ball around a perturbed centre does."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 4, 5, 17, 64, 300):
        pts = rng.normal(size=(n, 3)) * rng.uniform(0.1, 2)
        c, r2 = np.zeros(3), np.zeros(1)
        L.orc_minball(n, pts.reshape(-1).copy(), c, r2)
        d2 = ((pts - c) ** 2).sum(1)
        assert d2.max() <= r2[0] * (1 + 1e-12) + 1e-300
        # optimality: the farthest-point distance grows in every direction
        for _ in range(20):
            dirn = rng.normal(size=3)
            dirn /= np.linalg.norm(dirn)
            c2 = c + dirn * 1e-6 * max(np.sqrt(r2[0]), 1e-3)
            assert ((pts - c2) ** 2).sum(1).max() >= r2[0] * (1 - 1e-10)


def test_kroupa_binaries_generator():
    """C3 initial conditions: deterministic, total mass 1, masses inside the
    IMF range (up to the normalisation), every pair bound with semi-major
    axis in [1e-5, 1e-3] and 10% of the bodies in pairs."""
    x, v, m = ics.kroupa_binaries(20000, seed=3)
    x2, v2, m2 = ics.kroupa_binaries(20000, seed=3)
    assert _sha(x, v, m) == _sha(x2, v2, m2)
    assert abs(m.sum() - 1.0) < 1e-12
    assert m.max() / m.min() <= 100.0 / 0.08 * (1 + 1e-12)
    npair = 1000
    ns = 20000 - 2 * npair
    a, b = slice(ns, None, 2), slice(ns + 1, None, 2)
    mt = m[a] + m[b]
    r = np.linalg.norm(x[b] - x[a], axis=1)
    vr = np.linalg.norm(v[b] - v[a], axis=1)
    energy = 0.5 * vr ** 2 - mt / r
    assert np.all(energy < 0)
    sma = -mt / (2 * energy)
    assert sma.min() > 1e-5 * (1 - 1e-9) and sma.max() < 1e-3 * (1 + 1e-9)
    # the pair's centre of mass is the system's Plummer position
    com = (m[a, None] * x[a] + m[b, None] * x[b]) / mt[:, None]
    xs, _, _ = ics.plummer(20000 - npair, seed=3)
    np.testing.assert_allclose(com, xs[ns:], atol=1e-12)
