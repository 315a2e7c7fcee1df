"""Pinning of the oracle's master (non-jerk) variant: the master default
build's get_force_and_potential (direct over leaf pairs, P2P_simple / P2P with
the time-step estimate) against the compiled master reference's outputs
(golden fixtures), the integrator's own all-pairs time step, and brute force."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, golden

sys.path.insert(0, GOLDEN)
import make_master_golden as mmg  # noqa: E402


def _case(g, k):
    c = mmg.CASES[k]
    x, v, m, sink, src = mmg.inputs(c)
    assert mmg.sha(x, v, m, sink, src) == str(g["case%d_sha" % k]), "IC generator changed"
    return c, x, v, m, sink, src


def _rel(a, b, mask):
    a, b = a[mask], b[mask]
    if a.ndim == 2:
        return float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("k", range(len(mmg.CASES)))
def test_master_oracle_matches_reference(oracle_lib, k):
    g = golden("master_direct.npz")
    c, x, v, m, sink, src = _case(g, k)
    o = oracle_lib.Oracle(8, ncrit=c["ncrit"])
    pot, acc, ts = o.master_direct(x, v, m, sink, src, steps=c["steps"], dt_param=c["dt"])
    live = sink.astype(bool)
    assert np.all(pot[~live] == 0) and np.all(g["case%d_pot" % k][~live] == 0)
    assert _rel(pot, g["case%d_pot" % k], live) < 1e-13
    assert _rel(acc, g["case%d_acc" % k], live) < 1e-12
    if c["steps"]:
        ref = g["case%d_ts" % k]
        assert np.all(ref[live] > 0)
        assert _rel(ts, ref, live) < 1e-12
    else:
        assert np.all(ts == 0)


def test_master_pair_mass_form_is_the_integrator_step(oracle_lib):
    """With every body in one leaf, the scalar (pair-mass) form gives
    timestep = tau_min^-4 with tau_min the integrator's own all-pairs step
    (kick_cpu_with_steps, integrator.h:101-180)."""
    g = golden("master_direct.npz")
    x, v, m = mmg.ics.plummer(600, seed=15)
    assert mmg.sha(x, v, m) == str(g["kick_sha"])
    o = oracle_lib.Oracle(8, ncrit=1000)
    _, _, ts = o.master_direct(x, v, m, steps=True, pair_mass=True,
                               dt_param=float(g["kick_dt"]))
    tau = ts ** -0.25
    np.testing.assert_allclose(tau, g["kick_ts"], rtol=1e-13)


def test_master_vector_form_accumulates_mass_along_the_leaf(oracle_lib):
    """The vector form's mass term grows along the source leaf, so it differs
    from the pair-mass form; with one source per leaf (ncrit = 1 cannot hold
    coincident bodies, so use well-separated ones) both agree."""
    rng = np.random.default_rng(3)
    n = 400
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3)) * 0.3
    m = rng.uniform(0.5, 1.5, n) / n
    o = oracle_lib.Oracle(8, ncrit=64)
    _, _, tv = o.master_direct(x, v, m, steps=True, pair_mass=False)
    _, _, tp = o.master_direct(x, v, m, steps=True, pair_mass=True)
    assert np.max(np.abs(tv - tp) / tp) > 1e-2
    o1 = oracle_lib.Oracle(8, ncrit=1)
    _, _, tv1 = o1.master_direct(x, v, m, steps=True, pair_mass=False)
    _, _, tp1 = o1.master_direct(x, v, m, steps=True, pair_mass=True)
    # one body per leaf: mi = q_i + q_j in both forms; the vector form adds
    # 1e-20 to |dv|^2 only
    np.testing.assert_allclose(tv1, tp1, rtol=1e-12)


def test_master_forces_brute_force(oracle_lib):
    rng = np.random.default_rng(9)
    n = 900
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    src = (rng.random(n) < 0.7).astype(np.uint8)
    o = oracle_lib.Oracle(8, ncrit=50)
    pot, acc, _ = o.master_direct(x, v, m, None, src, steps=False)
    d = x[None, :, :] - x[:, None, :]
    r2 = np.sum(d * d, -1)
    np.fill_diagonal(r2, np.inf)
    w = m * src
    pb = np.sum(w[None, :] / np.sqrt(r2), 1)
    ab = np.sum(w[None, :, None] * d / r2[..., None] ** 1.5, 1)
    np.testing.assert_allclose(pot, pb, rtol=1e-13)
    assert np.max(np.linalg.norm(acc - ab, axis=1) / np.linalg.norm(ab, axis=1)) < 1e-12


def _fmm_case(g, k):
    c = mmg.FMM_CASES[k]
    x, v, m, sink, src, ao, idx = mmg.fmm_inputs(c)
    assert mmg.sha(x, v, m, sink, src, ao, idx) == str(g["fmm%d_sha" % k]), "IC generator changed"
    return c, x, v, m, sink, src, ao, idx


@pytest.mark.parametrize("k", [k for k, c in enumerate(mmg.FMM_CASES) if c["mode"] == 2])
def test_master_fmm_passes_match_reference(oracle_lib, k):
    """The -DFMM build after its pre-pass: symmetric error-controlled walk
    (both f1, f2 < 1; tiny cells to P2P; with steps no far field), symmetric
    M2L_rotate, P2P_simple / P2P with the time step, downward pass.  Same
    acc_old in both, so every decision is reproduced: round-off agreement."""
    g = golden("master_direct.npz")
    c, x, v, m, sink, src, ao, idx = _fmm_case(g, k)
    o = oracle_lib.Oracle(c["p"], ncrit=c["ncrit"], mac=1, force_accuracy=2e-7)
    pot, acc, ts, _ = o.master_fmm(x, v, m, sink, src, steps=c["steps"], acc_old=ao,
                                   prepass=False)
    live = sink[idx].astype(bool)
    assert np.all(pot[idx][~live] == 0)
    assert _rel(pot[idx], g["fmm%d_pot" % k], live) < 1e-12
    assert _rel(acc[idx], g["fmm%d_acc" % k], live) < 1e-11
    if c["steps"]:
        assert o.stats["n_m2l"] == 0
        assert _rel(ts[idx], g["fmm%d_ts" % k], live) < 1e-12
    else:
        assert o.stats["n_m2l"] > 0


def test_master_fmm_with_prepass_matches_reference(oracle_lib):
    """The whole -DFMM get_force_and_potential.  The pre-pass's acceptance test
    |dXg|^2 > 3 (Rg_i + Rg_j)^2 ties exactly for diagonal neighbours on the
    tree grid; the reference build's rounding of the left side (GCC contracts
    x*x + y*y + z*z into two FMAs) is reproduced, so every pre-pass decision,
    acc_old value and error-MAC decision is the reference's."""
    g = golden("master_direct.npz")
    k = [k for k, c in enumerate(mmg.FMM_CASES) if c["mode"] == 1][0]
    c, x, v, m, sink, src, ao, idx = _fmm_case(g, k)
    o = oracle_lib.Oracle(c["p"], ncrit=c["ncrit"], mac=1, force_accuracy=2e-7)
    pot, acc, ts, aco = o.master_fmm(x, v, m, sink, src, steps=c["steps"], acc_old=ao)
    e = np.abs(aco[idx] - g["fmm%d_acc_old" % k]) / g["fmm%d_acc_old" % k]
    assert e.max() < 1e-13
    live = sink[idx].astype(bool)
    ea = (np.linalg.norm(acc[idx][live] - g["fmm%d_acc" % k][live], axis=1) /
          np.linalg.norm(g["fmm%d_acc" % k][live], axis=1))
    assert ea.max() < 1e-12
    assert _rel(pot[idx], g["fmm%d_pot" % k], live) < 1e-12


def test_master_fmm_below_ncrit_is_direct(oracle_lib):
    x, v, m = mmg.ics.plummer(300, seed=41)
    o = oracle_lib.Oracle(8, ncrit=400, mac=1, force_accuracy=2e-7)
    pf, af, tf, ao = o.master_fmm(x, v, m, steps=True, acc_old=np.full(300, 3.0))
    pd, ad, td = oracle_lib.Oracle(8, ncrit=400).master_direct(x, v, m, steps=True)
    np.testing.assert_array_equal(pf, pd)
    np.testing.assert_array_equal(af, ad)
    np.testing.assert_array_equal(tf, td)
    assert np.all(ao == 3.0)


@pytest.mark.parametrize("k", range(len(mmg.EVOLVE_CASES)))
def test_master_step_loop_matches_reference(oracle_lib, k):
    """do_evolve with GRADIENT (evolve_split_omelyan + kick_gradient over a
    hierarchy of levels) for a few steps, through the default build's forces:
    the oracle's loop reproduces the compiled reference to round-off."""
    g = golden("master_direct.npz")
    c = mmg.EVOLVE_CASES[k]
    x, v, m = mmg.ics.plummer(c["n"], seed=c["seed"])
    assert mmg.sha(x, v, m) == str(g["evolve%d_sha" % k])
    o = oracle_lib.master_evolve(m, x, v, c["nsteps"], c["dt"], ncrit=1000)
    assert o["simtime"] == float(g["evolve%d_simtime" % k])
    np.testing.assert_allclose(o["timestep"], g["evolve%d_ts" % k], rtol=1e-12)
    assert np.max(np.abs(o["pos"] - g["evolve%d_pos" % k])) < 1e-13
    assert np.max(np.abs(o["vel"] - g["evolve%d_vel" % k])) < 1e-12
    assert np.max(np.abs(o["pos_e"] - g["evolve%d_pos_e" % k])) < 1e-15
    # the system moved: the comparison is not of unchanged inputs
    assert np.max(np.abs(o["pos"] - x)) > 1e-3


def test_master_step_loop_matches_reference_one_leaf_n5000(oracle_lib):
    """One step of N = 5000 (dt = 1/4096) with ncrit = 65536: every kick is
    the all-pairs branch and the time-step sum's pair mass runs over one leaf
    of all the sources (kernel_fast_lean.h:539); the oracle's loop against the
    compiled reference's step (about 20 s on 8 cores)."""
    g = golden("master_evolve_big.npz")
    c = mmg.EVOLVE_BIG
    x, v, m = mmg.ics.plummer(c["n"], seed=c["seed"])
    assert mmg.sha(x, v, m) == str(g["sha"])
    o = oracle_lib.master_evolve(m, x, v, c["nsteps"], c["dt"], ncrit=1 << 16)
    assert o["simtime"] == float(g["simtime"])
    np.testing.assert_allclose(o["timestep"], g["ts"], rtol=1e-12)
    assert np.max(np.abs(o["pos"] - g["pos"])) < 1e-13
    assert np.max(np.abs(o["vel"] - g["vel"])) < 1e-12
    assert np.max(np.abs(o["pos_e"] - g["pos_e"])) < 1e-15
    # the system moved (by about v dt ~ 2e-4)
    assert np.max(np.abs(o["pos"] - x)) > 1e-5


def test_master_step_loop_conserves_energy(oracle_lib):
    """A fourth-order scheme over the level hierarchy: energy error after a
    step small against the energy (a wrong kick weight fails this)."""
    x, v, m = mmg.ics.plummer(200, seed=71)
    def energy(pos, vel):
        d = pos[None] - pos[:, None]
        r = np.sqrt(np.sum(d * d, -1))
        np.fill_diagonal(r, np.inf)
        return 0.5 * np.sum(m * np.sum(vel * vel, 1)) - 0.5 * np.sum(m[:, None] * m[None] / r)
    e0 = energy(x, v)
    o = oracle_lib.master_evolve(m, x, v, 1, 1.0 / 64, ncrit=1000)
    e1 = energy(o["pos"], o["vel"])
    assert abs(e1 - e0) < 1e-8 * abs(e0)
