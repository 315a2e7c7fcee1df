"""GPU parity of the master (non-jerk) variant through the C ABI: the master
default build's get_force_and_potential (direct over leaf pairs, with and
without the time-step estimate) and the build with -DFMM (master pre-pass,
symmetric walk, symmetric M2L) against the oracle on the same inputs and
against the compiled master reference's outputs (golden fixtures).

Tolerances (FP64): forces and potentials are the same sums in another order
(max relative error per body); the time-step sum takes a minimum per source
leaf and the fourth power of it, so its relative error is about four times
that of one estimate."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, golden

sys.path.insert(0, GOLDEN)
import make_master_golden as mmg  # noqa: E402

pytestmark = pytest.mark.gpu

POT_TOL = 1e-12
ACC_TOL = 1e-11
TS_TOL = 1e-11


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _gpu(torch, x, v, m, sink=None, src=None, ncrit=64, steps=True, dt=0.025, mode=0):
    import taichi_collisional_amd as tca
    f = tca.Fmm(p=8, ncrit=ncrit, direct_below=0)
    dx, dv, dm = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m))
    fl = None
    if sink is not None or src is not None:
        n = m.shape[0]
        s1 = np.ones(n, np.uint8) if sink is None else sink
        s2 = np.ones(n, np.uint8) if src is None else src
        fl = torch.from_numpy((s1 | (s2 << 1)).astype(np.uint8)).cuda()
    pot, acc, ts = f.master_eval(dx, dv, dm, flags=fl, get_steps=steps, dt_param=dt, mode=mode)
    torch.cuda.synchronize()
    return (pot.cpu().numpy(), acc.cpu().numpy(), None if ts is None else ts.cpu().numpy(),
            f.stats())


def _rel(a, b, mask):
    a, b = a[mask], b[mask]
    if a.ndim == 2:
        return float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))
    return float(np.max(np.abs(a - b) / np.abs(b)))


def _check(got, want, live, steps, what):
    gp, ga, gt = got
    wp, wa, wt = want
    assert np.all(gp[~live] == 0) and np.all(ga[~live] == 0)
    ep, ea = _rel(gp, wp, live), _rel(ga, wa, live)
    et = _rel(gt, wt, live) if steps else 0.0
    print("%s pot %.2e acc %.2e ts %.2e" % (what, ep, ea, et))
    assert ep < POT_TOL and ea < ACC_TOL and et < TS_TOL, (what, ep, ea, et)
    if steps:
        assert np.all(gt[~live] == 0)


@pytest.mark.parametrize("k", range(len(mmg.CASES)))
def test_gpu_master_matches_oracle_and_reference(torch_cuda, oracle_lib, k):
    g = golden("master_direct.npz")
    c = mmg.CASES[k]
    x, v, m, sink, src = mmg.inputs(c)
    assert mmg.sha(x, v, m, sink, src) == str(g["case%d_sha" % k])
    gp, ga, gt, st = _gpu(torch_cuda, x, v, m, sink, src, ncrit=c["ncrit"], steps=c["steps"],
                          dt=c["dt"])
    o = oracle_lib.Oracle(8, ncrit=c["ncrit"])
    op, oa, ot = o.master_direct(x, v, m, sink, src, steps=c["steps"], dt_param=c["dt"])
    live = sink.astype(bool)
    _check((gp, ga, gt), (op, oa, ot), live, c["steps"], "case %d vs oracle" % k)
    _check((gp, ga, gt), (g["case%d_pot" % k], g["case%d_acc" % k], g["case%d_ts" % k]), live,
           c["steps"], "case %d vs reference" % k)


def test_gpu_master_pair_mass_form(torch_cuda, oracle_lib):
    """Scalar form: against the oracle with several leaves, and with one leaf
    against the integrator's own all-pairs step (kick_cpu_with_steps)."""
    import taichi_collisional_amd as tca
    g = golden("master_direct.npz")
    x, v, m = mmg.ics.plummer(600, seed=15)
    assert mmg.sha(x, v, m) == str(g["kick_sha"])
    _, _, gt, _ = _gpu(torch_cuda, x, v, m, ncrit=1000, steps=True, dt=float(g["kick_dt"]),
                       mode=tca.TC_MASTER_PAIR_MASS)
    np.testing.assert_allclose(gt ** -0.25, g["kick_ts"], rtol=1e-12)
    x, v, m = mmg.ics.plummer(3000, seed=21)
    gp, ga, gt, _ = _gpu(torch_cuda, x, v, m, ncrit=40, steps=True, mode=tca.TC_MASTER_PAIR_MASS)
    o = oracle_lib.Oracle(8, ncrit=40)
    op, oa, ot = o.master_direct(x, v, m, steps=True, pair_mass=True)
    _check((gp, ga, gt), (op, oa, ot), np.ones(3000, bool), True, "pair mass")


def test_gpu_master_body_records(torch_cuda):
    """The host Body-record drop-in equals the device-array form."""
    import taichi_collisional_amd as tca
    x, v, m = mmg.ics.plummer(4000, seed=22)
    rng = np.random.default_rng(22)
    sink = (rng.random(4000) < 0.9).astype(np.uint8)
    src = (rng.random(4000) < 0.9).astype(np.uint8)
    gp, ga, gt, _ = _gpu(torch_cuda, x, v, m, sink, src, ncrit=100, steps=True)
    b = np.zeros(4000, tca.master_body_dtype)
    b["X"], b["V"], b["q"], b["issink"], b["issource"] = x, v, m, sink, src
    b["index"] = np.arange(4000)
    b["acc_old"] = 7.0
    f = tca.Fmm(p=8, ncrit=100, direct_below=0)
    f.master_get_force_and_potential(b, get_steps=True, dt_param=0.025)
    np.testing.assert_array_equal(b["p"], gp)
    np.testing.assert_array_equal(b["F"], ga)
    np.testing.assert_array_equal(b["timestep"], gt)
    assert np.all(b["acc_old"] == 7.0) and np.array_equal(b["index"], np.arange(4000))
    np.testing.assert_array_equal(b["X"], x)


def test_gpu_master_full_size_forces(torch_cuda):
    """N = 2^16: forces equal the FP64 all-pairs sum of tc_direct_soa."""
    import taichi_collisional_amd as tca
    n = 1 << 16
    x, v, m = mmg.ics.plummer(n, seed=23)
    gp, ga, gt, st = _gpu(torch_cuda, x, v, m, ncrit=300, steps=True)
    assert np.all(np.isfinite(gt)) and np.all(gt > 0)
    f = tca.Fmm(p=8, direct_below=0)
    t = torch_cuda
    dx, dv, dm = (t.from_numpy(a).cuda() for a in (x, v, m))
    tg = t.arange(0, n, 7, dtype=t.int32)
    dp, da, _ = f.direct(dx, dv, dm, tg)
    idx = tg.numpy()
    assert _rel(gp[idx], dp.cpu().numpy(), np.ones(idx.size, bool)) < POT_TOL
    assert _rel(ga[idx], da.cpu().numpy(), np.ones(idx.size, bool)) < ACC_TOL
    assert st["nleaves"] > 200


def test_gpu_master_edges_and_errors(torch_cuda, oracle_lib):
    import taichi_collisional_amd as tca
    # one and two bodies
    x = np.array([[0.1, 0.2, 0.3]])
    v = np.array([[0.0, 1.0, 0.0]])
    gp, ga, gt, _ = _gpu(torch_cuda, x, v, np.array([1.0]), steps=True)
    assert gp[0] == 0 and np.all(ga == 0) and gt[0] < 1e-100
    x2 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    v2 = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.0]])
    m2 = np.array([1.0, 2.0])
    gp, ga, gt, _ = _gpu(torch_cuda, x2, v2, m2, steps=True)
    o = oracle_lib.Oracle(8, ncrit=64)
    op, oa, ot = o.master_direct(x2, v2, m2, steps=True)
    np.testing.assert_allclose(gp, [2.0, 1.0], rtol=1e-15)
    np.testing.assert_allclose(ga[:, 0], [2.0, -1.0], rtol=1e-15)
    np.testing.assert_allclose(gt, ot, rtol=1e-13)
    # errors come back as exceptions, not faults
    f = tca.Fmm(p=8, ncrit=64, direct_below=0)
    t = torch_cuda
    dx, dv, dm = (t.from_numpy(a).cuda() for a in mmg.ics.plummer(500, seed=1))
    with pytest.raises(RuntimeError, match="dt_param"):
        f.master_eval(dx, dv, dm, get_steps=True, dt_param=0.0)
    with pytest.raises(RuntimeError, match="mode"):
        f.master_eval(dx, dv, dm, mode=64)
    # below ncrit the FMM build runs direct(), as the reference does
    f2 = tca.Fmm(p=8, ncrit=600, direct_below=0)
    p1, a1, t1 = f2.master_eval(dx, dv, dm, get_steps=True, mode=tca.TC_MASTER_FMM)
    p0, a0, t0 = f2.master_eval(dx, dv, dm, get_steps=True)
    assert t.equal(p1, p0) and t.equal(a1, a0) and t.equal(t1, t0)


def _fmm_gpu(torch, c, x, v, m, sink, src, ao, mode_extra=0):
    import taichi_collisional_amd as tca
    f = tca.Fmm(p=c["p"], ncrit=c["ncrit"], mac=tca.TC_MAC_ERROR, force_accuracy=2e-7,
                direct_below=0)
    dx, dv, dm, dao = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m, ao))
    fl = torch.from_numpy((sink | (src << 1)).astype(np.uint8)).cuda()
    pot, acc, ts = f.master_eval(dx, dv, dm, flags=fl, get_steps=c["steps"],
                                 mode=tca.TC_MASTER_FMM | mode_extra, acc_old=dao)
    torch.cuda.synchronize()
    return (pot.cpu().numpy(), acc.cpu().numpy(), None if ts is None else ts.cpu().numpy(),
            dao.cpu().numpy(), f.stats())


@pytest.mark.parametrize("k", range(len(mmg.FMM_CASES)))
def test_gpu_master_fmm_matches_oracle(torch_cuda, oracle_lib, k):
    """The whole -DFMM path (pre-pass from the caller's acc_old included)
    against the oracle on the same inputs: the same tie decisions in the
    pre-pass and the walk, so results agree to round-off; the case with
    the mode-1 reference golden is also checked against it: the pre-pass ties
    are decided by the reference build's rounding (FMA-contracted norm), so
    acc_old and the forces agree with the reference to round-off."""
    g = golden("master_direct.npz")
    c = mmg.FMM_CASES[k]
    x, v, m, sink, src, ao, idx = mmg.fmm_inputs(c)
    assert mmg.sha(x, v, m, sink, src, ao, idx) == str(g["fmm%d_sha" % k])
    gp, ga, gt, gao, st = _fmm_gpu(torch_cuda, c, x, v, m, sink, src, ao)
    o = oracle_lib.Oracle(c["p"], ncrit=c["ncrit"], mac=1, force_accuracy=2e-7)
    op, oa, ot, oao = o.master_fmm(x, v, m, sink, src, steps=c["steps"], acc_old=ao)
    assert st["n_m2l"] == o.stats["n_m2l"] and st["n_p2p_cells"] == o.stats["n_p2p_cells"]
    np.testing.assert_allclose(gao, oao, rtol=1e-12)
    live = sink.astype(bool)
    _check((gp, ga, gt), (op, oa, ot), live, c["steps"], "fmm case %d vs oracle" % k)
    if c["mode"] == 1:
        e = np.abs(gao[idx] - g["fmm%d_acc_old" % k]) / g["fmm%d_acc_old" % k]
        assert e.max() < 1e-12
        lv = live[idx]
        ea = (np.linalg.norm(ga[idx][lv] - g["fmm%d_acc" % k][lv], axis=1) /
              np.linalg.norm(g["fmm%d_acc" % k][lv], axis=1))
        print("fmm case %d vs reference: acc_old max %.2e acc max %.2e" % (k, e.max(), ea.max()))
        assert ea.max() < 1e-10


def test_gpu_master_fmm_body_records(torch_cuda):
    """Body records: acc_old accumulates onto the records' values in the FMM
    build above ncrit, stays untouched below it and in the default build."""
    import taichi_collisional_amd as tca
    c = dict(p=8, ncrit=64, steps=True)
    x, v, m = mmg.ics.plummer(6000, seed=51)
    ones = np.ones(6000, np.uint8)
    ao = np.full(6000, 0.25)
    gp, ga, gt, gao, _ = _fmm_gpu(torch_cuda, c, x, v, m, ones, ones, ao.copy())
    b = np.zeros(6000, tca.master_body_dtype)
    b["X"], b["V"], b["q"], b["issink"], b["issource"] = x, v, m, 1, 1
    b["acc_old"] = 0.25
    f = tca.Fmm(p=8, ncrit=64, mac=tca.TC_MAC_ERROR, force_accuracy=2e-7, direct_below=0)
    f.master_get_force_and_potential(b, get_steps=True, mode=tca.TC_MASTER_FMM)
    np.testing.assert_allclose(b["acc_old"], gao, rtol=1e-13)
    assert np.all(b["acc_old"] > 0.25)
    np.testing.assert_allclose(b["p"], gp, rtol=1e-12)
    np.testing.assert_allclose(b["timestep"], gt, rtol=1e-12)
    b2 = b.copy()
    f.master_get_force_and_potential(b2, get_steps=False)
    assert np.array_equal(b2["acc_old"], b["acc_old"])


# ---------------------------------------------------------------------------
# Per-kernel entry points over the master Cell / Body records: the oracle's
# master tree and walk lists are replayed kernel by kernel through the C ABI.

def _master_records(tca, cl, x, v, m, sink, src, acc_old, p):
    order = cl["order"]
    n = order.size
    bodies = np.zeros(n, tca.master_body_dtype)
    bodies["X"], bodies["V"], bodies["q"] = x[order], v[order], m[order]
    bodies["issink"], bodies["issource"], bodies["index"] = sink[order], src[order], order
    bodies["acc_old"] = acc_old[order]
    nc = cl["level"].size
    cells = np.zeros(nc, tca.master_cell_dtype(p))
    cells["NCHILD"], cells["NBODY"] = cl["nchild"], cl["nbody"]
    csz = np.uint64(cells.dtype.itemsize)
    cells["CHILD"] = np.where(cl["nchild"] > 0,
                              np.uint64(cells.ctypes.data) + cl["child"].astype(np.uint64) * csz, 0)
    cells["BODY"] = np.uint64(bodies.ctypes.data) + cl["body"].astype(np.uint64) * np.uint64(128)
    cells["X"], cells["R"] = cl["Xg"], cl["Rg"]
    return bodies, cells


def test_gpu_master_per_kernel_replay_oracle(torch_cuda, oracle_lib):
    """P2M, M2M, symmetric M2L, L2L, L2P and P2P (with the time-step sum) on
    the master records reproduce the oracle's -DFMM evaluation: coefficients
    to 1e-12 / 1e-11, potentials, forces and time steps to the module's bars.
    The walk with steps accepts no far pairs and lists the same P2P pairs, so
    time steps come from that run and forces from the run without steps."""
    import taichi_collisional_amd as tca
    n, p = 20000, 8
    x, v, m = mmg.ics.plummer(n, seed=34)
    rng = np.random.default_rng(34)
    sink = (rng.random(n) < 0.7).astype(np.uint8)
    src = (rng.random(n) < 0.9).astype(np.uint8)
    o = oracle_lib.Oracle(p, ncrit=64, mac=1, force_accuracy=2e-7)
    _, _, ot, _ = o.master_fmm(x, v, m, sink, src, steps=True)
    tp1, sp1 = o.lists(1)
    assert o.lists(0)[0].size == 0
    op, oa, _, oao = o.master_fmm(x, v, m, sink, src, steps=False)
    cl = o.cells()
    nc = cl["level"].size
    OM, _, OL, _ = o.coefs(nc)
    tm, sm = o.lists(0)
    tp, sp = o.lists(1)
    assert np.array_equal(tp, tp1) and np.array_equal(sp, sp1)
    assert tm.size > 1000 and np.any(tm > sm) and np.any(tm < sm)
    bodies, cells = _master_records(tca, cl, x, v, m, sink, src, oao, p)
    f = tca.Fmm(p=p, ncrit=64, mac=tca.TC_MAC_ERROR, force_accuracy=2e-7, direct_below=0)
    leaves = np.where(cl["nchild"] == 0)[0]
    internal = np.where(cl["nchild"] > 0)[0]
    up = internal[np.argsort(-cl["level"][internal], kind="stable")]
    down = internal[np.argsort(cl["level"][internal], kind="stable")]

    def rel(a, b):
        return np.abs(a - b).max() / np.abs(b).max()

    f.master_P2M(bodies, cells, leaves)
    f.master_M2M(bodies, cells, up)
    assert np.abs(cells["X"] - cl["X"]).max() < 1e-12
    assert np.abs(cells["R"] - cl["R"]).max() < 1e-12
    assert rel(cells["M"], OM) < 1e-12
    # has_sink and min_acc over each cell's bodies
    cs = np.concatenate([[0], np.cumsum(bodies["issink"].astype(np.int64))])
    want_sink = cs[cl["body"] + cl["nbody"]] > cs[cl["body"]]
    assert np.array_equal(cells["has_sink"].astype(bool), want_sink)
    want_min = np.array([bodies["acc_old"][b:b + k].min() for b, k in zip(cl["body"], cl["nbody"])])
    np.testing.assert_array_equal(cells["min_acc"], want_min)
    f.master_M2L(cells, tm, sm)
    assert not cells["L"][~want_sink].any()  # locals of cells without sinks stay unwritten
    f.master_L2L(cells, down)
    assert rel(cells["L"], OL) < 1e-11
    f.master_L2P(bodies, cells, leaves)
    f.master_P2P(bodies, cells, tp, sp, mode=tca.TC_MASTER_STEPS, dt_param=0.025)
    order = cl["order"]
    live = sink[order].astype(bool)
    _check((bodies["p"], bodies["F"], bodies["timestep"]), (op[order], oa[order], ot[order]),
           live, True, "master per-kernel replay")
    assert not bodies["F"][~live].any() and not bodies["timestep"][~live].any()


def test_gpu_master_per_kernel_semantics(torch_cuda, oracle_lib):
    """List-order semantics of the symmetric M2L and the P2P modes: an entry
    with ci > cj does nothing, a repeated entry adds twice, (ci, cj) fills
    both locals; P2P without TC_MASTER_STEPS leaves timestep alone and the
    pair-mass form differs from the vector form; bad input is rejected."""
    import taichi_collisional_amd as tca
    n, p = 8000, 4
    x, v, m = mmg.ics.plummer(n, seed=35)
    ones = np.ones(n, np.uint8)
    o = oracle_lib.Oracle(p, ncrit=64, mac=1, force_accuracy=2e-7)
    _, _, _, oao = o.master_fmm(x, v, m, ones, ones, steps=False)
    cl = o.cells()
    tm, sm = o.lists(0)
    bodies, cells = _master_records(tca, cl, x, v, m, ones, ones, oao, p)
    f = tca.Fmm(p=p, ncrit=64, mac=tca.TC_MAC_ERROR, force_accuracy=2e-7, direct_below=0)
    leaves = np.where(cl["nchild"] == 0)[0]
    internal = np.where(cl["nchild"] > 0)[0]
    f.master_P2M(bodies, cells, leaves)
    f.master_M2M(bodies, cells, internal[np.argsort(-cl["level"][internal], kind="stable")])
    k = int(np.where(tm < sm)[0][0])
    i, j = int(tm[k]), int(sm[k])
    base = cells.copy()
    f.master_M2L(cells, [j], [i])                 # the pair in the other order: nothing
    assert np.array_equal(cells["L"], base["L"])
    f.master_M2L(cells, [i], [j])
    one = cells["L"].copy()
    assert one[i].any() and one[j].any()
    assert not np.delete(one, [i, j], axis=0).any()
    f.master_M2L(cells, [i, i], [j, j])
    np.testing.assert_allclose(cells["L"][[i, j]], 3 * one[[i, j]], rtol=1e-13, atol=0)
    with pytest.raises(RuntimeError, match="itself"):
        f.master_M2L(cells, [i], [i])
    tp, sp = o.lists(1)

    def p2p(mode):
        # the cells' BODY pointers address `bodies`: run on it, from zeroed outputs
        for k in ("p", "F", "timestep"):
            bodies[k] = 0
        f.master_P2P(bodies, cells, tp, sp, mode=mode)
        return bodies.copy()

    b0 = p2p(0)
    assert not b0["timestep"].any() and b0["p"].any()
    b1 = p2p(tca.TC_MASTER_STEPS)
    b2 = p2p(tca.TC_MASTER_STEPS | tca.TC_MASTER_PAIR_MASS)
    np.testing.assert_allclose(b1["p"], b0["p"], rtol=1e-13, atol=0)
    assert (b1["timestep"] > 0).all() and (b2["timestep"] > 0).all()
    assert not np.allclose(b1["timestep"], b2["timestep"], rtol=1e-6)
    with pytest.raises(RuntimeError, match="mode"):
        f.master_P2P(bodies, cells, tp, sp, mode=tca.TC_MASTER_FMM)
    bad = cells.copy()
    bad["BODY"][leaves[0]] += np.uint64(8)
    with pytest.raises(RuntimeError):
        f.master_P2M(bodies, bad, leaves[:1])
