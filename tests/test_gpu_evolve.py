"""GPU parity of the master variant's step loop (do_evolve with the GRADIENT
scheme: evolve_split_omelyan + kick_gradient over a hierarchy of time-step
levels, master integrator.h:720-884) through the C ABI: against the compiled
reference's outputs (golden fixtures) and the oracle's loop.

Tolerances (FP64): every force of the loop is the same sum in another order,
so after the few hundred sub-steps of a case positions agree to ~1e-14 and
velocities to ~1e-13; the bars are 1e-11 and 1e-10.  The level structure
(time steps, splits, simulation time) must be identical."""
import sys

import numpy as np
import pytest

from conftest import GOLDEN, golden

sys.path.insert(0, GOLDEN)
import make_master_golden as mmg  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _state(torch, x, v, m):
    n = m.shape[0]
    dev = torch.device("cuda")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)  # noqa: E731
    return dict(pos=t(x), pos_e=t(np.zeros((n, 3))), vel=t(v), mass=t(m),
                timestep=t(np.zeros(n)), acc=t(np.zeros((n, 3))), pot=t(np.zeros(n)),
                acc_old=t(np.zeros(n)))


def _run(torch, x, v, m, nsteps, dt, ncrit=1000, mode=0):
    import taichi_collisional_amd as tca
    f = tca.Fmm(p=8, ncrit=ncrit, mac=tca.TC_MAC_ERROR, force_accuracy=2e-7, direct_below=0)
    st = _state(torch, x, v, m)
    t = 0.0
    for _ in range(nsteps):
        t = f.master_evolve(st, dt, simtime=t, mode=mode)
    torch.cuda.synchronize()
    return {k: a.cpu().numpy() for k, a in st.items()}, t


@pytest.mark.parametrize("k", range(len(mmg.EVOLVE_CASES)))
def test_gpu_step_loop_matches_reference(torch_cuda, oracle_lib, k):
    g = golden("master_direct.npz")
    c = mmg.EVOLVE_CASES[k]
    x, v, m = mmg.ics.plummer(c["n"], seed=c["seed"])
    assert mmg.sha(x, v, m) == str(g["evolve%d_sha" % k])
    out, t = _run(torch_cuda, x, v, m, c["nsteps"], c["dt"])
    print("case %d: dpos %.2e dvel %.2e" % (k, np.abs(out["pos"] - g["evolve%d_pos" % k]).max(),
                                            np.abs(out["vel"] - g["evolve%d_vel" % k]).max()))
    assert t == float(g["evolve%d_simtime" % k])
    np.testing.assert_allclose(out["timestep"], g["evolve%d_ts" % k], rtol=1e-11)
    assert np.abs(out["pos"] - g["evolve%d_pos" % k]).max() < 1e-11
    assert np.abs(out["vel"] - g["evolve%d_vel" % k]).max() < 1e-10
    o = oracle_lib.master_evolve(m, x, v, c["nsteps"], c["dt"], ncrit=1000)
    assert np.abs(out["pos"] - o["pos"]).max() < 1e-11
    # the potential the last kick left (zero where that kick had no sources:
    # with every body on the fast side, kick(fast, slow) sees an empty slow set)
    assert np.abs(out["pot"] - o["pot"]).max() <= 1e-10 * max(1.0, np.abs(o["pot"]).max())


def test_gpu_step_loop_above_small_tree_size(torch_cuda, oracle_lib):
    """One step of N = 5000 (dt = 1/4096, smallest step 3.4e-6): above the 4096
    bodies of the one-launch time-step tree, so kick_naive runs the
    level-by-level tree, the leaf-aligned time-step sum and, with the split,
    the multi-launch partition.  The time-step sum depends on the leaf
    partition (the pair mass grows along each source leaf, kernel_fast_lean.h:
    539), hence on ncrit: the compiled reference's step was taken at ncrit =
    65536 (one leaf, the all-pairs branch of get_force_and_potential), and the
    multi-leaf partition (ncrit = 1000) is checked against the oracle's loop."""
    g = golden("master_evolve_big.npz")
    c = mmg.EVOLVE_BIG
    x, v, m = mmg.ics.plummer(c["n"], seed=c["seed"])
    assert mmg.sha(x, v, m) == str(g["sha"])
    out, t = _run(torch_cuda, x, v, m, c["nsteps"], c["dt"], ncrit=1 << 16)
    dp, dv = np.abs(out["pos"] - g["pos"]).max(), np.abs(out["vel"] - g["vel"]).max()
    print("N=5000, one leaf: dpos %.2e dvel %.2e" % (dp, dv))
    assert t == float(g["simtime"])
    np.testing.assert_allclose(out["timestep"], g["ts"], rtol=1e-11)
    assert dp < 1e-11 and dv < 1e-10
    out, t = _run(torch_cuda, x, v, m, c["nsteps"], c["dt"], ncrit=1000)
    o = oracle_lib.master_evolve(m, x, v, c["nsteps"], c["dt"], ncrit=1000)
    dp, dv = np.abs(out["pos"] - o["pos"]).max(), np.abs(out["vel"] - o["vel"]).max()
    print("N=5000, ncrit 1000: dpos %.2e dvel %.2e" % (dp, dv))
    assert t == o["simtime"]
    np.testing.assert_allclose(out["timestep"], o["timestep"], rtol=1e-11)
    assert dp < 1e-11 and dv < 1e-10


def test_gpu_step_loop_particle_records(torch_cuda):
    """The host particle-record form equals the device-array form."""
    import taichi_collisional_amd as tca
    x, v, m = mmg.ics.plummer(300, seed=81)
    out, t = _run(torch_cuda, x, v, m, 1, 1.0 / 128)
    part = np.zeros(300, tca.particle_dtype)
    part["pos"], part["vel"], part["mass"] = x, v, m
    part["id"] = np.arange(300)
    f = tca.Fmm(p=8, ncrit=1000, mac=tca.TC_MAC_ERROR, force_accuracy=2e-7, direct_below=0)
    t2 = f.master_evolve_particles(part, 1.0 / 128)
    assert t2 == t
    np.testing.assert_array_equal(part["pos"], out["pos"])
    np.testing.assert_array_equal(part["vel"], out["vel"])
    np.testing.assert_array_equal(part["timestep"], out["timestep"])
    assert np.all(part["postime"] == 1.0 / 128)
    assert np.array_equal(part["id"], np.arange(300))


def test_gpu_step_loop_fmm_forces(torch_cuda, oracle_lib):
    """The loop on the -DFMM build's forces (pre-pass, symmetric walk) against
    the oracle's loop on the same forces."""
    import taichi_collisional_amd as tca
    x, v, m = mmg.ics.plummer(1200, seed=82)
    out, t = _run(torch_cuda, x, v, m, 1, 1.0 / 256, ncrit=64, mode=tca.TC_MASTER_FMM)
    o = oracle_lib.master_evolve(m, x, v, 1, 1.0 / 256, ncrit=64, fmm=True)
    assert t == o["simtime"]
    np.testing.assert_allclose(out["timestep"], o["timestep"], rtol=1e-11)
    assert np.abs(out["pos"] - o["pos"]).max() < 1e-11
    assert np.abs(out["vel"] - o["vel"]).max() < 1e-10
    assert np.all(out["acc_old"] > 0)


def test_gpu_step_loop_fmm_kick_cpu_above_ncrit(torch_cuda, oracle_lib):
    """kick_cpu (integrator.h:65-98) is an exact all-pairs sum in every build.
    With ncrit = 16 and N = 150, the subsystems of <= 128 bodies take the
    kick_cpu branch with up to 256 bodies in one force call, above ncrit: the
    FMM build must still use the exact sum there, so the loop agrees with the
    oracle's (whose kick_cpu is the exact sum) to round-off."""
    import taichi_collisional_amd as tca
    x, v, m = mmg.ics.plummer(150, seed=84)
    out, t = _run(torch_cuda, x, v, m, 2, 1.0 / 64, ncrit=16, mode=tca.TC_MASTER_FMM)
    o = oracle_lib.master_evolve(m, x, v, 2, 1.0 / 64, ncrit=16, fmm=True)
    dp, dv = np.abs(out["pos"] - o["pos"]).max(), np.abs(out["vel"] - o["vel"]).max()
    print("kick_cpu above ncrit: dpos %.2e dvel %.2e levels %s" %
          (dp, dv, np.unique(np.round(np.log2(1.0 / 64 / out["timestep"])))))
    assert t == o["simtime"]
    np.testing.assert_allclose(out["timestep"], o["timestep"], rtol=1e-11)
    # some bodies must sit below the top level, or kick_cpu never runs
    assert np.any(out["timestep"] < 1.0 / 64)
    assert dp < 1e-13 and dv < 1e-12


def test_gpu_step_loop_edges(torch_cuda):
    import taichi_collisional_amd as tca
    x, v, m = mmg.ics.plummer(50, seed=83)
    f = tca.Fmm(p=8, ncrit=1000, direct_below=0)
    st = _state(torch_cuda, x, v, m)
    assert f.master_evolve(st, 0.0, simtime=1.5) == 1.5  # dt = 0 returns at once
    assert np.array_equal(st["pos"].cpu().numpy(), x)
    with pytest.raises(RuntimeError, match="dt_param"):
        f.master_evolve(st, 0.01, dt_param=0.0)


def test_gpu_step_loop_jerk_fmm_equals_master_loop_when_exact(torch_cuda):
    """TC_EVOLVE_JERK_FMM drives the same fourth-order loop with the jerk
    build's evaluation.  With all-pairs forces (direct_below > n) and a dt
    below every body's time step (no split), it must reproduce the master
    loop with its default (all-pairs) forces to round-off."""
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    n, dt = 2000, 2.0 ** -20
    x, v, m = ics.plummer(n, seed=17)
    fm = tca.Fmm(p=8, ncrit=64, direct_below=0)
    sm = _state(torch_cuda, x, v, m)
    tm = fm.master_evolve(sm, dt, mode=0)
    assert float(sm["timestep"].min()) >= dt  # single level in the master loop too
    fj = tca.Fmm(p=8, ncrit=64, direct_below=n + 1)
    sj = _state(torch_cuda, x, v, m)
    tj = fj.master_evolve(sj, dt, mode=tca.TC_EVOLVE_JERK_FMM)
    torch_cuda.cuda.synchronize()
    assert tm == tj == dt
    assert bool(torch_cuda.isinf(sj["timestep"]).all())
    dp = float((sj["pos"] - sm["pos"]).abs().max())
    dv = float((sj["vel"] - sm["vel"]).abs().max())
    print("jerk-FMM loop vs master loop: dpos %.1e dvel %.1e" % (dp, dv))
    assert dp < 1e-14 and dv < 1e-12
    fm.close()
    fj.close()


def test_gpu_c5_fourth_order_steps_dehnen(torch_cuda):
    """The learned-order workload's integration as a bounded test: Dehnen
    gamma = 1.5, ten fourth-order steps of dt = 1e-4 with every force the
    jerk build's theta-MAC FMM at the fallback p = 12 / theta = 0.4
    (N = 2^18 here; the benchmark runs N = 2^23).  Energy is conserved to the
    force error and the force error does not grow over the steps."""
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    from conftest import relerr_vec
    torch = torch_cuda
    n, dt = 1 << 18, 1e-4
    x, v, m = ics.dehnen(n, gamma=1.5, seed=23)
    f = tca.Fmm(p=12, ncrit=64, theta=0.4, direct_below=0)
    f.set_seed(23)
    st = _state(torch, x, v, m)

    def energy():
        kin = 0.5 * float(torch.sum(st["mass"][:, None] * st["vel"] ** 2))
        _, _, _ = f.eval(st["pos"], st["vel"], st["mass"], pot=st["pot"])
        return kin - 0.5 * float(torch.sum(st["mass"] * st["pot"]))

    rng = np.random.default_rng(3)
    tg = torch.from_numpy(np.sort(rng.choice(n, 2048, replace=False)).astype(np.int32)).cuda()

    def force_error():
        _, acc, _ = f.eval(st["pos"], st["vel"], st["mass"])
        _, da, _ = f.direct(st["pos"], st["vel"], st["mass"], tg)
        return relerr_vec(acc[tg.long()].cpu().numpy(), da.cpu().numpy())

    err0 = force_error()
    e0 = energy()
    t = 0.0
    for _ in range(10):
        t = f.master_evolve(st, dt, simtime=t, mode=tca.TC_EVOLVE_JERK_FMM)
    e1 = energy()
    err = force_error()
    print("C5-style: t=%.1e dE/E %.2e, force rms %.2e max %.2e (initial %.2e / %.2e)" % (
        t, abs(e1 - e0) / abs(e0), err[0], err[1], err0[0], err0[1]))
    assert abs(t - 10 * dt) < 1e-15
    assert abs(e1 - e0) < 1e-6 * abs(e0)
    # the FMM's own error (dominated by far-halo bodies of the cusp model)
    # does not grow over the steps
    assert err[0] <= 1.5 * err0[0] and err[1] <= 1.5 * err0[1]
    f.close()
