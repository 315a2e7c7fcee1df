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
