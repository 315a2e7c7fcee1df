"""Mixed-precision near field (tc_set_precision): FP32 body pairs on
group-relative coordinates, FP64 inside the close radius.

Checked against the FP64 FMM on the same tree (the difference is the near
field's FP32 rounding only: same lists, same far field) and against the FP64
all-pairs sum (the error may not grow much beyond the FMM's own).  The binary
workload puts pairs inside the close radius, so the FP64 branch is exercised;
with the radius set to zero those pairs go through FP32 and the binaries'
forces visibly degrade, which shows the branch is what keeps them exact.
"""
import numpy as np
import pytest

from conftest import relerr_vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _eval(torch, tca, x, v, m, mode, radius, p=12, theta=0.4, seed=3):
    f = tca.Fmm(p=p, ncrit=64, theta=theta, direct_below=0)
    f.set_seed(seed)
    f.set_precision(mode, close_radius=radius)
    dx, dv, dm = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m))
    pot, acc, jerk = f.eval(dx, dv, dm)
    torch.cuda.synchronize()
    out = (pot.cpu().numpy(), acc.cpu().numpy(), jerk.cpu().numpy())
    return f, (dx, dv, dm), out


@pytest.mark.parametrize("mode_name", ["TC_P2P_MIXED", "TC_P2P_MIXED_HILO"])
def test_mixed_p2p_plummer_gate(torch_cuda, mode_name):
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    n = 1 << 18
    x, v, m = ics.plummer(n, seed=4)
    f64, dev, (p0, a0, j0) = _eval(torch_cuda, tca, x, v, m, tca.TC_P2P_FP64, 0.0)
    fmx, _, (p1, a1, j1) = _eval(torch_cuda, tca, x, v, m, getattr(tca, mode_name), 1e-3)
    assert f64.stats()["n_p2p_pairs"] == fmx.stats()["n_p2p_pairs"]
    rng = np.random.default_rng(7)
    tg = np.sort(rng.choice(n, 4096, replace=False)).astype(np.int32)
    _, da, dj = f64.direct(*dev, torch_cuda.from_numpy(tg).cuda())
    da, dj = da.cpu().numpy(), dj.cpu().numpy()
    e64 = relerr_vec(a0[tg], da)
    emx = relerr_vec(a1[tg], da)
    jmx = relerr_vec(j1[tg], dj)
    dmx = relerr_vec(a1, a0)
    print("%s: force vs direct fp64 rms %.2e max %.2e | mixed rms %.2e max %.2e | "
          "jerk rms %.2e | mixed vs fp64 FMM rms %.2e max %.2e"
          % (mode_name, e64[0], e64[1], emx[0], emx[1], jmx[0], dmx[0], dmx[1]))
    # the FMM's own error against the direct sum is dominated by a few outer
    # halo bodies (relative errors ~1e-5 at r ~ 20 for p = 12, theta = 0.4 in
    # the oracle as in the reference); the near field's FP32 rounding adds
    # little to it: the single-float offsets lose digits for pairs much
    # closer than the leaf size, the hi + lo offsets keep them
    hilo = mode_name == "TC_P2P_MIXED_HILO"
    assert emx[0] <= (1.2 if hilo else 1.5) * e64[0]
    assert dmx[0] < (5e-8 if hilo else 2e-7) and dmx[1] < (5e-6 if hilo else 3e-5)
    assert relerr_vec(j1, j0)[0] < 1e-5
    assert np.max(np.abs(p1 - p0) / np.abs(p0)) < 1e-6


def test_mixed_p2p_close_pairs_in_fp64(torch_cuda):
    """Kroupa IMF + 10 % binaries (a in [1e-5, 1e-3]): the binary members'
    forces are dominated by their companion at r < 1e-3, evaluated in FP64."""
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    n = 1 << 16
    x, v, m = ics.kroupa_binaries(n, seed=8)
    _, _, (p0, a0, j0) = _eval(torch_cuda, tca, x, v, m, tca.TC_P2P_FP64, 0.0)
    _, _, (p1, a1, j1) = _eval(torch_cuda, tca, x, v, m, tca.TC_P2P_MIXED_HILO, 1e-3)
    _, _, (p2, a2, j2) = _eval(torch_cuda, tca, x, v, m, tca.TC_P2P_MIXED_HILO, 0.0)
    # binary members: a close companion
    from scipy.spatial import cKDTree
    d, _ = cKDTree(x).query(x, k=2)
    close = d[:, 1] < 1e-3
    assert close.sum() > 1000
    e_fp64 = relerr_vec(a1[close], a0[close])
    e_fp32 = relerr_vec(a2[close], a0[close])
    print("binary members %d: mixed (r<1e-3 in FP64) vs FP64 rms %.2e max %.2e; "
          "all pairs FP32 rms %.2e max %.2e" % (close.sum(), *e_fp64, *e_fp32))
    assert e_fp64[1] < 1e-6
    assert e_fp64[0] < e_fp32[0]
