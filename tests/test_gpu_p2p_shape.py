"""Launch shapes of the near field (tc_set_p2p_shape, tc_tune_p2p): every
tile size, list slice, target-group cut and number of targets per lane
gives the same forces,
potentials, jerks and pair counts as the default shape, to FP64 summation
round-off -- for leaves of <= 64 bodies (ncrit 64) and for leaves longer than
a tile (ncrit 300: groups of the 64-body cut, ranges streamed through the
tile), in FP64 and in the mixed near-field forms.  The default shape itself
is pinned against the oracle and the reference's goldens by the parity
tests."""
import numpy as np
import pytest

from conftest import relerr_scalar, relerr_vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run(torch, f, dev):
    pot, acc, jerk = f.eval(*dev)
    torch.cuda.synchronize()
    return pot.cpu().numpy(), acc.cpu().numpy(), jerk.cpu().numpy(), f.stats()


@pytest.mark.parametrize("ncrit,mode", [(64, "TC_P2P_FP64"), (300, "TC_P2P_FP64"),
                                        (64, "TC_P2P_MIXED_HILO")])
def test_shapes_agree(torch_cuda, ncrit, mode):
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    torch = torch_cuda
    x, v, m = ics.plummer(40000, seed=17)
    dev = tuple(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m))
    f = tca.Fmm(p=8, ncrit=ncrit, theta=0.5, direct_below=0)
    f.set_shift([0.15, -0.05, 0.2])
    f.set_precision(getattr(tca, mode), close_radius=1e-3)
    assert f.p2p_shape() == (128, 128, 8.0, 1)
    p0, a0, j0, s0 = _run(torch, f, dev)
    for tile, lst, stage, per_lane in [(64, 128, 4.0, 1), (192, 128, 4.0, 1), (128, 1, 0.0, 1),
                                       (128, 37, 8.0, 1), (64, 1000, 0.0, 1), (192, 256, 100.0, 1),
                                       (128, 128, 4.0, 2), (64, 64, 0.0, 2), (192, 1, 8.0, 2)]:
        f.set_p2p_shape(tile, lst, stage, per_lane)
        p1, a1, j1, s1 = _run(torch, f, dev)
        ea, ej, ep = relerr_vec(a1, a0)[1], relerr_vec(j1, j0)[1], relerr_scalar(p1, p0)[1]
        print("ncrit %d %s shape (%d, %d, %g, %d): acc %.1e jerk %.1e pot %.1e" % (
            ncrit, mode, tile, lst, stage, per_lane, ea, ej, ep))
        assert s1["n_p2p_pairs"] == s0["n_p2p_pairs"] and s1["n_m2l"] == s0["n_m2l"]
        if mode == "TC_P2P_FP64":
            assert ea < 1e-13 and ej < 1e-12 and ep < 1e-13
        else:
            # FP32 sums run within a tile, so the tile's contents set their
            # rounding: shapes differ by FP32 rounding of the near field
            assert relerr_vec(a1, a0)[0] < 5e-8 and ea < 5e-6 and ep < 1e-6
    f.close()


def test_bad_shapes_rejected(torch_cuda):
    import taichi_collisional_amd as tca
    f = tca.Fmm(p=4)
    for args in [(96, 128, 4.0, 1), (128, 0, 4.0, 1), (128, 70000, 4.0, 1), (128, 128, -1.0, 1),
                 (128, 128, 4.0, 3)]:
        with pytest.raises(RuntimeError):
            f.set_p2p_shape(*args)
    with pytest.raises(RuntimeError):  # nothing evaluated yet
        f.tune_p2p()
    f.close()


def test_tuner_keeps_results(torch_cuda):
    """tc_tune_p2p picks a shape of its grid no slower than the default, and
    the following evaluation gives the same results as before tuning."""
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    torch = torch_cuda
    x, v, m = ics.plummer(1 << 17, seed=21)
    dev = tuple(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m))
    f = tca.Fmm(p=8, ncrit=64, theta=0.5, direct_below=0)
    f.set_seed(4)
    p0, a0, j0, s0 = _run(torch, f, dev)
    shape, best, dflt = f.tune_p2p(reps=2)
    print("tuned shape %s: %.3f ms (default %.3f ms)" % (shape, best, dflt))
    assert shape[0] in (64, 128, 192) and shape[1] in (64, 128, 256)
    assert shape[2] in (0.0, 2.0, 4.0, 8.0) and shape[3] in (1, 2)
    assert 0 < best <= dflt
    f.set_seed(4)
    p1, a1, j1, s1 = _run(torch, f, dev)
    assert s1["n_p2p_pairs"] == s0["n_p2p_pairs"]
    assert relerr_vec(a1, a0)[1] < 1e-13 and relerr_vec(j1, j0)[1] < 1e-12
    f.close()
