"""Many independent clusters in one evaluation (tc_eval_clusters, a forest
of per-cluster trees): equal to evaluating each cluster on a fresh context,
with the same root shift, to summation round-off -- the same pair counts,
the same forces, potentials and jerks -- for the theta MAC and the
reference's error-controlled MAC, with sink/source flags."""
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


def _clusters(k, seed, lo=10, hi=13):
    from taichi_collisional_amd import ics
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(lo * np.log(2), hi * np.log(2), size=k)).astype(int)
    xs, vs, ms = [], [], []
    for i, n in enumerate(sizes):
        x, v, m = ics.plummer(int(n), seed=seed * 1000 + i)
        xs.append(x + rng.normal(scale=5.0, size=3))  # clusters apart (no interaction anyway)
        vs.append(v)
        ms.append(m)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return np.concatenate(xs), np.concatenate(vs), np.concatenate(ms), starts


@pytest.mark.parametrize("mac", ["theta", "error"])
def test_forest_equals_per_cluster_contexts(torch_cuda, mac):
    import taichi_collisional_amd as tca
    torch = torch_cuda
    x, v, m, starts = _clusters(9, seed=3 if mac == "theta" else 4)
    n = int(starts[-1])
    rng = np.random.default_rng(5)
    flags = rng.choice(np.array([3, 3, 3, 1, 2], np.uint8), size=n)
    kw = dict(p=8, ncrit=64, theta=0.5, direct_below=0)
    if mac == "error":
        kw.update(mac=tca.TC_MAC_ERROR, force_accuracy=1e-7)
    dx, dv, dm = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m))
    fl = torch.from_numpy(flags).cuda()
    f = tca.Fmm(**kw)
    f.set_shift([0.2, -0.1, 0.05])
    pot, acc, jerk = f.eval_clusters(starts, dx, dv, dm, flags=fl)
    st = f.stats()
    pot, acc, jerk = pot.cpu().numpy(), acc.cpu().numpy(), jerk.cpu().numpy()
    m2l = p2p = 0
    for k in range(len(starts) - 1):
        a, b = int(starts[k]), int(starts[k + 1])
        g = tca.Fmm(**kw)
        g.set_shift([0.2, -0.1, 0.05])
        p1, a1, j1 = g.eval(dx[a:b].contiguous(), dv[a:b].contiguous(), dm[a:b].contiguous(),
                            flags=fl[a:b].contiguous())
        s1 = g.stats()
        m2l += s1["n_m2l"]
        p2p += s1["n_p2p_pairs"]
        p1, a1, j1 = p1.cpu().numpy(), a1.cpu().numpy(), j1.cpu().numpy()
        sink = (flags[a:b] & 1).astype(bool)
        assert not np.any(acc[a:b][~sink]) and not np.any(pot[a:b][~sink])
        ea = relerr_vec(acc[a:b][sink], a1[sink])[1]
        ej = relerr_vec(jerk[a:b][sink], j1[sink])[1]
        ep = relerr_scalar(pot[a:b][sink], p1[sink])[1]
        assert ea < 1e-12 and ej < 1e-11 and ep < 1e-13, (k, ea, ej, ep)
        g.close()
    assert st["n_m2l"] == m2l and st["n_p2p_pairs"] == p2p
    f.close()


def test_forest_rejects_bad_starts(torch_cuda):
    import taichi_collisional_amd as tca
    torch = torch_cuda
    f = tca.Fmm(p=4)
    x = torch.zeros(10, 3, dtype=torch.float64, device="cuda")
    m = torch.ones(10, dtype=torch.float64, device="cuda")
    for bad in ([0, 5, 5, 10], [1, 5, 10], [0, 6, 5, 10]):
        with pytest.raises(RuntimeError):
            f.eval_clusters(np.array(bad, np.int32), x, x, m)
    f.close()
