"""tc_direct_soa_noalloc (fixed scratch, no device allocation) against
tc_direct_soa, through the C-ABI: two calls in a row, the second on a smaller
input, each equal to the allocating entry point's result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(fmm, torch, n, nt, seed):
    from taichi_collisional_amd import ics
    dev = torch.device("cuda:0")
    x, v, m = ics.plummer(n, seed=seed)
    rng = np.random.default_rng(seed)
    flags = rng.integers(1, 4, size=n).astype(np.uint8)  # sinks and/or sources
    tg = np.sort(rng.choice(n, size=nt, replace=False)).astype(np.int32)
    X, V, M = (torch.from_numpy(a).to(dev) for a in (x, v, m))
    F = torch.from_numpy(flags).to(dev)
    T = torch.from_numpy(tg).to(dev)
    return X, V, M, F, T


def test_noalloc_matches_direct_two_calls():
    torch = pytest.importorskip("torch")
    import taichi_collisional_amd as tca
    fmm = tca.Fmm(p=4, ncrit=64, theta=0.5)
    # the second call is smaller, so stale rows of the first stay in the
    # scratch array behind it
    for n, nt, seed in ((20000, 3000, 5), (3001, 700, 6)):
        print("noalloc case n=%d nt=%d seed=%d" % (n, nt, seed), flush=True)
        X, V, M, F, T = _case(fmm, torch, n, nt, seed)
        p0, a0, j0 = fmm.direct(X, V, M, T, flags=F)
        pot = torch.full((nt,), 7.0, dtype=torch.float64, device=X.device)
        acc = torch.full((nt, 3), 7.0, dtype=torch.float64, device=X.device)
        jerk = torch.full((nt, 3), 7.0, dtype=torch.float64, device=X.device)
        torch.cuda.synchronize()
        fmm.direct_noalloc(X, V, M, T, pot, acc, jerk, flags=F)
        torch.cuda.synchronize()
        for got, ref in ((pot, p0), (acc, a0), (jerk, j0)):
            g, r = got.cpu().numpy(), ref.cpu().numpy()
            # the same kernel over the same slabs; only the order of the
            # slabs' atomic adds may differ
            d = np.abs(g - r).reshape(nt, -1).max(axis=1)
            scale = np.abs(r).reshape(nt, -1).max(axis=1)
            assert np.all(d <= 1e-12 * scale)
    fmm.close()


def test_noalloc_rejects_too_many_bodies():
    torch = pytest.importorskip("torch")
    import taichi_collisional_amd as tca
    fmm = tca.Fmm(p=4, ncrit=64, theta=0.5)
    n = (1 << 18) + 1
    dev = torch.device("cuda:0")
    X = torch.zeros(n, 3, dtype=torch.float64, device=dev)
    M = torch.zeros(n, dtype=torch.float64, device=dev)
    T = torch.zeros(1, dtype=torch.int32, device=dev)
    out = [torch.zeros(1, dtype=torch.float64, device=dev), torch.zeros(1, 3, dtype=torch.float64, device=dev),
           torch.zeros(1, 3, dtype=torch.float64, device=dev)]
    with pytest.raises(RuntimeError, match="scratch"):
        fmm.direct_noalloc(X, X, M, T, *out)
    fmm.close()
