"""Pin the MLP oracle (oracle/mlp_oracle.py) by mathematics: its analytic
gradient equals central finite differences of its loss, and its forward pass
equals an independent per-sample loop."""
import numpy as np

import mlp_oracle as mo


def test_param_count_matches_library():
    import taichi_collisional_amd as tca
    assert mo.NPARAM == tca.mlp_param_count() == 9153


def test_forward_equals_per_sample_loop():
    rng = np.random.default_rng(1)
    P = mo.init_params(3) + rng.normal(scale=0.05, size=mo.NPARAM)
    X = rng.normal(size=(7, mo.NIN))
    w = mo.unpack(P)
    y = mo.forward(P, X)
    for b in range(7):
        h = X[b]
        for k in ("1", "2", "3"):
            W, bias = w["W" + k], w["b" + k]
            h = np.array([np.tanh(sum(W[j, i] * h[i] for i in range(W.shape[1])) + bias[j])
                          for j in range(W.shape[0])])
        assert abs(y[b] - (np.dot(w["W4"], h) + w["b4"][0])) < 1e-12


def test_gradient_equals_finite_differences():
    rng = np.random.default_rng(2)
    P = mo.init_params(4) + rng.normal(scale=0.05, size=mo.NPARAM)
    X = rng.normal(size=(33, mo.NIN))
    T = rng.normal(size=33)
    tau = 0.8  # the expectile weights: both branches of w are crossed
    _, g = mo.loss_and_grad(P, X, T, tau)
    # every parameter block, a few entries each (biases and weights alike)
    idx, o = [], 0
    for _, sh in mo.SIZES:
        n = int(np.prod(sh))
        idx += list(o + rng.choice(n, size=min(n, 6), replace=False))
        o += n
    h = 1e-6
    for k in idx:
        e = np.zeros(mo.NPARAM)
        e[k] = h
        fd = (mo.loss_and_grad(P + e, X, T, tau)[0] -
              mo.loss_and_grad(P - e, X, T, tau)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])
