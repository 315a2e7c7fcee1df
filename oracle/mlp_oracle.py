"""Test infrastructure only: CPU (numpy, FP64) restatement of the order
model's MLP, the checker for the HIP training and inference kernels
(taichi_collisional_amd/csrc/tc_learn.hip).

The network has no counterpart in the reference (which sets its error
control by a formula, taichi_with_jerks/traverse_eager.h:59-122); what pins
this restatement is mathematics: its analytic gradient is checked against
central finite differences (tests/test_mlp_oracle.py), and the forward pass
against a plain layer-by-layer evaluation.

Parameter layout (the kernels'): W1[64][11] b1[64] W2[64][64] b2[64]
W3[64][64] b3[64] W4[64] b4[1]; tanh hidden units, linear output.
"""
import numpy as np

NIN, NH = 11, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
         ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NH,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(P):
    out, o = {}, 0
    for k, sh in SIZES:
        n = int(np.prod(sh))
        out[k] = np.asarray(P[o:o + n], dtype=np.float64).reshape(sh)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[k], dtype=np.float64).reshape(-1) for k, _ in SIZES])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for k, sh in SIZES:
        if k.startswith("W"):
            fan_in = sh[1] if len(sh) == 2 else sh[0]
            fan_out = sh[0] if len(sh) == 2 else 1
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            d[k] = rng.uniform(-lim, lim, size=sh)
        else:
            d[k] = np.zeros(sh)
    return pack(d)


def forward(P, X, keep=False):
    w = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.tanh(X @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    h3 = np.tanh(h2 @ w["W3"].T + w["b3"])
    y = h3 @ w["W4"] + w["b4"][0]
    return (y, (h1, h2, h3)) if keep else y


def loss_and_grad(P, X, T, tau=0.5):
    """The expectile loss mean(w (y - T)^2), w = tau where y < T and 1 - tau
    elsewhere, and its gradient with respect to every parameter."""
    w = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    B = X.shape[0]
    y, (h1, h2, h3) = forward(P, X, keep=True)
    r = y - T
    wt = np.where(r < 0, tau, 1.0 - tau)
    loss = float(np.mean(wt * r * r))
    dy = 2.0 * wt * r / B
    g = {"W4": h3.T @ dy, "b4": np.array([dy.sum()])}
    d3 = dy[:, None] * w["W4"][None, :] * (1 - h3 * h3)
    g["W3"], g["b3"] = d3.T @ h2, d3.sum(0)
    d2 = (d3 @ w["W3"]) * (1 - h2 * h2)
    g["W2"], g["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ w["W2"]) * (1 - h1 * h1)
    g["W1"], g["b1"] = d1.T @ X, d1.sum(0)
    return loss, pack(g)
