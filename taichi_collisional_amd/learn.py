"""The learned per-cell expansion order and opening angle: measurement,
training and the model file.

    python -m taichi_collisional_amd.learn train [--snapshots 200] [--out PATH]
    python -m taichi_collisional_amd.learn calibrate [--out PATH]

1. Measurement.  For each seeded snapshot (Plummer and Dehnen gamma = 1.5 /
   1.0 spheres of 2^15 .. 2^17 bodies, some at 2^18 and 2^19) the FP64 direct sum
   gives the reference accelerations of a random quarter of the bodies; then
   the FMM runs once per candidate (q, theta) with every leaf at that pair
   (tc_set_adaptive mode 1) and tc_leaf_errors measures each leaf's rms
   relative force error on the sampled bodies.  Each (leaf, candidate) is one
   training row: the leaf's 9 features (tc_leaf_data), q / 20, theta ->
   log10 error.
2. Training.  The MLP (tc_mlp_*: 11 -> 64 -> 64 -> 64 -> 1, tanh, FP32) is
   fitted by Adam with the library's HIP kernels on an expectile loss
   (tau = 0.9: an under-predicted error weighs nine times an over-predicted
   one, so predictions lean to the safe side); inputs are standardized
   during training and the standardization is folded into the first layer
   afterwards.
3. Cost model.  A leaf's cost at (q, theta) is its mean M2L pairs at theta
   times the measured time of one M2L pair at order q, plus its mean P2P body
   pairs at theta times the measured time of one body pair (the times from
   one evaluation per order at N = 2^20 on the GPU doing the training: the
   per-pair time of k_m2l grows faster than its flops above p = 12).
4. Choice (in the library, tc_set_order_model): every leaf takes the cheapest
   candidate whose predicted error is at or below the target; internal cells
   inherit the largest order and smallest angle below them.  apply_model
   offers the network the angles at the context's order by default (see
   there for why mixed orders are not the default).

The fixed (p = 12, theta = 0.4) method is the fallback: a context without a
model (or with tc_set_adaptive mode 0) evaluates every pair at its p, theta.
"""
import argparse
import json
import os
import time

import numpy as np

from . import flops, ics

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "order_model.npz")
Q_GRID = (4, 6, 8, 10, 12, 14, 16, 18, 20)
THETA_GRID = (0.3, 0.4, 0.5, 0.6, 0.7)
P_MAX = 20
PROBE_THETA = 0.7


def candidates():
    q = np.array([q for q in Q_GRID for _ in THETA_GRID], dtype=np.int32)
    t = np.array([t for _ in Q_GRID for t in THETA_GRID], dtype=np.float32)
    return q, t


def snapshot_specs(count, seed=0):
    """(kind, n, seed) of the training snapshots: a seeded mix of models and
    sizes, 2^15 .. 2^17 bodies, every 6th at 2^18 (a gamma = 1.5 cusp) and
    every 40th at 2^19 (so that the deep cusp leaves and far halos of large
    systems are in the training range)."""
    rng = np.random.default_rng(seed)
    kinds = ("plummer", "dehnen15", "dehnen10")
    specs = []
    for i in range(count):
        n = int(1 << int(rng.integers(15, 18)))
        if i % 6 == 4:  # every 6th, a cusp model at 2^18
            n = 1 << 18
        if i % 40 == 39:
            n = 1 << 19
        specs.append((kinds[i % 3], n, int(seed * 100000 + i + 1)))
    return specs


def make_bodies(kind, n, seed):
    if kind == "plummer":
        return ics.plummer(n, seed=seed)
    if kind == "dehnen15":
        return ics.dehnen(n, gamma=1.5, seed=seed)
    if kind == "dehnen10":
        return ics.dehnen(n, gamma=1.0, seed=seed)
    raise ValueError(kind)


def measure_snapshot(torch, tca, fmm, x, v, m, seed, frac=0.25):
    """Training rows of one snapshot: features [R, 10], log10 errors [R], and
    per candidate the mean M2L pairs and P2P body pairs per leaf."""
    dev = torch.device("cuda", torch.cuda.current_device())
    n = m.size
    dx, dv, dm = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (x, v, m))
    rng = np.random.default_rng(seed + 7)
    tg = np.sort(rng.choice(n, max(1, int(n * frac)), replace=False)).astype(np.int32)
    tgd = torch.from_numpy(tg).to(dev)
    _, da, _ = fmm.direct(dx, dv, dm, tgd)
    ref = torch.zeros(n, 3, dtype=torch.float64, device=dev)
    ref[tgd.long()] = da
    # one tree for every candidate: a fixed root shift per snapshot
    fmm.set_shift(rng.uniform(-0.5, 0.49, size=3))
    qs, ts = candidates()
    acc = torch.empty(n, 3, dtype=torch.float64, device=dev)
    pot = torch.empty(n, dtype=torch.float64, device=dev)
    jerk = torch.empty(n, 3, dtype=torch.float64, device=dev)
    feats, rows, errs, work = None, [], [], []
    for q, t in zip(qs, ts):
        fmm.set_adaptive(1, int(q), float(t), PROBE_THETA)
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
        if feats is None:
            feats, _, _ = fmm.leaf_data()
        le = fmm.leaf_errors(acc, ref)
        ok = le > -29.0  # leaves with sampled bodies
        nl = int(le.numel())
        x_rows = torch.cat([feats, torch.full((nl, 1), float(q) / 20.0, device=dev),
                            torch.full((nl, 1), float(t), device=dev)], dim=1)
        rows.append(x_rows[ok])
        errs.append(le[ok])
        st = fmm.stats()
        work.append((st["n_m2l"] / nl, st["n_p2p_pairs"] / nl))
    fmm.set_adaptive(0)
    return torch.cat(rows), torch.cat(errs), np.array(work)


def _init_params(seed):
    """Glorot-uniform weights, zero biases (layout of tc_mlp_param_count)."""
    rng = np.random.default_rng(seed)
    nin, nh = 11, 64
    parts = []
    for fi, fo, shape in ((nin, nh, (nh, nin)), (nh, nh, (nh, nh)), (nh, nh, (nh, nh))):
        lim = np.sqrt(6.0 / (fi + fo))
        parts += [rng.uniform(-lim, lim, size=shape).ravel(), np.zeros(nh)]
    lim = np.sqrt(6.0 / (nh + 1))
    parts += [rng.uniform(-lim, lim, size=nh), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def fit(torch, tca, fmm, X, T, epochs=30, batch=16384, lr=2e-3, seed=0, tau=0.9, log=print):
    """Adam on the expectile loss mean(w (MLP(X) - T)^2) with the library's
    kernels (tc_mlp_grad, tc_mlp_adam); inputs standardized during training,
    the standardization folded into the first layer of the returned
    parameters."""
    dev = X.device
    mu = X.mean(0)
    sd = X.std(0).clamp_min(1e-3)
    Xn = ((X - mu) / sd).contiguous()
    Tn = T.contiguous()
    P = torch.from_numpy(_init_params(seed)).to(dev)
    mom = torch.zeros_like(P)
    vel = torch.zeros_like(P)
    g = torch.Generator(device="cpu").manual_seed(seed)
    step = 0
    hist = []
    B = X.shape[0]
    for ep in range(epochs):
        perm = torch.randperm(B, generator=g).to(dev)
        tot = 0.0
        nb = 0
        for k in range(0, B, batch):
            idx = perm[k:k + batch]
            xb = Xn[idx].contiguous()
            tb = Tn[idx].contiguous()
            G, loss = fmm.mlp_grad(P, xb, tb, tau)
            step += 1
            rate = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))  # cosine decay
            fmm.mlp_adam(P, G, mom, vel, max(rate, lr * 0.02), step)
            tot += loss
            nb += 1
        hist.append(tot / max(nb, 1))
        log("  epoch %d: train loss %.4f (rms %.3f dex)" % (ep, hist[-1], hist[-1] ** 0.5))
    # fold x_n = (x - mu) / sd into W1, b1
    Pc = P.detach().cpu().numpy().astype(np.float64)
    nin, nh = 11, 64
    W1 = Pc[:nh * nin].reshape(nh, nin)
    b1 = Pc[nh * nin:nh * nin + nh]
    mu_c, sd_c = mu.cpu().numpy().astype(np.float64), sd.cpu().numpy().astype(np.float64)
    W1r = W1 / sd_c[None, :]
    b1r = b1 - W1r @ mu_c
    Pc[:nh * nin] = W1r.ravel()
    Pc[nh * nin:nh * nin + nh] = b1r
    return Pc.astype(np.float32), hist


def pair_times(torch, tca, n=1 << 20, seed=1, theta=0.5):
    """Measured GPU time of one M2L pair per order in Q_GRID and of one P2P
    body pair: one evaluation per order of a Plummer N = n at theta, each in a
    context of that order (ns per pair)."""
    x, v, m = ics.plummer(n, seed=seed)
    dev = torch.device("cuda", torch.cuda.current_device())
    dx, dv, dm = (torch.from_numpy(a).to(dev) for a in (x, v, m))
    t_m2l, t_p2p = {}, []
    for q in Q_GRID:
        f = tca.Fmm(p=q, ncrit=64, theta=theta, direct_below=0)
        f.set_shift([0.1, 0.2, -0.3])
        f.eval(dx, dv, dm)
        ms_m2l = ms_p2p = 0.0
        for _ in range(2):
            f.eval(dx, dv, dm)
            st = f.stats()
            ms_m2l += st["ms_m2l"] / 2
            ms_p2p += st["ms_p2p"] / 2
        t_m2l[q] = ms_m2l * 1e6 / max(st["n_m2l"], 1)
        t_p2p.append(ms_p2p * 1e6 / max(st["n_p2p_pairs"], 1))
        f.close()
    return t_m2l, float(np.median(t_p2p))


def cost_table(work_rows, q, t, t_m2l, t_p2p):
    """Relative cost per candidate: mean M2L pairs per leaf x the time of one
    M2L pair at order q + mean P2P body pairs per leaf x the time of one body
    pair (pair counts averaged over the measured snapshots)."""
    w = np.mean(np.stack(work_rows), axis=0)  # [ncand, 2]
    c = np.array([w[k, 0] * t_m2l[int(q[k])] + w[k, 1] * t_p2p for k in range(len(q))])
    return (c / c.max()).astype(np.float32)


def save_model(path, params, q, t, cost, feat_lo, feat_hi, meta):
    np.savez(path, params=params, q=q, theta=t, cost=cost, feat_lo=feat_lo, feat_hi=feat_hi,
             meta=json.dumps(meta))


def load_model(path=MODEL_PATH):
    d = np.load(path, allow_pickle=False)
    return {"params": d["params"], "q": d["q"], "theta": d["theta"], "cost": d["cost"],
            "feat_lo": d["feat_lo"], "feat_hi": d["feat_hi"],
            "meta": json.loads(str(d["meta"]))}


MIN_ORDER = 8


def apply_model(fmm, model, target_rms=1e-7, margin=1.0, min_order=MIN_ORDER, orders="ctx"):
    """Give a context the learned choice, per-leaf target = margin x
    target_rms so that the rms over all bodies stays near target_rms.

    orders="ctx" (default): every leaf keeps the context's order and the
    network picks its opening angle.  For one leaf the predicted error falls
    monotonically with the angle, so the pick is reliable; measured on
    N = 2^20 Plummer and Dehnen cusps the result lands on the target (rms
    9.8e-8 and 9.6e-8 at target 1e-7).  orders=None: candidates of every
    order in [min_order, context order] -- the network then has to rank
    (q, theta) pairs whose errors differ by less than its ~0.3 dex scatter,
    and the measured rms misses the target by up to 5x; or a list of
    orders.  Orders below 8 never reach ~1e-7 on the angle grid and are left
    out."""
    if isinstance(orders, str) and orders == "ctx":
        orders = (fmm.p,)
    keep = (model["q"] <= fmm.p) & (model["q"] >= min_order)
    if orders is not None:
        keep &= np.isin(model["q"], np.asarray(orders))
    fmm.set_order_model(model["params"], model["q"][keep], model["theta"][keep],
                        model["cost"][keep], model["feat_lo"], model["feat_hi"],
                        target_rms * margin)
    fmm.set_adaptive(2, probe_theta=PROBE_THETA)


def calibrate(torch, tca, model, cases=(("plummer", 1 << 18, 9001), ("dehnen15", 1 << 18, 9002),
                                       ("plummer", 1 << 20, 1), ("dehnen15", 1 << 20, 2),
                                       ("dehnen10", 1 << 19, 77)),
              thetas=(0.4, 0.35, 0.3), margins=(1.0, 0.5), target=1e-7, p=12, log=print):
    """The learned per-leaf angle against fixed angles on snapshots outside the
    training set: rms relative force error (FP64 direct sum, seeded 4096-body
    subset), M2L + P2P work (flop) and ms of one evaluation per setting."""
    import time
    dev = torch.device("cuda", torch.cuda.current_device())
    rows = []
    for kind, n, seed in cases:
        x, v, m = make_bodies(kind, n, seed)
        dx, dv, dm = (torch.from_numpy(a).to(dev) for a in (x, v, m))
        rng = np.random.default_rng(seed)
        tg = torch.from_numpy(np.sort(rng.choice(n, 4096, replace=False)).astype(np.int32)).to(dev)
        f0 = tca.Fmm(p=4, direct_below=0)
        _, da, _ = f0.direct(dx, dv, dm, tg)
        f0.close()

        def run(f):
            f.set_seed(seed)
            f.eval(dx, dv, dm)
            torch.cuda.synchronize()
            t = time.perf_counter()
            _, acc, _ = f.eval(dx, dv, dm)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) * 1e3
            st = f.stats()
            c = f.m2l_order_counts()
            w = sum(c[q] * flops.c_m2l(q) for q in range(len(c))) + 44.0 * st["n_p2p_pairs"]
            e = torch.linalg.norm(acc[tg.long()] - da, dim=1) / torch.linalg.norm(da, dim=1)
            return {"rms": float(torch.sqrt((e ** 2).mean())), "max": float(e.max()),
                    "work": float(w), "ms": ms}
        row = {"kind": kind, "n": n, "seed": seed}
        for th in thetas:
            f = tca.Fmm(p=p, ncrit=64, theta=th, direct_below=0)
            row["fixed_%.2f" % th] = run(f)
            f.close()
        for mg in margins:
            f = tca.Fmm(p=p, ncrit=64, theta=0.4, direct_below=0)
            apply_model(f, model, target_rms=target, margin=mg)
            row["learned_margin_%.1f" % mg] = run(f)
            f.close()
        log(json.dumps(row))
        rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "calibrate"])
    ap.add_argument("--snapshots", type=int, default=200)
    ap.add_argument("--validation", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    import torch
    import taichi_collisional_amd as tca
    if args.cmd == "calibrate":
        calibrate(torch, tca, load_model(args.out), log=lambda r: print(r, flush=True))
        return
    t0 = time.time()
    fmm = tca.Fmm(p=P_MAX, ncrit=64, theta=0.4, direct_below=0)
    fmm.set_seed(1)
    Xs, Ts, work = [], [], []
    specs = snapshot_specs(args.snapshots + args.validation, seed=args.seed)
    for i, (kind, n, sd) in enumerate(specs):
        x, v, m = make_bodies(kind, n, sd)
        X, T, w = measure_snapshot(torch, tca, fmm, x, v, m, sd)
        Xs.append(X)
        Ts.append(T)
        work.append(w)
        if i % 10 == 0 or i == len(specs) - 1:
            print("measured %d/%d (%s n=%d): %d rows, %.0f s" % (
                i + 1, len(specs), kind, n, X.shape[0], time.time() - t0), flush=True)
    ntr = args.snapshots
    Xtr, Ttr = torch.cat(Xs[:ntr]), torch.cat(Ts[:ntr])
    Xva, Tva = torch.cat(Xs[ntr:]), torch.cat(Ts[ntr:])
    print("training rows %d, validation rows %d" % (Xtr.shape[0], Xva.shape[0]), flush=True)
    params, hist = fit(torch, tca, fmm, Xtr, Ttr, epochs=args.epochs, seed=args.seed,
                       log=lambda s: print(s, flush=True))
    Pd = torch.from_numpy(params).to(Xva.device)
    yva = fmm.mlp_forward(Pd, Xva.contiguous())
    va = float(torch.mean((yva - Tva) ** 2))
    ytr = fmm.mlp_forward(Pd, Xtr[:200000].contiguous())
    tr = float(torch.mean((ytr - Ttr[:200000]) ** 2))
    q, t = candidates()
    t_m2l, t_p2p = pair_times(torch, tca)
    print("ns per M2L pair by order %s, per P2P body pair %.4f" % (
        {k: round(v, 3) for k, v in t_m2l.items()}, t_p2p), flush=True)
    cost = cost_table(work[:ntr], q, t, t_m2l, t_p2p)
    meta = {"snapshots": ntr, "validation": args.validation, "rows": int(Xtr.shape[0]),
            "ns_per_m2l_pair": {str(k): v for k, v in t_m2l.items()}, "ns_per_p2p_pair": t_p2p,
            "epochs": args.epochs, "train_mse": tr, "validation_mse": va, "loss_history": hist,
            "probe_theta": PROBE_THETA, "p_max": P_MAX, "q_grid": list(Q_GRID),
            "theta_grid": list(THETA_GRID), "seconds": time.time() - t0}
    feats = Xtr[:, :9]
    save_model(args.out, params, q, t, cost, feats.min(0).values.cpu().numpy(),
               feats.max(0).values.cpu().numpy(), meta)
    print(json.dumps({k: v for k, v in meta.items() if k != "loss_history"}), flush=True)


if __name__ == "__main__":
    main()
