#!/usr/bin/env python3
"""Benchmark: particles/s per FMM force+jerk evaluation on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n N] [--p P] [--theta T]
                    [--ncrit C] [--no-cpu] [--gpus 1]

Workload (default): Plummer sphere N = 2^20 (Henon units, seed 1), expansion
order p = 12, opening angle theta = 0.4, ncrit = 64, FP64 throughout; a
"step" is one full evaluation (tree build through L2P) of potential,
acceleration and jerk for every body, with positions, velocities and masses
already resident in HBM.  Prints one JSON line.

The CPU baseline is the reference implementation itself (compiled from its
sources into oracle/_ref/), run in a child process on the host cores before
the GPU is touched, on the same workload (one evaluation, theta-MAC passes).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s per FMM force+jerk eval at N=1M, rel force err <=1e-7 vs FP64 direct"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def cpu_baseline(n, p, theta, ncrit, seed, threads, timeout, repeats=3, default_path=True):
    """Time the reference (oracle/_ref) on the host in a child process: one
    warm-up and `repeats` timed evaluations of the theta-MAC workload (median
    reported), then one evaluation of the reference's default path (p=2
    pre-pass + error-controlled MAC, get_force_and_potential as shipped)."""
    code = r"""
import sys, time, json, os
sys.path.insert(0, %r); sys.path.insert(0, os.path.join(%r, "oracle"))
import numpy as np, pyoracle as po
from taichi_collisional_amd import ics
n, p, theta, ncrit, seed, repeats, dflt = %d, %d, %r, %d, %d, %d, %d
x, v, m = ics.plummer(n, seed=seed)
if not po.ref_available(p):
    print(json.dumps({"error": "reference harness for p=%%d not built" %% p})); sys.exit(0)
L = po.reference(p)
po.ref_eval(p, x, v, m, seed, ncrit=ncrit, theta=theta, mode=0)  # warm-up
ts = []
for _ in range(repeats):
    t = time.time()
    pot, acc, jerk, acco, counts, shift = po.ref_eval(p, x, v, m, seed, ncrit=ncrit, theta=theta, mode=0)
    ts.append(time.time() - t)
out = {"seconds": sorted(ts)[len(ts) // 2], "all": ts, "isa": L.isa, "counts": [int(c) for c in counts]}
if dflt:
    t = time.time()
    po.ref_eval(p, x, v, m, seed, ncrit=ncrit, theta=theta, mode=1)
    out["default_seconds"] = time.time() - t
print(json.dumps(out))
""" % (ROOT, ROOT, n, p, theta, ncrit, seed, repeats, int(default_path))
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
               OMP_WAIT_POLICY="active", OMP_STACKSIZE="100M")
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=timeout, env=env)
        out = json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # noqa: BLE001 - reported in the JSON line
        return {"value": None, "unit": "particles/s", "cores": threads, "kind": "reference",
                "sample": "failed: %s" % str(e)[:200]}
    if "seconds" not in out:
        return {"value": None, "unit": "particles/s", "cores": threads, "kind": "reference",
                "sample": out.get("error", "unavailable")}
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    res = {"value": n / out["seconds"], "unit": "particles/s", "cores": threads,
           "kind": "reference",
           "sample": "median of %d full evaluations after one warm-up, same workload (theta-MAC "
                     "passes, N=%d, p=%d, theta=%g, ncrit=%d): %s s wall, %s build, OpenMP %d "
                     "threads, %s" % (len(out["all"]), n, p, theta, ncrit,
                                      ", ".join("%.2f" % t for t in out["all"]), out["isa"],
                                      threads, cpu_model),
           "seconds": out["seconds"], "m2l_pairs": out["counts"][0]}
    if "default_seconds" in out:
        res["default_path_seconds"] = out["default_seconds"]
        res["default_path_value"] = n / out["default_seconds"]
    return res


def master_cpu_baseline(n, seed, threads, timeout):
    """The master reference's default build (direct() with get_steps) on the
    host, one evaluation after a warm-up on a small input."""
    code = r"""
import sys, time, json, os
sys.path.insert(0, %r); sys.path.insert(0, os.path.join(%r, "oracle"))
import numpy as np, pyoracle as po
from taichi_collisional_amd import ics
if not po.master_ref_available(8):
    print(json.dumps({})); sys.exit(0)
x, v, m = ics.plummer(%d, seed=%d)
po.master_ref_eval(8, x[:4096], v[:4096], m[:4096], ncrit=300, mode=0, get_steps=True)
t = time.time()
po.master_ref_eval(8, x, v, m, ncrit=300, mode=0, get_steps=True)
print(json.dumps({"cpu_reference_seconds": time.time() - t}))
""" % (ROOT, ROOT, n, seed)
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
               OMP_WAIT_POLICY="active", OMP_STACKSIZE="100M")
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=timeout, env=env)
        out = json.loads(r.stdout.strip().splitlines()[-1])
    except Exception:  # noqa: BLE001 - the row just lacks the CPU figure
        return {}
    out["cpu_cores"] = threads
    return out


def master_evolve_cpu_baseline(n, dt, seed, threads, timeout):
    """The master reference's do_evolve (GRADIENT) on the host, one step."""
    code = r"""
import sys, time, json, os
sys.path.insert(0, %r); sys.path.insert(0, os.path.join(%r, "oracle"))
import numpy as np, pyoracle as po
from taichi_collisional_amd import ics
if not po.master_ref_available(8):
    print(json.dumps({})); sys.exit(0)
x, v, m = ics.plummer(%d, seed=%d)
t = time.time()
po.master_ref_evolve(8, m, x, v, 1, %r, ncrit=300)
print(json.dumps({"cpu_reference_seconds": time.time() - t}))
""" % (ROOT, ROOT, n, seed, dt)
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
               OMP_WAIT_POLICY="active", OMP_STACKSIZE="100M")
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=timeout, env=env)
        out = json.loads(r.stdout.strip().splitlines()[-1])
    except Exception:  # noqa: BLE001 - the row just lacks the CPU figure
        return {}
    out["cpu_cores"] = threads
    return out


def load_traffic(path):
    """HBM bytes and executed FP64 flops per k_m2l launch, from the committed
    rocprofv3 counter summary (None where absent)."""
    try:
        d = json.load(open(path))
        return d.get("hbm_bytes_per_launch"), d.get("hw_fp64_flops_per_launch")
    except (OSError, ValueError):
        return None, None


def master_evolve_row(tca, ics, torch, dev, p, seed):
    """do_evolve of the master step loop (GRADIENT scheme) on a Plummer N =
    4096, dt = 1/1024: after one untimed warm-up on a copy, three timed calls
    in a row from the initial state; ms_per_step is their median."""
    import numpy as np
    ne, dte = 4096, 1.0 / 1024
    xe, ve, me = ics.plummer(ne, seed=seed)
    est = {k: torch.zeros(ne, 3, dtype=torch.float64, device=dev) for k in ("pos_e", "acc")}
    est.update({k: torch.zeros(ne, dtype=torch.float64, device=dev)
                for k in ("timestep", "pot", "acc_old")})
    est["pos"], est["vel"], est["mass"] = (torch.from_numpy(a).to(dev) for a in (xe, ve, me))
    fe = tca.Fmm(p=p, ncrit=300, direct_below=0)
    warm = {k: a.clone() for k, a in est.items()}
    fe.master_evolve(warm, dte)  # warm-up on a copy (first launches, allocations)
    del warm
    calls = []
    for c in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fe.master_evolve(est, dte)
        torch.cuda.synchronize()
        calls.append((time.perf_counter() - t) * 1e3)
        if c == 0:
            tau = float(est["timestep"][0])
    ems = float(np.median(calls))
    levels = 0
    while dte / 2 ** levels > tau:
        levels += 1
    fe.close()
    return {"what": "master step loop: one do_evolve(sys, dt) of the GRADIENT scheme, "
                    "Plummer N=4096, dt=1/1024, dt_param 0.025, default-build forces "
                    "(median of three calls in a row)",
            "ms_per_step": ems, "ms_calls": calls, "levels": levels, "substeps": 2 ** levels}


def timed_row(fmm, torch, dx, dv, dm, out, idx, da, dj, reps=5, warm=2):
    """ms per evaluation of a configured context (after warm-up), its phase
    times and the force / jerk error against the direct sums da, dj of the
    bodies idx (not timed)."""
    pot, acc, jerk = out
    for _ in range(warm):
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
    torch.cuda.synchronize()
    t = time.perf_counter()
    sts = []
    for _ in range(reps):
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
        sts.append(fmm.stats())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / reps
    ea = torch.linalg.norm(acc[idx] - da, dim=1) / torch.linalg.norm(da, dim=1)
    ej = torch.linalg.norm(jerk[idx] - dj, dim=1) / torch.linalg.norm(dj, dim=1)
    return {"ms_per_eval": ms, "particles_per_s": dx.shape[0] / (ms * 1e-3),
            "ms_p2p": sum(s_["ms_p2p"] for s_ in sts) / reps,
            "ms_m2l": sum(s_["ms_m2l"] for s_ in sts) / reps,
            "force_rms": float(torch.sqrt(torch.mean(ea ** 2))), "force_max": float(ea.max()),
            "jerk_rms": float(torch.sqrt(torch.mean(ej ** 2))), "jerk_max": float(ej.max()),
            "m2l_pairs": sts[-1]["n_m2l"], "p2p_body_pairs": sts[-1]["n_p2p_pairs"]}


def gate_row(tca, torch, dx, dv, dm, idx, da, p=12, theta=0.4, ncrit=64, nshift=8, seed=2024):
    """The error gate across root shifts: each method evaluated at the same
    nshift root shifts (fixed, seeded), with its time per evaluation (the
    library's own event timing of the whole evaluation, median) and the force
    rms against the FP64 direct sum of the bodies idx at every shift; how many
    shifts meet rms <= 1e-7.  Methods: the fixed p / theta of the benchmark, a
    smaller fixed angle (0.35, also with the mixed-precision near field: hi +
    lo offsets, FP64 pairs inside r = 1e-3), the order p + 2 at the same angle,
    and the learned per-leaf angle (target 1e-7), in FP64 and mixed."""
    import numpy as np
    from taichi_collisional_amd import learn
    rng = np.random.default_rng(seed)
    shifts = rng.uniform(-0.5, 0.5, size=(nshift, 3))
    model = learn.load_model() if os.path.exists(learn.MODEL_PATH) else None
    methods = [("fixed_p%d_theta%.2f" % (p, theta), p, theta, False, False),
               ("fixed_p%d_theta0.35" % p, p, 0.35, False, False),
               ("fixed_p%d_theta0.35_mixed" % p, p, 0.35, False, True),
               ("fixed_p%d_theta%.2f" % (p + 2, theta), p + 2, theta, False, False)]
    if model is not None:
        methods += [("learned_target1e-7", p, theta, True, False),
                    ("learned_target1e-7_mixed", p, theta, True, True)]
    out = {"what": "force rms vs the FP64 direct sum at %d fixed root shifts per method; "
                   "ms_per_eval is the median over them" % nshift}
    for name, pm, th, learned, mixed in methods:
        f = tca.Fmm(p=pm, ncrit=ncrit, theta=th, direct_below=0)
        if learned:
            learn.apply_model(f, model, target_rms=1e-7)
        if mixed:
            f.set_precision(tca.TC_P2P_MIXED_HILO, close_radius=1e-3)
        f.set_shift(shifts[0])
        f.eval(dx, dv, dm)  # warm-up
        ms, rms = [], []
        for sh in shifts:
            f.set_shift(sh)
            _, acc, _ = f.eval(dx, dv, dm)
            torch.cuda.synchronize()
            ms.append(f.stats()["ms_total"])
            ea = torch.linalg.norm(acc[idx] - da, dim=1) / torch.linalg.norm(da, dim=1)
            rms.append(float(torch.sqrt(torch.mean(ea ** 2))))
        out[name] = {"ms_per_eval": float(np.median(ms)), "force_rms_median": float(np.median(rms)),
                     "force_rms_max": float(max(rms)), "force_rms": rms,
                     "shifts_meeting_1e-7": "%d of %d" % (sum(r <= 1e-7 for r in rms), nshift)}
        f.close()
    return out


def clusters_row(tca, ics, torch, dev, seed=7, k=256, reps=3):
    """Many independent Plummer clusters (N log-uniform in 2^10..2^16, equal
    masses within a cluster, p = 8, theta = 0.5): one forest evaluation of all
    of them (tc_eval_clusters) against one evaluation per cluster."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(10 * np.log(2), 16 * np.log(2), size=k)).astype(int)
    xs, vs, ms = [], [], []
    for i, nn in enumerate(sizes):
        x, v, m = ics.plummer(int(nn), seed=seed * 1000 + i)
        xs.append(x + rng.normal(scale=20.0, size=3))
        vs.append(v)
        ms.append(m)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(starts[-1])
    dx, dv, dm = (torch.from_numpy(np.concatenate(a)).to(dev) for a in (xs, vs, ms))
    out = (torch.empty(n, dtype=torch.float64, device=dev),
           torch.empty(n, 3, dtype=torch.float64, device=dev),
           torch.empty(n, 3, dtype=torch.float64, device=dev))
    f = tca.Fmm(p=8, ncrit=64, theta=0.5, direct_below=0)
    f.set_seed(seed)
    f.eval_clusters(starts, dx, dv, dm, pot=out[0], acc=out[1], jerk=out[2])
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        f.eval_clusters(starts, dx, dv, dm, pot=out[0], acc=out[1], jerk=out[2])
    torch.cuda.synchronize()
    forest_ms = (time.perf_counter() - t) * 1e3 / reps
    st = f.stats()
    # one evaluation per cluster on the same context (after a warm-up pass)
    views = [(dx[a:b], dv[a:b], dm[a:b], out[0][a:b], out[1][a:b], out[2][a:b])
             for a, b in zip(starts[:-1], starts[1:])]
    for vw in views:
        f.eval(*vw[:3], pot=vw[3], acc=vw[4], jerk=vw[5])
    torch.cuda.synchronize()
    t = time.perf_counter()
    for vw in views:
        f.eval(*vw[:3], pot=vw[3], acc=vw[4], jerk=vw[5])
    torch.cuda.synchronize()
    each_ms = (time.perf_counter() - t) * 1e3
    f.close()
    return {"what": "%d independent Plummer clusters (N log-uniform 2^10..2^16, %d bodies), "
                    "p=8, theta=0.5: all clusters in one evaluation (each phase one launch "
                    "for all clusters) vs one evaluation per cluster" % (k, n),
            "bodies": n, "forest_ms": forest_ms, "per_cluster_ms": each_ms,
            "speedup": each_ms / forest_ms, "particles_per_s": n / (forest_ms * 1e-3),
            "m2l_pairs": st["n_m2l"], "p2p_body_pairs": st["n_p2p_pairs"]}


def cusp_steps_row(tca, ics, torch, dev, n, learned=False, steps=10, dt=1e-4, seed=5):
    """Dehnen gamma = 1.5 (N bodies): ten fourth-order steps of dt = 1e-4
    (the master GRADIENT loop with the jerk build's theta-MAC FMM forces,
    TC_EVOLVE_JERK_FMM) at the fallback p = 12 / theta = 0.4, or with the
    learned per-leaf choice; energy change and the final force error."""
    import numpy as np
    from taichi_collisional_amd import flops, learn
    x, v, m = ics.dehnen(n, gamma=1.5, seed=seed)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    st = dict(pos=t(x), pos_e=torch.zeros(n, 3, dtype=torch.float64, device=dev), vel=t(v),
              mass=t(m), timestep=torch.zeros(n, dtype=torch.float64, device=dev),
              acc=torch.zeros(n, 3, dtype=torch.float64, device=dev),
              pot=torch.zeros(n, dtype=torch.float64, device=dev),
              acc_old=torch.zeros(n, dtype=torch.float64, device=dev))
    if learned:
        # candidates up to the fallback's order (the context's order)
        f = tca.Fmm(p=12, ncrit=64, theta=0.4, direct_below=0)
        learn.apply_model(f, learn.load_model(), target_rms=1e-7)
    else:
        f = tca.Fmm(p=12, ncrit=64, theta=0.4, direct_below=0)
    f.set_seed(seed)

    def energy():
        f.eval(st["pos"], st["vel"], st["mass"], pot=st["pot"])
        kin = 0.5 * float(torch.sum(st["mass"][:, None] * st["vel"] ** 2))
        return kin - 0.5 * float(torch.sum(st["mass"] * st["pot"]))

    e0 = energy()
    ev_ms = f.stats()["ms_total"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim = 0.0
    for _ in range(steps):
        sim = f.master_evolve(st, dt, simtime=sim, mode=tca.TC_EVOLVE_JERK_FMM)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    e1 = energy()
    rng = np.random.default_rng(seed)
    tg = torch.from_numpy(np.sort(rng.choice(n, 4096, replace=False)).astype(np.int32)).to(dev)
    _, acc, _ = f.eval(st["pos"], st["vel"], st["mass"])
    c = f.m2l_order_counts()
    work = sum(c[q] * flops.c_m2l(q) for q in range(len(c))) + 44.0 * f.stats()["n_p2p_pairs"]
    _, da, _ = f.direct(st["pos"], st["vel"], st["mass"], tg)
    ea = torch.linalg.norm(acc[tg.long()] - da, dim=1) / torch.linalg.norm(da, dim=1)
    f.close()
    return {"ms_per_step": ms_step, "evals_per_step": 5, "ms_per_eval_first": ev_ms,
            "energy_rel_change": abs(e1 - e0) / abs(e0),
            "final_force_rms": float(torch.sqrt(torch.mean(ea ** 2))),
            "final_force_max": float(ea.max()), "m2l_plus_p2p_flops_final_eval": work,
            "m2l_pairs_by_order": {q: cc for q, cc in enumerate(c) if cc}}


def kroupa_mixed_row(tca, ics, torch, dev, n=1 << 22, seed=3, p=12, theta=0.4):
    """Plummer N with Kroupa IMF masses and 10 % of the bodies in binaries
    (a in [1e-5, 1e-3]): the FP64 near field against the mixed forms (FP64
    pairs inside r = 1e-3, which holds the binaries); error against the FP64
    direct sum on a seeded 4096-body subset."""
    import numpy as np
    x, v, m = ics.kroupa_binaries(n, seed=seed)
    dx, dv, dm = (torch.from_numpy(a).to(dev) for a in (x, v, m))
    out = (torch.empty(n, dtype=torch.float64, device=dev),
           torch.empty(n, 3, dtype=torch.float64, device=dev),
           torch.empty(n, 3, dtype=torch.float64, device=dev))
    rng = np.random.default_rng(seed)
    tg = torch.from_numpy(np.sort(rng.choice(n, 4096, replace=False)).astype(np.int32)).to(dev)
    f0 = tca.Fmm(p=4, direct_below=0)
    _, da, dj = f0.direct(dx, dv, dm, tg)
    f0.close()
    row = {"what": "Plummer N=%d, Kroupa IMF, 10%% of the bodies in binaries (a 1e-5..1e-3), "
                   "p=%d, theta=%g: FP64 near field vs the mixed forms (FP64 pairs r < 1e-3)"
                   % (n, p, theta)}
    for name, mode in (("fp64", tca.TC_P2P_FP64), ("fp32_offsets", tca.TC_P2P_MIXED),
                       ("fp32_hilo_offsets", tca.TC_P2P_MIXED_HILO)):
        f = tca.Fmm(p=p, ncrit=64, theta=theta, direct_below=0)
        f.set_seed(seed)
        f.set_precision(mode, close_radius=1e-3)
        row[name] = timed_row(f, torch, dx, dv, dm, out, tg.long(), da, dj, reps=2, warm=1)
        f.close()
    return row


def pareto(args):
    """Cost and accuracy of the ways to trade one for the other on the workload
    (Plummer N, seed): fixed (p, theta) on a grid, the error-controlled MAC at
    several force_accuracy values, and the learned per-leaf choice at several
    targets; force error against the FP64 direct sum on a 4096-body subset."""
    import numpy as np
    import torch
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import flops, ics, learn
    x, v, m = ics.plummer(args.n, seed=args.seed)
    dev = torch.device("cuda:0")
    dx, dv, dm = (torch.from_numpy(a).to(dev) for a in (x, v, m))
    out = (torch.empty(args.n, dtype=torch.float64, device=dev),
           torch.empty(args.n, 3, dtype=torch.float64, device=dev),
           torch.empty(args.n, 3, dtype=torch.float64, device=dev))
    rng = np.random.default_rng(12345)
    tg = torch.from_numpy(rng.choice(args.n, 4096, replace=False).astype(np.int32)).to(dev)
    idx = tg.long()
    f0 = tca.Fmm(p=4, ncrit=args.ncrit, direct_below=0)
    _, da, dj = f0.direct(dx, dv, dm, tg)
    f0.close()

    def emit(kind, f, **kw):
        r = timed_row(f, torch, dx, dv, dm, out, idx, da, dj, reps=3, warm=1)
        c = f.m2l_order_counts()
        r["m2l_flops"] = float(sum(c[q] * flops.c_m2l(q) for q in range(len(c))))
        r["p2p_flops"] = 44.0 * r["p2p_body_pairs"]
        r.update(kind=kind, **kw)
        r["meets_1e-7"] = r["force_rms"] <= 1e-7
        print(json.dumps(r), flush=True)

    for p in (6, 8, 10, 12, 14):
        for th in (0.3, 0.4, 0.5, 0.6, 0.7):
            f = tca.Fmm(p=p, ncrit=args.ncrit, theta=th, direct_below=0)
            f.set_seed(args.seed)
            emit("fixed", f, p=p, theta=th)
            f.close()
    for facc in (1e-7, 3e-8, 1e-8, 3e-9):
        for p in (10, 12):
            f = tca.Fmm(p=p, ncrit=args.ncrit, mac=tca.TC_MAC_ERROR, force_accuracy=facc,
                        direct_below=0)
            f.set_seed(args.seed)
            emit("error_mac", f, p=p, force_accuracy=facc)
            f.close()
    if os.path.exists(learn.MODEL_PATH):
        model = learn.load_model()
        for pc in (12, 14):
            for target in (2e-7, 1e-7, 5e-8):
                f = tca.Fmm(p=pc, ncrit=args.ncrit, direct_below=0)
                f.set_seed(args.seed)
                learn.apply_model(f, model, target_rms=target)
                emit("learned", f, target_rms=target, p_max=pc)
                f.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--p", type=int, default=12)
    ap.add_argument("--theta", type=float, default=0.4)
    ap.add_argument("--ncrit", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-timeout", type=float, default=240.0)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default near-field launch shape (no tc_tune_p2p after the "
                         "warm-up)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the default-path and host-copy rows")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "m2l_traffic.json"))
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size once (a plain copy kernel "
                         "of known size for scaling HBM byte counters)")
    ap.add_argument("--only-master-evolve", action="store_true",
                    help="time only the master step loop row (for profiling it) and print it")
    ap.add_argument("--only", choices=["clusters_c4", "cusp_steps_c5", "kroupa_binaries_c3",
                                       "learned_c2", "gate_c2"],
                    help="time only this row of the other workloads and print it (for "
                         "profiling it)")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the rows of the other workloads (many clusters, the "
                         "fourth-order steps on the cusp, the learned choice)")
    ap.add_argument("--c5-n", type=int, default=1 << 23,
                    help="bodies of the fourth-order-steps row (Dehnen gamma=1.5)")
    ap.add_argument("--pareto", action="store_true",
                    help="instead of the benchmark line: time and error of a grid of fixed "
                         "(p, theta), of the error-controlled MAC and of the learned choice "
                         "on the workload; one JSON line per configuration")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (the evaluation "
              "is not sharded across GPUs)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.only_master_evolve:
        import torch
        import taichi_collisional_amd as tca
        from taichi_collisional_amd import ics
        print(json.dumps({"master_evolve": master_evolve_row(
            tca, ics, torch, torch.device("cuda:0"), args.p, args.seed)}), flush=True)
        return

    if args.pareto:
        pareto(args)
        return
    if args.only:
        import numpy as np
        import torch
        import taichi_collisional_amd as tca
        from taichi_collisional_amd import ics
        dev = torch.device("cuda:0")
        if args.only == "clusters_c4":
            row = clusters_row(tca, ics, torch, dev)
        elif args.only == "kroupa_binaries_c3":
            row = kroupa_mixed_row(tca, ics, torch, dev)
        elif args.only == "gate_c2":
            x, v, m = ics.plummer(args.n, seed=args.seed)
            dx, dv, dm = (torch.from_numpy(a).to(dev) for a in (x, v, m))
            rng = np.random.default_rng(12345)
            tg = torch.from_numpy(rng.choice(args.n, min(4096, args.n),
                                             replace=False).astype(np.int32)).to(dev)
            f0 = tca.Fmm(p=4, direct_below=0)
            _, da, _ = f0.direct(dx, dv, dm, tg)
            f0.close()
            row = gate_row(tca, torch, dx, dv, dm, tg.long(), da, args.p, args.theta, args.ncrit)
        elif args.only == "learned_c2":
            from taichi_collisional_amd import learn
            x, v, m = ics.plummer(args.n, seed=args.seed)
            dx, dv, dm = (torch.from_numpy(a).to(dev) for a in (x, v, m))
            f = tca.Fmm(p=args.p, ncrit=args.ncrit, direct_below=0)
            f.set_seed(args.seed)
            learn.apply_model(f, learn.load_model(), target_rms=1e-7)
            for _ in range(4):
                f.eval(dx, dv, dm)
            torch.cuda.synchronize()
            st = f.stats()
            row = {k: v for k, v in st.items() if k.startswith("ms_") or k.startswith("n_")}
            f.close()
        else:
            row = cusp_steps_row(tca, ics, torch, dev, args.c5_n, steps=2)
        print(json.dumps({args.only: row}), flush=True)
        return

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.n, args.p, args.theta, args.ncrit, args.seed, args.cpu_threads,
                           args.cpu_timeout, repeats=args.cpu_repeats,
                           default_path=not args.no_extra)

    import numpy as np
    import torch
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import flops, ics

    x, v, m = ics.plummer(args.n, seed=args.seed)
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    dv = torch.from_numpy(v).to(dev)
    dm = torch.from_numpy(m).to(dev)
    fmm = tca.Fmm(p=args.p, ncrit=args.ncrit, theta=args.theta, direct_below=0)
    fmm.set_seed(args.seed)
    pot = torch.empty(args.n, dtype=torch.float64, device=dev)
    acc = torch.empty(args.n, 3, dtype=torch.float64, device=dev)
    jerk = torch.empty(args.n, 3, dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
    tuning = None
    if not args.no_tune and args.warmup > 0:
        # launch shape of the near field, timed on this workload's lists (not
        # part of the timed evaluations, which each do the whole work)
        shape, best, dflt = fmm.tune_p2p(reps=2)
        tuning = {"shape": {"tile": shape[0], "list_slice": shape[1], "stage_rounds": shape[2],
                            "targets_per_lane": shape[3]},
                  "p2p_ms_tuned": best, "p2p_ms_default_shape": dflt}
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
    torch.cuda.synchronize()
    stats = []
    shifts = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
        stats.append(fmm.stats())
        shifts.append(fmm.last_shift())  # host copy of the three values drawn (no GPU work)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = args.n / (ms * 1e-3)

    st = stats[-1]
    n_edges = st["ncells"] - 1
    fl = flops.eval_flops(args.p, args.n, st["n_p2p_pairs"], st["n_m2l"], n_edges)
    m2l_ms = sum(s["ms_m2l"] for s in stats) / len(stats)
    m2l_tflops = fl["M2L"] / (m2l_ms * 1e-3) / 1e12
    phases = {k[3:]: round(sum(s[k] for s in stats) / len(stats), 3)
              for k in ("ms_tree", "ms_upward", "ms_traverse", "ms_lists", "ms_m2l", "ms_p2p",
                        "ms_downward", "ms_total")}

    # accuracy vs the FP64 direct sum on a seeded 4096-body subset (not timed),
    # for every timed evaluation: each drew its own root shift (the reference's
    # three drand48() values), which moves the cell boundaries and so the
    # error; the evaluations are repeated with those shifts
    rng = np.random.default_rng(12345)
    tg = torch.from_numpy(rng.choice(args.n, min(4096, args.n), replace=False).astype(np.int32))
    dp, da, dj = fmm.direct(dx, dv, dm, tg.to(dev))
    idx = tg.long().to(dev)
    per_eval = []
    for sh in shifts:
        fmm.set_shift(sh)
        fmm.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
        ea = torch.linalg.norm(acc[idx] - da, dim=1) / torch.linalg.norm(da, dim=1)
        ej = torch.linalg.norm(jerk[idx] - dj, dim=1) / torch.linalg.norm(dj, dim=1)
        per_eval.append((float(torch.sqrt(torch.mean(ea ** 2))), float(ea.max()),
                         float(torch.sqrt(torch.mean(ej ** 2))), float(ej.max())))
    pe = np.array(per_eval)
    err = {"force_rms": float(np.median(pe[:, 0])), "force_rms_max": float(pe[:, 0].max()),
           "force_max": float(pe[:, 1].max()), "jerk_rms": float(np.median(pe[:, 2])),
           "jerk_max": float(pe[:, 3].max()), "subset": int(tg.numel()),
           "force_rms_per_eval": [float(q) for q in pe[:, 0]],
           "root_shifts": [[float(c) for c in sh] for sh in shifts],
           "gate_force_rms_1e-7": "%d of %d timed evaluations" % (int(np.sum(pe[:, 0] <= 1e-7)),
                                                                   len(pe)),
           "note": "force_rms and jerk_rms are medians over the timed evaluations, the maxima "
                   "over them; the error moves with the root shift (at this p and theta about "
                   "one shift in twelve gives rms above 1e-7)"}

    extra = {}
    if not args.no_extra:
        # the reference's default path on the GPU: p=2 pre-pass + error-controlled MAC
        fd = tca.Fmm(p=args.p, ncrit=args.ncrit, theta=args.theta, mac=tca.TC_MAC_ERROR,
                     force_accuracy=2e-8, direct_below=0)
        fd.set_seed(args.seed)
        for _ in range(2):  # warm-up: buffers settle over the first evaluations
            fd.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            fd.eval(dx, dv, dm, pot=pot, acc=acc, jerk=jerk)
        torch.cuda.synchronize()
        dms = (time.perf_counter() - t) * 1e3 / 5
        sd = fd.stats()
        ea = torch.linalg.norm(acc[idx] - da, dim=1) / torch.linalg.norm(da, dim=1)
        extra["default_path"] = {
            "what": "reference default pass: p=2 pre-pass (acc_old), error-controlled MAC "
                    "(force_accuracy=2e-8), same bodies and p",
            "ms_per_eval": dms, "particles_per_s": args.n / (dms * 1e-3),
            "force_rms": float(torch.sqrt(torch.mean(ea ** 2))), "force_max": float(ea.max()),
            "m2l_pairs": sd["n_m2l"], "p2p_leaf_pairs": sd["n_p2p_cells"],
            "ms_prepass": sd["ms_prepass"]}
        if cpu and cpu.get("default_path_seconds"):
            extra["default_path"]["cpu_reference_seconds"] = cpu["default_path_seconds"]
            extra["default_path"]["gpu_over_cpu"] = cpu["default_path_seconds"] / (dms * 1e-3)
        del fd
        # the mixed-precision near field: FP32 body pairs on group-relative
        # coordinates (one float each, or hi + lo), FP64 inside r < 1e-3
        rows = {}
        for name, mode in (("fp32_offsets", tca.TC_P2P_MIXED),
                           ("fp32_hilo_offsets", tca.TC_P2P_MIXED_HILO)):
            fx = tca.Fmm(p=args.p, ncrit=args.ncrit, theta=args.theta, direct_below=0)
            fx.set_seed(args.seed)
            fx.set_precision(mode, close_radius=1e-3)
            rows[name] = timed_row(fx, torch, dx, dv, dm, (pot, acc, jerk), idx, da, dj)
            fx.close()
        extra["mixed_p2p"] = {
            "what": "same workload with the P2P near field in FP32 (pairs closer than r=1e-3 "
                    "in FP64, FP32 sums per LDS tile, FP64 across tiles); M2L/L2L/M2M in FP64",
            "fp64_ms_p2p": phases["p2p"], **rows}
        # host Body records in and out (AoS <-> device SoA, PCIe both ways)
        bodies = np.zeros(args.n, tca.body_dtype)
        bodies["X"], bodies["V"], bodies["q"] = x, v, m
        bodies["issink"], bodies["issource"] = 1, 1
        fmm.get_force_and_potential(bodies)
        t = time.perf_counter()
        for _ in range(2):
            fmm.get_force_and_potential(bodies)
        hms = (time.perf_counter() - t) * 1e3 / 2
        extra["host_path"] = {"what": "tc_eval_bodies on host Body records: host AoS to device "
                                      "SoA, evaluation, results back (PCIe both ways)",
                              "ms_per_eval": hms, "particles_per_s": args.n / (hms * 1e-3)}

    if not args.no_extra:
        # the master variant's default build: direct over leaf pairs with the
        # time-step estimate (N = 2^16, ncrit 300, dt_param 0.025)
        nm = 1 << 16
        xm, vm, mm = ics.plummer(nm, seed=args.seed)
        dxm, dvm, dmm = (torch.from_numpy(a).to(dev) for a in (xm, vm, mm))
        fm = tca.Fmm(p=args.p, ncrit=300, direct_below=0)
        fm.master_eval(dxm, dvm, dmm, get_steps=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            fm.master_eval(dxm, dvm, dmm, get_steps=True)
        torch.cuda.synchronize()
        mms = (time.perf_counter() - t) * 1e3 / 3
        row = {"what": "master default build: direct() over every leaf pair, P2P with the "
                       "per-leaf time-step estimate, N=65536, ncrit=300",
               "ms_per_eval": mms, "body_pairs_per_s": nm * nm / (mms * 1e-3),
               "ms_kernel": fm.stats()["ms_p2p"]}
        del fm
        if cpu is not None:
            row.update(master_cpu_baseline(nm, args.seed, args.cpu_threads, args.cpu_timeout))
            if row.get("cpu_reference_seconds"):
                row["gpu_over_cpu"] = row["cpu_reference_seconds"] / (mms * 1e-3)
        extra["master_direct"] = row
        # the master step loop: one do_evolve (GRADIENT scheme) on N = 4096
        row = master_evolve_row(tca, ics, torch, dev, args.p, args.seed)
        if cpu is not None:
            row.update(master_evolve_cpu_baseline(4096, 1.0 / 1024, args.seed, args.cpu_threads,
                                                  args.cpu_timeout))
            if row.get("cpu_reference_seconds"):
                row["gpu_over_cpu"] = row["cpu_reference_seconds"] / (row["ms_per_step"] * 1e-3)
        extra["master_evolve"] = row

    if not args.no_extra and not args.no_workloads:
        from taichi_collisional_amd import learn
        extra["clusters_c4"] = clusters_row(tca, ics, torch, dev)
        row = {"what": "Dehnen gamma=1.5 N=%d, 10 fourth-order steps dt=1e-4 (GRADIENT loop, "
                       "jerk-build theta-MAC FMM forces, 5 evaluations per step)" % args.c5_n,
               "fallback_p12_theta04": cusp_steps_row(tca, ics, torch, dev, args.c5_n)}
        if os.path.exists(learn.MODEL_PATH):
            row["learned"] = cusp_steps_row(tca, ics, torch, dev, args.c5_n, learned=True)
            row["learned_speedup"] = (row["fallback_p12_theta04"]["ms_per_step"] /
                                      row["learned"]["ms_per_step"])
        extra["cusp_steps_c5"] = row
        extra["kroupa_binaries_c3"] = kroupa_mixed_row(tca, ics, torch, dev)
        if os.path.exists(learn.MODEL_PATH):
            flr = tca.Fmm(p=args.p, ncrit=args.ncrit, direct_below=0)
            flr.set_seed(args.seed)
            learn.apply_model(flr, learn.load_model(), target_rms=1e-7)
            r = timed_row(flr, torch, dx, dv, dm, (pot, acc, jerk), idx, da, dj)
            c = flr.m2l_order_counts()
            r["m2l_flops"] = float(sum(c[q] * flops.c_m2l(q) for q in range(len(c))))
            r["speedup_vs_fixed_p12"] = ms / r["ms_per_eval"]
            r["what"] = ("same workload, learned per-leaf opening angle at order %d, target rms "
                         "1e-7 (FP64)" % args.p)
            extra["learned_c2"] = r
            flr.close()
        extra["gate_c2"] = gate_row(tca, torch, dx, dv, dm, idx, da, args.p, args.theta,
                                    args.ncrit)

    if args.calibrate_copy_mib > 0:
        src = torch.ones(args.calibrate_copy_mib << 17, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
    traffic, hw_flops = load_traffic(args.traffic_file)
    line = {
        "metric": METRIC, "value": value, "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Plummer N=%d seed=%d (Henon units), p=%d, theta=%g, ncrit=%d, "
                               "FP64 force+potential+jerk" % (args.n, args.seed, args.p,
                                                             args.theta, args.ncrit),
                   "n": args.n, "p": args.p, "theta": args.theta, "ncrit": args.ncrit,
                   "mac": "theta", "precision": "fp64"},
        "roofline": {"bound": "fp64-valu", "achieved": m2l_tflops, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": m2l_tflops / PEAK_FP64_TFLOPS,
                     "traffic": traffic, "kernel": "k_m2l (M2L, FP64)",
                     "flops_per_launch": fl["M2L"], "ms_per_launch": m2l_ms,
                     "hw_fp64_flops_per_launch": hw_flops,
                     "hw_frac": (hw_flops / (m2l_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
                                 if hw_flops else None),
                     "note": "k_m2l issues no MFMA: it is bound by FP64 VALU issue "
                             "(v_fma_f64; 78.6 TFLOP/s vector peak). frac counts the "
                             "reference's algorithmic flops; hw_frac the FP64 flops the "
                             "kernel executes (SQ_INSTS_VALU_{FMA,MUL,ADD}_F64 counters)"},
        "cpu_baseline": cpu,
        "near_field": {"kernel": "k_p2p (P2P, FP64)", "ms_per_launch": phases["p2p"],
                       "flops_per_launch": 44.0 * st["n_p2p_pairs"],
                       "achieved": 44.0 * st["n_p2p_pairs"] / (phases["p2p"] * 1e-3) / 1e12,
                       "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                       "frac": 44.0 * st["n_p2p_pairs"] / (phases["p2p"] * 1e-3) / 1e12
                       / PEAK_FP64_TFLOPS,
                       "note": "44 flops per body pair (the reference's count); the time is "
                               "the launch's own, on the second stream beside the downward "
                               "pass", "tuning": tuning},
        "error_vs_direct": err,
        "phases_ms": phases,
        "counts": {"m2l_pairs": st["n_m2l"], "p2p_leaf_pairs": st["n_p2p_cells"],
                   "p2p_body_pairs": st["n_p2p_pairs"], "cells": st["ncells"],
                   "leaves": st["nleaves"], "depth": st["depth"]},
        "algorithmic_tflops_per_step": fl["total"] / 1e12,
        "achieved_tflops_whole_step": fl["total"] / (ms * 1e-3) / 1e12,
    }
    if cpu and cpu.get("value"):
        line["gpu_over_cpu"] = value / cpu["value"]
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
