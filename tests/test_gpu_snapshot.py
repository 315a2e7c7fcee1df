"""The snapshot workflow of plot_jerk_error.py on the GPU path: one evaluation
of the HIP FMM and one FP64 direct sum on the GPU, each written as a
snapshot (libtcio), compared through the files.

Plummer N = 16,384 (seed 21, root shift from srand48(21)), p = 10,
theta = 0.5 (the C1 expansion order and MAC).  The oracle with the same root
shift, against an FP64 direct sum on the CPU, has relative acceleration
errors of median 5.81e-7 and maximum 1.29e-3 on these bodies; the GPU FMM
reproduces the oracle to round-off, so its CDF read from the files must show
the same two figures (to 2 %).  Every other dataset of the two files
(positions, velocities, masses, ids) is identical."""
import numpy as np
import pytest

from conftest import relerr_vec

pytestmark = pytest.mark.gpu


def test_fmm_and_direct_snapshots(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import taichi_collisional_amd as tca
    from taichi_collisional_amd import ics
    try:
        from taichi_collisional_amd import snapshot as snap
        snap.lib()
    except (RuntimeError, OSError) as e:
        pytest.skip("snapshot library unavailable: %s" % e)
    n = 16384
    x, v, m = ics.plummer(n, seed=21)
    print("snapshot test: plummer n=%d seed=21 p=10 theta=0.5" % n, flush=True)
    f = tca.Fmm(p=10, ncrit=64, theta=0.5, direct_below=0)
    f.set_seed(21)
    dx, dv, dm = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, v, m))
    pot, acc, jerk = f.eval(dx, dv, dm)
    tg = torch.arange(n, dtype=torch.int32, device="cuda")
    dpot, dacc, djerk = f.direct(dx, dv, dm, tg)
    torch.cuda.synchronize()
    f.close()
    rng = np.random.default_rng(5)
    ids = rng.permutation(n).astype(np.uint32)
    paths = []
    for tag, (p_, a_, j_) in (("", (pot, acc, jerk)), ("_ref", (dpot, dacc, djerk))):
        part = np.zeros(n, dtype=tca.particle_dtype)
        part["pos"], part["vel"], part["mass"], part["id"] = x, v, m, ids
        part["pot"], part["acc"], part["jerk"] = (t.cpu().numpy() for t in (p_, a_, j_))
        paths.append(tmp_path / ("snapshot_0%s.hdf5" % tag))
        snap.write_snapshot(paths[-1], part, 0.0)
    ea, frac = snap.error_cdf(paths[0], paths[1], "acc")
    ej, _ = snap.error_cdf(paths[0], paths[1], "jerk")
    print("acc error: median %.2e max %.2e; jerk error: median %.2e max %.2e"
          % (np.median(ea), ea[-1], np.median(ej), ej[-1]), flush=True)
    assert ea.size == n and frac[0] == 1.0
    assert abs(np.median(ea) / 5.81e-7 - 1) < 0.02 and abs(ea[-1] / 1.29e-3 - 1) < 0.02
    # the files agree with the arrays they were written from
    order = np.argsort(ids)
    a_fmm = acc.cpu().numpy()[order]
    a_dir = dacc.cpu().numpy()[order]
    back = np.stack([snap.read_dataset(paths[0], "Acc" + c) for c in "xyz"], axis=1)
    np.testing.assert_array_equal(back, a_fmm)
    assert np.isclose(ea[-1], relerr_vec(a_fmm, a_dir)[1], rtol=1e-12)
    for name in ("Posx", "Velz", "Mass"):
        np.testing.assert_array_equal(snap.read_dataset(paths[0], name),
                                      snap.read_dataset(paths[1], name))
