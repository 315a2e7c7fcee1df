"""Pin each oracle kernel: against the reference's own kernels (golden KATs)
and against what the mathematics fixes (translation exactness, the two M2L
formulations, brute force, finite differences)."""
import numpy as np
import pytest

from conftest import golden

ORDERS = (4, 5, 8, 10, 12, 16, 20, 24)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))




@pytest.mark.parametrize("p", ORDERS)
def test_p2m_matches_reference(oracle_lib, p):
    g = golden("kat_p%d.npz" % p)
    L = oracle_lib.lib()
    nt = (p + 1) ** 2
    for k in range(3):
        x, v, m = g["p2m%d_x" % k], g["p2m%d_v" % k], g["p2m%d_m" % k]
        X, R, M, Md = np.zeros(3), np.zeros(1), np.zeros(nt), np.zeros(nt)
        L.orc_k_p2m(p, m.size, x.reshape(-1).copy(), v.reshape(-1).copy(), m, X, R, M, Md)
        # geometry: smallest enclosing ball (independent algorithm) to round-off
        assert np.allclose(X, g["p2m%d_X" % k], rtol=0, atol=1e-13)
        assert abs(R[0] - g["p2m%d_R" % k][0]) < 1e-13
        assert _rel(M, g["p2m%d_M" % k]) < 1e-12
        assert _rel(Md, g["p2m%d_Md" % k]) < 1e-12


@pytest.mark.parametrize("p", ORDERS)
def test_m2l_matches_reference(oracle_lib, p):
    g = golden("kat_p%d.npz" % p)
    L = oracle_lib.lib()
    nt = (p + 1) ** 2
    for k in range(g["m2l_Xi"].shape[0]):
        args = (g["m2l_Xi"][k], g["m2l_Xj"][k], g["m2l_M"][k], g["m2l_Md"][k])
        # the rotation form's own round-off on this input: rotate against the
        # O(p^4) form without the factorial quirk (mathematically equal); it
        # grows steeply with p (7e-14 at p = 12, 7e-11 at p = 24) through the
        # (n+|m|)!(n-|m|)! scaling, so two correct rotation codes may differ
        # by that much
        Lr, Ldr, Ld4, Ldd4 = (np.zeros(nt) for _ in range(4))
        L.orc_k_m2l(p, 0, 1, *args, Lr, Ldr)
        L.orc_k_m2l(p, 0, 0, *args, Ld4, Ldd4)
        rot_tol = max(1e-12, 2.0 * max(_rel(Lr, Ld4), _rel(Ldr, Ldd4)))
        for rot, key in ((1, "m2l"), (0, "m2l4")):
            Lo, Ldo = np.zeros(nt), np.zeros(nt)
            L.orc_k_m2l(p, 1, rot, *args, Lo, Ldo)
            tol = rot_tol if rot else 1e-12
            assert _rel(Lo, g[key + "_L"][k]) < tol, (rot, k)
            assert _rel(Ldo, g[key + "_Ld"][k]) < tol, (rot, k)


@pytest.mark.parametrize("p", ORDERS)
def test_m2l_rotate_equals_direct_form(oracle_lib, p):
    """The rotation-based M2L equals the O(p^4) M2L to round-off (quirk off);
    with the reference's factorial quirk only the order-0 term differs."""
    L = oracle_lib.lib()
    nt = (p + 1) ** 2
    rng = np.random.default_rng(p)
    for _ in range(6):
        Xi = rng.normal(size=3)
        Xj = Xi + rng.normal(size=3) * 3
        M, Md = rng.normal(size=nt), rng.normal(size=nt)
        a, ad, b, bd, q, qd = (np.zeros(nt) for _ in range(6))
        L.orc_k_m2l(p, 0, 1, Xi, Xj, M, Md, a, ad)
        L.orc_k_m2l(p, 0, 0, Xi, Xj, M, Md, b, bd)
        assert _rel(a, b) < 1e-12 and _rel(ad, bd) < 1e-12
        L.orc_k_m2l(p, 1, 1, Xi, Xj, M, Md, q, qd)
        assert np.allclose(q[1:], a[1:], rtol=1e-12, atol=1e-12 * np.abs(a).max())
        assert q[0] != a[0]


@pytest.mark.parametrize("p", ORDERS)
def test_m2m_matches_reference_and_is_exact(oracle_lib, p):
    g = golden("kat_p%d.npz" % p)
    L = oracle_lib.lib()
    nt = (p + 1) ** 2
    nch = g["m2m_Xc"].shape[0]
    M, Md = np.zeros(nt), np.zeros(nt)
    L.orc_k_m2m(p, nch, g["m2m_Xc"].reshape(-1).copy(), g["m2m_Mc"].reshape(-1).copy(),
                g["m2m_Mdc"].reshape(-1).copy(), g["m2m_Xp"], M, Md)
    assert _rel(M, g["m2m_M"]) < 1e-12 and _rel(Md, g["m2m_Md"]) < 1e-12
    # parent ball of the leaf children's bodies
    c, r2 = np.zeros(3), np.zeros(1)
    L.orc_minball(g["m2m_bx"].shape[0], g["m2m_bx"].reshape(-1).copy(), c, r2)
    assert np.allclose(c, g["m2m_Xp"], atol=1e-13)
    assert abs(np.sqrt(r2[0]) + 1e-6 - g["m2m_Rp"][0]) < 1e-13
    # exactness: moments about the parent == translated moments about the children
    rng = np.random.default_rng(p + 50)
    x = rng.normal(size=(30, 3)) * 0.2
    v = rng.normal(size=(30, 3))
    m = rng.uniform(0.1, 1.0, 30)
    Xc = np.array([0.05, -0.02, 0.03])
    Xp = np.array([-0.1, 0.2, 0.15])
    Mc, Mdc, Mp, Mdp = (np.zeros(nt) for _ in range(4))
    L.orc_k_p2m_at(p, 30, x.reshape(-1).copy(), v.reshape(-1).copy(), m, Xc, Mc, Mdc)
    L.orc_k_p2m_at(p, 30, x.reshape(-1).copy(), v.reshape(-1).copy(), m, Xp, Mp, Mdp)
    Mt, Mdt = np.zeros(nt), np.zeros(nt)
    L.orc_k_m2m(p, 1, Xc, Mc, Mdc, Xp, Mt, Mdt)
    assert _rel(Mt, Mp) < 1e-12 and _rel(Mdt, Mdp) < 1e-12


@pytest.mark.parametrize("p", ORDERS)
def test_l2l_l2p_match_reference_and_are_exact(oracle_lib, p):
    g = golden("kat_p%d.npz" % p)
    L = oracle_lib.lib()
    nt = (p + 1) ** 2
    Lc, Ldc = np.zeros(nt), np.zeros(nt)
    L.orc_k_l2l(p, g["l2l_Xp"], g["l2l_L"], g["l2l_Ld"], g["l2l_Xc"], Lc, Ldc)
    assert _rel(Lc, g["l2l_Lc"]) < 1e-12 and _rel(Ldc, g["l2l_Ldc"]) < 1e-12
    nb = g["l2p_x"].shape[0]
    pot, acc, jerk = np.zeros(nb), np.zeros(3 * nb), np.zeros(3 * nb)
    L.orc_k_l2p(p, g["l2p_X"], g["l2l_L"], g["l2l_Ld"], nb, g["l2p_x"].reshape(-1).copy(),
                g["l2p_v"].reshape(-1).copy(), pot, acc, jerk)
    assert _rel(pot, g["l2p_pot"]) < 1e-12
    assert _rel(acc.reshape(nb, 3), g["l2p_acc"]) < 1e-12
    assert _rel(jerk.reshape(nb, 3), g["l2p_jerk"]) < 1e-12
    # exactness: a local expansion is a polynomial, so evaluating it directly or
    # after L2L to a child centre gives the same field (up to terms of order p
    # that L2P at the child drops for n <= 1, which vanish for n + k > p)
    x = g["l2l_Xc"] + np.random.default_rng(3).normal(size=(5, 3)) * 0.05
    v = np.random.default_rng(4).normal(size=(5, 3))
    p1, a1, j1 = np.zeros(5), np.zeros(15), np.zeros(15)
    p2, a2, j2 = np.zeros(5), np.zeros(15), np.zeros(15)
    L.orc_k_l2p(p, g["l2l_Xp"], g["l2l_L"], g["l2l_Ld"], 5, x.reshape(-1).copy(),
                v.reshape(-1).copy(), p1, a1, j1)
    L.orc_k_l2p(p, g["l2l_Xc"], Lc, Ldc, 5, x.reshape(-1).copy(), v.reshape(-1).copy(), p2, a2, j2)
    assert _rel(p2, p1) < 1e-11 and _rel(a2, a1) < 1e-11 and _rel(j2, j1) < 1e-11


@pytest.mark.parametrize("p", ORDERS)
def test_p2p_matches_reference_and_brute_force(oracle_lib, p):
    g = golden("kat_p%d.npz" % p)
    L = oracle_lib.lib()
    xi, vi, xj, vj, mj = (g["p2p_" + k] for k in ("xi", "vi", "xj", "vj", "mj"))
    ni, nj = xi.shape[0], xj.shape[0]
    pot, acc, jerk = np.zeros(ni), np.zeros(3 * ni), np.zeros(3 * ni)
    L.orc_k_p2p(ni, xi.reshape(-1).copy(), vi.reshape(-1).copy(), nj, xj.reshape(-1).copy(),
                vj.reshape(-1).copy(), mj, pot, acc, jerk)
    assert _rel(pot, g["p2p_pot"]) < 1e-14
    assert _rel(acc.reshape(ni, 3), g["p2p_acc"]) < 1e-14
    assert _rel(jerk.reshape(ni, 3), g["p2p_jerk"]) < 1e-14
    # brute force in long double (independent formula, coincident pair skipped)
    d = xj[None, :, :].astype(np.longdouble) - xi[:, None, :]
    dv = vj[None, :, :].astype(np.longdouble) - vi[:, None, :]
    r2 = (d ** 2).sum(-1)
    ok = r2 > 0
    r2s = np.where(ok, r2, 1)
    ir = np.where(ok, 1 / np.sqrt(r2s), 0)
    mm = mj[None, :] * ir
    bp = mm.sum(1)
    ba = (d * (mm * ir ** 2)[..., None]).sum(1)
    rv = (d * dv).sum(-1)
    bj = (dv * (mm * ir ** 2)[..., None] - 3 * d * (mm * ir ** 4 * rv)[..., None]).sum(1)
    assert _rel(pot, bp.astype(float)) < 1e-14
    assert _rel(acc.reshape(ni, 3), ba.astype(float)) < 1e-14
    assert _rel(jerk.reshape(ni, 3), bj.astype(float)) < 1e-13


def test_far_field_chain_vs_direct_and_jerk(oracle_lib):
    """L2P(M2L(P2M)) approximates the direct sum to ~theta^(p+1); its jerk is
    the time derivative of its acceleration (central difference, fixed centres)."""
    L = oracle_lib.lib()
    p = 12
    nt = (p + 1) ** 2
    rng = np.random.default_rng(9)
    xs = rng.normal(size=(40, 3)) * 0.1
    vs = rng.normal(size=(40, 3))
    ms = rng.uniform(0.1, 1, 40)
    Xj = np.zeros(3)
    Xi = np.array([1.5, -1.0, 0.8])
    xt = Xi + rng.normal(size=(6, 3)) * 0.1
    vt = rng.normal(size=(6, 3))

    def far(xs_):
        M, Md, Lc, Ldc = (np.zeros(nt) for _ in range(4))
        L.orc_k_p2m_at(p, 40, xs_.reshape(-1).copy(), vs.reshape(-1).copy(), ms, Xj, M, Md)
        L.orc_k_m2l(p, 0, 1, Xi, Xj, M, Md, Lc, Ldc)
        return Lc, Ldc

    Lc, Ldc = far(xs)
    pot, acc, jerk = np.zeros(6), np.zeros(18), np.zeros(18)
    L.orc_k_l2p(p, Xi, Lc, Ldc, 6, xt.reshape(-1).copy(), vt.reshape(-1).copy(), pot, acc, jerk)
    dp, da, dj = np.zeros(6), np.zeros(18), np.zeros(18)
    L.orc_k_p2p(6, xt.reshape(-1).copy(), vt.reshape(-1).copy(), 40, xs.reshape(-1).copy(),
                vs.reshape(-1).copy(), ms, dp, da, dj)
    assert _rel(acc, da) < 1e-8 and _rel(jerk, dj) < 1e-6
    h = 1e-6
    accs = []
    for sgn in (1, -1):
        Lh, Ldh = far(xs + sgn * h * vs)
        pp, aa, jj = np.zeros(6), np.zeros(18), np.zeros(18)
        L.orc_k_l2p(p, Xi, Lh, Ldh, 6, (xt + sgn * h * vt).reshape(-1).copy(),
                    vt.reshape(-1).copy(), pp, aa, jj)
        accs.append(aa)
    fd = (accs[0] - accs[1]) / (2 * h)
    assert _rel(jerk, fd) < 1e-6
