"""ctypes access to the oracle (liboracle.so) and the compiled reference
(_ref/libtcref_p<P>_<isa>.so).  TEST INFRASTRUCTURE ONLY -- imported by
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the
product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class OrcParams(C.Structure):
    _fields_ = [("p", C.c_int), ("ncrit", C.c_int), ("theta", C.c_double), ("mac", C.c_int),
                ("force_accuracy", C.c_double), ("quirk", C.c_int),
                ("shift_frac", C.c_double * 3), ("direct_below", C.c_int),
                ("nthreads", C.c_int)]


class OrcStats(C.Structure):
    _fields_ = [("n_m2l", C.c_longlong), ("n_p2p_cells", C.c_longlong),
                ("n_p2p_pairs", C.c_longlong), ("ncells", C.c_int), ("nleaves", C.c_int),
                ("depth", C.c_int)]


def _opt(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_eval.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_void_p,
                               _dp, _dp, _dp, _dp, C.POINTER(OrcStats)]
        L.orc_cells.argtypes = [C.c_void_p, C.c_int] + [_ip] * 5 + [_dp] * 4 + [_ip]
        L.orc_coefs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_lists.restype = C.c_longlong
        L.orc_lists.argtypes = [C.c_void_p, C.c_int, C.c_longlong, _ip, _ip]
        L.orc_direct.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, C.c_int, C.c_int,
                                 _dp, _dp, _dp]
        L.orc_k_p2m.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_k_p2m_at.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_k_m2l.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_k_m2m.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_k_l2l.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_k_l2p.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_k_p2p.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_minball.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_harmonics.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_swap.argtypes = [C.c_int, _dp]
        L.orc_shift_from_seed.argtypes = [C.c_long, C.c_int, _dp]
        L.orc_master_direct.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_void_p,
                                        C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_master_evolve.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                        C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                        C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.orc_master_fmm.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_void_p,
                                     C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp,
                                     C.POINTER(OrcStats)]
        _lib = L
    return _lib


def shift_from_seed(seed, nskip=0):
    u = np.zeros(3)
    lib().orc_shift_from_seed(seed, nskip, u)
    return u


class Oracle:
    """One FMM configuration of the CPU restatement."""

    def __init__(self, p, ncrit=64, theta=0.5, mac=0, force_accuracy=2e-8, quirk=1,
                 shift_frac=(0.0, 0.0, 0.0), direct_below=0, nthreads=0):
        prm = OrcParams(p, ncrit, theta, mac, force_accuracy, quirk,
                        (C.c_double * 3)(*shift_frac), direct_below, nthreads)
        self.p = p
        self.h = lib().orc_create(C.byref(prm))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def eval(self, x, v, m, issink=None, issource=None):
        n = m.size
        pot, acc, jerk, acco = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
        st = OrcStats()
        s = _opt(issink, np.uint8)
        so = _opt(issource, np.uint8)
        rc = lib().orc_eval(self.h, n, np.ascontiguousarray(x).reshape(-1),
                            np.ascontiguousarray(v).reshape(-1), np.ascontiguousarray(m),
                            None if s is None else s.ctypes.data,
                            None if so is None else so.ctypes.data,
                            pot, acc, jerk, acco, C.byref(st))
        assert rc == 0
        self._n = n
        self.stats = {k: getattr(st, k) for k, _ in OrcStats._fields_}
        return pot, acc.reshape(n, 3), jerk.reshape(n, 3), acco

    def master_direct(self, x, v, m, issink=None, issource=None, steps=True, pair_mass=False,
                      dt_param=0.025):
        """Master default build's get_force_and_potential (direct over leaf pairs).
        Construct the Oracle with shift_frac=(0,0,0): the master tree has no root shift.
        Returns pot, acc, timestep (the per-body sum of leaf-minimum tau^-4)."""
        n = m.size
        pot, acc, ts = np.zeros(n), np.zeros(3 * n), np.zeros(n)
        s = _opt(issink, np.uint8)
        so = _opt(issource, np.uint8)
        rc = lib().orc_master_direct(self.h, n, np.ascontiguousarray(x).reshape(-1),
                                     np.ascontiguousarray(v).reshape(-1), np.ascontiguousarray(m),
                                     None if s is None else s.ctypes.data,
                                     None if so is None else so.ctypes.data, int(steps),
                                     int(pair_mass), dt_param, pot, acc, ts)
        assert rc == 0
        self._n = n
        return pot, acc.reshape(n, 3), ts

    def master_fmm(self, x, v, m, issink=None, issource=None, steps=False, pair_mass=False,
                   dt_param=0.025, acc_old=None, prepass=True):
        """Master get_force_and_potential built with -DFMM (the handle's p, ncrit,
        force_accuracy; construct with shift_frac=(0,0,0)).  acc_old: the caller's
        values the pre-pass adds to (prepass=False: used as they are).
        Returns pot, acc, timestep, acc_old."""
        n = m.size
        pot, acc, ts = np.zeros(n), np.zeros(3 * n), np.zeros(n)
        ao = np.zeros(n) if acc_old is None else np.array(acc_old, dtype=np.float64)
        st = OrcStats()
        s = _opt(issink, np.uint8)
        so = _opt(issource, np.uint8)
        rc = lib().orc_master_fmm(self.h, n, np.ascontiguousarray(x).reshape(-1),
                                  np.ascontiguousarray(v).reshape(-1), np.ascontiguousarray(m),
                                  None if s is None else s.ctypes.data,
                                  None if so is None else so.ctypes.data, int(steps),
                                  int(pair_mass), int(prepass), dt_param, ao, pot, acc, ts,
                                  C.byref(st))
        assert rc == 0
        self._n = n
        self.stats = {k: getattr(st, k) for k, _ in OrcStats._fields_}
        return pot, acc.reshape(n, 3), ts, ao

    def cells(self):
        n = -lib().orc_cells(self.h, 0, *([np.zeros(1, np.int32)] * 5),
                             *([np.zeros(1)] * 4), np.zeros(1, np.int32))
        nb = self._n
        out = dict(level=np.zeros(n, np.int32), nchild=np.zeros(n, np.int32),
                   child=np.zeros(n, np.int32), body=np.zeros(n, np.int32),
                   nbody=np.zeros(n, np.int32), Xg=np.zeros(3 * n), Rg=np.zeros(n),
                   X=np.zeros(3 * n), R=np.zeros(n), order=np.zeros(nb, np.int32))
        lib().orc_cells(self.h, n, out["level"], out["nchild"], out["child"], out["body"],
                        out["nbody"], out["Xg"], out["Rg"], out["X"], out["R"], out["order"])
        out["Xg"] = out["Xg"].reshape(n, 3)
        out["X"] = out["X"].reshape(n, 3)
        return out

    def coefs(self, ncells):
        nt = (self.p + 1) ** 2
        M, Md, L, Ld = (np.zeros(ncells * nt) for _ in range(4))
        lib().orc_coefs(self.h, M, Md, L, Ld)
        return [a.reshape(ncells, nt) for a in (M, Md, L, Ld)]

    def lists(self, which):
        n = -lib().orc_lists(self.h, which, 0, np.zeros(1, np.int32), np.zeros(1, np.int32))
        t, s = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        lib().orc_lists(self.h, which, n, t, s)
        return t[:n], s[:n]


def direct(x, v, m, targets=None, long_double=False, nthreads=0):
    """All-pairs potential/acceleration/jerk for the given targets (default all)."""
    n = m.size
    tg = np.arange(n, dtype=np.int32) if targets is None else np.ascontiguousarray(targets, np.int32)
    k = tg.size
    pot, acc, jerk = np.zeros(k), np.zeros(3 * k), np.zeros(3 * k)
    lib().orc_direct(n, np.ascontiguousarray(x).reshape(-1), np.ascontiguousarray(v).reshape(-1),
                     np.ascontiguousarray(m), k, tg, int(long_double), nthreads, pot, acc, jerk)
    return pot, acc.reshape(k, 3), jerk.reshape(k, 3)


# ----------------------------------------------------------- reference ----
def _cpu_has_avx512():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f in flags for f in ("avx512f", "avx512dq", "avx512bw", "avx512vl"))


_refs = {}


def ref_available(p):
    d = os.path.join(HERE, "_ref")
    return any(os.path.exists(os.path.join(d, "libtcref_p%d_%s.so" % (p, v))) for v in ("v4", "v3"))


def reference(p):
    """ctypes handle on the reference harness built for expansion order p."""
    if p in _refs:
        return _refs[p]
    d = os.path.join(HERE, "_ref")
    isa = "v4" if _cpu_has_avx512() else "v3"
    path = os.path.join(d, "libtcref_p%d_%s.so" % (p, isa))
    if not os.path.exists(path):
        raise RuntimeError("reference harness for p=%d not built (%s)" % (p, path))
    L = C.CDLL(path)
    L.ref_eval.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_void_p, C.c_long, C.c_int,
                           C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                           np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"), _dp]
    L.ref_cells.argtypes = [C.c_int] + [_ip] * 4 + [_dp, _dp, _ip]
    L.ref_k_p2m.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.ref_k_m2l.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.ref_k_m2m.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _ip, _dp, _dp, _dp, _dp]
    L.ref_k_l2l.argtypes = [_dp] * 6
    L.ref_k_l2p.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.ref_k_p2p.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.ref_layout.argtypes = [np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]
    L.isa = isa
    L.path = path
    _refs[p] = L
    return L


def ref_eval(p, x, v, m, seed, ncrit=64, theta=0.5, mode=0, nthreads=0):
    """Run the compiled reference once; returns pot, acc, jerk, acc_old, counts, shift."""
    L = reference(p)
    n = m.size
    pot, acc, jerk, acco = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
    counts = np.zeros(3, np.int64)
    shift = np.zeros(3)
    rc = L.ref_eval(n, np.ascontiguousarray(x).reshape(-1), np.ascontiguousarray(v).reshape(-1),
                    np.ascontiguousarray(m), None, None, seed, ncrit, theta, mode, nthreads,
                    pot, acc, jerk, acco, counts, shift)
    assert rc == 0
    return pot, acc.reshape(n, 3), jerk.reshape(n, 3), acco, counts, shift


def master_evolve(m, pos, vel, nsteps, dt, p=8, ncrit=300, force_accuracy=2e-7, fmm=False,
                  pair_mass=False, dt_param=0.025, simtime=0.0):
    """The oracle's master step loop (do_evolve, GRADIENT) for nsteps steps.
    Returns a dict of pos, vel, pos_e, timestep, acc, pot, acc_old, simtime."""
    n = m.size
    out = dict(pos=np.array(pos, dtype=np.float64).reshape(-1),
               vel=np.array(vel, dtype=np.float64).reshape(-1), pos_e=np.zeros(3 * n),
               timestep=np.zeros(n), acc=np.zeros(3 * n), pot=np.zeros(n), acc_old=np.zeros(n))
    t = C.c_double(simtime)
    rc = lib().orc_master_evolve(n, np.ascontiguousarray(m, dtype=np.float64), out["pos"],
                                 out["pos_e"], out["vel"], out["timestep"], out["acc"],
                                 out["pot"], out["acc_old"], p, ncrit, force_accuracy, int(fmm),
                                 int(pair_mass), dt_param, nsteps, dt, C.byref(t))
    assert rc == 0
    for k in ("pos", "vel", "pos_e", "acc"):
        out[k] = out[k].reshape(n, 3)
    out["simtime"] = t.value
    return out


_mrefs = {}


def master_ref_available(p):
    d = os.path.join(HERE, "_ref")
    return any(os.path.exists(os.path.join(d, "libtcrefm_p%d_%s.so" % (p, v))) for v in ("v4", "v3"))


def master_reference(p):
    """ctypes handle on the master-variant reference harness (ref_master_harness.cpp)."""
    if p in _mrefs:
        return _mrefs[p]
    isa = "v4" if _cpu_has_avx512() else "v3"
    path = os.path.join(HERE, "_ref", "libtcrefm_p%d_%s.so" % (p, isa))
    if not os.path.exists(path):
        raise RuntimeError("master reference harness for p=%d not built (%s)" % (p, path))
    L = C.CDLL(path)
    L.refm_eval.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                            C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    L.refm_kick_steps.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, _dp]
    L.refm_layout.argtypes = [C.c_void_p]
    L.refm_evolve.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                              C.c_double, C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.isa = isa
    _mrefs[p] = L
    return L


def master_ref_layout(p):
    """Sizes and byte offsets of the master Body and Cell as the reference
    compiles them at order p: dict of 'Body.X' ... and 'Cell.M' ... ."""
    out = np.zeros(32, np.int64)
    n = master_reference(p).refm_layout(out.ctypes.data)
    names = ["Body", "Body.X", "Body.q", "Body.F", "Body.p", "Body.V", "Body.acc_old",
             "Body.Morton", "Body.timestep", "Body.index", "Body.issink", "Body.issource",
             "Cell", "Cell.M", "Cell.L", "Cell.Pn", "Cell.NCHILD", "Cell.NBODY", "Cell.CHILD",
             "Cell.BODY", "Cell.X", "Cell.R", "Cell.min_acc", "Cell.cell_mass", "Cell.has_sink"]
    assert n == len(names)
    return dict(zip(names, (int(v) for v in out[:n])))


def master_ref_eval(p, x, v, m, issink=None, issource=None, ncrit=300, dt_param=0.025,
                    force_accuracy=2e-7, mode=0, get_steps=True, nthreads=0, acc_old=None):
    """Run the compiled master reference once (mode 0: direct, 1: FMM
    get_force_and_potential, 2: the FMM passes after the pre-pass with the given
    acc_old); returns pot, acc, timestep, acc_old."""
    L = master_reference(p)
    n = m.size
    pot, acc, ts = np.zeros(n), np.zeros(3 * n), np.zeros(n)
    acco = np.zeros(n) if acc_old is None else np.array(acc_old, dtype=np.float64)
    s = _opt(issink, np.uint8)
    so = _opt(issource, np.uint8)
    rc = L.refm_eval(n, np.ascontiguousarray(x).reshape(-1), np.ascontiguousarray(v).reshape(-1),
                     np.ascontiguousarray(m), None if s is None else s.ctypes.data,
                     None if so is None else so.ctypes.data, ncrit, dt_param, force_accuracy,
                     mode, int(get_steps), nthreads, pot, acc, ts, acco)
    assert rc == 0
    return pot, acc.reshape(n, 3), ts, acco


def master_ref_evolve(p, m, pos, vel, nsteps, dt, ncrit=300, dt_param=0.025,
                      force_accuracy=2e-7, simtime=0.0):
    """nsteps calls of the compiled master do_evolve; outputs by particle id."""
    L = master_reference(p)
    n = m.size
    out = dict(pos=np.array(pos, dtype=np.float64).reshape(-1),
               vel=np.array(vel, dtype=np.float64).reshape(-1), pos_e=np.zeros(3 * n),
               pot=np.zeros(n), timestep=np.zeros(n), acc=np.zeros(3 * n))
    t = C.c_double(simtime)
    rc = L.refm_evolve(n, np.ascontiguousarray(m, dtype=np.float64), out["pos"], out["vel"],
                       out["pos_e"], out["pot"], out["timestep"], out["acc"], ncrit, dt_param,
                       force_accuracy, nsteps, dt, C.byref(t))
    assert rc == 0
    for k in ("pos", "vel", "pos_e", "acc"):
        out[k] = out[k].reshape(n, 3)
    out["simtime"] = t.value
    return out


def ref_cells(p, ncells_max, nbodies):
    L = reference(p)
    nc = ncells_max
    out = dict(nchild=np.zeros(nc, np.int32), child=np.zeros(nc, np.int32),
               body=np.zeros(nc, np.int32), nbody=np.zeros(nc, np.int32), X=np.zeros(3 * nc),
               R=np.zeros(nc), order=np.zeros(nbodies, np.int32))
    n = L.ref_cells(nc, out["nchild"], out["child"], out["body"], out["nbody"], out["X"],
                    out["R"], out["order"])
    if n < 0:
        return ref_cells(p, -n, nbodies)
    for k in ("nchild", "child", "body", "nbody", "R"):
        out[k] = out[k][:n]
    out["X"] = out["X"][:3 * n].reshape(n, 3)
    return out
