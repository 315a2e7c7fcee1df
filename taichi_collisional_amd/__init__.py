"""taichi_collisional_amd -- MI355X (gfx950) FMM force + potential + jerk.

Python side of the C ABI in include/tcfmm.h (libtcfmm.so, built in-tree by
``python -m taichi_collisional_amd.build``).  The interface mirrors the
reference's force path (taichi_with_jerks/integrator.h:308-432):

    fmm = Fmm(p=12, ncrit=64, theta=0.4)          # P/EXPANSION, ncrit, theta
    fmm.set_seed(1)                               # srand48(seed) of taichi.cxx:32
    fmm.get_force_and_potential(bodies)           # host Body records (body_dtype)
    pot, acc, jerk = fmm.eval(x, v, m)            # device tensors (torch, cuda)

The product path is the HIP library only; if it is missing the import of
this module's ``lib()`` raises, there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtcfmm.so")

TC_MAC_THETA = 0
TC_MAC_ERROR = 1
TC_P2P_FP64, TC_P2P_MIXED, TC_P2P_MIXED_HILO = 0, 1, 2

# struct Body of the reference (taichi_with_jerks/exafmm.h:22-46), 152 bytes.
body_dtype = np.dtype({
    "names": ["X", "q", "F", "p", "V", "acc_old", "Morton", "timestep", "J", "index",
              "issink", "issource"],
    "formats": [("<f8", 3), "<f8", ("<f8", 3), "<f8", ("<f8", 3), "<f8", ("<u4", 3), "<f8",
                ("<f8", 3), "<i4", "u1", "u1"],
    "offsets": [0, 24, 32, 56, 64, 88, 96, 112, 120, 144, 148, 149],
    "itemsize": 152,
})

# struct Body of the reference's master variant (exafmm.h:21-48), 128 bytes.
master_body_dtype = np.dtype({
    "names": ["X", "q", "F", "p", "V", "acc_old", "Morton", "timestep", "index", "issink",
              "issource"],
    "formats": [("<f8", 3), "<f8", ("<f8", 3), "<f8", ("<f8", 3), "<f8", ("<u4", 3), "<f8",
                "<i4", "u1", "u1"],
    "offsets": [0, 24, 32, 56, 64, 88, 96, 112, 120, 124, 125],
    "itemsize": 128,
})

# struct particle of the reference (master exafmm.h:73-86), 168 bytes.
particle_dtype = np.dtype({
    "names": ["pos", "pos_e", "mass", "vel", "timestep", "acc", "pot", "jerk", "postime",
              "acc_old", "id"],
    "formats": [("<f8", 3), ("<f8", 3), "<f8", ("<f8", 3), "<f8", ("<f8", 3), "<f8",
                ("<f8", 3), "<f8", "<f8", "<u4"],
    "offsets": [0, 24, 48, 56, 80, 88, 112, 120, 144, 152, 160],
    "itemsize": 168,
})

# mode bits of tc_master_eval_* (include/tcfmm.h)
TC_MASTER_STEPS = 1
TC_MASTER_FMM = 2
TC_MASTER_PAIR_MASS = 4
TC_EVOLVE_JERK_FMM = 8


class TcConfig(C.Structure):
    _fields_ = [("p", C.c_int), ("ncrit", C.c_int), ("theta", C.c_double), ("mac", C.c_int),
                ("force_accuracy", C.c_double), ("direct_below", C.c_int),
                ("quirk_factorial", C.c_int)]


class TcStats(C.Structure):
    _fields_ = [("n_m2l", C.c_longlong), ("n_p2p_cells", C.c_longlong),
                ("n_p2p_pairs", C.c_longlong), ("ncells", C.c_int), ("nleaves", C.c_int),
                ("depth", C.c_int), ("nlevels", C.c_int), ("ms_tree", C.c_double),
                ("ms_upward", C.c_double), ("ms_traverse", C.c_double),
                ("ms_lists", C.c_double), ("ms_m2l", C.c_double), ("ms_p2p", C.c_double),
                ("ms_downward", C.c_double), ("ms_total", C.c_double),
                ("n_m2l_low", C.c_longlong), ("n_p2p_low", C.c_longlong),
                ("ms_prepass", C.c_double)]


class TcParticles(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.c_void_p) for k in (
        "pos", "pos_e", "vel", "mass", "timestep", "acc", "pot", "acc_old")]


class TcCellLayout(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("size", "nchild", "nbody", "child", "body", "X", "R",
                                       "min_acc", "cell_mass", "has_sink", "M", "L", "Mdot",
                                       "Ldot", "Pn")]


_lib = None


def lib():
    """Load libtcfmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtcfmm.so is not built (%s); run "
                               "`python -m taichi_collisional_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.tc_create.argtypes = [C.POINTER(TcConfig), vp, C.POINTER(vp)]
        L.tc_destroy.argtypes = [vp]
        L.tc_set_seed.argtypes = [vp, C.c_long]
        L.tc_set_shift.argtypes = [vp, C.POINTER(C.c_double)]
        L.tc_last_shift.argtypes = [vp, C.POINTER(C.c_double)]
        L.tc_set_precision.argtypes = [vp, C.c_int, C.c_double]
        L.tc_set_p2p_shape.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int]
        L.tc_set_small_tree.argtypes = [vp, C.c_int]
        L.tc_get_p2p_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.tc_tune_p2p.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
        L.tc_set_adaptive.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.tc_eval_clusters.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        L.tc_set_order_model.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp,
                                         C.c_double]
        L.tc_leaf_count.argtypes = [vp]
        L.tc_leaf_data.argtypes = [vp, vp, vp, vp]
        L.tc_leaf_errors.argtypes = [vp, vp, vp, vp]
        L.tc_m2l_order_counts.argtypes = [vp, C.POINTER(C.c_longlong), C.c_int]
        L.tc_mlp_forward.argtypes = [vp, vp, vp, C.c_int, vp]
        L.tc_mlp_grad.argtypes = [vp, vp, vp, vp, C.c_int, C.c_float, vp, vp]
        L.tc_mlp_adam.argtypes = [vp, vp, vp, vp, vp, C.c_float, C.c_int]
        L.tc_eval_bodies.argtypes = [vp, vp, C.c_int]
        L.tc_eval_soa.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.tc_direct_soa.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]
        L.tc_direct_soa_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp,
                                            vp]
        L.tc_get_acc_old.argtypes = [vp, vp]
        L.tc_get_cell_layout.argtypes = [C.c_int, C.POINTER(TcCellLayout)]
        for k in ("tc_P2M", "tc_M2M", "tc_L2P"):
            getattr(L, k).argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int]
        L.tc_L2L.argtypes = [vp, vp, C.c_int, vp, C.c_int]
        L.tc_M2L.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int]
        L.tc_P2P.argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int]
        L.tc_get_master_cell_layout.argtypes = [C.c_int, C.POINTER(TcCellLayout)]
        for k in ("tc_master_P2M", "tc_master_M2M", "tc_master_L2P"):
            getattr(L, k).argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int]
        L.tc_master_L2L.argtypes = [vp, vp, C.c_int, vp, C.c_int]
        L.tc_master_M2L.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int]
        L.tc_master_P2P.argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_int,
                                    C.c_double]
        L.tc_get_stats.argtypes = [vp, C.POINTER(TcStats)]
        L.tc_last_error.argtypes = [vp]
        L.tc_last_error.restype = C.c_char_p
        L.tc_supported_orders.argtypes = [C.POINTER(C.c_int), C.c_int]
        L.tc_master_eval_bodies.argtypes = [vp, vp, C.c_int, C.c_int, C.c_double]
        L.tc_master_evolve.argtypes = [vp, C.POINTER(TcParticles), C.POINTER(C.c_double),
                                       C.c_double, C.c_int, C.c_double]
        L.tc_master_evolve_particles.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_double),
                                                 C.c_double, C.c_int, C.c_double]
        L.tc_master_eval_soa.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_double, vp,
                                         vp, vp, vp]
        _lib = L
    return _lib


EXPORTED_SYMBOLS = ["tc_create", "tc_destroy", "tc_set_seed", "tc_set_shift", "tc_last_shift",
                    "tc_set_precision", "tc_set_p2p_shape", "tc_get_p2p_shape", "tc_tune_p2p",
                    "tc_set_small_tree",
                    "tc_eval_clusters", "tc_set_adaptive", "tc_set_order_model", "tc_leaf_count",
                    "tc_leaf_data", "tc_leaf_errors", "tc_m2l_order_counts", "tc_mlp_param_count",
                    "tc_mlp_input_count", "tc_mlp_forward", "tc_mlp_grad", "tc_mlp_adam",
                    "tc_eval_bodies", "tc_eval_soa", "tc_direct_soa", "tc_direct_soa_noalloc",
                    "tc_get_acc_old",
                    "tc_get_cell_layout", "tc_P2M", "tc_M2M", "tc_M2L", "tc_L2L", "tc_L2P",
                    "tc_P2P", "tc_master_eval_bodies", "tc_master_eval_soa",
                    "tc_master_evolve", "tc_master_evolve_particles", "tc_get_stats",
                    "tc_last_error", "tc_supported_orders", "tc_get_master_cell_layout",
                    "tc_master_P2M", "tc_master_M2M", "tc_master_M2L", "tc_master_L2L",
                    "tc_master_L2P", "tc_master_P2P"]


def cell_dtype(p):
    """numpy dtype of the reference's struct Cell for order p (exafmm.h:50-71)."""
    lay = TcCellLayout()
    if lib().tc_get_cell_layout(int(p), C.byref(lay)) != 0:
        raise ValueError("no cell layout for p=%r" % p)
    nt = (p + 1) ** 2
    return np.dtype({"names": ["NCHILD", "NBODY", "CHILD", "BODY", "X", "R", "min_acc",
                               "cell_mass", "has_sink", "M", "L", "Mdot", "Ldot", "Pn"],
                     "formats": ["<i4", "<i4", "<u8", "<u8", ("<f8", 3), "<f8", "<f8", "<f8",
                                 "u1", ("<f8", nt), ("<f8", nt), ("<f8", nt), ("<f8", nt),
                                 ("<f8", p + 1)],
                     "offsets": [lay.nchild, lay.nbody, lay.child, lay.body, lay.X, lay.R,
                                 lay.min_acc, lay.cell_mass, lay.has_sink, lay.M, lay.L,
                                 lay.Mdot, lay.Ldot, lay.Pn],
                     "itemsize": lay.size})


def master_cell_dtype(p):
    """numpy dtype of the master sources' struct Cell for order p (master
    exafmm.h:50-71: M, L, Pn first; no Mdot / Ldot)."""
    lay = TcCellLayout()
    if lib().tc_get_master_cell_layout(int(p), C.byref(lay)) != 0:
        raise ValueError("no cell layout for p=%r" % p)
    nt = (p + 1) ** 2
    return np.dtype({"names": ["M", "L", "Pn", "NCHILD", "NBODY", "CHILD", "BODY", "X", "R",
                               "min_acc", "cell_mass", "has_sink"],
                     "formats": [("<f8", nt), ("<f8", nt), ("<f8", p + 1), "<i4", "<i4", "<u8",
                                 "<u8", ("<f8", 3), "<f8", "<f8", "<f8", "u1"],
                     "offsets": [lay.M, lay.L, lay.Pn, lay.nchild, lay.nbody, lay.child,
                                 lay.body, lay.X, lay.R, lay.min_acc, lay.cell_mass,
                                 lay.has_sink],
                     "itemsize": lay.size})


def supported_orders():
    buf = (C.c_int * 64)()
    n = lib().tc_supported_orders(buf, 64)
    return list(buf[:n])


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def check_tensors(specs, n, device_type="cuda"):
    """Validate the device tensors an entry point hands to the kernels as raw
    pointers.  specs: (name, tensor, dtype, trailing shape) with shape (3,) for
    an (n, 3) array and () for an (n,) array; tensor None is skipped.  Every
    tensor must be a contiguous `device_type` tensor of that dtype and shape,
    all on one device.  Raises TypeError (type, dtype, device, layout) or
    ValueError (shape)."""
    import torch
    dev = None
    for name, t, dtype, tail in specs:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError("%s must be a torch tensor" % name)
        if t.device.type != device_type:
            raise TypeError("%s must be a %s tensor (got %s)" % (name, device_type, t.device))
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise TypeError("%s is on %s, the other arrays on %s" % (name, t.device, dev))
        if t.dtype != dtype:
            raise TypeError("%s must be %s (got %s)" % (name, dtype, t.dtype))
        if not t.is_contiguous():
            raise TypeError("%s must be contiguous" % name)
        if tuple(t.shape) != (n,) + tuple(tail):
            raise ValueError("%s must have shape %s (got %s)" % (name, (n,) + tuple(tail),
                                                                 tuple(t.shape)))
    return dev


def mlp_param_count():
    return lib().tc_mlp_param_count()


def _mlp_check(params, X):
    import torch
    if params.dtype != torch.float32 or params.shape != (mlp_param_count(),) or \
            not params.is_contiguous() or params.device.type != "cuda":
        raise TypeError("params must be a contiguous float32 CUDA tensor of %d values"
                        % mlp_param_count())
    if X.dtype != torch.float32 or X.dim() != 2 or X.shape[1] != lib().tc_mlp_input_count() \
            or not X.is_contiguous() or X.device != params.device:
        raise TypeError("X must be a contiguous float32 (B, %d) tensor on the params' device"
                        % lib().tc_mlp_input_count())


class Fmm:
    """One FMM configuration on the current GPU (see include/tcfmm.h)."""

    def __init__(self, p=12, ncrit=64, theta=0.4, mac=TC_MAC_THETA, force_accuracy=2e-8,
                 direct_below=8000, quirk_factorial=True, stream=None):
        cfg = TcConfig(p, ncrit, theta, mac, force_accuracy, direct_below, int(quirk_factorial))
        h = C.c_void_p()
        s = None if stream is None else C.c_void_p(int(stream))
        rc = lib().tc_create(C.byref(cfg), s, C.byref(h))
        if rc != 0:
            raise ValueError("tc_create failed (%d): unsupported configuration %r; compiled "
                             "orders %s" % (rc, dict(p=p, ncrit=ncrit, theta=theta),
                                            supported_orders()))
        self.h = h
        self.p, self.ncrit, self.theta = p, ncrit, theta
        self._last_n = 0

    def close(self):
        # at interpreter exit the module globals may already be cleared to None
        L = globals().get("_lib")
        if getattr(self, "h", None) and L is not None:
            L.tc_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d): %s" % (what, rc,
                                                       lib().tc_last_error(self.h).decode()))

    def set_precision(self, p2p_mode=TC_P2P_FP64, close_radius=1e-3):
        """Near-field precision of the following evaluations (tc_set_precision):
        TC_P2P_FP64, TC_P2P_MIXED or TC_P2P_MIXED_HILO with FP64 pairs below
        close_radius."""
        self._check(lib().tc_set_precision(self.h, int(p2p_mode), float(close_radius)),
                    "tc_set_precision")

    def set_p2p_shape(self, tile=128, list_slice=128, stage_rounds=8.0, targets_per_lane=1):
        """Launch shape of the near field (tc_set_p2p_shape): source bodies per
        LDS tile (64, 128, 192), list entries per work item, the staging cost
        that sets how leaves are cut into target groups, and targets per lane
        (1 or 2)."""
        self._check(lib().tc_set_p2p_shape(self.h, int(tile), int(list_slice),
                                           float(stage_rounds), int(targets_per_lane)),
                    "tc_set_p2p_shape")

    def set_small_tree(self, max_bodies=4096):
        """Master time-step evaluations of at most max_bodies bodies build their
        tree in one launch (tc_set_small_tree); 0: always level by level."""
        self._check(lib().tc_set_small_tree(self.h, int(max_bodies)), "tc_set_small_tree")

    def p2p_shape(self):
        """(tile, list_slice, stage_rounds, targets_per_lane)"""
        t, l, g, r = C.c_int(), C.c_int(), C.c_double(), C.c_int()
        self._check(lib().tc_get_p2p_shape(self.h, C.byref(t), C.byref(l), C.byref(g),
                                           C.byref(r)), "tc_get_p2p_shape")
        return t.value, l.value, g.value, r.value

    P2P_TUNE_GRID = [(t, l, g, r) for r in (1, 2) for t in (64, 128, 192) for l in (64, 128, 256)
                     for g in (0.0, 2.0, 4.0, 8.0)]

    def tune_p2p(self, reps=2, grid=False):
        """After an evaluation: time its near field under a grid of launch
        shapes (P2P_TUNE_GRID, as (tile, list_slice, stage_rounds,
        targets_per_lane)) and keep the fastest (tc_tune_p2p).  Returns
        (shape, best ms, ms of the shape set before), plus {shape: ms} for the
        whole grid when grid=True."""
        b, d = C.c_double(), C.c_double()
        g = (C.c_double * len(self.P2P_TUNE_GRID))()
        self._check(lib().tc_tune_p2p(self.h, int(reps), C.byref(b), C.byref(d), g),
                    "tc_tune_p2p")
        out = (self.p2p_shape(), b.value, d.value)
        if grid:
            out += ({sh: g[k] for k, sh in enumerate(self.P2P_TUNE_GRID)},)
        return out

    def set_seed(self, seed):
        self._check(lib().tc_set_seed(self.h, int(seed)), "tc_set_seed")

    def set_shift(self, frac):
        a = (C.c_double * 3)(*[float(f) for f in frac])
        self._check(lib().tc_set_shift(self.h, a), "tc_set_shift")

    def last_shift(self):
        a = (C.c_double * 3)()
        lib().tc_last_shift(self.h, a)
        return list(a)

    def get_force_and_potential(self, bodies):
        """Drop-in for the reference call on a host array of body_dtype records."""
        if bodies.dtype != body_dtype or not bodies.flags.c_contiguous:
            raise TypeError("bodies must be a C-contiguous array of body_dtype")
        self._check(lib().tc_eval_bodies(self.h, bodies.ctypes.data, bodies.size),
                    "tc_eval_bodies")
        return bodies

    def eval(self, x, v, m, flags=None, pot=None, acc=None, jerk=None):
        """Device tensors: x, v (n,3) float64, m (n,) float64, flags (n,) uint8."""
        import torch
        n = m.shape[0]
        f8, u1 = torch.float64, torch.uint8
        dev = check_tensors([("x", x, f8, (3,)), ("v", v, f8, (3,)), ("m", m, f8, ()),
                             ("flags", flags, u1, ()), ("pot", pot, f8, ()),
                             ("acc", acc, f8, (3,)), ("jerk", jerk, f8, (3,))], n)
        if pot is None:
            pot = torch.empty(n, dtype=torch.float64, device=dev)
        if acc is None:
            acc = torch.empty(n, 3, dtype=torch.float64, device=dev)
        if jerk is None:
            jerk = torch.empty(n, 3, dtype=torch.float64, device=dev)
        self._last_n = n
        self._check(lib().tc_eval_soa(self.h, n, _ptr(x), _ptr(v), _ptr(m), _ptr(flags),
                                      _ptr(pot), _ptr(acc), _ptr(jerk)), "tc_eval_soa")
        return pot, acc, jerk

    def eval_clusters(self, starts, x, v, m, flags=None, pot=None, acc=None, jerk=None):
        """tc_eval_clusters: independent clusters [starts[k], starts[k+1]) of the
        device arrays in one evaluation (starts: host ints, starts[0] = 0,
        starts[-1] = n)."""
        import numpy as np
        import torch
        st = np.ascontiguousarray(starts, dtype=np.int32)
        n = m.shape[0]
        if st.ndim != 1 or st.size < 2 or int(st[-1]) != n:
            raise ValueError("starts must hold ncl + 1 offsets ending at n = %d" % n)
        f8, u1 = torch.float64, torch.uint8
        dev = check_tensors([("x", x, f8, (3,)), ("v", v, f8, (3,)), ("m", m, f8, ()),
                             ("flags", flags, u1, ()), ("pot", pot, f8, ()),
                             ("acc", acc, f8, (3,)), ("jerk", jerk, f8, (3,))], n)
        pot = torch.empty(n, dtype=f8, device=dev) if pot is None else pot
        acc = torch.empty(n, 3, dtype=f8, device=dev) if acc is None else acc
        jerk = torch.empty(n, 3, dtype=f8, device=dev) if jerk is None else jerk
        self._last_n = n
        self._check(lib().tc_eval_clusters(self.h, st.size - 1, st.ctypes.data, _ptr(x), _ptr(v),
                                           _ptr(m), _ptr(flags), _ptr(pot), _ptr(acc),
                                           _ptr(jerk)), "tc_eval_clusters")
        return pot, acc, jerk

    def acc_old(self):
        """|a| from the last error-MAC evaluation's pre-pass (caller order), cuda float64."""
        import torch
        n = self.stats_n()
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        self._check(lib().tc_get_acc_old(self.h, _ptr(out)), "tc_get_acc_old")
        return out

    def stats_n(self):
        return self._last_n

    def direct(self, x, v, m, targets, flags=None):
        """FP64 all-pairs potential/acceleration/jerk for the given target indices."""
        import torch
        nt = targets.shape[0]
        f8 = torch.float64
        dev = check_tensors([("x", x, f8, (3,)), ("v", v, f8, (3,)), ("m", m, f8, ()),
                             ("flags", flags, torch.uint8, ())], m.shape[0])
        tg = targets.to(device=dev, dtype=torch.int32).contiguous()
        pot = torch.empty(nt, dtype=torch.float64, device=dev)
        acc = torch.empty(nt, 3, dtype=torch.float64, device=dev)
        jerk = torch.empty(nt, 3, dtype=torch.float64, device=dev)
        self._check(lib().tc_direct_soa(self.h, m.shape[0], _ptr(x), _ptr(v), _ptr(m),
                                        _ptr(flags), nt, _ptr(tg), _ptr(pot), _ptr(acc),
                                        _ptr(jerk)), "tc_direct_soa")
        return pot, acc, jerk

    def direct_noalloc(self, x, v, m, targets, pot, acc, jerk, flags=None):
        """tc_direct_soa_noalloc: the same sum into caller-given outputs (pot (nt,),
        acc and jerk (nt, 3), float64 on the bodies' device), staged in the
        library's fixed scratch; queued on the context's stream, not waited for."""
        import torch
        nt = targets.shape[0]
        f8 = torch.float64
        dev = check_tensors([("x", x, f8, (3,)), ("v", v, f8, (3,)), ("m", m, f8, ()),
                             ("flags", flags, torch.uint8, ())], m.shape[0])
        odev = check_tensors([("pot", pot, f8, ()), ("acc", acc, f8, (3,)),
                              ("jerk", jerk, f8, (3,))], nt)
        if odev != dev:
            raise TypeError("the outputs are on %s, the bodies on %s" % (odev, dev))
        if targets.dtype != torch.int32 or not targets.is_contiguous() or targets.device != dev:
            raise TypeError("targets must be a contiguous int32 tensor on %s" % dev)
        self._check(lib().tc_direct_soa_noalloc(self.h, m.shape[0], _ptr(x), _ptr(v), _ptr(m),
                                                _ptr(flags), nt, _ptr(targets), _ptr(pot),
                                                _ptr(acc), _ptr(jerk)),
                    "tc_direct_soa_noalloc")

    # ---- master (non-jerk) variant ----------------------------------------
    def master_get_force_and_potential(self, bodies, get_steps=False, dt_param=0.025, mode=0):
        """Drop-in for the master variant's get_force_and_potential(bodies, get_steps)
        on a host array of master_body_dtype records (F, p and, with get_steps,
        timestep are overwritten)."""
        if bodies.dtype != master_body_dtype or not bodies.flags.c_contiguous:
            raise TypeError("bodies must be a C-contiguous array of master_body_dtype")
        mode = int(mode) | (TC_MASTER_STEPS if get_steps else 0)
        self._check(lib().tc_master_eval_bodies(self.h, bodies.ctypes.data, bodies.size, mode,
                                                float(dt_param)), "tc_master_eval_bodies")
        return bodies

    def master_eval(self, x, v, m, flags=None, get_steps=False, dt_param=0.025, mode=0,
                    acc_old=None):
        """Device tensors in, (pot, acc, timestep-or-None) cuda float64 tensors out.
        acc_old (n,) cuda float64, TC_MASTER_FMM only: read, then overwritten."""
        import torch
        n = m.shape[0]
        f8 = torch.float64
        dev = check_tensors([("x", x, f8, (3,)), ("v", v, f8, (3,)), ("m", m, f8, ()),
                             ("flags", flags, torch.uint8, ()), ("acc_old", acc_old, f8, ())], n)
        mode = int(mode) | (TC_MASTER_STEPS if get_steps else 0)
        pot = torch.empty(n, dtype=torch.float64, device=dev)
        acc = torch.empty(n, 3, dtype=torch.float64, device=dev)
        ts = torch.empty(n, dtype=torch.float64, device=dev) if get_steps else None
        self._last_n = n
        self._check(lib().tc_master_eval_soa(self.h, n, _ptr(x), _ptr(v), _ptr(m), _ptr(flags),
                                             mode, float(dt_param), _ptr(pot), _ptr(acc),
                                             _ptr(ts), _ptr(acc_old)), "tc_master_eval_soa")
        return pot, acc, ts

    def master_evolve(self, state, dt, simtime=0.0, mode=0, dt_param=0.025):
        """One do_evolve(sys, dt) of the master step loop (GRADIENT scheme) on
        device tensors: state is a dict with pos, pos_e, vel, acc (n,3), mass,
        timestep, pot, acc_old (n,) float64 cuda tensors, updated in place.
        Returns the new simulation time."""
        import torch
        keys = ("pos", "pos_e", "vel", "mass", "timestep", "acc", "pot", "acc_old")
        missing = [k for k in keys if state.get(k) is None]
        if missing:
            raise TypeError("state lacks %s" % ", ".join(missing))
        n = state["mass"].shape[0]
        vec = ("pos", "pos_e", "vel", "acc")
        check_tensors([(k, state[k], torch.float64, (3,) if k in vec else ()) for k in keys], n)
        sp = TcParticles(n, *[state[k].data_ptr() for k in keys])
        t = C.c_double(simtime)
        self._check(lib().tc_master_evolve(self.h, C.byref(sp), C.byref(t), float(dt), int(mode),
                                           float(dt_param)), "tc_master_evolve")
        return t.value

    def master_evolve_particles(self, part, dt, simtime=0.0, mode=0, dt_param=0.025):
        """The same on a host array of particle_dtype records (the reference's
        struct particle); returns the new simulation time."""
        if part.dtype != particle_dtype or not part.flags.c_contiguous:
            raise TypeError("part must be a C-contiguous array of particle_dtype")
        t = C.c_double(simtime)
        self._check(lib().tc_master_evolve_particles(self.h, part.ctypes.data, part.size,
                                                     C.byref(t), float(dt), int(mode),
                                                     float(dt_param)),
                    "tc_master_evolve_particles")
        return t.value

    # ---- per-kernel entry points over the reference's Cell / Body records ----
    def _recs(self, bodies, cells):
        if cells.dtype != cell_dtype(self.p) or not cells.flags.c_contiguous:
            raise TypeError("cells must be a C-contiguous array of cell_dtype(p)")
        if bodies is not None and (bodies.dtype != body_dtype or not bodies.flags.c_contiguous):
            raise TypeError("bodies must be a C-contiguous array of body_dtype")

    @staticmethod
    def _ids(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        return a, a.ctypes.data

    def P2M(self, bodies, cells, ids):
        self._recs(bodies, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_P2M(self.h, bodies.ctypes.data, bodies.size, cells.ctypes.data,
                                 cells.size, pa, a.size), "tc_P2M")

    def M2M(self, bodies, cells, ids):
        self._recs(bodies, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_M2M(self.h, bodies.ctypes.data, bodies.size, cells.ctypes.data,
                                 cells.size, pa, a.size), "tc_M2M")

    def M2L(self, cells, targets, sources):
        self._recs(None, cells)
        t, pt = self._ids(targets)
        s, ps = self._ids(sources)
        self._check(lib().tc_M2L(self.h, cells.ctypes.data, cells.size, pt, ps, t.size),
                    "tc_M2L")

    def L2L(self, cells, ids):
        self._recs(None, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_L2L(self.h, cells.ctypes.data, cells.size, pa, a.size), "tc_L2L")

    def L2P(self, bodies, cells, ids):
        self._recs(bodies, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_L2P(self.h, bodies.ctypes.data, bodies.size, cells.ctypes.data,
                                 cells.size, pa, a.size), "tc_L2P")

    def P2P(self, bodies, cells, targets, sources):
        self._recs(bodies, cells)
        t, pt = self._ids(targets)
        s, ps = self._ids(sources)
        self._check(lib().tc_P2P(self.h, bodies.ctypes.data, bodies.size, cells.ctypes.data,
                                 cells.size, pt, ps, t.size), "tc_P2P")

    # ---- the same over the master sources' Cell / Body records ----
    def _mrecs(self, bodies, cells):
        if cells.dtype != master_cell_dtype(self.p) or not cells.flags.c_contiguous:
            raise TypeError("cells must be a C-contiguous array of master_cell_dtype(p)")
        if bodies is not None and (bodies.dtype != master_body_dtype
                                   or not bodies.flags.c_contiguous):
            raise TypeError("bodies must be a C-contiguous array of master_body_dtype")

    def master_P2M(self, bodies, cells, ids):
        self._mrecs(bodies, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_master_P2M(self.h, bodies.ctypes.data, bodies.size,
                                        cells.ctypes.data, cells.size, pa, a.size),
                    "tc_master_P2M")

    def master_M2M(self, bodies, cells, ids):
        self._mrecs(bodies, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_master_M2M(self.h, bodies.ctypes.data, bodies.size,
                                        cells.ctypes.data, cells.size, pa, a.size),
                    "tc_master_M2M")

    def master_M2L(self, cells, ci, cj):
        """Symmetric M2L_rotate(Ci, Cj) per entry (only ci < cj acts)."""
        self._mrecs(None, cells)
        t, pt = self._ids(ci)
        s, ps = self._ids(cj)
        if t.size != s.size:
            raise ValueError("ci and cj differ in length")
        self._check(lib().tc_master_M2L(self.h, cells.ctypes.data, cells.size, pt, ps, t.size),
                    "tc_master_M2L")

    def master_L2L(self, cells, ids):
        self._mrecs(None, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_master_L2L(self.h, cells.ctypes.data, cells.size, pa, a.size),
                    "tc_master_L2L")

    def master_L2P(self, bodies, cells, ids):
        self._mrecs(bodies, cells)
        a, pa = self._ids(ids)
        self._check(lib().tc_master_L2P(self.h, bodies.ctypes.data, bodies.size,
                                        cells.ctypes.data, cells.size, pa, a.size),
                    "tc_master_L2P")

    def master_P2P(self, bodies, cells, targets, sources, mode=0, dt_param=0.025):
        """P2P_simple, or with TC_MASTER_STEPS in mode P2P with the time-step sum."""
        self._mrecs(bodies, cells)
        t, pt = self._ids(targets)
        s, ps = self._ids(sources)
        if t.size != s.size:
            raise ValueError("targets and sources differ in length")
        self._check(lib().tc_master_P2P(self.h, bodies.ctypes.data, bodies.size,
                                        cells.ctypes.data, cells.size, pt, ps, t.size,
                                        int(mode), float(dt_param)), "tc_master_P2P")

    # ---- per-cell order and angle (learned choice) --------------------------
    def set_adaptive(self, mode, q=0, theta=0.0, probe_theta=0.7):
        """tc_set_adaptive: 0 off, 1 every leaf at (q, theta), 2 the learned choice."""
        self._check(lib().tc_set_adaptive(self.h, int(mode), int(q), float(theta),
                                          float(probe_theta)), "tc_set_adaptive")

    def set_order_model(self, params, q, theta, cost, feat_lo, feat_hi, target_rms):
        """The MLP parameters (float32, mlp_param_count() values), the candidate
        grid (orders q, angles theta, relative costs: one each per candidate)
        and the features' training range [feat_lo, feat_hi] (9 values each)."""
        import numpy as np
        P = np.ascontiguousarray(params, dtype=np.float32)
        qa = np.ascontiguousarray(q, dtype=np.int32)
        ta = np.ascontiguousarray(theta, dtype=np.float32)
        ca = np.ascontiguousarray(cost, dtype=np.float32)
        lo = np.ascontiguousarray(feat_lo, dtype=np.float32)
        hi = np.ascontiguousarray(feat_hi, dtype=np.float32)
        if not (qa.size == ta.size == ca.size) or lo.size != 9 or hi.size != 9:
            raise ValueError("q, theta and cost need one entry per candidate, the feature "
                             "range 9 values each")
        self._check(lib().tc_set_order_model(self.h, P.ctypes.data, P.size, qa.ctypes.data,
                                             ta.ctypes.data, ca.ctypes.data, qa.size,
                                             lo.ctypes.data, hi.ctypes.data, float(target_rms)),
                    "tc_set_order_model")

    def leaf_data(self):
        """After an evaluation with per-cell orders: (features [nleaves, 9],
        leaf index of every body [n], predicted log10 error [nleaves] or None)."""
        import torch
        nl = lib().tc_leaf_count(self.h)
        dev = torch.device("cuda", torch.cuda.current_device())
        f = torch.empty(nl, 9, dtype=torch.float32, device=dev)
        lob = torch.empty(self._last_n, dtype=torch.int32, device=dev)
        pred = torch.empty(nl, dtype=torch.float32, device=dev)
        self._check(lib().tc_leaf_data(self.h, _ptr(f), _ptr(lob), _ptr(pred)), "tc_leaf_data")
        return f, lob, pred

    def leaf_errors(self, acc, acc_ref):
        """log10 rms relative force error per leaf of the last evaluation's tree."""
        import torch
        dev = check_tensors([("acc", acc, torch.float64, (3,)),
                             ("acc_ref", acc_ref, torch.float64, (3,))], self._last_n)
        out = torch.empty(lib().tc_leaf_count(self.h), dtype=torch.float32, device=dev)
        self._check(lib().tc_leaf_errors(self.h, _ptr(acc), _ptr(acc_ref), _ptr(out)),
                    "tc_leaf_errors")
        return out

    def m2l_order_counts(self):
        """M2L pairs of the last evaluation per truncation order (index = order)."""
        buf = (C.c_longlong * 25)()
        lib().tc_m2l_order_counts(self.h, buf, 24)
        return list(buf)

    def mlp_forward(self, params, X):
        import torch
        Y = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        _mlp_check(params, X)
        self._check(lib().tc_mlp_forward(self.h, _ptr(params), _ptr(X), X.shape[0], _ptr(Y)),
                    "tc_mlp_forward")
        return Y

    def mlp_grad(self, params, X, T, tau=0.5):
        """(gradient of the expectile loss mean(w (Y - T)^2) w.r.t. params, loss);
        w = tau where Y < T, 1 - tau elsewhere."""
        import torch
        _mlp_check(params, X)
        if T.dtype != torch.float32 or T.shape != (X.shape[0],) or not T.is_contiguous():
            raise TypeError("T must be a contiguous float32 tensor of shape (B,)")
        G = torch.empty_like(params)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        self._check(lib().tc_mlp_grad(self.h, _ptr(params), _ptr(X), _ptr(T), X.shape[0],
                                      float(tau), _ptr(G), _ptr(loss)), "tc_mlp_grad")
        return G, float(loss.item())

    def mlp_adam(self, params, G, m, v, lr, step):
        self._check(lib().tc_mlp_adam(self.h, _ptr(params), _ptr(G), _ptr(m), _ptr(v),
                                      float(lr), int(step)), "tc_mlp_adam")

    def stats(self):
        st = TcStats()
        lib().tc_get_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in TcStats._fields_}
