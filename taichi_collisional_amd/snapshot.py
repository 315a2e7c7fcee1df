"""The taichi_collisional HDF5 snapshot format (libtcio.so, include/tcio.h).

Mirrors the reference's write_snapshot / read_snapshot
(taichi_with_jerks/io.h:11-334, master io.h:11-303) over arrays of
``particle_dtype`` records (the reference's struct particle).  The files are
the reference's: a /Header group with NumPartTotal and Time, one 1-D dataset
per field in increasing particle id.

    write_snapshot("snapshot_3.hdf5", part, t_now)       # the jerk build's fields
    part, t_now = read_snapshot("snapshot_3.hdf5")         # pos, vel, mass, id
    acc_x = read_dataset("snapshot_3.hdf5", "Accx")
"""
import ctypes as C
import os

import numpy as np

from . import particle_dtype

HERE = os.path.dirname(os.path.abspath(__file__))

SNAP_POT = 1    # Potential (OUTPUTPOT)
SNAP_DEBUG = 2  # Stepsize, Scalarforce (DEBUG)
SNAP_ACC = 4    # Accx, Accy, Accz (OUTPUTACC)
SNAP_JERK = 8   # Jerkx, Jerky, Jerkz (the jerk build's OUTPUTACC)
SNAP_MASTER_BUILD = SNAP_POT | SNAP_DEBUG | SNAP_ACC
SNAP_JERK_BUILD = SNAP_MASTER_BUILD | SNAP_JERK

EXPORTED_SYMBOLS = ["tc_write_snapshot", "tc_read_snapshot_header", "tc_read_snapshot",
                    "tc_snapshot_number", "tc_read_snapshot_dataset", "tc_snapshot_describe",
                    "tc_snapshot_list", "tc_io_last_error"]

_lib = None


def lib():
    """Load libtcio.so (raises if it has not been built: it needs libhdf5)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(HERE, "libtcio.so")
    if not os.path.isfile(path):
        raise RuntimeError("libtcio.so is not built (python -m taichi_collisional_amd.build; "
                           "it needs libhdf5)")
    L = C.CDLL(path)
    cp, vp, u, i = C.c_char_p, C.c_void_p, C.c_uint, C.c_int
    L.tc_write_snapshot.argtypes = [cp, vp, u, C.c_double, i]
    L.tc_read_snapshot_header.argtypes = [cp, C.POINTER(u), C.POINTER(C.c_double)]
    L.tc_read_snapshot.argtypes = [cp, vp, u]
    L.tc_snapshot_number.argtypes = [cp]
    L.tc_read_snapshot_dataset.argtypes = [cp, cp, vp, u, i]
    L.tc_snapshot_describe.argtypes = [cp, cp, C.POINTER(i), C.POINTER(i), C.POINTER(i),
                                       C.POINTER(C.c_longlong)]
    L.tc_snapshot_list.argtypes = [cp, C.c_char_p, i]
    L.tc_io_last_error.restype = C.c_char_p
    for f in EXPORTED_SYMBOLS[:-1]:
        getattr(L, f).restype = C.c_int
    _lib = L
    return L


def _b(path):
    return os.fsencode(path)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, lib().tc_io_last_error().decode()))


def write_snapshot(path, part, t_now, fields=SNAP_JERK_BUILD):
    """write_snapshot (io.h:106-334): part is a C-contiguous particle_dtype array."""
    part = np.ascontiguousarray(part)
    if part.dtype != particle_dtype or part.ndim != 1:
        raise TypeError("part must be a 1-D array of particle_dtype")
    _check(lib().tc_write_snapshot(_b(path), part.ctypes.data, part.size, float(t_now),
                                   int(fields)), "tc_write_snapshot")


def read_header(path):
    n, t = C.c_uint(), C.c_double()
    _check(lib().tc_read_snapshot_header(_b(path), C.byref(n), C.byref(t)),
           "tc_read_snapshot_header")
    return n.value, t.value


def read_snapshot(path):
    """read_snapshot (io.h:11-103): (records with pos, vel, mass, id set, Time)."""
    n, t = read_header(path)
    part = np.zeros(n, dtype=particle_dtype)
    _check(lib().tc_read_snapshot(_b(path), part.ctypes.data, n), "tc_read_snapshot")
    return part, t


def read_dataset(path, name, as_uint=False):
    """One 1-D dataset of the file as float64 (or uint32)."""
    n, _ = read_header(path)
    out = np.empty(n, dtype=np.uint32 if as_uint else np.float64)
    _check(lib().tc_read_snapshot_dataset(_b(path), name.encode(), out.ctypes.data, n,
                                          int(bool(as_uint))), "tc_read_snapshot_dataset")
    return out


def describe(path, name):
    """(class "float"/"int", size in bytes, signed, number of elements) of a
    dataset, or of an attribute named "<group>/<attribute>"."""
    c, s, sg, ln = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
    _check(lib().tc_snapshot_describe(_b(path), name.encode(), C.byref(c), C.byref(s),
                                      C.byref(sg), C.byref(ln)), "tc_snapshot_describe")
    return ("float", "int", "other")[c.value], s.value, bool(sg.value), ln.value


def list_objects(path):
    """Names of the root-level objects, in name order."""
    buf = C.create_string_buffer(1 << 14)
    _check(lib().tc_snapshot_list(_b(path), buf, len(buf)), "tc_snapshot_list")
    return buf.value.decode().split("\n") if buf.value else []


def snapshot_number(path):
    """The number in "snapshot_<num>.hdf5" (io.h:15-17), or -1."""
    return lib().tc_snapshot_number(_b(path))


def snapshot_name(num):
    """The file name write_snapshot uses for snapshot `num` (io.h:112)."""
    return "snapshot_%d.hdf5" % num


def error_cdf(path, ref_path, quantity="acc"):
    """Relative errors of one snapshot against a reference snapshot of the same
    particles, as plot_jerk_error.py (taichi_with_jerks/plot_jerk_error.py:13-45)
    plots them: |d a| / |a_ref| per particle (quantity "acc", from Acc*; "jerk"
    from Jerk*), sorted, and the fraction of particles above each, 1 - i/n.
    Both files hold the particles in id order, so the rows correspond."""
    stem = {"acc": "Acc", "jerk": "Jerk"}[quantity]
    if not np.array_equal(read_dataset(path, "ID", True), read_dataset(ref_path, "ID", True)):
        raise ValueError("the snapshots hold different particle ids")
    a = np.stack([read_dataset(path, stem + c) for c in "xyz"], axis=1)
    r = np.stack([read_dataset(ref_path, stem + c) for c in "xyz"], axis=1)
    err = np.sort(np.linalg.norm(a - r, axis=1) / np.linalg.norm(r, axis=1))
    return err, 1.0 - np.arange(err.size) / err.size
