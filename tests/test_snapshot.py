"""HDF5 snapshot format (libtcio) against the reference's own write_snapshot /
read_snapshot (taichi_with_jerks/io.h, master io.h), compiled from its headers
into oracle/_ref/libtcrefio_*.so by oracle/build_ref.sh.  CPU only."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import taichi_collisional_amd as tca
from taichi_collisional_amd import snapshot as snap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

if not os.path.isfile(os.path.join(os.path.dirname(tca.__file__), "libtcio.so")):
    pytest.skip("libtcio.so not built (no libhdf5)", allow_module_level=True)

# reference build -> (harness library, field set its Makefile flags select)
BUILDS = {
    "jerk": ("libtcrefio_jerk.so", snap.SNAP_JERK_BUILD),
    "jerk_plain": ("libtcrefio_jerk_plain.so", 0),
    "master": ("libtcrefio_master.so", snap.SNAP_MASTER_BUILD),
}


def _ref(build):
    path = os.path.join(REF, BUILDS[build][0])
    if not os.path.isfile(path):
        pytest.skip("reference snapshot harness %s not built" % path)
    L = C.CDLL(path)
    L.refio_write.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_uint, C.c_double]
    L.refio_read.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_uint,
                             C.POINTER(C.c_uint), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    assert L.refio_particle_size() == tca.particle_dtype.itemsize == 168
    return L


def _particles(n, seed, unique=True):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, dtype=tca.particle_dtype)
    for f in ("pos", "pos_e", "vel", "acc", "jerk"):
        p[f] = rng.standard_normal((n, 3))
    for f in ("mass", "timestep", "pot", "postime", "acc_old"):
        p[f] = rng.standard_normal(n)
    p["id"] = rng.permutation(n) * 3 + 5 if unique else rng.integers(0, max(n // 4, 1), n)
    return p


def _dump(path):
    """Every root object and header attribute: (description, values)."""
    out = {}
    for name in snap.list_objects(path):
        if name == "Header":
            continue
        d = snap.describe(path, name)
        out[name] = (d, snap.read_dataset(path, name, as_uint=d[0] == "int"))
    for a in ("Header/NumPartTotal", "Header/Time"):
        out[a] = (snap.describe(path, a), None)
    return out


@pytest.mark.parametrize("build", list(BUILDS))
def test_write_matches_reference(build, tmp_path):
    """Our file holds the datasets the reference's write_snapshot writes, with
    the same names, on-disk types, lengths and values (io.h:106-334)."""
    L = _ref(build)
    fields = BUILDS[build][1]
    part = _particles(3000, seed=1)
    da, db = tmp_path / "ref", tmp_path / "ours"
    da.mkdir()
    db.mkdir()
    assert L.refio_write(os.fsencode(da), 7, part.ctypes.data, part.size, 0.625) == 0
    snap.write_snapshot(db / snap.snapshot_name(7), part, 0.625, fields)
    fa, fb = str(da / "snapshot_7.hdf5"), str(db / "snapshot_7.hdf5")
    assert snap.list_objects(fa) == snap.list_objects(fb)
    a, b = _dump(fa), _dump(fb)
    assert a.keys() == b.keys()
    for k in a:
        assert a[k][0] == b[k][0], k
        if a[k][1] is not None:
            np.testing.assert_array_equal(a[k][1], b[k][1], err_msg=k)
    assert snap.read_header(fa) == snap.read_header(fb) == (3000, 0.625)
    # the ids come out sorted and carry their records along
    order = np.argsort(part["id"])
    np.testing.assert_array_equal(snap.read_dataset(fb, "ID", as_uint=True), part["id"][order])
    np.testing.assert_array_equal(snap.read_dataset(fb, "Posy"), part["pos"][order, 1])
    if fields & snap.SNAP_JERK:
        np.testing.assert_array_equal(snap.read_dataset(fb, "Jerkz"), part["jerk"][order, 2])
    if fields & snap.SNAP_DEBUG:
        np.testing.assert_array_equal(snap.read_dataset(fb, "Scalarforce"),
                                      part["acc_old"][order])


@pytest.mark.parametrize("build", ["jerk", "master"])
@pytest.mark.parametrize("num", [0, 12, 250])
def test_reference_reads_our_file(build, num, tmp_path):
    """The reference's read_snapshot reads our file into the same records and
    the same snapshot number as our reader (io.h:11-103)."""
    L = _ref(build)
    part = _particles(777, seed=num + 3)
    name = snap.snapshot_name(num)
    snap.write_snapshot(tmp_path / name, part, 3.25, BUILDS[build][1])
    out = np.zeros(part.size, dtype=tca.particle_dtype)
    n, t, k = C.c_uint(), C.c_double(), C.c_int()
    assert L.refio_read(os.fsencode(tmp_path), name.encode(), out.ctypes.data, out.size,
                        C.byref(n), C.byref(t), C.byref(k)) == 0
    assert (n.value, t.value, k.value) == (777, 3.25, num)
    ours, t2 = snap.read_snapshot(tmp_path / name)
    assert t2 == 3.25 and snap.snapshot_number(str(tmp_path / name)) == num
    for f in ("pos", "vel", "mass", "id"):
        np.testing.assert_array_equal(out[f], ours[f], err_msg=f)
    order = np.argsort(part["id"])
    np.testing.assert_array_equal(ours["pos"], part["pos"][order])
    # the reader sets only what the reference reads
    for f in ("acc", "jerk", "pot", "timestep", "acc_old"):
        assert not np.any(ours[f])


def test_we_read_the_reference_file(tmp_path):
    """Our reader on the reference writer's file returns its records in id order."""
    L = _ref("jerk")
    part = _particles(2048, seed=9)
    assert L.refio_write(os.fsencode(tmp_path), 3, part.ctypes.data, part.size, -1.5) == 0
    ours, t = snap.read_snapshot(tmp_path / "snapshot_3.hdf5")
    order = np.argsort(part["id"])
    assert t == -1.5
    for f in ("pos", "vel", "mass", "id"):
        np.testing.assert_array_equal(ours[f], part[f][order], err_msg=f)


def test_equal_ids_keep_input_order(tmp_path):
    """Equal ids keep their input order (the reference's std::sort leaves it
    unspecified), so a file written twice is the same."""
    part = _particles(5000, seed=4, unique=False)
    path = tmp_path / "snapshot_1.hdf5"
    snap.write_snapshot(path, part, 0.0)
    order = np.argsort(part["id"], kind="stable")
    np.testing.assert_array_equal(snap.read_dataset(path, "Mass"), part["mass"][order])
    np.testing.assert_array_equal(snap.read_dataset(path, "Accx"), part["acc"][order, 0])


@pytest.mark.parametrize("n", [0, 1])
def test_tiny_snapshots(n, tmp_path):
    part = _particles(n, seed=2)
    path = tmp_path / "snapshot_5.hdf5"
    snap.write_snapshot(path, part, 1.0)
    back, t = snap.read_snapshot(path)
    assert back.size == n and t == 1.0
    np.testing.assert_array_equal(back["pos"], part["pos"])
    assert snap.describe(path, "Posx") == ("float", 8, True, n)
    assert snap.describe(path, "ID") == ("int", 4, False, n)


def test_errors(tmp_path):
    part = _particles(10, seed=0)
    path = tmp_path / "snapshot_2.hdf5"
    with pytest.raises(RuntimeError, match="field bits"):
        snap.write_snapshot(path, part, 0.0, fields=16)
    snap.write_snapshot(path, part, 0.0, fields=0)
    with pytest.raises(RuntimeError, match="no dataset Potential"):
        snap.read_dataset(path, "Potential")
    with pytest.raises(RuntimeError, match="cannot open"):
        snap.read_snapshot(tmp_path / "missing.hdf5")
    with pytest.raises(RuntimeError, match="cannot create"):
        snap.write_snapshot(tmp_path / "no_such_dir" / "x.hdf5", part, 0.0)
    out = np.zeros(9, dtype=tca.particle_dtype)
    rc = snap.lib().tc_read_snapshot(os.fsencode(path), out.ctypes.data, 9)
    assert rc == -1 and b"NumPartTotal is 10" in snap.lib().tc_io_last_error()
    for name, k in [("snapshot_12.hdf5", 12), ("/a/b/snapshot_0.hdf5", 0),
                    ("snapshot_.hdf5", -1), ("snap_1.hdf5", -1), ("snapshot_1x.hdf5", -1)]:
        assert snap.snapshot_number(name) == k, name


def test_error_cdf(tmp_path):
    """plot_jerk_error.py's CDF: zero against itself; a known relative
    perturbation of one particle shows up as that particle's error only."""
    part = _particles(1000, seed=6)
    a, b = tmp_path / "snapshot_0_ref.hdf5", tmp_path / "snapshot_0.hdf5"
    snap.write_snapshot(a, part, 0.0)
    err, frac = snap.error_cdf(a, a)
    assert not np.any(err) and frac[0] == 1.0 and abs(frac[-1] - 1e-3) < 1e-15
    k = int(np.argmax(part["id"]))  # the last row of the files
    q = part.copy()
    q["jerk"][k] *= 1.0 + 1e-6
    snap.write_snapshot(b, q, 0.0)
    assert not np.any(snap.error_cdf(b, a, "acc")[0])
    ej = snap.error_cdf(b, a, "jerk")[0]
    assert np.count_nonzero(ej) == 1 and abs(ej[-1] / 1e-6 - 1) < 1e-9
    q["id"][0] += 10**6
    snap.write_snapshot(b, q, 0.0)
    with pytest.raises(ValueError, match="different particle ids"):
        snap.error_cdf(b, a)


def test_header_declares_exports():
    """libtcio exports every function include/tcio.h declares."""
    hdr = open(os.path.join(ROOT, "include", "tcio.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(tc_\w+)\(", hdr, re.M))
    assert declared == set(snap.EXPORTED_SYMBOLS)
    L = snap.lib()
    for f in declared:
        assert hasattr(L, f), f
