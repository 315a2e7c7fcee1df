"""The C-ABI library loads and exports what include/tcfmm.h declares; the
Body record layout equals the reference's (no GPU calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import taichi_collisional_amd as tca

HEADER = os.path.join(ROOT, "include", "tcfmm.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tc_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = tca.lib()
    names = declared_functions()
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(tca.EXPORTED_SYMBOLS) == names


def test_body_layout_matches_reference_struct():
    assert tca.body_dtype.itemsize == 152
    offs = {k: tca.body_dtype.fields[k][1] for k in tca.body_dtype.names}
    want = dict(X=0, q=24, F=32, p=56, V=64, acc_old=88, Morton=96, timestep=112, J=120,
                index=144, issink=148, issource=149)
    assert offs == want
    import pyoracle
    if not pyoracle.ref_available(12):
        pytest.skip("reference harness not built")
    out = np.zeros(28, np.int64)
    pyoracle.reference(12).ref_layout(out)
    ref = dict(zip(["size", "X", "q", "F", "p", "V", "acc_old", "Morton", "timestep", "J", "index",
                    "issink", "issource"], out[:13]))
    assert ref["size"] == 152
    for k, v in want.items():
        assert ref[k] == v, k
    # Cell layout of the reference at p=12 (documented for per-kernel callers)
    cell = dict(zip(["size", "NCHILD", "NBODY", "CHILD", "BODY", "X", "R", "min_acc", "cell_mass",
                     "has_sink", "M", "L", "Mdot", "Ldot", "Pn"], out[13:28]))
    assert cell["size"] == 80 + 4 * 169 * 8 + 13 * 8
    assert (cell["M"], cell["L"], cell["Mdot"], cell["Ldot"], cell["Pn"]) == (
        80, 80 + 1352, 80 + 2 * 1352, 80 + 3 * 1352, 80 + 4 * 1352)


def test_header_compiles_as_c_and_cpp(tmp_path):
    import subprocess
    c = tmp_path / "t.c"
    c.write_text('#include "tcfmm.h"\nint main(void){return sizeof(tc_body)==152?0:1;}\n')
    for comp, src in (("gcc", c),):
        r = subprocess.run([comp, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                            str(tmp_path / "a.out")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert subprocess.run([str(tmp_path / "a.out")]).returncode == 0
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "tcfmm.h"\nint main(){return sizeof(tc_body)==152?0:1;}\n')
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(cpp), "-o",
                        str(tmp_path / "b.out")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_compiled_orders_and_loud_failure_without_library(monkeypatch, tmp_path):
    # the product path refuses to run without the HIP library (no CPU fallback)
    monkeypatch.setattr(tca, "_lib", None)
    monkeypatch.setattr(tca, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError):
        tca.lib()


def test_cell_layout_matches_reference_cell():
    """struct Cell (exafmm.h:50-71) at p = 12 on x86-64: 5592 bytes, M@80,
    L@1432, Mdot@2784, Ldot@4136, Pn@5488."""
    import taichi_collisional_amd as tca
    dt = tca.cell_dtype(12)
    assert dt.itemsize == 5592
    off = {k: dt.fields[k][1] for k in dt.names}
    assert (off["NCHILD"], off["NBODY"], off["CHILD"], off["BODY"]) == (0, 4, 8, 16)
    assert (off["X"], off["R"], off["min_acc"], off["cell_mass"], off["has_sink"]) == (24, 48, 56, 64, 72)
    assert (off["M"], off["L"], off["Mdot"], off["Ldot"], off["Pn"]) == (80, 1432, 2784, 4136, 5488)


def test_master_cell_layout_matches_reference_cell():
    """The master struct Cell (master exafmm.h:50-71: M, L, Pn first, no
    Mdot / Ldot): at p = 12, 2888 bytes, L@1352, Pn@2704, NCHILD@2808; and
    the layout the compiled reference reports at p = 8 and 12."""
    import taichi_collisional_amd as tca
    dt = tca.master_cell_dtype(12)
    off = {k: dt.fields[k][1] for k in dt.names}
    assert dt.itemsize == 2888
    assert (off["M"], off["L"], off["Pn"], off["NCHILD"], off["has_sink"]) == (0, 1352, 2704, 2808,
                                                                            2880)
    lay = tca.TcCellLayout()
    assert tca.lib().tc_get_master_cell_layout(12, ctypes.byref(lay)) == 0
    assert lay.Mdot == -1 and lay.Ldot == -1
    import pyoracle
    if not pyoracle.master_ref_available(12):
        pytest.skip("master reference harness not built")
    for p in (8, 12):
        ref = pyoracle.master_ref_layout(p)
        dt = tca.master_cell_dtype(p)
        assert dt.itemsize == ref["Cell"]
        for k in dt.names:
            assert dt.fields[k][1] == ref["Cell." + k], (p, k)
        assert tca.master_body_dtype.itemsize == ref["Body"]
        for k in tca.master_body_dtype.names:
            assert tca.master_body_dtype.fields[k][1] == ref["Body." + k], k


def test_every_expansion_order_is_compiled():
    """The reference builds EXPANSION=24 and 20 (taichi_with_jerks/Makefile:7,
    10); every order 2..24 has kernels."""
    import taichi_collisional_amd as tca
    assert tca.supported_orders() == list(range(2, 25))
