"""Build libtcfmm.so (HIP, gfx950) in-tree.

    python -m taichi_collisional_amd.build [--jobs N]

Each csrc/*.hip translation unit is compiled with hipcc for gfx950 into
build/*.o, then linked into taichi_collisional_amd/libtcfmm.so.  The HDF5 snapshot library
(csrc/tc_io.cpp, host code) is built with g++ into libtcio.so when libhdf5 is
present.  Packed-FP32
instructions are disabled for all device code.  Always recompiles every unit
(the driver's build check expects that); object files go to
taichi_collisional_amd/build/ (git-ignored).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "libtcfmm.so")
ARCH = os.environ.get("TC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
# Units whose kernels are instantiated for every expansion order are compiled
# once per order shard (-DTC_SHARD=k, see TC_NSHARDS in csrc/tc_kernels.h), so
# the copies build in parallel.
NSHARDS = 3
SHARDED = ["tc_m2l.hip", "tc_upward.hip", "tc_downward.hip"]
PLAIN = ["tc_tree.hip", "tc_p2p.hip", "tc_low.hip", "tc_batch.hip", "tc_master.hip",
         "tc_evolve.hip", "tc_noalloc.hip", "tc_dispatch.hip", "tc_learn.hip", "tc_engine.hip"]
UNITS = [(u, k) for u in SHARDED for k in range(NSHARDS)] + [(u, None) for u in PLAIN]


def _obj(unit):
    u, k = unit
    return os.path.join(BUILD, u.replace(".hip", "" if k is None else "_s%d" % k) + ".o")


def _compile(unit):
    u, k = unit
    src = os.path.join(CSRC, u)
    obj = _obj(unit)
    t = time.time()
    defs = [] if k is None else ["-DTC_SHARD=%d" % k, "-DTC_NSHARDS=%d" % NSHARDS]
    cmd = [HIPCC] + CFLAGS + defs + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (u, " ".join(cmd), r.stderr[-6000:]))
    return os.path.basename(obj), time.time() - t


HDF5_DIR = os.environ.get("TC_HDF5_DIR", "/opt/conda")
IO_LIB = os.path.join(HERE, "libtcio.so")


def build_io(verbose=True):
    """Build libtcio.so (the HDF5 snapshot format, host C++) with g++ against
    libhdf5 under HDF5_DIR.  Returns None, with a message, when HDF5 is absent."""
    inc = os.path.join(HDF5_DIR, "include")
    libdir = os.path.join(HDF5_DIR, "lib")
    if not os.path.isfile(os.path.join(inc, "hdf5.h")):
        if verbose:
            print("  libtcio: no hdf5.h under %s; snapshot library not built" % inc, flush=True)
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + inc,
           os.path.join(CSRC, "tc_io.cpp"), "-o", tmp, "-L" + libdir, "-lhdf5",
           "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for tc_io.cpp:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, IO_LIB)
    if verbose:
        print("  linked", IO_LIB, flush=True)
    return IO_LIB


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    # slowest first, so the long shards start at once
    order = sorted(UNITS, key=lambda u: u[1] is None)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for unit, dt in ex.map(_compile, order):
            if verbose:
                print("  hipcc %-22s %6.1fs" % (unit, dt), flush=True)
    objs = [_obj(u) for u in UNITS]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("  linked", LIB, flush=True)
    build_io(verbose)
    return LIB


if __name__ == "__main__":
    j = None
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(j)
