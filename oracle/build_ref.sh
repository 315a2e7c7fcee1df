#!/usr/bin/env bash
# Build the reference harness (test infrastructure only) into oracle/_ref/.
#
# Compiles oracle/ref_harness.cpp against the taichi_collisional reference
# headers where they lie (/root/reference/taichi_with_jerks), with the jerk
# build's own feature flags (taichi_with_jerks/Makefile:1-9, minus DEBUG and
# the HDF5 I/O, which the harness does not include).  glibc >= 2.27 dropped
# the SVID macro HUGE the reference uses; the historical value is passed in.  One library per
# expansion order; two ISA variants so the GPU hosts (AMD EPYC) can load one:
#   libtcref_p<P>_v4.so  x86-64-v4 (AVX-512), libtcref_p<P>_v3.so  x86-64-v3 (AVX2)
# Nothing is built when the reference is absent (e.g. on the GPU hosts, which
# only use the libraries built here).
set -euo pipefail
REF=${TC_REFERENCE_DIR:-/root/reference/taichi_with_jerks}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
ORDERS=${TC_REF_ORDERS:-"4 5 8 10 12 16 20 24"}
if [ ! -d "$REF" ]; then
  echo "build_ref: reference not present at $REF; skipping"
  exit 0
fi
mkdir -p "$OUT"
FLAGS="-O3 -fopenmp -fcx-limited-range -funroll-loops -Wno-format -w -fPIC -shared \
  -DFMM -DSIMD_P2P -DDOUBLE_P2P -DSIMD_M2L -DMINIBALL -DHUGE=3.40282347e+38F"
pids=()
for P in $ORDERS; do
  for V in v4 v3; do
    tgt="$OUT/libtcref_p${P}_${V}.so"
    if [ -f "$tgt" ] && [ "$tgt" -nt "$HERE/ref_harness.cpp" ]; then continue; fi
    g++ -march=x86-64-${V} $FLAGS -DEXPANSION=$P -I"$REF" "$HERE/ref_harness.cpp" -o "$tgt.tmp" \
      && mv "$tgt.tmp" "$tgt" &
    pids+=($!)
  done
done
# Master (non-jerk) variant: the top-level sources with the master Makefile's
# flags (Makefile:1-12, minus DEBUG and the HDF5 I/O) plus -DFMM so that both
# of get_force_and_potential's branches can be driven (ref_master_harness.cpp).
MREF=${TC_REFERENCE_MASTER_DIR:-$(dirname "$REF")}
MORDERS=${TC_REFM_ORDERS:-"8 12"}
MFLAGS="-O3 -fopenmp -fcx-limited-range -funroll-loops -Wno-format -w -fPIC -shared \
  -DFMM -DGRADIENT -DOUTPUTPOT -DOUTPUTACC -DSIMD_P2P -DDOUBLE_P2P -DMINIBALL -DHUGE=3.40282347e+38F"
if [ -f "$MREF/kernel_fast_lean.h" ]; then
  for P in $MORDERS; do
    for V in v4 v3; do
      tgt="$OUT/libtcrefm_p${P}_${V}.so"
      if [ -f "$tgt" ] && [ "$tgt" -nt "$HERE/ref_master_harness.cpp" ]; then continue; fi
      g++ -march=x86-64-${V} $MFLAGS -DEXPANSION=$P -I"$MREF" "$HERE/ref_master_harness.cpp" \
        -o "$tgt.tmp" && mv "$tgt.tmp" "$tgt" &
      pids+=($!)
    done
  done
fi
# The snapshot I/O (io.h) over libhdf5, when it is installed: the jerk build
# with the jerk Makefile's output flags (Makefile:2-3,9), the same without
# them, and the master io.h with the master Makefile's (Makefile:5-6,12).
# EXPANSION only sizes Cell arrays that io.h does not touch.
H5=${TC_HDF5_DIR:-/opt/conda}
if [ -f "$H5/include/hdf5.h" ]; then
  IOFLAGS="-O2 -fPIC -shared -w -DEXPANSION=4 -DH5_USE_16_API -DHUGE=3.40282347e+38F \
    -I$H5/include -L$H5/lib -lhdf5 -Wl,-rpath,$H5/lib"
  OUTS="-DOUTPUTPOT -DOUTPUTACC -DDEBUG"
  for spec in "jerk:$REF:$OUTS" "jerk_plain:$REF:" "master:$MREF:$OUTS"; do
    IFS=: read -r name dir defs <<< "$spec"
    tgt="$OUT/libtcrefio_${name}.so"
    if [ -f "$tgt" ] && [ "$tgt" -nt "$HERE/ref_io_harness.cpp" ]; then continue; fi
    g++ $defs -I"$dir" "$HERE/ref_io_harness.cpp" $IOFLAGS -o "$tgt.tmp" && mv "$tgt.tmp" "$tgt" &
    pids+=($!)
  done
fi
for p in "${pids[@]}"; do wait "$p"; done
echo "build_ref: libraries in $OUT"
