#!/bin/bash
# Reference builds, compiled from the reference's own C sources where they lie under
# $REF_DIR (never copied into this repository).  Outputs only under oracle/_ref/, which is
# git-ignored and travels with the tree to the machine that runs the benchmark.
#
# Every binary is the reference's timing driver linked against one backend, the pattern of
# bench/geom/Makefile:47-48 (driver.o + one implementation):
#
#   _ref/<isa>/st/<kernel>   bench/geom/driver.c + one single-core kernel of the reference's
#                            best-exemplar set (bench/geom/best/kernels/<kernel>.c)
#   _ref/<isa>/mt/<kernel>   bench/mt/driver.c + one OpenMP kernel (bench/mt/impl_4/<kernel>.c)
#   _ref/geom_driver_mi355x  bench/geom/driver.c + this repository's libfft3d_mi355x.so: the
#                            drop-in boundary exercised by the reference's own harness
#
# <isa>: v4 = x86-64-v4 (AVX-512), v3 = x86-64-v3 (AVX2); the caller picks by CPU flags.
# FFTW (ext/fftw-3.3.10) is not built: it needs its configure-generated config.h.
set -euo pipefail
REF_DIR=${REF_DIR:-/root/reference}
CC=${CC:-gcc}
HERE=$(cd "$(dirname "$0")" && pwd)
GEOM=$REF_DIR/bench/geom
MT=$REF_DIR/bench/mt
if [ ! -f "$GEOM/driver.c" ]; then
    echo "build_ref.sh: $REF_DIR absent; keeping prebuilt _ref/ (if any)"
    exit 0
fi
mkdir -p "$HERE/_ref"

LIB=$HERE/../opt-fft_amd/lib
if [ -f "$LIB/libfft3d_mi355x.so" ]; then
    # The driver includes its own fft3d_api.h; the library exports the same six symbols.
    $CC -O2 -std=gnu11 -w -I"$GEOM" -o "$HERE/_ref/geom_driver_mi355x" "$GEOM/driver.c" \
        -L"$LIB" -lfft3d_mi355x -Wl,-rpath,'$ORIGIN/../../opt-fft_amd/lib' -lm
    echo "build_ref.sh: built $HERE/_ref/geom_driver_mi355x"
else
    echo "build_ref.sh: $LIB/libfft3d_mi355x.so not built; skipping geom_driver_mi355x"
fi

BASE="-O3 -std=gnu11 -fno-math-errno -funroll-loops -w"
# Single-core kernels: the best-exemplar library's winner and alternate per geometry
# (bench/geom/best/fft3d_best.c's table).
ST_K="L6_pfa L6_unrolled L8_batchsimd L8_fusedaxes L17_matrixsimd L17_rader
      L36_mixedradix L36_pfa L36_pencilfused"
# All-core kernels of the reference's multi-threaded round (bench/mt/impl_4).
MT_K="L6_pfa L8_fusedaxes L8_batchsimd L17_winograd L17_matrixsimd L36_mixedradix L36_pfa"

for isa in ${ISAS:-v4 v3}; do
    out=$HERE/_ref/$isa
    mkdir -p "$out/st" "$out/mt"
    march="-march=x86-64-$isa"
    $CC $BASE $march -I"$GEOM" -c -o "$out/st/driver.o" "$GEOM/driver.c"
    $CC $BASE $march -fopenmp -I"$MT" -c -o "$out/mt/driver.o" "$MT/driver.c"
    for k in $ST_K; do
        $CC $BASE $march -I"$GEOM/best" -I"$GEOM/best/kernels" -I"$GEOM" \
            -o "$out/st/$k" "$GEOM/best/kernels/$k.c" "$out/st/driver.o" -lm &
    done
    for k in $MT_K; do
        $CC $BASE $march -fopenmp -I"$MT" -I"$MT/impl_4" \
            -o "$out/mt/$k" "$MT/impl_4/$k.c" "$out/mt/driver.o" -lm -lpthread &
    done
    wait
    echo "build_ref.sh: built $out/st/{$ST_K} $out/mt/{$MT_K}" | tr -s ' \n' ' '
    echo
done
