#!/bin/bash
# Reference CPU baseline, compiled from the reference's own C sources where they lie under
# $REF_DIR (never copied into this repository).  Outputs only under oracle/_ref/, which is
# git-ignored and travels with the tree to the machine that runs the benchmark.
#
#   _ref/<isa>/ref_bench_st  single thread: the reference's best-exemplar library
#                            bench/geom/best/fft3d_best.c (race-at-create over its kernels)
#   _ref/<isa>/ref_bench_mt  all cores: the reference's OpenMP kernels, bench/mt/impl_4
#                            (L6_pfa, L17_winograd, L36_mixedradix) and bench/mt/impl_3
#                            L8_fusedaxes (impl_4 has no L=8 kernel)
# <isa>: v4 = x86-64-v4 (AVX-512), v3 = x86-64-v3 (AVX2); the caller picks by CPU flags.
# The timing harness ref_bench.c is this repository's own code.
set -euo pipefail
REF_DIR=${REF_DIR:-/root/reference}
CC=${CC:-gcc}
HERE=$(cd "$(dirname "$0")" && pwd)
BEST=$REF_DIR/bench/geom/best
if [ ! -d "$BEST/kernels" ]; then
    echo "build_ref.sh: $REF_DIR absent; keeping prebuilt _ref/ (if any)"
    exit 0
fi
BASE="-O3 -std=gnu11 -fno-math-errno -funroll-loops -w"
rename() {  # prefix -> -D flags renaming the six ABI symbols
    local k=$1
    echo "-Dfft3d_name=${k}_name -Dfft3d_description=${k}_description" \
         "-Dfft3d_supports=${k}_supports -Dfft3d_create=${k}_create" \
         "-Dfft3d_execute=${k}_execute -Dfft3d_destroy=${k}_destroy"
}
# prefixes named by fft3d_best.c's geometry table
BEST_K="k6:L6_pfa k8:L8_batchsimd k13:L13_direct k17:L17_matrixsimd k23:L23_rader
        k36:L36_mixedradix k45:L45_pfa k64:L64_blocked k8b:L8_fusedaxes
        k36b:L36_pencilfused k64b:L64_radix8 k6b:L6_unrolled k13b:L13_rader
        k17b:L17_rader k23b:L23_matrixsimd k36c:L36_pfa k45b:L45_mixedradix"
MT_K="m6:$REF_DIR/bench/mt/impl_4/L6_pfa m8:$REF_DIR/bench/mt/impl_3/L8_fusedaxes
      m17:$REF_DIR/bench/mt/impl_4/L17_winograd m36:$REF_DIR/bench/mt/impl_4/L36_mixedradix"

for isa in ${ISAS:-v4 v3}; do
    out=$HERE/_ref/$isa
    mkdir -p "$out/best" "$out/mt"
    march="-march=x86-64-$isa"
    objs=()
    for kv in $BEST_K; do
        k=${kv%%:*}; src=${kv#*:}
        $CC $BASE $march -I"$BEST" -I"$BEST/kernels" -I"$REF_DIR/bench/geom" $(rename "$k") \
            -c -o "$out/best/$k.o" "$BEST/kernels/$src.c" &
        objs+=("$out/best/$k.o")
    done
    $CC $BASE $march -I"$BEST" -c -o "$out/best/fft3d_best.o" "$BEST/fft3d_best.c" &
    wait
    $CC $BASE $march -I"$BEST" -DREF_BEST -o "$out/ref_bench_st" "$HERE/ref_bench.c" \
        "$out/best/fft3d_best.o" "${objs[@]}" -lm
    mobjs=()
    for kv in $MT_K; do
        k=${kv%%:*}; src=${kv#*:}
        $CC $BASE -fopenmp $march -I"$REF_DIR/bench/mt" -I"$(dirname "$src")" $(rename "$k") \
            -c -o "$out/mt/$k.o" "$src.c" &
        mobjs+=("$out/mt/$k.o")
    done
    wait
    $CC $BASE -fopenmp $march -DREF_MT -o "$out/ref_bench_mt" "$HERE/ref_bench.c" \
        "${mobjs[@]}" -lm -lpthread
    echo "build_ref.sh: built $out/ref_bench_st $out/ref_bench_mt"
done
