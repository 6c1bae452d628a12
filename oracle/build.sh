#!/bin/sh
# Build recipe for the CPU oracle (test infrastructure only).
# Flags follow the reference build (tools/build_fortran_reference.sh:29-30:
# -O2 -fno-tree-vectorize) plus -ffp-contract=off, so float32 expressions are
# evaluated exactly as written.  Outputs go to oracle/build/.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/build"
mkdir -p "$OUT"
CFLAGS="-O2 -fno-tree-vectorize -ffp-contract=off -fno-fast-math -fPIC -std=c11 -D_GNU_SOURCE -Wall -Wno-unused-function"
SRC="$HERE/ora_core.c $HERE/ora_energy.c $HERE/ora_ep.c $HERE/ora_surrogate.c $HERE/ora_press.c $HERE/ora_sums.c $HERE/ora_weights.c $HERE/ora_intra.c $HERE/ora_data.c"
gcc $CFLAGS -shared -o "$OUT/liboracle_glibc.so.tmp" $SRC -lm
mv "$OUT/liboracle_glibc.so.tmp" "$OUT/liboracle_glibc.so"
gcc $CFLAGS -DORACLE_CR_MATH -shared -o "$OUT/liboracle_cr.so.tmp" $SRC -lm
mv "$OUT/liboracle_cr.so.tmp" "$OUT/liboracle_cr.so"
# all-cores variant of the glibc flavour for the CPU baseline (OpenMP over the
# loops the Fortran marks !$omp parallel do)
gcc $CFLAGS -fopenmp -DORACLE_OMP -shared -o "$OUT/liboracle_omp.so.tmp" $SRC -lm
mv "$OUT/liboracle_omp.so.tmp" "$OUT/liboracle_omp.so"
