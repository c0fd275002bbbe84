#!/usr/bin/env bash
# Builds the reference's single-core acoustic solver from its own sources
# (singlecore/src/main_cpp/cpu_acoustic.cpp + singlecore/src/cpp/*.hpp) into
# oracle/_ref/, one binary per precision, and optionally runs them to produce
# golden receiver traces and energy histories.
#
# Only this path of the reference compiles from its own few source files:
#   * singlecore/  : header-only C++20, builds with clang++ (g++ 11 rejects the
#                    _Float16 mention at cpu_acoustic.cpp:367).
#   * mpi/         : needs an MPI library -> treated as unbuildable.
#   * cuda/        : needs cuda_fp16.h / cuda_bf16.h (namespace_type.cuh) even
#                    for the host twin -> unbuildable without stand-in headers.
#
# Flags: the reference makefile uses -O3 -march=native (singlecore/build/makefile:44).
# -march=native lets the compiler contract a*b+c into FMA, which makes the
# rounding depend on the build host. The golden build pins IEEE semantics with
# -ffp-contract=off and no -march, so the oracle can be compared bitwise.
# A second "native" build keeps the reference's own flags for timing.
#
# Usage: oracle/ref_build.sh            build only
#        oracle/ref_build.sh run        build, then run f64/f32/f16 (60000 steps
#                                       each, ~15-40 min) and collect outputs
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
CXX=${CXX:-/opt/rocm/lib/llvm/bin/clang++}

if [ ! -d "$REF/singlecore" ]; then
  echo "reference tree not present at $REF; nothing to build" >&2
  exit 0
fi
mkdir -p "$OUT"
SRC="$REF/singlecore/src/main_cpp/cpu_acoustic.cpp"
INC="$REF/singlecore/src/cpp"
for P in 64 32 16; do
  "$CXX" -O3 -std=c++20 -ffp-contract=off -DNDEBUG -DPRCSFLAG=$P -I "$INC" "$SRC" \
         -o "$OUT/singlecore_acoustic_f$P"
done
"$CXX" -O3 -std=c++20 -march=native -DPRCSFLAG=32 -I "$INC" "$SRC" \
       -o "$OUT/singlecore_acoustic_f32_native"

if [ "${1:-}" = "run" ]; then
  for P in 64 32 16; do
    RUN="$OUT/run_f$P"
    mkdir -p "$RUN/a"
    ( cd "$RUN/a" && "$OUT/singlecore_acoustic_f$P" > "$RUN/stdout.txt" 2>&1 ) &
  done
  wait
  for P in 64 32 16; do
    RUN="$OUT/run_f$P"
    D=$(dirname "$(find "$RUN/result" -name E.txt | head -1)")
    for f in Vx Vy P E; do cp "$D/$f.txt" "$OUT/golden_f${P}_$f.txt"; done
  done
fi
