#!/usr/bin/env bash
# Builds the reference's acoustic CPU solvers from their own sources into
# oracle/_ref/ and optionally runs them to produce golden receiver traces and
# energy histories (test infrastructure: golden vectors for the oracle and the CPU baseline).
#
# What builds from the reference's own few source files:
#   * singlecore/  : header-only C++20, builds with clang++ (g++ 11 rejects the
#                    _Float16 mention at cpu_acoustic.cpp:367).
#                    singlecore_acoustic_f*  : the program as written (its main
#                                              never parses the command line:
#                                              600^2, Nt = 60000, no compensation)
#                    singlecore_cli_f*       : the same program behind
#                                              oracle/ref_cli.cpp, which calls the
#                                              reference's own option parser first
#                                              (-Nt, -cpst, -multiple)
#   * mpi/         : the same scheme over MPI ranks; it links the MPI library of
#                    the image (MPICH under /opt/conda) when present and is used
#                    only as the multi-core CPU baseline, never as an oracle.
#   * cuda/        : needs cuda_fp16.h / cuda_bf16.h (namespace_type.cuh) even
#                    for the host twin -> unbuildable without stand-in headers.
#
# Flags: the reference makefile uses -O3 -march=native (singlecore/build/makefile:44)
# with asserts on.  -march=native lets the compiler contract a*b+c into FMA, which
# makes the rounding depend on the build host, so the golden builds pin IEEE
# semantics with -ffp-contract=off, no -march and -DNDEBUG; the "native" builds
# keep the reference's own flags and are used for timing only.
#
# Usage: oracle/ref_build.sh            build only
#        oracle/ref_build.sh run        build, then run f64/f32/f16 of the program
#                                       as written (60000 steps each, ~15-40 min)
#        oracle/ref_build.sh run-cpst   build, then run f64/f32 with -cpst 3 and
#                                       -cpst 6 for 20000 steps (~3 min)
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
CXX=${CXX:-/opt/rocm/lib/llvm/bin/clang++}
MPI_ROOT=${MPI_ROOT:-/opt/conda}

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
for P in 64 32; do
  "$CXX" -O3 -std=c++20 -ffp-contract=off -DNDEBUG -DPRCSFLAG=$P -I "$INC" \
         -I "$REF/singlecore/src/main_cpp" "$HERE/ref_cli.cpp" -o "$OUT/singlecore_cli_f$P"
done
# timing builds with the reference's own flags (asserts on); -march is left to
# the compiler's default x86-64 so the binary runs on any host it travels to
"$CXX" -O3 -std=c++20 -DPRCSFLAG=32 -I "$INC" -I "$REF/singlecore/src/main_cpp" \
       "$HERE/ref_cli.cpp" -o "$OUT/singlecore_cli_f32_refflags"
if [ -f "$MPI_ROOT/include/mpi.h" ] && [ -f "$MPI_ROOT/lib/libmpi.so" ]; then
  # the system libstdc++ must precede the one under $MPI_ROOT/lib
  "$CXX" -O3 -std=c++20 -DPRCSFLAG=32 -I "$REF/mpi/src/cpp" -I "$MPI_ROOT/include" \
         "$REF/mpi/src/main_cpp/cpu_acoustic.cpp" "$MPI_ROOT/lib/libmpi.so" \
         -Wl,-rpath,/usr/lib/x86_64-linux-gnu -Wl,-rpath,"$MPI_ROOT/lib" \
         -o "$OUT/mpi_acoustic_f32_refflags"
else
  echo "no MPI library under $MPI_ROOT: the MPI solver is not built" >&2
fi

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

if [ "${1:-}" = "run-cpst" ]; then
  NT=${NT:-20000}
  for P in 64 32; do
    for C in 3 6; do
      RUN="$OUT/run_f${P}_cpst$C"
      rm -rf "$RUN"
      mkdir -p "$RUN/a"
      ( cd "$RUN/a" && "$OUT/singlecore_cli_f$P" -Nt "$NT" -cpst "$C" > "$RUN/stdout.txt" 2>&1 ) &
    done
  done
  wait
  for P in 64 32; do
    for C in 3 6; do
      RUN="$OUT/run_f${P}_cpst$C"
      D=$(dirname "$(find "$RUN/result" -name E.txt | head -1)")
      for f in Vx Vy P E; do cp "$D/$f.txt" "$OUT/golden_f${P}_cpst${C}_$f.txt"; done
    done
  done
fi
