#!/usr/bin/env bash
# Builds the oracle: the C++ restatement (oracle/liboracle.so) and, when the
# reference sources are present, the reference harness oracle/_ref/ref_harness
# compiled from the reference's headers where they lie (nothing is copied).
# The reference CPU path is header-only C++17 + rocThrust host containers, so
# hipcc compiles it host-side; clang keeps fill_vertices' (x, y) order.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ref="${DGC_REFERENCE:-/root/reference}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"

g++ -O3 -std=c++17 -fPIC -shared -Wall -Wno-unknown-pragmas -o "$here/liboracle.so.tmp" "$here/dg_oracle.cpp"
mv "$here/liboracle.so.tmp" "$here/liboracle.so"
# the same restatement with its cell loops threaded (all-core CPU baseline)
g++ -O3 -std=c++17 -fPIC -shared -Wall -fopenmp -o "$here/liboracle_omp.so.tmp" "$here/dg_oracle.cpp"
mv "$here/liboracle_omp.so.tmp" "$here/liboracle_omp.so"

if [[ -d "$ref/src" ]]; then
  mkdir -p "$here/_ref"
  inc=(-I"$ref/src" -I"$ref/src/mesh" -I"$ref/examples/maxwell_2d" -I"$ref/examples/advection_2d" -I"$ref/examples/advection_1d"
       -I"$ref/examples/poisson_2d")
  # checked build (reference asserts on) for golden vectors
  "$HIPCC" -x hip --offload-arch=gfx950 -std=c++17 -O2 -w "${inc[@]}" \
      -o "$here/_ref/ref_harness.tmp" "$here/ref_harness.cpp"
  mv "$here/_ref/ref_harness.tmp" "$here/_ref/ref_harness"
  # timing build (-O3 -DNDEBUG) for the CPU baseline
  "$HIPCC" -x hip --offload-arch=gfx950 -std=c++17 -O3 -DNDEBUG -w "${inc[@]}" \
      -o "$here/_ref/ref_bench.tmp" "$here/ref_harness.cpp"
  mv "$here/_ref/ref_bench.tmp" "$here/_ref/ref_bench"
fi
