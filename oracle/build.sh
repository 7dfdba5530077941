#!/bin/bash
# Builds the CPU oracle (test infrastructure) into oracle/_build/.
# -ffp-contract=off and no fast-math: the reference's host build uses the
# default x86-64 ISA without FMA and without fast-math (gcc.cmake:7-10).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
g++ -O2 -std=c++17 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall \
    "$here/fp_oracle.cpp" -o "$here/_build/libfp_oracle.so.tmp"
mv "$here/_build/libfp_oracle.so.tmp" "$here/_build/libfp_oracle.so"
