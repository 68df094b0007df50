#!/usr/bin/env bash
# Builds the oracle (CPU restatement of the reference, test infrastructure
# only) into oracle/_build/liboracle.so. No FMA contraction, no fast-math:
# Rust never fuses multiply-add, and flattening decisions must match bit for bit.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
g++ -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -pthread \
    -Wall -Wextra -Wno-unused-parameter \
    -o "$here/_build/liboracle.so.tmp" "$here/orc_capi.cpp"
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
