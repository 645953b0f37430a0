#!/bin/bash
# Builds oracle/_ref/ref_netgen from the reference's own host model-layer
# sources (macro.cpp, model/model.cpp, model/netgen.cpp) where they lie under
# /root/reference, plus the harness oracle/ref_netgen_main.cpp.  Nothing is
# copied from the reference tree; outputs go to oracle/_ref/ only.
# TEST INFRASTRUCTURE ONLY.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${MGB_REFERENCE_DIR:-/root/reference}"
if [ ! -d "$REF/src/model" ]; then
    echo "build_ref: reference sources not found at $REF; skipping" >&2
    exit 0
fi
OUT="$HERE/_ref"
mkdir -p "$OUT"
g++ -std=c++14 -O2 -w -I"$REF/src" \
    "$HERE/ref_netgen_main.cpp" \
    "$REF/src/macro.cpp" "$REF/src/model/model.cpp" "$REF/src/model/netgen.cpp" \
    -o "$OUT/ref_netgen"
echo "build_ref: built $OUT/ref_netgen"
