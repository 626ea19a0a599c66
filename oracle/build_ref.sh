#!/bin/bash
# Test infrastructure: compile the reference's CPU FDK sources in place into
# oracle/_ref/libfdkref.so through oracle/ref_harness.cpp.  Needs the reference
# checkout (default /root/reference); exits 0 with a message when it is absent.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${FDK_REFERENCE_DIR:-/root/reference}"
if [ ! -f "$REF/common/type.h" ] || [ ! -f "$REF/3dct/3dct.cpp" ]; then
  echo "build_ref: reference sources not found at $REF; skipping oracle/_ref"
  exit 0
fi
CXX="${REF_CXX:-/opt/rocm/lib/llvm/bin/clang++}"
mkdir -p "$HERE/_ref"
"$CXX" -std=c++11 -O2 -fPIC -shared -fdelayed-template-parsing -fopenmp -msse4.1 \
  -w -I"$REF" -I"$HERE" "$HERE/ref_harness.cpp" \
  -Wl,-rpath,/opt/rocm/lib/llvm/lib -o "$HERE/_ref/libfdkref.so"
echo "build_ref: wrote $HERE/_ref/libfdkref.so"
