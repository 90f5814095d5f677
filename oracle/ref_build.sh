#!/bin/bash
# Compiles the reference's pure-host headers (core/FlashGraph.h, core/flash_compactor.h,
# core/reporting.h), where they lie under the reference checkout, into oracle/_ref/, wrapped
# by oracle/ref_harness.cpp. Nothing is copied: g++ reads the headers in place. Test
# infrastructure only (the tests skip what needs it when oracle/_ref is absent).
# Usage: bash oracle/ref_build.sh [reference_root]   (default /root/reference)
set -euo pipefail
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -f "$REF/core/FlashGraph.h" ]; then
  echo "ref_build: $REF/core/FlashGraph.h not found; oracle/_ref not built" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O2 -std=c++17 -fPIC -shared -w -I "$REF/core" -o "$HERE/_ref/libref_core.so" "$HERE/ref_harness.cpp"
echo "ref_build: built $HERE/_ref/libref_core.so"
