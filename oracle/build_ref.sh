#!/bin/bash
# Builds oracle/_ref/libgeoref.so: geogram's own Delaunay / ConvexCell / kd-tree
# code (the "Delaunay" pluggable software module that geogram's sources list in
# src/lib/geogram/delaunay/Delaunay.psm) with the driver oracle/ref_driver.cpp.
#
# Test infrastructure only (the reference, as checker and CPU baseline).  The
# module's files are assembled from $REF (default /root/reference) in a
# temporary directory — headers and sources in the manifest's order, geogram
# includes removed, GEOGRAM_PSM defined, as geogram's tools/extract_psm.sh
# assembles a module — compiled with the flags of the reference build
# (cmake/platforms/Linux-gcc.cmake: -frounding-math -ffp-contract=off), and
# only the shared library is kept.  Nothing from $REF is written to the
# repository.  The compiler is ROCm's clang++: in one translation unit the
# module expands GEO_FP_CONTRACT_OFF (basic/numeric.h:415-437) once per
# source file, which under GCC defines the same file-scope struct each time,
# and under clang is a pragma.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
PSM_DIR=$REF/src/lib/geogram/delaunay
if [ ! -f "$PSM_DIR/Delaunay.psm" ]; then
    echo "build_ref: $PSM_DIR/Delaunay.psm not found (REF=$REF)" >&2
    exit 1
fi
# the manifest is a list of shell variables (HEADERS, SOURCES, ...)
HEADERS=""
SOURCES=""
# shellcheck disable=SC1090
. "$PSM_DIR/Delaunay.psm"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
strip_file() {
    # drop the leading copyright comment and every geogram include
    awk 'BEGIN { s = 0 }
         { if (s == 0) { if ($0 ~ /\/\*.*/) s = 1 }
           else if (s == 1) { if ($0 ~ / \*\/.*/) s = 2 }
           else print }' "$PSM_DIR/$1" | grep -v '#include "' | grep -v '#include <geogram/' || true
}
{
    echo '#define GEOGRAM_PSM'
    echo '#ifndef GEO_STATIC_LIBS'
    echo '#define GEO_DYNAMIC_LIBS'
    echo '#endif'
    for f in $HEADERS; do echo "/* $f */"; strip_file "$f"; done
} > "$TMP/Delaunay_psm.h"
{
    echo '#include "Delaunay_psm.h"'
    for f in $SOURCES; do echo "/* $f */"; strip_file "$f"; done
} > "$TMP/Delaunay_psm.cpp"
mkdir -p "$OUT"
${CXX:-/opt/rocm/llvm/bin/clang++} -std=c++17 -O3 -fPIC -shared -fopenmp -frounding-math -ffp-contract=off \
    -w -I"$TMP" -o "$OUT/libgeoref.so.tmp" "$TMP/Delaunay_psm.cpp" "$HERE/ref_driver.cpp" \
    -ldl -lm -lpthread
mv "$OUT/libgeoref.so.tmp" "$OUT/libgeoref.so"
echo "build_ref: $OUT/libgeoref.so"
