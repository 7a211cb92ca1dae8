#!/bin/bash
# Builds the reference indexer (Converter/src/indexer.cpp with its samplers,
# HierarchyBuilder and vendored brotli encoder) straight from the read-only
# reference checkout, plus our ref_driver.cpp, into oracle/_ref/, and the
# vendored LASzip raw item readers behind ref_point_decode.cpp.
# The reference chunker is not built: it includes laszip_api.h, which needs
# the CMake-generated laszip_api_version.h that the checkout does not hold.
# Without TBB headers libstdc++ uses its serial PSTL backend, so the
# reference's std::sort(par_unseq, ...) is std::sort.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/Converter" ]; then
  echo "reference checkout not found at $REF; skipping"
  exit 0
fi
mkdir -p "$OUT/obj"
CONV="$REF/Converter"
INC="-I$CONV/include -I$CONV/modules -I$CONV/libs -I$CONV/libs/brotli/c/include"
CXXFLAGS="-O3 -std=gnu++20 -w $INC"
CFLAGS="-O3 -w -I$CONV/libs/brotli/c/include"
objs=()
for f in "$CONV"/libs/brotli/c/common/*.c "$CONV"/libs/brotli/c/enc/*.c; do
  o="$OUT/obj/$(basename "$f" .c).o"
  if [ ! -f "$o" ] || [ "$f" -nt "$o" ]; then gcc $CFLAGS -c "$f" -o "$o"; fi
  objs+=("$o")
done
for f in "$CONV/src/indexer.cpp" "$CONV/src/logger.cpp" "$CONV/modules/unsuck/unsuck_platform_specific.cpp" "$HERE/ref_driver.cpp"; do
  o="$OUT/obj/$(basename "$f" .cpp).o"
  if [ ! -f "$o" ] || [ "$f" -nt "$o" ]; then g++ $CXXFLAGS -c "$f" -o "$o" & fi
  objs+=("$o")
done
wait
g++ -o "$OUT/ref_indexer" "${objs[@]}" -lpthread
echo "built $OUT/ref_indexer"
# LASzip's raw point readers (header-only part of the vendored laszip that
# needs no generated header) behind a small decode driver
g++ -O2 -std=gnu++17 -w -I"$CONV/libs/laszip/src" -o "$OUT/ref_point_decode" "$HERE/ref_point_decode.cpp"
echo "built $OUT/ref_point_decode"
