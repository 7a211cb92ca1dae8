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
# The samplers alone (header-only) behind a subtree driver: one
# Sampler::sample call on a subtree read from a file, callbacks recorded
g++ $CXXFLAGS -o "$OUT/ref_sampler" "$HERE/ref_sampler.cpp" "$OUT/obj/unsuck_platform_specific.o" \
  "$OUT/obj/logger.o" -lpthread
echo "built $OUT/ref_sampler"
# The reference indexer with the GPU samplers of INTEGRATION.md plugged in
# (its PotreeGpuSampler.h block, extracted here), linked to the library
LIBDIR="$HERE/../potreeconverter_amd"
awk '/^\/\/ PotreeGpuSampler.h/{on=1} on&&/^```/{on=0} on' "$HERE/../INTEGRATION.md" > "$OUT/PotreeGpuSampler.h"
if [ -f "$LIBDIR/libpotree_gpu.so" ]; then
  g++ $CXXFLAGS -DREF_GPU_SAMPLERS -I"$OUT" -I"$HERE/../include" -c "$HERE/ref_driver.cpp" -o "$OUT/obj/ref_driver_gpu.o"
  gpu_objs=()
  for o in "${objs[@]}"; do [ "$(basename "$o")" = ref_driver.o ] || gpu_objs+=("$o"); done
  g++ -o "$OUT/ref_indexer_gpu" "${gpu_objs[@]}" "$OUT/obj/ref_driver_gpu.o" -L"$LIBDIR" -lpotree_gpu \
    -Wl,-rpath,'$ORIGIN/../../potreeconverter_amd' -lpthread
  echo "built $OUT/ref_indexer_gpu"
fi
# The same glue over a CPU stand-in of pcg_sample_subtree made of the
# reference samplers (host code only, AddressSanitizer)
g++ -O1 -g -std=gnu++20 -w -fsanitize=address $INC -DREF_GPU_SAMPLERS -I"$OUT" -I"$HERE/../include" \
  -c "$HERE/ref_driver.cpp" -o "$OUT/obj/ref_driver_stub.o"
g++ -O1 -g -std=gnu++20 -w -fsanitize=address $INC -I"$HERE/../include" -c "$HERE/ref_sampler_stub.cpp" \
  -o "$OUT/obj/ref_sampler_stub.o"
stub_objs=()
for o in "${objs[@]}"; do [ "$(basename "$o")" = ref_driver.o ] || stub_objs+=("$o"); done
g++ -fsanitize=address -o "$OUT/ref_indexer_stub" "${stub_objs[@]}" "$OUT/obj/ref_driver_stub.o" \
  "$OUT/obj/ref_sampler_stub.o" -lpthread
echo "built $OUT/ref_indexer_stub"
