// BROTLI node encoding of octree.bin (Converter/src/indexer.cpp:1110-1401,
// selected by --encoding BROTLI, Converter/src/main.cpp:60,125).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "attributes.h"

namespace pcg {

// One node with points, in completion order: its records in HBM.
struct EncodeInput {
  const uint8_t* d_records;
  int64_t count;
};

// Writer::writeAndUnload with encoding "BROTLI" (indexer.cpp:1453-1516):
// the GPU lays every node out as attribute arrays in position-code order,
// the host compresses the nodes (brotli quality 6, default window and
// mode) on a thread pool and appends them to `image` in input order.
// `sizes` receives each node's compressed byte size.
int encode_brotli(const Attributes& attributes, int32_t bpp, const std::vector<EncodeInput>& nodes,
                  std::vector<uint8_t>& image, std::vector<int64_t>& sizes, hipStream_t s);

// The system's libbrotlienc (resolved at run time); false with the error set
// when it is missing.
bool brotli_available();

}  // namespace pcg
