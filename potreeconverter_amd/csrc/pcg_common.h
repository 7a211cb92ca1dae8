// Shared host/device definitions for the gfx950 chunker + indexer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#define PCG_CHECK(expr)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      pcg::set_error(std::string(#expr) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + \
                     ":" + std::to_string(__LINE__));                                         \
      return 1;                                                                               \
    }                                                                                         \
  } while (0)

namespace pcg {

void set_error(const std::string& msg);
const std::string& last_error();

constexpr int kWave = 64;

// Record transcoding op (LAS record -> output record), one per attribute.
enum TranscodeMode : int32_t {
  TC_COPY = 0,        // copy len bytes
  TC_LOW3 = 1,        // src byte & 7            (return number)
  TC_MID3 = 2,        // (src byte >> 3) & 7     (number of returns)
  TC_LOW5 = 3,        // src byte & 31           (classification, formats 0-5)
};

// Attribute statistic kinds (min/max in double of one element).
enum StatKind : int32_t {
  ST_U8 = 0, ST_I8 = 1, ST_U16 = 2, ST_I16 = 3, ST_U32 = 4, ST_I32 = 5, ST_U64 = 6, ST_I64 = 7,
  ST_F32 = 8, ST_F64 = 9, ST_LOW3 = 10, ST_MID3 = 11, ST_LOW5 = 12, ST_NONE = 13
};

constexpr int kMaxOps = 24;
constexpr int kMaxStats = 32;

struct TranscodeOp {
  int32_t src;   // byte offset in the LAS record
  int32_t dst;   // byte offset in the output record
  int32_t len;   // bytes
  int32_t mode;  // TranscodeMode
};

struct StatOp {
  int32_t src;   // byte offset in the LAS record
  int32_t kind;  // StatKind
};

struct ChunkerParams {
  double las_scale[3];
  double las_offset[3];
  double pos_scale[3];
  double pos_offset[3];
  double min[3];
  double cube;
  double g_over_cube;  // grid_size / cube (set by chunker_set_params)
  int32_t grid_size;
  int32_t bpp_in;
  int32_t bpp_out;
  int32_t num_ops;
  int32_t num_stats;
  int32_t class_src;  // byte offset of the classification byte (histogram), -1 none
  TranscodeOp ops[kMaxOps];
  StatOp stats[kMaxStats];
};

// order-preserving map double -> uint64 (for atomic min/max)
__host__ __device__ inline uint64_t ordered_bits(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double from_ordered_bits(uint64_t b) {
  b = (b & 0x8000000000000000ull) ? (b & 0x7FFFFFFFFFFFFFFFull) : ~b;
  double v;
  __builtin_memcpy(&v, &b, 8);
  return v;
}

}  // namespace pcg
