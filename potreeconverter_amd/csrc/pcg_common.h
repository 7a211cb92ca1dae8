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
  TC_RET14 = 4,       // legacy return number of a format 6-10 returns byte
  TC_NRET14 = 5,      // legacy number of returns of a format 6-10 returns byte
  TC_LOW4 = 6,        // src byte & 15           (classification flags, formats 6-10)
};

// Attribute statistic kinds (min/max in double of one element).
enum StatKind : int32_t {
  ST_U8 = 0, ST_I8 = 1, ST_U16 = 2, ST_I16 = 3, ST_U32 = 4, ST_I32 = 5, ST_U64 = 6, ST_I64 = 7,
  ST_F32 = 8, ST_F64 = 9, ST_LOW3 = 10, ST_MID3 = 11, ST_LOW5 = 12, ST_RET14 = 13, ST_NRET14 = 14,
  ST_LOW4 = 15, ST_NONE = 16
};

// The statistics of the standard attributes of each point format, in the
// order of ChunkerParams::stats (the count kernel folds them with
// compile-time field offsets; extra-byte statistics follow them).
struct StdStat {
  int32_t src;
  int32_t kind;
};
__host__ __device__ constexpr int32_t std_stat_count(int32_t fmt) {
  return fmt == 0 ? 10 : fmt == 1 ? 11 : fmt == 2 ? 13 : fmt == 3 ? 14 : fmt == 4 ? 11 : fmt == 5 ? 14
       : fmt == 6 ? 12 : fmt == 7 ? 15 : 0;
}
__host__ __device__ constexpr StdStat std_stat(int32_t fmt, int32_t i) {
  // position X, Y, Z (the stored integers; monotone in the coordinate)
  if (i < 3) return StdStat{4 * i, ST_I32};
  if (fmt <= 5) {
    // legacy record (formats 0-5): intensity, returns byte, class byte,
    // scan angle rank, user data, point source id, [gps time], [rgb]
    constexpr StdStat legacy[7] = {{12, ST_U16}, {14, ST_LOW3}, {14, ST_MID3}, {15, ST_LOW5},
                                   {16, ST_I8},  {17, ST_U8},   {18, ST_U16}};
    if (i < 10) return legacy[i - 3];
    const bool gps = fmt == 1 || fmt == 3 || fmt == 4 || fmt == 5;
    if (gps && i == 10) return StdStat{20, ST_F64};
    const int32_t rgb0 = fmt == 2 ? 20 : 28;
    const int32_t k = i - (gps ? 11 : 10);
    return StdStat{rgb0 + 2 * k, ST_U16};
  }
  // formats 6-7: intensity, returns byte (legacy return number and number
  // of returns), classification flags, class, user data, scan angle, point
  // source id, gps time, [rgb]
  constexpr StdStat ext[9] = {{12, ST_U16}, {14, ST_RET14}, {14, ST_NRET14}, {15, ST_LOW4}, {16, ST_U8},
                              {17, ST_U8},  {18, ST_I16},   {20, ST_U16},    {22, ST_F64}};
  if (i < 12) return ext[i - 3];
  return StdStat{30 + 2 * (i - 12), ST_U16};
}
// bytes of the standard part of a record (formats 4-5: without the wave
// packet, which carries no statistic)
__host__ __device__ constexpr int32_t std_record_bytes(int32_t fmt) {
  return fmt == 0 ? 20 : fmt == 1 ? 28 : fmt == 2 ? 26 : fmt == 3 ? 34 : fmt == 4 ? 28 : fmt == 5 ? 34
       : fmt == 6 ? 30 : fmt == 7 ? 36 : 12;
}

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
  int32_t class_src;     // byte offset of the classification byte (histogram), -1 none
  int32_t class_mask;    // 31 (formats 0-5: 5-bit class) or 255 (formats 6-10)
  int32_t point_format;  // 0-7: stats[0 .. std_stat_count) are that format's; -1: none
  int32_t pad0;
  TranscodeOp ops[kMaxOps];
  StatOp stats[kMaxStats];
};

// Legacy 3-bit return number / number of returns that laszip derives from
// the 4-bit fields of a point format 6-10 record
// (laszip/src/lasreaditemraw.hpp:220-243, LASreadItemRaw_POINT14_LE).
__host__ __device__ inline uint32_t legacy_return_number(uint32_t b) {
  const uint32_t r = b & 15, nr = (b >> 4) & 15;
  if (nr > 7) return r > 6 ? (r >= nr ? 7u : 6u) : r;
  return r & 7;  // stored into a 3-bit field
}
__host__ __device__ inline uint32_t legacy_number_of_returns(uint32_t b) {
  const uint32_t nr = (b >> 4) & 15;
  return nr > 7 ? 7u : nr;
}

// order-preserving map double -> uint64 (for atomic min/max)
__host__ __device__ inline uint64_t ordered_bits(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline uint64_t ordered_bits_from_raw(uint64_t b) {
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double from_ordered_bits(uint64_t b) {
  b = (b & 0x8000000000000000ull) ? (b & 0x7FFFFFFFFFFFFFFFull) : ~b;
  double v;
  __builtin_memcpy(&v, &b, 8);
  return v;
}

}  // namespace pcg
