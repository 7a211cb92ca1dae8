// Seeded synthetic pre-chunked point sets: the input of doIndexing
// (Converter/src/indexer.cpp:53-179, getChunks) without a chunking pass.
//
// The cube [0, S)^3 (S = cube_size metres) is split into the 16^3 cells of
// octree level 4; a chunk is one cell, named "r" + four octal digits (x -> 4,
// y -> 2, z -> 1 per level, chunker_countsort_laszip.cpp:95-108), so the
// ascending chunk name is the ascending Morton index m of the cell.  Chunk i
// of num_chunks (a power of two) is the cell m = i * 4096 / num_chunks.
// Each chunk is a local patch: X, Y uniform over the cell, Z a value-noise
// heightfield inside the cell's height range, and on a quarter of the
// patches one box "building" whose roof lifts the points above it.  Records
// are in the converter's output layout for LAS point format 2 plus three
// extra-bytes attributes (34 bytes): int32 X, Y, Z (scale `scale`, offset
// 0), intensity, return number, number of returns, classification
// (2 ground, 6 roof), scan angle rank, user data, point source id, rgb
// (height colour map + N(0, 2000)), float32 N(0, 1), uint16 uniform, int8
// uniform; N(0, 1) is Irwin-Hall (sum of 12 uniforms - 6).
//
// Point counts are log-normal, round(median * exp(sigma * g)) clipped to
// [min_count, max_count], g ~ N(0, 1) by Box-Muller from the chunk's hash;
// they are drawn on the host (libm exp/log/cos).  Every record is a pure
// function of (seed, chunk, index) through the counter hash of lasgen.h
// and IEEE binary64 + - * / (every translation unit is compiled with
// -ffp-contract=off), so the host and the GPU write the same bytes.
#pragma once

#include <stdint.h>

#include "lasgen.h"

#if defined(__HIPCC__)
#define CHUNKGEN_HD __host__ __device__ inline
#else
#define CHUNKGEN_HD inline
#endif

struct ChunkGenParams {
  uint64_t seed;
  double cube_size;   // metres
  double scale;       // position scale of the records (offset 0)
  double median;      // point count median
  double sigma;       // log-normal shape
  int64_t min_count;  // count clip range
  int64_t max_count;
  int32_t num_chunks;  // power of two, 1 .. 4096
  int32_t pad;
};

constexpr int kChunkGenBytes = 34;
constexpr int kChunkGenCells = 4096;  // 16^3 cells of level 4

// Morton index of chunk i
CHUNKGEN_HD int32_t chunkgen_morton(const ChunkGenParams& p, int32_t i) {
  return i * (kChunkGenCells / p.num_chunks);
}

// cell (x, y, z) of Morton index m: level l (1..4) digit = (m >> 3 (4 - l)) & 7
CHUNKGEN_HD void chunkgen_cell(int32_t m, int32_t& x, int32_t& y, int32_t& z) {
  x = y = z = 0;
  for (int l = 0; l < 4; l++) {
    const int32_t d = (m >> (3 * (3 - l))) & 7;
    x = (x << 1) | (d >> 2);
    y = (y << 1) | ((d >> 1) & 1);
    z = (z << 1) | (d & 1);
  }
}

// chunk name "rABCD" of Morton index m (6 bytes with the terminator)
CHUNKGEN_HD void chunkgen_name(int32_t m, char name[6]) {
  name[0] = 'r';
  for (int l = 0; l < 4; l++) name[1 + l] = (char)('0' + ((m >> (3 * (3 - l))) & 7));
  name[5] = 0;
}

CHUNKGEN_HD void chunkgen_store16(uint8_t* d, uint32_t v) {
  d[0] = (uint8_t)v;
  d[1] = (uint8_t)(v >> 8);
}
CHUNKGEN_HD void chunkgen_store32(uint8_t* d, uint32_t v) {
  d[0] = (uint8_t)v;
  d[1] = (uint8_t)(v >> 8);
  d[2] = (uint8_t)(v >> 16);
  d[3] = (uint8_t)(v >> 24);
}

// Irwin-Hall N(0, 1) from 12 hashes (channel c .. c + 11)
CHUNKGEN_HD double chunkgen_normal(uint64_t s, uint64_t key, uint64_t c) {
  double g = -6.0;
  for (uint64_t q = 0; q < 12; q++) g = g + lasgen_u01(lasgen_hash(s, key, c + q));
  return g;
}

// record i of the chunk with Morton index m into out (kChunkGenBytes bytes)
CHUNKGEN_HD void chunkgen_point(const ChunkGenParams& p, int32_t m, int64_t i, uint8_t* out) {
  const uint64_t s = p.seed ^ 0xC2C2000000000000ull;
  const uint64_t key = ((uint64_t)(uint32_t)m << 40) ^ (uint64_t)i;
  int32_t cx, cy, cz;
  chunkgen_cell(m, cx, cy, cz);
  const double w = p.cube_size / 16.0;
  const double x0 = w * cx, y0 = w * cy, z0 = w * cz;
  LasGenParams lp;  // lasgen_terrain reads seed, extent and amplitude only
  lp.seed = s ^ (uint64_t)m;
  lp.extent = w;
  lp.amplitude = 0.3 * w;
  const double lx = lasgen_u01(lasgen_hash(s, key, 1)) * (w * 0.999);
  const double ly = lasgen_u01(lasgen_hash(s, key, 2)) * (w * 0.999);
  double h = 0.5 * w + 0.5 * lasgen_terrain(lp, lx, ly);
  uint8_t cls = 2;
  if ((lasgen_hash(s, (uint64_t)m, 9) & 3) == 0) {
    const double bx = w * (0.3 + 0.4 * lasgen_u01(lasgen_hash(s, (uint64_t)m, 10)));
    const double by = w * (0.3 + 0.4 * lasgen_u01(lasgen_hash(s, (uint64_t)m, 11)));
    const double hw = 0.1 * w;
    const double dx = lx - bx, dy = ly - by;
    if (dx < hw && -dx < hw && dy < hw && -dy < hw) {
      h = h + 0.15 * w;
      cls = 6;
    }
  }
  const double top = 0.999 * w;
  if (h < 0.0) h = 0.0;
  if (h > top) h = top;
  const double inv = 1.0 / p.scale;
  const int32_t X = (int32_t)lasgen_floor((x0 + lx) * inv);
  const int32_t Y = (int32_t)lasgen_floor((y0 + ly) * inv);
  const int32_t Z = (int32_t)lasgen_floor((z0 + h) * inv);
  const uint64_t ha = lasgen_hash(s, key, 5);
  const uint32_t nret = 1 + (uint32_t)((ha >> 16) % 3);
  const uint32_t ret = 1 + (uint32_t)((ha >> 24) % nret);
  const double tz = h / w;
  double mid = 2.0 * tz - 1.0;
  if (mid < 0.0) mid = -mid;
  chunkgen_store32(out + 0, (uint32_t)X);
  chunkgen_store32(out + 4, (uint32_t)Y);
  chunkgen_store32(out + 8, (uint32_t)Z);
  chunkgen_store16(out + 12, (uint32_t)(ha & 0xFFFF));
  out[14] = (uint8_t)ret;
  out[15] = (uint8_t)nret;
  out[16] = cls;
  out[17] = 0;  // scan angle rank
  out[18] = 0;  // user data
  chunkgen_store16(out + 19, (uint32_t)m);
  chunkgen_store16(out + 21, lasgen_clamp_u16(65535.0 * tz + 2000.0 * chunkgen_normal(s, key, 16)));
  chunkgen_store16(out + 23, lasgen_clamp_u16(65535.0 * (1.0 - mid) + 2000.0 * chunkgen_normal(s, key, 32)));
  chunkgen_store16(out + 25, lasgen_clamp_u16(65535.0 * (1.0 - tz) + 2000.0 * chunkgen_normal(s, key, 48)));
  const float gf = (float)chunkgen_normal(s, key, 64);
  uint32_t gb;
  __builtin_memcpy(&gb, &gf, 4);
  chunkgen_store32(out + 27, gb);
  const uint64_t hb = lasgen_hash(s, key, 80);
  chunkgen_store16(out + 31, (uint32_t)(hb & 0xFFFF));
  out[33] = (uint8_t)((hb >> 16) & 0xFF);
}

// Statistics of a chunk set, folded as ordered bits (pcg::ordered_bits) of
// the attribute values: [0..2] X, Y, Z (integers), [3] intensity, [4]
// return number, [5] number of returns, [6] classification, [7] scan angle
// rank, [8] user data, [9] point source id, [10..12] rgb, [13] gauss,
// [14] uniform16, [15] uniform8 (signed); min at [k], max at [16 + k].
constexpr int kChunkGenStats = 16;

CHUNKGEN_HD void chunkgen_values(const uint8_t* r, double v[kChunkGenStats]) {
  auto u16 = [&](int o) { return (double)((uint32_t)r[o] | ((uint32_t)r[o + 1] << 8)); };
  auto i32 = [&](int o) {
    return (double)(int32_t)((uint32_t)r[o] | ((uint32_t)r[o + 1] << 8) | ((uint32_t)r[o + 2] << 16) |
                             ((uint32_t)r[o + 3] << 24));
  };
  v[0] = i32(0);
  v[1] = i32(4);
  v[2] = i32(8);
  v[3] = u16(12);
  v[4] = r[14];
  v[5] = r[15];
  v[6] = r[16];
  v[7] = r[17];
  v[8] = r[18];
  v[9] = u16(19);
  v[10] = u16(21);
  v[11] = u16(23);
  v[12] = u16(25);
  uint32_t gb = (uint32_t)r[27] | ((uint32_t)r[28] << 8) | ((uint32_t)r[29] << 16) | ((uint32_t)r[30] << 24);
  float gf;
  __builtin_memcpy(&gf, &gb, 4);
  v[13] = (double)gf;
  v[14] = u16(31);
  v[15] = (double)(int8_t)r[33];
}
