// Seeded synthetic LAS 1.4 point generator (PDRF 2, optional extra bytes).
//
// One header shared by the host (g++) and the gfx950 kernels (hipcc), so the
// CPU and the GPU produce byte-identical point records from a point index:
// every value is a pure function of (params, index) built from a counter hash
// and IEEE binary64 + - * only (no libm calls, no division on the point path),
// and every translation unit that includes it is compiled with
// -ffp-contract=off.
//
// Scene: X,Y uniform over [0, extent)^2 in acquisition tiles of
// `tile_points` consecutive points; Z = 6-octave value-noise fBm terrain of
// the given amplitude plus `num_buildings` axis-aligned boxes with points on
// roofs and walls; intensity uniform u16; RGB from a height colour map plus
// Irwin-Hall(12) noise with sigma 2000; classification 2 (ground) or 6
// (building).  The output record is a LAS point data record of format
// `point_format` (0-7; 2 unless changed: 26 B), followed by 7 extra bytes
// (float32 N(0,1), uint16, int8) when `extra_bytes` is set.  Format 2
// records carry no user data, no classification flags and no scan
// direction / edge bits; the other formats set them from the point hash,
// plus gps time, a wave packet (4-5) and the 4-bit return fields, scan
// angle and full classification byte of formats 6-7.
//
// Precision scenes (`span_units` > 0) add, in LAS integer units:
//   - two anchor points (records 0 and 1) at the corners q0 and q0 + span of
//     the X/Y range, so the cube edge is exactly `span_units` units;
//   - near-duplicate clusters: a `dup_fraction` share of the points sits on
//     one of `num_clusters` quantized centres with an integer jitter of
//     -1, 0 or +1 unit per axis (exact duplicates are frequent);
//   - a flat square lattice: a `lattice_fraction` share of the points sits on
//     the nodes of a `lattice_side`^2 grid of pitch `lattice_pitch_units`,
//     chosen equal to a node spacing so that neighbour distances tie with it
//     up to binary64 rounding.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define LASGEN_HD __host__ __device__ inline
#else
#define LASGEN_HD inline
#endif

struct LasGenParams {
  uint64_t seed;
  int64_t n;              // number of points
  double extent;          // metres covered in X and Y
  double origin[3];       // world position of local (0,0,0)
  double offset[3];       // LAS header offset
  double inv_scale[3];    // 1 / LAS header scale
  double amplitude;       // fBm amplitude in metres
  int32_t num_buildings;  // <= 64
  int32_t tile_points;    // points per acquisition tile
  int32_t extra_bytes;    // 0: 26-byte records, 1: 33-byte records
  int32_t tiles_per_side; // filled by lasgen_prepare
  int64_t num_tiles;      // filled by lasgen_prepare
  double wall_fraction;   // fraction of points on building walls
  // precision scene (all zero for the plain scenes)
  double dup_fraction;          // share of points in near-duplicate clusters
  double lattice_fraction;      // share of points on the lattice
  int64_t span_units;           // X/Y range in LAS units (0: no anchors)
  int64_t num_clusters;         // filled by lasgen_prepare
  int64_t lattice_pitch_units;  // lattice pitch in LAS units
  int64_t lattice_side;         // filled by lasgen_prepare
  int32_t point_format;         // LAS point data record format 0-7 (presets: 2)
  int32_t pad;
};

struct LasGenBuilding {
  double cx, cy, hx, hy, base, height;
};

LASGEN_HD uint64_t lasgen_mix(uint64_t z) {
  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

LASGEN_HD uint64_t lasgen_hash(uint64_t seed, uint64_t a, uint64_t b) {
  return lasgen_mix(lasgen_mix(seed ^ (a * 0xD1B54A32D192ED03ull)) + b);
}

// uniform double in [0,1) with 53 random bits
LASGEN_HD double lasgen_u01(uint64_t h) {
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

LASGEN_HD double lasgen_floor(double v) {
  // exact floor for |v| < 2^52 without libm: truncate then correct
  double t = (double)(int64_t)v;
  return (t > v) ? t - 1.0 : t;
}

// value noise lattice value in [-1,1]
LASGEN_HD double lasgen_lattice(uint64_t seed, int64_t ix, int64_t iy, int32_t octave) {
  uint64_t h = lasgen_hash(seed ^ 0x5bd1e995ull, (uint64_t)ix * 73856093ull ^ (uint64_t)iy * 19349663ull,
                           (uint64_t)octave + 977ull);
  return lasgen_u01(h) * 2.0 - 1.0;
}

LASGEN_HD double lasgen_smooth(double t) { return t * t * (3.0 - 2.0 * t); }

LASGEN_HD double lasgen_terrain(const LasGenParams& p, double x, double y) {
  double freq = 4.0 / p.extent;
  double amp = p.amplitude / 1.96875;  // sum of 0.5^k for k = 0..5
  double z = 0.0;
  for (int32_t o = 0; o < 6; o++) {
    double fx = x * freq;
    double fy = y * freq;
    double flx = lasgen_floor(fx);
    double fly = lasgen_floor(fy);
    int64_t ix = (int64_t)flx;
    int64_t iy = (int64_t)fly;
    double tx = lasgen_smooth(fx - flx);
    double ty = lasgen_smooth(fy - fly);
    double v00 = lasgen_lattice(p.seed, ix, iy, o);
    double v10 = lasgen_lattice(p.seed, ix + 1, iy, o);
    double v01 = lasgen_lattice(p.seed, ix, iy + 1, o);
    double v11 = lasgen_lattice(p.seed, ix + 1, iy + 1, o);
    double a = v00 + (v10 - v00) * tx;
    double b = v01 + (v11 - v01) * tx;
    z = z + amp * (a + (b - a) * ty);
    freq = freq * 2.0;
    amp = amp * 0.5;
  }
  return z;
}

LASGEN_HD LasGenBuilding lasgen_building(const LasGenParams& p, int32_t b) {
  LasGenBuilding r;
  uint64_t s = p.seed ^ 0xB0D1ull;
  r.cx = (0.05 + 0.9 * lasgen_u01(lasgen_hash(s, (uint64_t)b, 0))) * p.extent;
  r.cy = (0.05 + 0.9 * lasgen_u01(lasgen_hash(s, (uint64_t)b, 1))) * p.extent;
  r.hx = 5.0 + 15.0 * lasgen_u01(lasgen_hash(s, (uint64_t)b, 2));
  r.hy = 5.0 + 15.0 * lasgen_u01(lasgen_hash(s, (uint64_t)b, 3));
  r.height = 5.0 + 25.0 * lasgen_u01(lasgen_hash(s, (uint64_t)b, 4));
  r.base = lasgen_terrain(p, r.cx, r.cy);
  return r;
}

// integer square root ceiling, exact
LASGEN_HD int64_t lasgen_isqrt_ceil(int64_t v) {
  int64_t r = 0;
  while (r * r < v) r++;
  return r;
}

inline void lasgen_prepare(LasGenParams& p) {
  p.num_tiles = (p.n + p.tile_points - 1) / p.tile_points;
  if (p.num_tiles < 1) p.num_tiles = 1;
  int64_t lo = 0, hi = 1;
  while (hi * hi < p.num_tiles) hi *= 2;
  while (lo < hi) {  // smallest t with t*t >= num_tiles
    int64_t mid = (lo + hi) / 2;
    if (mid * mid >= p.num_tiles) hi = mid; else lo = mid + 1;
  }
  p.tiles_per_side = (int32_t)lo;
  p.num_clusters = 0;
  p.lattice_side = 0;
  if (p.span_units > 0) {
    int64_t nd = (int64_t)((double)p.n * p.dup_fraction);
    p.num_clusters = nd / 16 > 0 ? nd / 16 : 1;
    int64_t nl = (int64_t)((double)p.n * p.lattice_fraction);
    int64_t side = 1;
    while (side * side < nl) side++;
    int64_t cap = p.lattice_pitch_units > 0 ? (p.span_units / 2) / p.lattice_pitch_units : 1;
    p.lattice_side = side < cap ? side : (cap > 0 ? cap : 1);
  }
}

LASGEN_HD int32_t lasgen_quantize(double world, double offset, double inv_scale) {
  return (int32_t)(int64_t)lasgen_floor((world - offset) * inv_scale + 0.5);
}

LASGEN_HD uint16_t lasgen_clamp_u16(double v) {
  if (v < 0.0) return 0;
  if (v > 65535.0) return 65535;
  return (uint16_t)(int32_t)v;
}

LASGEN_HD void lasgen_store16(uint8_t* d, uint16_t v) { d[0] = (uint8_t)v; d[1] = (uint8_t)(v >> 8); }
LASGEN_HD void lasgen_store32(uint8_t* d, uint32_t v) {
  d[0] = (uint8_t)v; d[1] = (uint8_t)(v >> 8); d[2] = (uint8_t)(v >> 16); d[3] = (uint8_t)(v >> 24);
}

// Writes record `i` (lasgen_record_length bytes) to `out`.  `bld` holds
// num_buildings entries.
LASGEN_HD void lasgen_point(const LasGenParams& p, const LasGenBuilding* bld, int64_t i, uint8_t* out) {
  uint64_t s = p.seed;
  double x, y, z;
  uint8_t cls = 2;
  uint64_t ui = (uint64_t)i;
  int64_t tile = i / p.tile_points;
  // precision scenes: 1 = anchor, 2 = near-duplicate cluster, 3 = lattice
  int32_t mode = 0;
  if (p.span_units > 0) {
    double sel = lasgen_u01(lasgen_hash(s, ui, 6));
    if (i < 2) mode = 1;
    else if (sel < p.dup_fraction) mode = 2;
    else if (sel < p.dup_fraction + p.lattice_fraction) mode = 3;
  }
  int32_t X, Y, Z;
  int32_t qx0 = lasgen_quantize(p.origin[0], p.offset[0], p.inv_scale[0]);
  int32_t qy0 = lasgen_quantize(p.origin[1], p.offset[1], p.inv_scale[1]);
  if (mode == 1) {
    int32_t d = (int32_t)(i ? p.span_units : 0);
    x = i ? p.extent : 0.0;
    y = x;
    z = lasgen_terrain(p, x, y);
    X = qx0 + d;
    Y = qy0 + d;
    Z = lasgen_quantize(p.origin[2] + z, p.offset[2], p.inv_scale[2]);
  } else if (mode == 2) {
    uint64_t c = (uint64_t)(lasgen_u01(lasgen_hash(s, ui, 7)) * (double)p.num_clusters);
    x = lasgen_u01(lasgen_hash(s ^ 0xC1A5ull, c, 0)) * p.extent * 0.999;
    y = lasgen_u01(lasgen_hash(s ^ 0xC1A5ull, c, 1)) * p.extent * 0.999;
    z = lasgen_terrain(p, x, y);
    uint64_t hj = lasgen_hash(s, ui, 8);
    X = lasgen_quantize(p.origin[0] + x, p.offset[0], p.inv_scale[0]) + (int32_t)(hj % 3) - 1;
    Y = lasgen_quantize(p.origin[1] + y, p.offset[1], p.inv_scale[1]) + (int32_t)((hj >> 8) % 3) - 1;
    Z = lasgen_quantize(p.origin[2] + z, p.offset[2], p.inv_scale[2]) + (int32_t)((hj >> 16) % 3) - 1;
    if (X < qx0 + 1) X = qx0 + 1;
    if (Y < qy0 + 1) Y = qy0 + 1;
  } else if (mode == 3) {
    int64_t M = p.lattice_side;
    int64_t k = (int64_t)(lasgen_u01(lasgen_hash(s, ui, 7)) * (double)(M * M));
    int64_t a = k % M, b = k / M;
    int64_t x0 = p.span_units / 4, y0 = p.span_units / 4;
    x = (double)x0 / p.inv_scale[0];
    y = (double)y0 / p.inv_scale[1];
    z = lasgen_terrain(p, x, y);
    X = qx0 + (int32_t)(x0 + a * p.lattice_pitch_units);
    Y = qy0 + (int32_t)(y0 + b * p.lattice_pitch_units);
    Z = lasgen_quantize(p.origin[2] + z, p.offset[2], p.inv_scale[2]);
  }
  if (mode != 0) {
    // colours and attributes below follow z like the plain scene
  } else if (lasgen_u01(lasgen_hash(s, ui, 0)) < p.wall_fraction && p.num_buildings > 0) {
    int32_t b = (int32_t)(lasgen_u01(lasgen_hash(s, ui, 1)) * (double)p.num_buildings);
    int32_t side = (int32_t)(lasgen_u01(lasgen_hash(s, ui, 2)) * 4.0);
    double t = lasgen_u01(lasgen_hash(s, ui, 3));
    const LasGenBuilding& B = bld[b];
    double h = lasgen_u01(lasgen_hash(s, ui, 4)) * B.height;
    if (side == 0) { x = B.cx - B.hx; y = B.cy - B.hy + t * 2.0 * B.hy; }
    else if (side == 1) { x = B.cx + B.hx; y = B.cy - B.hy + t * 2.0 * B.hy; }
    else if (side == 2) { y = B.cy - B.hy; x = B.cx - B.hx + t * 2.0 * B.hx; }
    else { y = B.cy + B.hy; x = B.cx - B.hx + t * 2.0 * B.hx; }
    z = B.base + h;
    cls = 6;
  } else {
    int64_t T = p.tiles_per_side;
    double tsz = p.extent / (double)T;
    int64_t tx = tile % T;
    int64_t ty = tile / T;
    x = ((double)tx + lasgen_u01(lasgen_hash(s, ui, 1))) * tsz;
    y = ((double)ty + lasgen_u01(lasgen_hash(s, ui, 2))) * tsz;
    z = lasgen_terrain(p, x, y);
    for (int32_t b = 0; b < p.num_buildings; b++) {
      const LasGenBuilding& B = bld[b];
      double dx = x - B.cx, dy = y - B.cy;
      if (dx < B.hx && -dx < B.hx && dy < B.hy && -dy < B.hy) {
        z = B.base + B.height;
        cls = 6;
        break;
      }
    }
  }
  if (mode == 0) {
    double lim = p.extent * 0.999999;
    if (x < 0.0) x = 0.0;
    if (x > lim) x = lim;
    if (y < 0.0) y = 0.0;
    if (y > lim) y = lim;
    X = lasgen_quantize(p.origin[0] + x, p.offset[0], p.inv_scale[0]);
    Y = lasgen_quantize(p.origin[1] + y, p.offset[1], p.inv_scale[1]);
    Z = lasgen_quantize(p.origin[2] + z, p.offset[2], p.inv_scale[2]);
  }

  uint64_t ha = lasgen_hash(s, ui, 5);
  uint16_t intensity = (uint16_t)(ha & 0xFFFF);
  int32_t nret = 1 + (int32_t)((ha >> 16) % 3);
  int32_t ret = 1 + (int32_t)((ha >> 24) % (uint64_t)nret);
  int32_t sar = (int32_t)((ha >> 32) % 61) - 30;

  double tz = (z + 40.0) * (1.0 / 110.0);
  if (tz < 0.0) tz = 0.0;
  if (tz > 1.0) tz = 1.0;
  double mid = 2.0 * tz - 1.0;
  if (mid < 0.0) mid = -mid;
  double cr = 65535.0 * tz, cg = 65535.0 * (1.0 - mid), cb = 65535.0 * (1.0 - tz);
  double nr = -6.0, ng = -6.0, nb = -6.0;
  for (int32_t k = 0; k < 12; k++) {
    nr = nr + lasgen_u01(lasgen_hash(s, ui, 16 + k));
    ng = ng + lasgen_u01(lasgen_hash(s, ui, 32 + k));
    nb = nb + lasgen_u01(lasgen_hash(s, ui, 48 + k));
  }

  const int32_t fmt = p.point_format;
  const uint16_t red = lasgen_clamp_u16(cr + 2000.0 * nr);
  const uint16_t green = lasgen_clamp_u16(cg + 2000.0 * ng);
  const uint16_t blue = lasgen_clamp_u16(cb + 2000.0 * nb);
  // fields of the other point formats (format 2 uses none of them)
  const uint64_t hf = lasgen_hash(s, ui, 96);
  const double gps = 400000000.0 + (double)i * 0.0001 + lasgen_u01(lasgen_hash(s, ui, 97)) * 0.00005;
  lasgen_store32(out + 0, (uint32_t)X);
  lasgen_store32(out + 4, (uint32_t)Y);
  lasgen_store32(out + 8, (uint32_t)Z);
  lasgen_store16(out + 12, intensity);
  int32_t core = 26;
  if (fmt <= 5) {
    out[14] = (uint8_t)((ret & 7) | ((nret & 7) << 3) | (fmt == 2 ? 0 : (int32_t)((hf >> 40) & 3) << 6));
    // formats other than 2 also set the synthetic / keypoint / withheld bits
    out[15] = (uint8_t)(cls | (fmt == 2 ? 0 : (int32_t)((hf >> 8) & 7) << 5));
    out[16] = (uint8_t)(int8_t)sar;
    out[17] = fmt == 2 ? 0 : (uint8_t)(hf >> 16);  // user data
    lasgen_store16(out + 18, (uint16_t)(tile & 0xFFFF));
    int32_t o = 20;
    if (fmt == 1 || fmt == 3 || fmt == 4 || fmt == 5) {
      uint64_t gb;
      __builtin_memcpy(&gb, &gps, 8);
      lasgen_store32(out + o, (uint32_t)gb);
      lasgen_store32(out + o + 4, (uint32_t)(gb >> 32));
      o += 8;
    }
    if (fmt == 2 || fmt == 3 || fmt == 5) {
      lasgen_store16(out + o, red);
      lasgen_store16(out + o + 2, green);
      lasgen_store16(out + o + 4, blue);
      o += 6;
    }
    if (fmt == 4 || fmt == 5) {
      // wave packet: descriptor index, byte offset, size, location, xyz(t)
      const uint64_t hw = lasgen_hash(s, ui, 98);
      out[o] = (uint8_t)hw;
      lasgen_store32(out + o + 1, (uint32_t)(hw >> 8));
      lasgen_store32(out + o + 5, (uint32_t)(hw >> 40));
      lasgen_store32(out + o + 9, (uint32_t)(hw >> 3));
      for (int32_t k = 0; k < 4; k++) lasgen_store32(out + o + 13 + 4 * k, (uint32_t)(hw >> (7 * k)) & 0x3F7FFFFFu);
      o += 29;
    }
    core = o;
  } else {
    // formats 6-7: 4-bit return number / number of returns (any pair, so
    // that laszip's legacy mapping sees every case), classification flags,
    // scanner channel, scan direction, edge of flight line; a full
    // classification byte that is sometimes above 31
    const int32_t nret4 = 1 + (int32_t)((hf >> 16) % 15);
    const int32_t ret4 = 1 + (int32_t)((hf >> 24) % 15);
    out[14] = (uint8_t)(ret4 | (nret4 << 4));
    out[15] = (uint8_t)((hf >> 32) & 0xFF);
    const uint32_t pick = (uint32_t)((hf >> 40) & 63);
    out[16] = (uint8_t)(pick == 0 ? 32 + (int32_t)((hf >> 48) % 224) : (pick == 1 ? (int32_t)((hf >> 48) & 31) : cls));
    out[17] = (uint8_t)(hf >> 56);
    const int32_t ang = (int32_t)((hf >> 4) % 30001) - 15000;
    lasgen_store16(out + 18, (uint16_t)(int16_t)ang);
    lasgen_store16(out + 20, (uint16_t)(tile & 0xFFFF));
    uint64_t gb;
    __builtin_memcpy(&gb, &gps, 8);
    lasgen_store32(out + 22, (uint32_t)gb);
    lasgen_store32(out + 26, (uint32_t)(gb >> 32));
    core = 30;
    if (fmt == 7) {
      lasgen_store16(out + 30, red);
      lasgen_store16(out + 32, green);
      lasgen_store16(out + 34, blue);
      core = 36;
    }
  }
  if (p.extra_bytes) {
    double g = -6.0;
    for (int32_t k = 0; k < 12; k++) g = g + lasgen_u01(lasgen_hash(s, ui, 64 + k));
    float f = (float)g;
    uint32_t fb;
    __builtin_memcpy(&fb, &f, 4);
    lasgen_store32(out + core, fb);
    uint64_t hb = lasgen_hash(s, ui, 80);
    lasgen_store16(out + core + 4, (uint16_t)(hb & 0xFFFF));
    out[core + 6] = (uint8_t)((hb >> 16) & 0xFF);
  }
}

// record length of a generated point (core record of the format + 7 extra bytes)
LASGEN_HD int32_t lasgen_record_length(const LasGenParams& p) {
  const int32_t core[8] = {20, 28, 26, 34, 57, 63, 30, 36};
  return core[p.point_format] + (p.extra_bytes ? 7 : 0);
}
