// Counting-sort chunker on gfx950: LAS records in HBM -> output records
// grouped by chunk, in input order inside each chunk.
//
// Restates Converter/src/chunker_countsort_laszip.cpp:
//   k_count           :131-361  per-cell counts (+ attribute min/max and the
//                               classification histogram of :380-658, which
//                               depend only on the input records)
//   k_pyramid_level   :1101-1183 image-pyramid merge
//   k_chunk_cells     :1144-1172 cells that become chunks
//   k_lut_fill        :1187-1199 LUT fill
//   k_tile_hist       :817-876  chunk of every point (LUT), per-tile counts
//   k_distribute      :770-904  transcode + stable scatter (point formats 0-3,
//                               6-7; runs as aligned 16-byte stores)
//   k_scatter         :770-904  the same for any record layout, byte-wise
// All geometry is IEEE binary64 with the reference's evaluation order; the
// TU is compiled with -ffp-contract=off.
#include <mutex>

#include "chunker.h"
#include "lasgen.h"
#include "pcg_common.h"
#include "scan.h"

namespace pcg {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kCountSteps = 16;             // 64-point steps per wave in k_count
constexpr int kHashSize = 2048;             // LDS hash slots in k_count
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kMaxBppIn = 128;
constexpr int kMaxStdStats = 15;  // std_stat_count of point format 7

__constant__ ChunkerParams c_params;

// Loads `count` records (bpp bytes each) starting at record `first` into a
// 16-byte aligned LDS area using this wave's lanes; returns the byte shift
// of the first record inside the area.  Reads never pass `total_bytes`.
__device__ inline uint32_t wave_load_records(const uint8_t* __restrict__ src, uint64_t total_bytes, int64_t first,
                                             int count, int bpp, uint8_t* lds) {
  const int lane = threadIdx.x & 63;
  uint64_t b0 = (uint64_t)first * bpp;
  uint64_t a0 = b0 & ~15ull;
  uint64_t end = b0 + (uint64_t)count * bpp;
  int nvec = (int)((end - a0 + 15) >> 4);
  for (int v = lane; v < nvec; v += 64) {
    uint64_t addr = a0 + (uint64_t)v * 16;
    if (addr + 16 <= total_bytes) {
      *(uint4*)(lds + v * 16) = *(const uint4*)(src + addr);
    } else {
      for (int b = 0; b < 16; b++) lds[v * 16 + b] = (addr + b < total_bytes) ? src[addr + b] : 0;
    }
  }
  return (uint32_t)(b0 - a0);
}

// 16 bytes of the LAS buffer at a 16-byte aligned offset; bytes at or past
// `total` read as 0
__device__ inline uint4 load_vec16(const uint8_t* __restrict__ src, uint64_t total, uint64_t addr) {
  if (addr + 16 <= total) return *(const uint4*)(src + addr);
  uint32_t w[4] = {0u, 0u, 0u, 0u};
  for (int b = 0; b < 16; b++)
    if (addr + b < total) w[b >> 2] |= (uint32_t)src[addr + b] << (8 * (b & 3));
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// 16-byte vectors per lane that hold one step's records of up to `bpp`
// bytes each (a step starts anywhere in a 16-byte vector)
__host__ __device__ constexpr int prefetch_vecs(int bpp) { return (64 * bpp + 30) / (16 * 64) + 1; }

__device__ inline int32_t ld_i32(const uint8_t* p) {
  return (int32_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
}

// 4 bytes at any address inside an LDS staging area (two aligned dword
// reads and a byte align); the area is padded so the second read stays in it
// (the aligned pointer is derived from p by pointer arithmetic so that the
// compiler keeps the LDS address space and emits ds_read, not flat loads)
__device__ inline uint32_t lds_u32(const uint8_t* p) {
  const uint32_t s = (uint32_t)((uintptr_t)p & 3);
  const uint32_t* w = (const uint32_t*)(p - s);
  return __builtin_amdgcn_alignbyte(w[1], w[0], s);
}
__device__ inline uint64_t lds_u64(const uint8_t* p) {
  const uint32_t s = (uint32_t)((uintptr_t)p & 3);
  const uint32_t* w = (const uint32_t*)(p - s);
  return (uint64_t)__builtin_amdgcn_alignbyte(w[1], w[0], s) |
         ((uint64_t)__builtin_amdgcn_alignbyte(w[2], w[1], s) << 32);
}

// Attribute statistics as order-preserving integer keys: unsigned values as
// they are, signed ones with the sign bit flipped, floating point through
// ordered_bits.  The key order equals the order of the values converted to
// double, so min/max over keys converted once give the double min/max.
struct StatLane {
  int src;
  uint32_t shift;
  uint64_t mask;
  int sign_bits;  // width of a signed field, 0 unsigned
  int fp;         // 0 integer, 4 float, 8 double
  int legacy;     // 1: legacy return number, 2: legacy number of returns (formats 6-10)
};

__device__ inline StatLane stat_lane(const StatOp& op) {
  StatLane s{op.src, 0, 0xFFull, 0, 0, 0};
  switch (op.kind) {
    case ST_U8: break;
    case ST_I8: s.sign_bits = 8; break;
    case ST_U16: s.mask = 0xFFFFull; break;
    case ST_I16: s.mask = 0xFFFFull; s.sign_bits = 16; break;
    case ST_U32: s.mask = 0xFFFFFFFFull; break;
    case ST_I32: s.mask = 0xFFFFFFFFull; s.sign_bits = 32; break;
    case ST_U64: s.mask = ~0ull; break;
    case ST_I64: s.mask = ~0ull; s.sign_bits = 64; break;
    case ST_F32: s.mask = 0xFFFFFFFFull; s.fp = 4; break;
    case ST_F64: s.mask = ~0ull; s.fp = 8; break;
    case ST_LOW3: s.mask = 7; break;
    case ST_MID3: s.mask = 7; s.shift = 3; break;
    case ST_LOW5: s.mask = 31; break;
    case ST_RET14: s.legacy = 1; break;
    case ST_NRET14: s.legacy = 2; break;
    case ST_LOW4: s.mask = 15; break;
    default: s.mask = 0; break;
  }
  return s;
}

__device__ inline uint64_t stat_key(const StatLane& s, uint64_t raw) {
  uint64_t v = (raw >> s.shift) & s.mask;
  if (s.legacy) return s.legacy == 1 ? legacy_return_number((uint32_t)v) : legacy_number_of_returns((uint32_t)v);
  if (s.fp == 4) {
    float f;
    uint32_t b = (uint32_t)v;
    __builtin_memcpy(&f, &b, 4);
    return ordered_bits((double)f);
  }
  if (s.fp == 8) {
    double d;
    __builtin_memcpy(&d, &v, 8);
    return ordered_bits(d);
  }
  if (s.sign_bits) {
    const int up = 64 - s.sign_bits;
    v = (uint64_t)(((int64_t)(v << up)) >> up) ^ 0x8000000000000000ull;
  }
  return v;
}

// a key back to the ordered bits of the double value
__device__ inline uint64_t stat_key_bits(const StatLane& s, uint64_t key) {
  if (s.fp) return key;
  if (s.sign_bits) return ordered_bits((double)(int64_t)(key ^ 0x8000000000000000ull));
  return ordered_bits((double)key);
}

// laszip coordinate (scale * X + offset), requantization and the cell of
// one point: chunker_countsort_laszip.cpp:224-257.  The grid has at most
// 512^3 cells, so the cell index and its per-axis parts are int32 (one
// conversion instruction each; double -> int64 is a long sequence on the
// GPU).  In the box every per-axis part lies in [0, G - 1]; outside it the
// point is an error and its cell unused, the clamp only keeps the
// conversion defined.
__device__ inline int32_t cell_of_record(const uint8_t* rec, const ChunkerParams& P, bool& inBox, int32_t q[3]) {
  const int32_t G = P.grid_size;
  const double dG = (double)G;
  int32_t idx = 0;
  inBox = true;
  int32_t mul = 1;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int32_t X = (int32_t)lds_u32(rec + 4 * a);
    double x = P.las_scale[a] * X + P.las_offset[a];
    int32_t Xq = (int32_t)((x - P.pos_offset[a]) / P.pos_scale[a]);
    const double v = (double)Xq * P.pos_scale[a] + P.pos_offset[a] - P.min[a];
    // G * (v / cube) without the division: v * (G / cube) differs from it
    // by a few 1e-13 at most, so its floor is the cell unless it lies within
    // 1e-9 of a cell boundary or of the box faces; those take the reference
    // expression
    const double t = v * P.g_over_cube;
    const double ft = __builtin_floor(t);
    const double fr = t - ft;
    int32_t ic;
    if (fr > 1e-9 && fr < 1.0 - 1e-9 && t > 1e-9 && t < dG - 1e-9) {
      ic = (int32_t)ft;
    } else {
      double u = v / P.cube;
      inBox = inBox && u >= 0.0 && u <= 1.0;
      double gu = dG * u;
      double lim = dG - 1.0;
      const double c = lim < gu ? lim : gu;
      ic = (int32_t)(c > 0.0 ? c : 0.0);
    }
    idx += ic * mul;
    mul *= G;
    q[a] = Xq;
  }
  return idx;
}


// requantized position of one point (the q[] of cell_of_record)
__device__ inline void quantize_record(const uint8_t* rec, const ChunkerParams& P, int32_t q[3]) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int32_t X = (int32_t)lds_u32(rec + 4 * a);
    double x = P.las_scale[a] * X + P.las_offset[a];
    q[a] = (int32_t)((x - P.pos_offset[a]) / P.pos_scale[a]);
  }
}

// peers of this lane with the same key in the wave (match-any by key bits)
__device__ inline uint64_t wave_match(uint32_t key, int bits, uint64_t active) {
  uint64_t peers = active;
  for (int b = 0; b < bits; b++) {
    uint64_t m = __ballot((key >> b) & 1);
    peers &= ((key >> b) & 1) ? m : ~m;
  }
  return peers;
}

// The distinct keys of one 64-point step, in the order of their lowest
// lane.  For every valid lane: `peers`, the valid lanes with its key, and
// `run`, the start of its key's run when the runs are laid out in that
// order.  The first kMatchRounds keys are taken one per round (spatially
// coherent input has one or two chunks per step); lanes still left after
// that take the bitwise match and a prefix sum over their leaders.
constexpr int kMatchRounds = 6;
__device__ inline void step_runs(uint32_t key, bool valid, int key_bits, uint64_t lt_mask, uint64_t& peers,
                                 uint32_t& run) {
  const int lane = threadIdx.x & 63;
  uint64_t left = __ballot(valid);
  uint32_t start = 0;
  peers = 0;
  run = 0;
  for (int r = 0; r < kMatchRounds && left; r++) {
    const int L = __builtin_ctzll(left);
    const uint32_t v = __builtin_amdgcn_readlane(key, L);
    const uint64_t m = __ballot(key == v) & left;
    if ((m >> lane) & 1) {
      peers = m;
      run = start;
    }
    start += (uint32_t)__popcll(m);
    left &= ~m;
  }
  if (left) {
    const bool mine = (left >> lane) & 1;
    const uint64_t p = wave_match(key, key_bits, left);
    const bool lead = mine && (p & lt_mask) == 0;
    const uint32_t len = lead ? (uint32_t)__popcll(p) : 0u;
    uint32_t incl = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d);
      if (lane >= d) incl += o;
    }
    const uint32_t rs = start + incl - len;  // valid on leader lanes
    const uint32_t r2 = __shfl(rs, mine ? __builtin_ctzll(p) : lane);
    if (mine) {
      peers = p;
      run = r2;
    }
  }
}

// bytes of one wave's record staging area in LDS
__host__ __device__ inline int record_area_bytes(int bpp) { return (64 * bpp + 48 + 15) & ~15; }

// 4 bytes of a record at compile-time byte offset o from its dwords
// (w[k] = bytes 4k .. 4k+3 of the record)
template <int O>
__device__ inline uint32_t rec_u32(const uint32_t* w) {
  if constexpr ((O & 3) == 0) return w[O >> 2];
  else return __builtin_amdgcn_alignbyte(w[(O >> 2) + 1], w[O >> 2], O & 3);
}

// order-preserving key of one standard statistic (compile-time kind)
template <int FMT, int I>
__device__ inline uint64_t std_key(const uint32_t* w) {
  constexpr StdStat st = std_stat(FMT, I);
  const uint32_t v = rec_u32<st.src>(w);
  if constexpr (st.kind == ST_F64) {
    return ordered_bits_from_raw((uint64_t)v | ((uint64_t)rec_u32<st.src + 4>(w) << 32));
  } else if constexpr (st.kind == ST_I32) {
    return v ^ 0x80000000u;
  } else if constexpr (st.kind == ST_I16) {
    return (uint32_t)(int32_t)(int16_t)(v & 0xFFFFu) ^ 0x80000000u;
  } else if constexpr (st.kind == ST_I8) {
    return (uint32_t)(int32_t)(int8_t)(v & 0xFFu) ^ 0x80000000u;
  } else if constexpr (st.kind == ST_U16) {
    return v & 0xFFFFu;
  } else if constexpr (st.kind == ST_LOW3) {
    return v & 7u;
  } else if constexpr (st.kind == ST_MID3) {
    return (v >> 3) & 7u;
  } else if constexpr (st.kind == ST_LOW5) {
    return v & 31u;
  } else if constexpr (st.kind == ST_LOW4) {
    return v & 15u;
  } else if constexpr (st.kind == ST_RET14) {
    return legacy_return_number(v & 0xFFu);
  } else if constexpr (st.kind == ST_NRET14) {
    return legacy_number_of_returns(v & 0xFFu);
  } else {
    return v & 0xFFu;  // ST_U8
  }
}

// a standard statistic's key back to the ordered bits of the double value
template <int FMT, int I>
__device__ inline uint64_t std_key_bits(uint64_t key) {
  constexpr StdStat st = std_stat(FMT, I);
  if constexpr (st.kind == ST_F64) return key;
  else if constexpr (st.kind == ST_I32 || st.kind == ST_I16 || st.kind == ST_I8)
    return ordered_bits((double)(int32_t)((uint32_t)key ^ 0x80000000u));
  else return ordered_bits((double)(uint32_t)key);
}

// per-lane running keys of the standard statistics: 32-bit keys, and the one
// 64-bit (gps time) key of the formats that have it
struct StdAcc {
  uint32_t mn[kMaxStdStats], mx[kMaxStdStats];
  uint64_t mn64, mx64;
};

template <int FMT, int I>
__device__ inline void fold_std(const uint32_t* w, StdAcc& a) {
  if constexpr (I < std_stat_count(FMT)) {
    const uint64_t k = std_key<FMT, I>(w);
    if constexpr (std_stat(FMT, I).kind == ST_F64) {
      a.mn64 = k < a.mn64 ? k : a.mn64;
      a.mx64 = k > a.mx64 ? k : a.mx64;
    } else {
      const uint32_t k32 = (uint32_t)k;
      a.mn[I] = k32 < a.mn[I] ? k32 : a.mn[I];
      a.mx[I] = k32 > a.mx[I] ? k32 : a.mx[I];
    }
    fold_std<FMT, I + 1>(w, a);
  }
}

// wave minimum / maximum of every standard key, as ordered double bits, into
// this wave's row (neutral ~0 / 0 for a wave without records)
template <int FMT, int I>
__device__ inline void publish_std(const StdAcc& acc, bool wave_seen, int lane, uint64_t* s_min_row,
                                   uint64_t* s_max_row) {
  if constexpr (I < std_stat_count(FMT)) {
    constexpr bool wide = std_stat(FMT, I).kind == ST_F64;
    uint64_t a = wide ? acc.mn64 : acc.mn[I], b = wide ? acc.mx64 : acc.mx[I];
    for (int d = 32; d >= 1; d >>= 1) {
      const uint64_t oa = __shfl_xor(a, d), ob = __shfl_xor(b, d);
      a = oa < a ? oa : a;
      b = ob > b ? ob : b;
    }
    if (lane == 0) {
      s_min_row[I] = wave_seen ? std_key_bits<FMT, I>(a) : ~0ull;
      s_max_row[I] = wave_seen ? std_key_bits<FMT, I>(b) : 0ull;
    }
    publish_std<FMT, I + 1>(acc, wave_seen, lane, s_min_row, s_max_row);
  }
}

// K1: per-cell counts, the cell of every point, attribute min/max (ordered
// uint64 bits) and the classification histogram.  Each wave walks a
// contiguous range of 64-point steps (consecutive LAS points are spatially
// coherent): runs of equal cells inside a step collapse to one LDS hash
// update, the LDS hash merges repeats across the block's steps before the
// global atomics.  The standard attributes of point format FMT are folded
// by every lane for its own record with compile-time field offsets (the
// record's leading dwords are read from LDS once); the extra-byte statistics
// are spread over the lanes (lane l folds extra statistic l % E of every
// (64 / E)-th record of the step).  FMT 8: no standard statistics.
template <int FMT, int PV>
__global__ __launch_bounds__(kBlock) void k_count(const uint8_t* __restrict__ las, int64_t n,
                                                   int32_t* __restrict__ grid, uint32_t* __restrict__ cells,
                                                   uint64_t* __restrict__ stat_min,
                                                   uint64_t* __restrict__ stat_max,
                                                   unsigned long long* __restrict__ class_hist,
                                                   int32_t* __restrict__ err) {
  constexpr int NS = std_stat_count(FMT);
  constexpr int NW = (std_record_bytes(FMT) + 3) / 4 + 1;  // dwords read per record (+1 for alignbyte)
  const ChunkerParams& P = c_params;
  __shared__ uint32_t s_key[kHashSize];
  __shared__ uint32_t s_cnt[kHashSize];
  __shared__ uint32_t s_hist[256];
  __shared__ uint64_t s_smin[kBlock];
  __shared__ uint64_t s_smax[kBlock];
  __shared__ uint64_t s_rmin[kWavesPerBlock][kMaxStdStats];
  __shared__ uint64_t s_rmax[kWavesPerBlock][kMaxStdStats];
  extern __shared__ __attribute__((aligned(16))) uint8_t s_dyn[];  // record staging, one area per wave
  for (int i = threadIdx.x; i < kHashSize; i += kBlock) { s_key[i] = kEmpty; s_cnt[i] = 0; }
  for (int i = threadIdx.x; i < 256; i += kBlock) s_hist[i] = 0;
  __syncthreads();
  // the wave index is uniform: readfirstlane keeps it (and every range and
  // address derived from it) in scalar registers
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int bpp = P.bpp_in;
  const uint64_t total = (uint64_t)n * bpp;
  uint8_t* rec_area = s_dyn + wave * record_area_bytes(bpp);
  const int E = P.num_stats - NS;  // extra-byte statistics
  const int R = E > 0 ? 64 / E : 0;  // lanes per extra statistic
  const int my_stat = (E > 0 && lane < R * E) ? NS + lane % E : -1;
  const int my_r0 = E > 0 ? lane / E : 0;
  const StatLane sl = stat_lane(my_stat >= 0 ? P.stats[my_stat] : StatOp{0, ST_NONE});
  uint64_t my_min = ~0ull, my_max = 0ull;  // keys
  bool seen = false;
  StdAcc acc;
#pragma unroll
  for (int k = 0; k < kMaxStdStats; k++) { acc.mn[k] = ~0u; acc.mx[k] = 0u; }
  acc.mn64 = ~0ull;
  acc.mx64 = 0ull;
  const int64_t steps_total = (n + 63) / 64;
  const int64_t wave_global = (int64_t)blockIdx.x * kWavesPerBlock + wave;
  const int64_t waves_total = (int64_t)gridDim.x * kWavesPerBlock;
  const int64_t per_wave = (steps_total + waves_total - 1) / waves_total;
  const int64_t step0 = wave_global * per_wave;
  const int64_t step1 = (step0 + per_wave) < steps_total ? (step0 + per_wave) : steps_total;
  const uint64_t low_incl = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
  bool bad = false;
  // PV > 0: the next step's records as PV 16-byte vectors per lane, loaded
  // without a branch per lane (lanes past the step reload its last vector)
  // except in a step that reaches the end of the input, in flight during
  // the current step and put into the record area after it
  uint4 pf[PV > 0 ? PV : 1];
  int pnvec = 0;
  auto issue = [&](int64_t f) {
    const int m = (int)((n - f) < 64 ? (n - f) : 64);
    const uint64_t b0 = (uint64_t)f * bpp, a0 = b0 & ~15ull;
    const int nvec = (int)((b0 + (uint64_t)m * bpp - a0 + 15) >> 4);
    pnvec = nvec;
    if (a0 + 16ull * (uint64_t)nvec <= total) {
#pragma unroll
      for (int j = 0; j < PV; j++) {
        const int v = lane + 64 * j;
        pf[j] = *(const uint4*)(las + a0 + 16ull * (uint64_t)(v < nvec ? v : nvec - 1));
      }
    } else {
#pragma unroll
      for (int j = 0; j < PV; j++) {
        const int v = lane + 64 * j;
        pf[j] = v < nvec ? load_vec16(las, total, a0 + 16ull * v) : make_uint4(0u, 0u, 0u, 0u);
      }
    }
  };
  auto put = [&]() {
#pragma unroll
    for (int j = 0; j < PV; j++) {
      const int v = lane + 64 * j;
      if (v < pnvec) *(uint4*)(rec_area + 16 * v) = pf[j];
    }
  };
  if constexpr (PV > 0) {
    if (step0 < step1) {
      issue(step0 * 64);
      put();
    }
  }
  for (int64_t step = step0; step < step1; step++) {
    const int64_t first = step * 64;
    const int cnt = (int)((n - first) < 64 ? (n - first) : 64);
    uint32_t shift;
    if constexpr (PV > 0) {
      if (step + 1 < step1) issue(first + 64);
      shift = (uint32_t)(((uint64_t)first * bpp) & 15u);
    } else {
      shift = wave_load_records(las, total, first, cnt, bpp, rec_area);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const bool valid = lane < cnt;
    const uint8_t* rec = rec_area + shift + lane * bpp;
    uint32_t ekey = kEmpty;  // cell, kEmpty for lanes without an in-box point
    if (valid) {
      bool inBox;
      int32_t q[3];
      uint32_t key = (uint32_t)cell_of_record(rec, P, inBox, q);
      if (inBox) ekey = key; else bad = true;
      if (cells) cells[first + lane] = ekey;
      if constexpr (NS > 0) {
        // the record's leading dwords: aligned LDS reads and byte aligns
        const uint32_t sh = (uint32_t)((uintptr_t)rec & 3);
        const uint32_t* a = (const uint32_t*)(rec - sh);
        uint32_t raw[NW + 1], w[NW];
#pragma unroll
        for (int k = 0; k <= NW; k++) raw[k] = a[k];
#pragma unroll
        for (int k = 0; k < NW; k++) w[k] = __builtin_amdgcn_alignbyte(raw[k + 1], raw[k], sh);
        fold_std<FMT, 0>(w, acc);
      }
    }
    // runs of equal cells: the first lane of a run adds the run length
    const uint32_t prev = __shfl_up(ekey, 1);
    const bool head = lane == 0 || prev != ekey;
    const uint64_t heads = __ballot(head);
    if (head && ekey != kEmpty) {
      const uint64_t after = heads & ~low_incl;
      const int end = after ? __builtin_ctzll(after) : 64;
      const uint32_t run = (uint32_t)(end - lane);
      uint32_t h = (ekey * 2654435761u) >> 21;  // 11 bits
      bool done = false;
      for (int probe = 0; probe < 8 && !done; probe++) {
        uint32_t slot = (h + probe) & (kHashSize - 1);
        uint32_t cur = s_key[slot];
        if (cur == kEmpty) {
          cur = atomicCAS(&s_key[slot], kEmpty, ekey);
          if (cur == kEmpty) cur = ekey;
        }
        if (cur == ekey) {
          atomicAdd(&s_cnt[slot], run);
          done = true;
        }
      }
      if (!done) atomicAdd(&grid[ekey], (int32_t)run);
    }
    if (P.class_src >= 0) {
      // one round per distinct class of the step for the first
      // kMatchRounds classes (a step rarely holds more), the bitwise match
      // over the lanes still left after that
      const uint32_t cm = (uint32_t)P.class_mask;
      const uint32_t c = valid ? ((uint32_t)rec[P.class_src] & cm) : cm + 1;
      uint64_t left = __ballot(valid);
      for (int r = 0; r < kMatchRounds && left; r++) {
        const uint32_t v = __builtin_amdgcn_readlane(c, __builtin_ctzll(left));
        const uint64_t m = __ballot(c == v) & left;
        if (lane == 0) atomicAdd(&s_hist[v], (uint32_t)__popcll(m));
        left &= ~m;
      }
      if (left) {
        const uint64_t peers = wave_match(c, cm == 31 ? 6 : 9, left);
        if (((left >> lane) & 1) && (peers & (low_incl >> 1)) == 0) atomicAdd(&s_hist[c], (uint32_t)__popcll(peers));
      }
    }
    if (my_stat >= 0) {
      for (int r = my_r0; r < cnt; r += R) {
        const uint64_t kk = stat_key(sl, lds_u64(rec_area + shift + r * bpp + sl.src));
        my_min = kk < my_min ? kk : my_min;
        my_max = kk > my_max ? kk : my_max;
        seen = true;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if constexpr (PV > 0) {
      if (step + 1 < step1) put();
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
  }
  if (bad) atomicOr(err, 1);
  // keys back to ordered double bits (a lane that saw no record keeps the
  // neutral ~0 / 0)
  s_smin[threadIdx.x] = seen ? stat_key_bits(sl, my_min) : ~0ull;
  s_smax[threadIdx.x] = seen ? stat_key_bits(sl, my_max) : 0ull;
  if constexpr (NS > 0) publish_std<FMT, 0>(acc, step0 < step1, lane, s_rmin[wave], s_rmax[wave]);
  __syncthreads();
  for (int i = threadIdx.x; i < kHashSize; i += kBlock) {
    uint32_t k = s_key[i];
    if (k != kEmpty && s_cnt[i]) atomicAdd(&grid[k], (int32_t)s_cnt[i]);
  }
  for (int i = threadIdx.x; i < 256; i += kBlock)
    if (s_hist[i]) atomicAdd(&class_hist[i], (unsigned long long)s_hist[i]);
  if ((int)threadIdx.x < NS) {
    uint64_t mn = ~0ull, mx = 0;
    for (int w = 0; w < kWavesPerBlock; w++) {
      mn = s_rmin[w][threadIdx.x] < mn ? s_rmin[w][threadIdx.x] : mn;
      mx = s_rmax[w][threadIdx.x] > mx ? s_rmax[w][threadIdx.x] : mx;
    }
    atomicMin((unsigned long long*)&stat_min[threadIdx.x], (unsigned long long)mn);
    atomicMax((unsigned long long*)&stat_max[threadIdx.x], (unsigned long long)mx);
  }
  if ((int)threadIdx.x < E) {
    uint64_t mn = ~0ull, mx = 0;
    for (int t = 0; t < kBlock; t++) {
      const int l = t & 63;
      if (l >= R * E || l % E != (int)threadIdx.x) continue;
      mn = s_smin[t] < mn ? s_smin[t] : mn;
      mx = s_smax[t] > mx ? s_smax[t] : mx;
    }
    atomicMin((unsigned long long*)&stat_min[NS + threadIdx.x], (unsigned long long)mn);
    atomicMax((unsigned long long*)&stat_max[NS + threadIdx.x], (unsigned long long)mx);
  }
}

// K2a: one pyramid level (chunker_countsort_laszip.cpp:1112-1180)
__global__ void k_pyramid_level(const int32_t* __restrict__ high, int32_t* __restrict__ low, int64_t gl,
                                int64_t cap) {
  int64_t cells = gl * gl * gl;
  int64_t gh = gl * 2;
  for (int64_t il = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; il < cells; il += (int64_t)gridDim.x * blockDim.x) {
    int64_t x = il % gl, y = (il / gl) % gl, z = il / (gl * gl);
    int64_t sum = 0;
    bool unmergeable = false;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      int64_t nx = 2 * x + ((j >> 2) & 1), ny = 2 * y + ((j >> 1) & 1), nz = 2 * z + (j & 1);
      int32_t v = high[nx + ny * gh + nz * gh * gh];
      if (v == -1) unmergeable = true; else sum += v;
    }
    low[il] = (unmergeable || sum > cap) ? -1 : (int32_t)sum;
  }
}

// K2b: cells of level l (values `cur`) with count > 0 under an unmergeable
// parent (`parent`, level l-1) become chunks (:1144-1172)
__global__ void k_chunk_cells(const int32_t* __restrict__ cur, const int32_t* __restrict__ parent, int64_t g,
                              int32_t level, int64_t* __restrict__ out, uint32_t* __restrict__ counter,
                              uint32_t max_out) {
  int64_t cells = g * g * g;
  int64_t gp = g / 2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cells; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = cur[i];
    if (v <= 0) continue;
    int64_t x = i % g, y = (i / g) % g, z = i / (g * g);
    int32_t pv = parent[(x / 2) + (y / 2) * gp + (z / 2) * gp * gp];
    if (pv != -1) continue;
    uint32_t slot = atomicAdd(counter, 1u);
    if (slot < max_out) {
      out[5 * slot + 0] = level; out[5 * slot + 1] = x; out[5 * slot + 2] = y; out[5 * slot + 3] = z;
      out[5 * slot + 4] = v;
    }
  }
}

// K2c: LUT[cell] = chunk index for every fine cell covered by a chunk
__global__ void k_lut_fill(const int64_t* __restrict__ chunks, int64_t num_chunks, int32_t level_max, int64_t G,
                           int32_t* __restrict__ lut) {
  for (int64_t c = blockIdx.x; c < num_chunks; c += gridDim.x) {
    int64_t level = chunks[4 * c + 0], x = chunks[4 * c + 1], y = chunks[4 * c + 2], z = chunks[4 * c + 3];
    int64_t s = (int64_t)1 << (level_max - level);
    int64_t cells = s * s * s;
    for (int64_t k = threadIdx.x; k < cells; k += blockDim.x) {
      int64_t ox = k % s, oy = (k / s) % s, oz = k / (s * s);
      lut[(s * x + ox) + (s * y + oy) * G + (s * z + oz) * G * G] = (int32_t)c;
    }
  }
}

// K3a: the chunk of every point (LUT entry of its cell, written over the
// cell; kEmpty where there is none) and the per-tile chunk histogram,
// chunk-major: M[c * tiles + t].  Each thread takes 4 consecutive points per
// pass (tile_size is a multiple of 4 * kBlock, so tiles start 16-byte
// aligned) and adds runs of equal chunks with one LDS atomic.
// IdT = uint32_t writes the chunks over the cells (ids_out == cells);
// uint16_t writes them to a separate array (kEmpty becomes 0xFFFF).
template <typename IdT>
__global__ __launch_bounds__(kBlock) void k_tile_hist(const uint32_t* cells, IdT* ids_out,
                                                       int64_t n, int32_t tile_size,
                                                       const int32_t* __restrict__ lut, int32_t num_chunks,
                                                       int64_t tiles, uint32_t* __restrict__ M,
                                                       int32_t* __restrict__ err) {
  extern __shared__ uint32_t s_hist[];  // num_chunks entries (or none: then global atomics)
  const bool lds_hist = num_chunks <= 8192;
  const int64_t t = blockIdx.x;
  if (lds_hist)
    for (int i = threadIdx.x; i < num_chunks; i += kBlock) s_hist[i] = 0;
  __syncthreads();
  const int64_t t0 = t * tile_size;
  const int64_t t1 = (t0 + tile_size) < n ? (t0 + tile_size) : n;
  bool bad = false;
  for (int64_t i = t0 + 4 * (int64_t)threadIdx.x; i < t1; i += 4 * kBlock) {
    const int m = (t1 - i) < 4 ? (int)(t1 - i) : 4;
    uint32_t v[4];
    if (m == 4) {
      const uint4 q = *(const uint4*)(cells + i);
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
      for (int j = 0; j < 4; j++) v[j] = j < m ? cells[i + j] : kEmpty;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int32_t c = -1;
      if (j < m && v[j] != kEmpty) c = lut[v[j]];
      if (j < m && c < 0) bad = true;
      v[j] = c < 0 ? kEmpty : (uint32_t)c;
    }
    if constexpr (sizeof(IdT) == 4) {
      if (m == 4) {
        *(uint4*)(ids_out + i) = make_uint4(v[0], v[1], v[2], v[3]);
      } else {
        for (int j = 0; j < m; j++) ids_out[i + j] = v[j];
      }
    } else {
      if (m == 4) {
        *(uint2*)(ids_out + i) = make_uint2((v[0] & 0xFFFFu) | (v[1] << 16), (v[2] & 0xFFFFu) | (v[3] << 16));
      } else {
        for (int j = 0; j < m; j++) ids_out[i + j] = (IdT)v[j];
      }
    }
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (j >= m || v[j] == kEmpty) continue;
      run++;
      if (j + 1 < m && v[j + 1] == v[j]) continue;
      if (lds_hist) atomicAdd(&s_hist[v[j]], run);
      else atomicAdd(&M[(int64_t)v[j] * tiles + t], run);
      run = 0;
    }
  }
  if (bad) atomicOr(err, 2);
  __syncthreads();
  if (lds_hist)
    for (int i = threadIdx.x; i < num_chunks; i += kBlock) M[(int64_t)i * tiles + t] = s_hist[i];
}


// bytes of one wave's output staging area in LDS
__host__ __device__ inline int out_area_bytes(int bpo) { return (64 * bpo + 16 + 15) & ~15; }

// dynamic LDS of k_scatter: per-wave chunk counters, record staging, output
// staging
__host__ __device__ inline size_t scatter_lds_bytes(int num_chunks, int bpp, int bpo) {
  return (((size_t)kWavesPerBlock * num_chunks * 4 + 15) & ~(size_t)15) +
         (size_t)kWavesPerBlock * (record_area_bytes(bpp) + out_area_bytes(bpo));
}

// Writes the step's runs from the LDS staging area to their places in the
// output: aligned dwords inside a run, bytes where a run starts or ends
// inside a word.  Leader lanes hold the run's first output record (`base`),
// its length and its start in the staging area.
__device__ inline void write_runs(uint64_t leaders, uint32_t base, uint32_t len, uint32_t run_start, int bpo,
                                  const uint8_t* out_area, uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  while (leaders) {
    const int L = __builtin_ctzll(leaders);
    leaders &= leaders - 1;
    const uint64_t g0 = (uint64_t)__builtin_amdgcn_readlane(base, L) * (uint64_t)bpo;
    const uint32_t head = (uint32_t)(g0 & 3);
    const uint32_t nb = (uint32_t)__builtin_amdgcn_readlane(len, L) * (uint32_t)bpo + head;
    uint8_t* gb = out + (g0 - head);  // 4-byte aligned; byte i <-> staging byte i - head
    const uint8_t* sb = out_area + (size_t)__builtin_amdgcn_readlane(run_start, L) * bpo;
    const uint32_t nw = (nb + 3) >> 2;
    for (uint32_t u = lane; u < nw; u += 64) {
      const uint32_t b = u << 2;
      if (b >= head && b + 4 <= nb) {
        *(uint32_t*)(gb + b) = lds_u32(sb + (b - head));
      } else {
        for (uint32_t k = b; k < b + 4; k++)
          if (k >= head && k < nb) gb[k] = sb[k - head];
      }
    }
  }
}

// K3b: transcode + stable scatter.  Each wave owns a contiguous quarter of
// the tile; pass 1 counts per chunk per wave, pass 2 places each point at
// M[c][t] + (earlier waves' count) + (running count) + (rank in step).  The
// 64 output records of a step are assembled in LDS grouped into runs of one
// chunk (a run is contiguous in the output) and written by write_runs.
// This kernel transcodes with the byte operations of ChunkerParams::ops (any
// record layout); k_distribute takes the layouts of point formats 0-3 and
// 6-7.
__global__ __launch_bounds__(kBlock) void k_scatter(const uint8_t* __restrict__ las,
                                                     const uint32_t* __restrict__ ids, int64_t n,
                                                     int32_t tile_size, int32_t num_chunks, int32_t key_bits,
                                                     int64_t tiles, const uint32_t* __restrict__ M,
                                                     uint8_t* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint8_t s_dyn[];
  const ChunkerParams& P = c_params;
  const int bpp = P.bpp_in, bpo = P.bpp_out;
  uint32_t* s_cnt = (uint32_t*)s_dyn;  // [kWavesPerBlock][num_chunks]
  uint8_t* s_stage = s_dyn + (((size_t)kWavesPerBlock * num_chunks * 4 + 15) & ~(size_t)15);
  const int64_t t = blockIdx.x;
  for (int i = threadIdx.x; i < kWavesPerBlock * num_chunks; i += kBlock) s_cnt[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint8_t* rec_area = s_stage + (size_t)wave * (record_area_bytes(bpp) + out_area_bytes(bpo));
  uint8_t* out_area = rec_area + record_area_bytes(bpp);
  const uint64_t total = (uint64_t)n * bpp;
  const int64_t t0 = t * tile_size;
  const int64_t t1 = (t0 + tile_size) < n ? (t0 + tile_size) : n;
  const int64_t per_wave = (tile_size + kWavesPerBlock - 1) / kWavesPerBlock;
  const int64_t w0 = t0 + wave * per_wave;
  const int64_t w1 = (w0 + per_wave) < t1 ? (w0 + per_wave) : t1;
  uint32_t* my_cnt = s_cnt + wave * num_chunks;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // pass 1: counts from the chunk ids
  for (int64_t first = w0; first < w1; first += 64) {
    int cnt = (int)((w1 - first) < 64 ? (w1 - first) : 64);
    uint32_t c = 0;
    bool valid = lane < cnt;
    if (valid) {
      const uint32_t cc = ids[first + lane];
      valid = cc < (uint32_t)num_chunks;
      c = valid ? cc : 0u;
    }
    uint64_t peers;
    uint32_t run;
    step_runs(c, valid, key_bits, lt_mask, peers, run);
    if (valid && (peers & lt_mask) == 0) my_cnt[c] += (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // exclusive prefix over waves, seeded with the global tile offset
  for (int c = threadIdx.x; c < num_chunks; c += kBlock) {
    uint32_t run = M[(int64_t)c * tiles + t];
    for (int w = 0; w < kWavesPerBlock; w++) {
      uint32_t v = s_cnt[w * num_chunks + c];
      s_cnt[w * num_chunks + c] = run;
      run += v;
    }
  }
  __syncthreads();
  // pass 2: transcode + scatter
  const int out_words = (64 * bpo + 3) / 4;
  for (int64_t first = w0; first < w1; first += 64) {
    int cnt = (int)((w1 - first) < 64 ? (w1 - first) : 64);
    uint32_t shift = wave_load_records(las, total, first, cnt, bpp, rec_area);
    for (int i = lane; i < out_words; i += 64) ((uint32_t*)out_area)[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    bool valid = lane < cnt;
    const uint8_t* rec = rec_area + shift + lane * bpp;
    uint32_t c = 0;
    int32_t q[3] = {0, 0, 0};
    if (valid) {
      const uint32_t cc = ids[first + lane];
      valid = cc < (uint32_t)num_chunks;
      c = valid ? cc : 0u;
      quantize_record(rec, P, q);
    }
    uint64_t peers;
    uint32_t my_run;
    step_runs(c, valid, key_bits, lt_mask, peers, my_run);
    const uint32_t base = valid ? my_cnt[c] : 0;
    __builtin_amdgcn_wave_barrier();
    const bool leader = valid && (peers & lt_mask) == 0;
    const uint32_t len = leader ? (uint32_t)__popcll(peers) : 0u;
    if (leader) my_cnt[c] = base + len;
    if (valid) {
      uint8_t* dst = out_area + (size_t)(my_run + (uint32_t)__popcll(peers & lt_mask)) * bpo;
      for (int a = 0; a < 3; a++) {
        const uint32_t v = (uint32_t)q[a];
        dst[4 * a + 0] = (uint8_t)v; dst[4 * a + 1] = (uint8_t)(v >> 8);
        dst[4 * a + 2] = (uint8_t)(v >> 16); dst[4 * a + 3] = (uint8_t)(v >> 24);
      }
      for (int o = 0; o < P.num_ops; o++) {
        const TranscodeOp op = P.ops[o];
        if (op.mode == TC_COPY) {
          for (int b = 0; b < op.len; b++) dst[op.dst + b] = rec[op.src + b];
        } else {
          const uint32_t v = rec[op.src];
          uint32_t o;
          switch (op.mode) {
            case TC_LOW3: o = v & 7; break;
            case TC_MID3: o = (v >> 3) & 7; break;
            case TC_RET14: o = legacy_return_number(v); break;
            case TC_NRET14: o = legacy_number_of_returns(v); break;
            case TC_LOW4: o = v & 15; break;
            default: o = v & 31; break;
          }
          dst[op.dst] = (uint8_t)o;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    write_runs(__ballot(leader), base, len, my_run, bpo, out_area, out);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
}

// K3b for the record layouts of point formats 0-3 and 6-7 (with or without
// extra bytes), whose output record is: the requantized position (12
// bytes), intensity (LAS bytes 12-13), three bytes derived from LAS bytes 14
// and 15 (formats 0-3: return number, number of returns, 5-bit class;
// formats 6-7, EXT: legacy return number, legacy number of returns,
// classification flags), then LAS bytes 16 .. bpp - 1 one byte later
// (chunker_countsort_laszip.cpp:391-651 writes them attribute by attribute;
// every one of them is a plain copy).  Each lane builds its output record as
// dwords: one byte align per dword from the staged LAS record, one byte
// permute to the record's alignment in the staging area, plain LDS stores
// for the dwords that lie inside the record and LDS ORs (the area is zeroed
// per step) for the two shared with the neighbouring records.

// ------------------------------------------------------ tile transcode
// Transcode + scatter for the record layouts of point formats 0-3 and 6-7
// (chunker_countsort_laszip.cpp:838-904: records grouped by chunk, input
// order inside a chunk), after k_tile_hist<uint16_t> wrote every record's
// chunk and the scan gave every (chunk, tile) its first output record.
// Each wave counts its quarter of the tile per chunk, then transcodes it 64
// records per step.  The records of one chunk in a step form one run of the
// output; a run is staged in LDS at the 16-byte phase of its destination
// and leaves as aligned 16-byte stores (aligned pieces of 8, 4, 2, 1 bytes
// at its two ends).
//
// gfx950 counts loads and stores of a wave on one in-order counter (vmcnt),
// so waiting for a load also waits for every store issued before it.  A
// step's stores are therefore issued at the start of the next iteration,
// ahead of that iteration's prefetch, and the wait for the prefetch comes
// after a whole transcode: both have that long to complete.
constexpr int kDistTile = 4096;
constexpr int kDistPerWave = kDistTile / kWavesPerBlock;  // 1024 records
constexpr int kDistSteps = kDistPerWave / 64;              // 16 steps
constexpr uint32_t kNoChunk16 = 0xFFFFu;

__host__ __device__ inline size_t align16z(size_t v) { return (v + 15) & ~(size_t)15; }

// one wave's output staging: 64 records, each run shifted to the phase of
// its destination (at most 15 bytes of padding per run), the dword after
// the last record
__host__ __device__ inline int stage_area_bytes(int bpo) { return (64 * bpo + 16 * 64 + 32 + 15) & ~15; }

// dynamic LDS of k_distribute: per-wave chunk bases, per-wave run table,
// the tile's chunk ids, per wave the record buffer and the output staging
__host__ __device__ inline size_t distribute_lds_bytes(int num_chunks, int bpp, int bpo) {
  return align16z((size_t)kWavesPerBlock * num_chunks * 4) + (size_t)kWavesPerBlock * 64 * 16 +
         (size_t)kDistTile * 2 + (size_t)kWavesPerBlock * (record_area_bytes(bpp) + stage_area_bytes(bpo));
}

template <bool EXT, int PV>
__global__ __launch_bounds__(kBlock) void k_distribute(const uint8_t* __restrict__ las,
                                                        const uint16_t* __restrict__ ids, int64_t n,
                                                        int32_t num_chunks, int64_t tiles,
                                                        const uint32_t* __restrict__ M,
                                                        uint8_t* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint8_t s_dyn[];
  const ChunkerParams& P = c_params;
  const int bpp = P.bpp_in, bpo = P.bpp_out;  // bpo = bpp + 1
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  uint32_t* s_base = (uint32_t*)s_dyn;  // [wave][num_chunks]: counts, then next output record
  uint4* s_runs = (uint4*)(s_dyn + align16z((size_t)kWavesPerBlock * num_chunks * 4));  // [wave][64]
  uint16_t* s_ids = (uint16_t*)(s_runs + kWavesPerBlock * 64);  // [kDistTile]: chunk of each record
  uint8_t* s_stage = (uint8_t*)(s_ids + kDistTile);
  const int rab = record_area_bytes(bpp);
  uint8_t* rec_area = s_stage + (size_t)wave * (rab + stage_area_bytes(bpo));
  uint8_t* out_area = rec_area + rab;
  for (int i = tid; i < kWavesPerBlock * num_chunks; i += kBlock) s_base[i] = 0;
  const int64_t t = blockIdx.x;
  const int64_t t0 = t * kDistTile;
  const int64_t t1 = (t0 + kDistTile) < n ? (t0 + kDistTile) : n;
  const int64_t w0 = (t0 + wave * kDistPerWave) < t1 ? (t0 + wave * kDistPerWave) : t1;
  const int64_t w1 = (w0 + kDistPerWave) < t1 ? (w0 + kDistPerWave) : t1;
  const int steps = (int)((w1 - w0 + 63) / 64);
  uint32_t* my_base = s_base + wave * num_chunks;
  uint4* my_runs = s_runs + wave * 64;
  uint16_t* my_ids = s_ids + wave * kDistPerWave;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t total = (uint64_t)n * bpp;

  // a step's records as 16-byte vectors, without a branch per lane (lanes
  // past the step reload its last vector) except in the one step that
  // reaches the end of the input
  uint4 pf[PV];
  int pnvec = 0;  // vectors of the step in pf
  auto issue = [&](int64_t f) {
    const int m = (int)((w1 - f) < 64 ? (w1 - f) : 64);
    const uint64_t b0 = (uint64_t)f * bpp, a0 = b0 & ~15ull;
    const int nvec = (int)((b0 + (uint64_t)m * bpp - a0 + 15) >> 4);
    pnvec = nvec;
    if (a0 + 16ull * (uint64_t)nvec <= total) {
#pragma unroll
      for (int j = 0; j < PV; j++) {
        const int v = lane + 64 * j;
        pf[j] = *(const uint4*)(las + a0 + 16ull * (uint64_t)(v < nvec ? v : nvec - 1));
      }
    } else {
#pragma unroll
      for (int j = 0; j < PV; j++) {
        const int v = lane + 64 * j;
        pf[j] = v < nvec ? load_vec16(las, total, a0 + 16ull * v) : make_uint4(0u, 0u, 0u, 0u);
      }
    }
  };
  // the step's vectors only: the record area holds record_area_bytes(bpp)
  auto put = [&]() {
#pragma unroll
    for (int j = 0; j < PV; j++) {
      const int v = lane + 64 * j;
      if (v < pnvec) *(uint4*)(rec_area + 16 * v) = pf[j];
    }
  };
  // the step's runs (staged in out_area, listed in my_runs) to the output
  auto write_runs = [&](int nr) {
    for (int r = 0; r < nr; r++) {
      const uint4 R = my_runs[r];
      const uint32_t ph = R.z & 15u;  // == R.x & 15
      const uint8_t* sv = out_area + (R.x - ph);
      uint8_t* gv = out + ((((uint64_t)R.w << 32) | R.z) - ph);
      const uint32_t span = ph + R.y;
      const uint32_t nvec = (span + 15) >> 4;
      for (uint32_t k = lane; k < nvec; k += 64) {
        const uint32_t b0 = k << 4;
        if (b0 >= ph && b0 + 16 <= span) {
          *(uint4*)(gv + b0) = *(const uint4*)(sv + b0);
        } else {
          // the run's bytes [lo, hi) of an end vector in naturally aligned
          // pieces: up to 8-byte alignment, then descending sizes
          uint32_t lo = b0 < ph ? ph : b0;
          const uint32_t hi = b0 + 16 < span ? b0 + 16 : span;
          if ((lo & 1) && lo < hi) { gv[lo] = sv[lo]; lo += 1; }
          if ((lo & 2) && lo + 2 <= hi) { *(uint16_t*)(gv + lo) = *(const uint16_t*)(sv + lo); lo += 2; }
          if ((lo & 4) && lo + 4 <= hi) { *(uint32_t*)(gv + lo) = *(const uint32_t*)(sv + lo); lo += 4; }
          if ((lo & 8) && lo + 8 <= hi) { *(uint2*)(gv + lo) = *(const uint2*)(sv + lo); lo += 8; }
          if (lo + 8 <= hi) { *(uint2*)(gv + lo) = *(const uint2*)(sv + lo); lo += 8; }
          if (lo + 4 <= hi) { *(uint32_t*)(gv + lo) = *(const uint32_t*)(sv + lo); lo += 4; }
          if (lo + 2 <= hi) { *(uint16_t*)(gv + lo) = *(const uint16_t*)(sv + lo); lo += 2; }
          if (lo < hi) gv[lo] = sv[lo];
        }
      }
    }
  };

  // the wave's chunk ids into LDS (read there by the transcode: a global
  // load inside a step would wait for the step's prefetch and stores), the
  // first step's records, the counts per chunk (one round per chunk)
  {
    uint32_t cv[kDistSteps];
#pragma unroll
    for (int j = 0; j < kDistSteps; j++) {
      const int64_t i = w0 + 64 * j + lane;
      cv[j] = i < w1 ? ids[i] : kNoChunk16;
    }
#pragma unroll
    for (int j = 0; j < kDistSteps; j++) my_ids[64 * j + lane] = (uint16_t)cv[j];
  }
  if (steps > 0) {
    issue(w0);
    put();
  }
  __syncthreads();
  for (int st = 0; st < steps; st++) {
    const uint32_t c = my_ids[64 * st + lane];
    for (uint64_t left = __ballot(c != kNoChunk16); left;) {
      const uint32_t v = __builtin_amdgcn_readlane(c, __builtin_ctzll(left));
      const uint64_t m = __ballot(c == v) & left;
      if (lane == 0) my_base[v] += (uint32_t)__popcll(m);
      left &= ~m;
    }
  }
  __syncthreads();
  // first output record of every chunk for each wave: the tile's (scanned
  // matrix) plus the earlier waves' counts
  for (int c = tid; c < num_chunks; c += kBlock) {
    uint32_t run = M[(int64_t)c * tiles + t];
#pragma unroll
    for (int q = 0; q < kWavesPerBlock; q++) {
      const uint32_t v = s_base[q * num_chunks + c];
      s_base[q * num_chunks + c] = run;
      run += v;
    }
  }
  __syncthreads();

  const int ow = (bpo + 3) / 4;     // output dwords of a record
  const int inner = (bpo - 4) / 4;  // dwords 1 .. inner lie inside the record at any alignment
  const uint32_t last_mask = (bpo & 3) ? ((1u << (8 * (bpo & 3))) - 1u) : 0xFFFFFFFFu;
  int nr_prev = 0;  // runs of the previous step, still to be written
  for (int st = 0; st < steps; st++) {
    const int64_t first = w0 + 64 * (int64_t)st;
    const int cnt = (int)((w1 - first) < 64 ? (w1 - first) : 64);
    // the previous step's stores, then the next step's records
    write_runs(nr_prev);
    if (st + 1 < steps) issue(first + 64);
    const uint32_t shift = (uint32_t)(((uint64_t)first * bpp) & 15u);
    const uint32_t c = lane < cnt ? (uint32_t)my_ids[64 * st + lane] : kNoChunk16;
    const bool valid = c != kNoChunk16;
    // runs in the order of their lowest lane, each staged at the 16-byte
    // phase of its first output byte
    uint32_t sofs = 0, my_so = 0;
    int nr = 0;
    for (uint64_t left = __ballot(valid); left; nr++) {
      const uint32_t v = __builtin_amdgcn_readlane(c, __builtin_ctzll(left));
      const uint64_t m = __ballot(c == v) & left;
      const uint32_t len = (uint32_t)__popcll(m);
      const uint32_t b = my_base[v];
      const uint64_t g0 = (uint64_t)b * (uint32_t)bpo;
      sofs += (uint32_t)(g0 - sofs) & 15u;
      if ((m >> lane) & 1) my_so = sofs + (uint32_t)__popcll(m & lt_mask) * (uint32_t)bpo;
      if (lane == 0) {
        my_base[v] = b + len;
        my_runs[nr] = make_uint4(sofs, len * (uint32_t)bpo, (uint32_t)g0, (uint32_t)(g0 >> 32));
      }
      sofs += len * (uint32_t)bpo;
      left &= ~m;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (uint32_t i = lane; i < (sofs + 8 + 15) / 16; i += 64) ((uint4*)out_area)[i] = make_uint4(0u, 0u, 0u, 0u);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (valid) {
      // the LAS record's dwords from aligned LDS reads
      const uint8_t* rec = rec_area + shift + lane * bpp;
      const uint32_t sh = (uint32_t)((uintptr_t)rec & 3);
      const uint32_t* ra = (const uint32_t*)(rec - sh);
      const uint32_t r0 = ra[0], r1 = ra[1], r2 = ra[2], r3 = ra[3], r4 = ra[4];
      const uint32_t wd[4] = {__builtin_amdgcn_alignbyte(r1, r0, sh), __builtin_amdgcn_alignbyte(r2, r1, sh),
                              __builtin_amdgcn_alignbyte(r3, r2, sh), __builtin_amdgcn_alignbyte(r4, r3, sh)};
      uint32_t o[4];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const double x = P.las_scale[a] * (int32_t)wd[a] + P.las_offset[a];
        o[a] = (uint32_t)(int32_t)((x - P.pos_offset[a]) / P.pos_scale[a]);
      }
      const uint32_t b14 = (wd[3] >> 16) & 0xFFu, b15 = wd[3] >> 24;
      uint32_t ret, nret, cls;
      if constexpr (EXT) {
        ret = legacy_return_number(b14);
        nret = legacy_number_of_returns(b14);
        cls = b15 & 15u;
      } else {
        ret = b14 & 7u;
        nret = (b14 >> 3) & 7u;
        cls = b15 & 31u;
      }
      o[3] = (wd[3] & 0xFFFFu) | (ret << 16) | (nret << 24);
      // output dword k >= 4 holds LAS bytes 4k-1 .. 4k+2 (byte 16 of the
      // output is the class byte): LAS byte offset sh + 4k - 1 from ra
      const uint32_t* rb = ra + ((sh + 3) >> 2) - 1;
      const uint32_t s2 = (sh + 3) & 3;
      const uint32_t off = my_so;
      const uint32_t sel = 0x07060504u - (off & 3) * 0x01010101u;
      uint32_t* dw = (uint32_t*)out_area + (off >> 2);
      atomicOr(dw, __builtin_amdgcn_perm(o[0], 0u, sel));
      dw[1] = __builtin_amdgcn_perm(o[1], o[0], sel);
      dw[2] = __builtin_amdgcn_perm(o[2], o[1], sel);
      dw[3] = __builtin_amdgcn_perm(o[3], o[2], sel);
      uint32_t prev = o[3];
      uint32_t lo = rb[4];
      for (int k = 4; k < ow; k++) {
        const uint32_t hi = rb[k + 1];
        uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, s2);
        lo = hi;
        if (k == 4) v = (v & 0xFFFFFF00u) | cls;
        if (k == ow - 1) v &= last_mask;
        const uint32_t x = __builtin_amdgcn_perm(v, prev, sel);
        if (k <= inner) dw[k] = x;
        else atomicOr(dw + k, x);
        prev = v;
      }
      atomicOr(dw + ow, __builtin_amdgcn_perm(0u, prev, sel));
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // the next step's records into the record area (read by this step's
    // transcode above), after a whole step of latency
    if (st + 1 < steps) put();
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    nr_prev = nr;
  }
  write_runs(nr_prev);
}

__global__ void k_lasgen(LasGenParams p, const LasGenBuilding* __restrict__ bld, int64_t first, int64_t count,
                         uint8_t* __restrict__ out, int32_t* __restrict__ qmin, int32_t* __restrict__ qmax) {
  __shared__ LasGenBuilding s_b[64];
  for (int i = threadIdx.x; i < p.num_buildings && i < 64; i += blockDim.x) s_b[i] = bld[i];
  __syncthreads();
  const int rl = lasgen_record_length(p);
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  uint8_t rec[72];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    lasgen_point(p, s_b, first + i, rec);
    uint8_t* d = out + i * rl;
    for (int b = 0; b < rl; b++) d[b] = rec[b];
    for (int a = 0; a < 3; a++) {
      int32_t v = ld_i32(rec + 4 * a);
      mn[a] = v < mn[a] ? v : mn[a];
      mx[a] = v > mx[a] ? v : mx[a];
    }
  }
  for (int a = 0; a < 3; a++) {
    for (int d = 32; d >= 1; d >>= 1) {
      int32_t o1 = __shfl_xor(mn[a], d, 64), o2 = __shfl_xor(mx[a], d, 64);
      mn[a] = o1 < mn[a] ? o1 : mn[a];
      mx[a] = o2 > mx[a] ? o2 : mx[a];
    }
  }
  if ((threadIdx.x & 63) == 0) {
    for (int a = 0; a < 3; a++) { atomicMin(&qmin[a], mn[a]); atomicMax(&qmax[a], mx[a]); }
  }
}

int g_bpp_in = 0;   // host copies of the record lengths in c_params
int g_bpp_out = 0;
int g_point_format = 8;  // k_count instantiation (8: no standard statistics)
int g_tail_layout = -1;  // k_distribute<false> (0), <true> (1), or k_scatter (-1)

// The layout k_distribute writes: intensity copied, the three derived
// bytes at 14-16, then LAS bytes 16 .. bpp - 1 copied one byte later.
int tail_layout(const ChunkerParams& p) {
  if (p.bpp_out != p.bpp_in + 1 || p.bpp_in < 17 || p.num_ops < 4) return -1;
  const TranscodeOp& i0 = p.ops[0];
  if (i0.src != 12 || i0.dst != 12 || i0.len != 2 || i0.mode != TC_COPY) return -1;
  int layout = -1;
  const int legacy[3] = {TC_LOW3, TC_MID3, TC_LOW5}, ext[3] = {TC_RET14, TC_NRET14, TC_LOW4};
  for (int v = 0; v < 2; v++) {
    const int* modes = v ? ext : legacy;
    bool ok = true;
    for (int k = 0; k < 3; k++) {
      const TranscodeOp& o = p.ops[1 + k];
      ok = ok && o.src == (k < 2 ? 14 : 15) && o.dst == 14 + k && o.len == 1 && o.mode == modes[k];
    }
    if (ok) layout = v;
  }
  if (layout < 0) return -1;
  int next = 16;
  for (int k = 4; k < p.num_ops; k++) {
    const TranscodeOp& o = p.ops[k];
    if (o.mode != TC_COPY || o.src != next || o.dst != o.src + 1 || o.len <= 0) return -1;
    next += o.len;
  }
  return next == p.bpp_in ? layout : -1;
}

// Fills n 32-bit words with v (a kernel rather than hipMemsetAsync, so that
// callers may capture the no-allocation entry point in a graph).
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Scratch of the no-allocation counting entry point: k_count's statistics
// minima and maxima (kMaxStats each), the classification histogram (256)
// and the error flags.  Every GPU the library runs on holds one copy until
// the process ends.  Host code never passes this name to a kernel: it is a
// host-side stand-in, and its device address comes from
// hipGetSymbolAddress (noalloc_scratch below).
constexpr size_t kNoAllocStatMin = 0;
constexpr size_t kNoAllocStatMax = kNoAllocStatMin + 8 * kMaxStats;
constexpr size_t kNoAllocHist = kNoAllocStatMax + 8 * kMaxStats;
constexpr size_t kNoAllocErr = kNoAllocHist + 8 * 256;
constexpr size_t kNoAllocBytes = 4096;
static_assert(kNoAllocErr + 16 <= kNoAllocBytes, "no-allocation scratch too small");
__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
uint8_t* g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mutex;

void fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace

int noalloc_scratch(uint8_t** out) {
  int dev = 0;
  PCG_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index outside the no-allocation scratch table");
    return 1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PCG_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (uint8_t*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int chunker_count_noalloc(const uint8_t* d_las, int64_t n, int32_t* d_grid, int64_t cells, int num_cus,
                          hipStream_t stream, int32_t* h_err) {
  uint8_t* scr = nullptr;
  if (noalloc_scratch(&scr)) return 1;
  uint64_t* smin = (uint64_t*)(scr + kNoAllocStatMin);
  uint64_t* smax = (uint64_t*)(scr + kNoAllocStatMax);
  unsigned long long* hist = (unsigned long long*)(scr + kNoAllocHist);
  int32_t* err = (int32_t*)(scr + kNoAllocErr);
  // the previous call's contents stay in the scratch: reset what starts at
  // a neutral value, and the caller's grid
  fill_u32((uint32_t*)smin, 2 * kMaxStats, 0xFFFFFFFFu, stream);
  fill_u32((uint32_t*)smax, 2 * kMaxStats, 0u, stream);
  fill_u32((uint32_t*)hist, 2 * 256, 0u, stream);
  fill_u32((uint32_t*)err, 4, 0u, stream);
  fill_u32((uint32_t*)d_grid, cells, 0u, stream);
  PCG_CHECK(hipGetLastError());
  if (chunker_count(d_las, n, d_grid, nullptr, smin, smax, hist, err, num_cus, stream)) return 1;
  PCG_CHECK(hipMemcpyAsync(h_err, err, 4, hipMemcpyDeviceToHost, stream));
  PCG_CHECK(hipStreamSynchronize(stream));
  return 0;
}

int chunker_set_params(const ChunkerParams& p, hipStream_t stream) {
  if (p.bpp_in > kMaxBppIn) {
    set_error("LAS record length above 128 bytes is not supported");
    return 1;
  }
  g_bpp_in = p.bpp_in;
  g_bpp_out = p.bpp_out;
  g_tail_layout = tail_layout(p);
  static ChunkerParams staged;  // the async copy reads it after this call returns
  PCG_CHECK(hipStreamSynchronize(stream));
  staged = p;
  staged.g_over_cube = (double)p.grid_size / p.cube;
  if (staged.class_mask == 0) staged.class_mask = 31;
  g_point_format = (p.point_format >= 0 && p.point_format <= 7) ? p.point_format : 8;
  const int ns = std_stat_count(g_point_format);
  if (p.num_stats < ns || p.num_stats - ns > 64) {
    set_error("statistics table does not start with the point format's standard statistics");
    return 1;
  }
  for (int i = 0; i < ns; i++) {
    const StdStat st = std_stat(g_point_format, i);
    if (p.stats[i].src != st.src || p.stats[i].kind != st.kind) {
      set_error("statistics table does not start with the point format's standard statistics");
      return 1;
    }
  }
  PCG_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_params), &staged, sizeof(staged), 0, hipMemcpyHostToDevice, stream));
  return 0;
}

int chunker_count(const uint8_t* d_las, int64_t n, int32_t* d_grid, uint32_t* d_cells, uint64_t* d_stat_min,
                  uint64_t* d_stat_max, unsigned long long* d_class_hist, int32_t* d_err, int num_cus,
                  hipStream_t stream) {
  int64_t steps = (n + 63) / 64;
  int64_t blocks = (steps + kWavesPerBlock * kCountSteps - 1) / (kWavesPerBlock * kCountSteps);
  size_t lds = (size_t)kWavesPerBlock * record_area_bytes(g_bpp_in);
  // records of up to 32 or 64 bytes take the prefetching forms; the grid is
  // what stays resident at once (every wave walks an equal range of steps,
  // so a second round of blocks would run on a partly filled GPU)
  const int pv = g_bpp_in <= 32 ? prefetch_vecs(32) : (g_bpp_in <= 64 ? prefetch_vecs(64) : 0);
#define PCG_COUNT_FN(F, V) (const void*)k_count<F, V>
#define PCG_COUNT_PV(F) (pv == prefetch_vecs(32) ? PCG_COUNT_FN(F, prefetch_vecs(32)) \
                         : (pv == prefetch_vecs(64) ? PCG_COUNT_FN(F, prefetch_vecs(64)) : PCG_COUNT_FN(F, 0)))
  const void* fn = nullptr;
  switch (g_point_format) {
    case 0: fn = PCG_COUNT_PV(0); break;
    case 1: fn = PCG_COUNT_PV(1); break;
    case 2: fn = PCG_COUNT_PV(2); break;
    case 3: fn = PCG_COUNT_PV(3); break;
    case 4: fn = PCG_COUNT_PV(4); break;
    case 5: fn = PCG_COUNT_PV(5); break;
    case 6: fn = PCG_COUNT_PV(6); break;
    case 7: fn = PCG_COUNT_PV(7); break;
    default: fn = PCG_COUNT_PV(8); break;
  }
#undef PCG_COUNT_PV
#undef PCG_COUNT_FN
  int per_cu = 0;
  PCG_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, lds));
  int64_t max_blocks = (int64_t)num_cus * (per_cu > 0 ? per_cu : 1);
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks < 1) blocks = 1;
#define PCG_COUNT(F)                                                                                      \
  do {                                                                                                    \
    if (pv == prefetch_vecs(32))                                                                          \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_count<F, prefetch_vecs(32)>), dim3((unsigned)blocks),          \
                         dim3(kBlock), lds, stream, d_las, n, d_grid, d_cells, d_stat_min, d_stat_max,      \
                         d_class_hist, d_err);                                                            \
    else if (pv == prefetch_vecs(64))                                                                     \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_count<F, prefetch_vecs(64)>), dim3((unsigned)blocks),          \
                         dim3(kBlock), lds, stream, d_las, n, d_grid, d_cells, d_stat_min, d_stat_max,      \
                         d_class_hist, d_err);                                                            \
    else                                                                                                  \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_count<F, 0>), dim3((unsigned)blocks), dim3(kBlock), lds,       \
                         stream, d_las, n, d_grid, d_cells, d_stat_min, d_stat_max, d_class_hist, d_err);  \
  } while (0)
  switch (g_point_format) {
    case 0: PCG_COUNT(0); break;
    case 1: PCG_COUNT(1); break;
    case 2: PCG_COUNT(2); break;
    case 3: PCG_COUNT(3); break;
    case 4: PCG_COUNT(4); break;
    case 5: PCG_COUNT(5); break;
    case 6: PCG_COUNT(6); break;
    case 7: PCG_COUNT(7); break;
    default: PCG_COUNT(8); break;
  }
#undef PCG_COUNT
  PCG_CHECK(hipGetLastError());
  return 0;
}

int chunker_pyramid(int32_t* const* d_levels, int32_t level_max, int64_t cap, hipStream_t stream) {
  for (int32_t l = level_max - 1; l >= 0; l--) {
    int64_t gl = (int64_t)1 << l;
    int64_t cells = gl * gl * gl;
    int64_t blocks = (cells + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pyramid_level, dim3((unsigned)blocks), dim3(256), 0, stream, d_levels[l + 1], d_levels[l],
                       gl, cap);
    PCG_CHECK(hipGetLastError());
  }
  return 0;
}

int chunker_chunk_cells(int32_t* const* d_levels, int32_t level_max, int64_t* d_out, uint32_t* d_counter,
                        uint32_t max_out, hipStream_t stream) {
  for (int32_t l = 1; l <= level_max; l++) {
    int64_t g = (int64_t)1 << l;
    int64_t cells = g * g * g;
    int64_t blocks = (cells + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_chunk_cells, dim3((unsigned)blocks), dim3(256), 0, stream, d_levels[l], d_levels[l - 1],
                       g, l, d_out, d_counter, max_out);
    PCG_CHECK(hipGetLastError());
  }
  return 0;
}

int chunker_lut_fill(const int64_t* d_chunks, int64_t num_chunks, int32_t level_max, int32_t* d_lut,
                     hipStream_t stream) {
  int64_t G = (int64_t)1 << level_max;
  PCG_CHECK(hipMemsetAsync(d_lut, 0xFF, sizeof(int32_t) * G * G * G, stream));
  if (num_chunks == 0) return 0;
  int64_t blocks = num_chunks < 4096 ? num_chunks : 4096;
  hipLaunchKernelGGL(k_lut_fill, dim3((unsigned)blocks), dim3(256), 0, stream, d_chunks, num_chunks, level_max, G,
                     d_lut);
  PCG_CHECK(hipGetLastError());
  return 0;
}

bool distribute_tail_fits(int32_t num_chunks) {
  return (g_tail_layout == 0 || g_tail_layout == 1) && g_bpp_in <= 64 && num_chunks < (int32_t)kNoChunk16 &&
         distribute_lds_bytes(num_chunks, g_bpp_in, g_bpp_out) <= 150 * 1024;
}

int chunker_tile_hist(uint32_t* d_cells, uint16_t* d_ids16, int64_t n, const int32_t* d_lut, int32_t num_chunks,
                      uint32_t* d_M, int64_t tile_base, int64_t stride, int32_t* d_err, hipStream_t stream) {
  const int64_t tiles = (n + kDistTile - 1) / kDistTile;
  if (tiles == 0) return 0;
  if (tile_base < 0 || tile_base + tiles > stride || ((uintptr_t)d_cells & 15) != 0 ||
      (d_ids16 && ((uintptr_t)d_ids16 & 7) != 0)) {
    set_error("tile histogram: tiles outside the matrix or misaligned id arrays");
    return 1;
  }
  const size_t hist_lds = num_chunks <= 8192 ? sizeof(uint32_t) * num_chunks : 0;
  if (d_ids16) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile_hist<uint16_t>), dim3((unsigned)tiles), dim3(kBlock), hist_lds, stream,
                       d_cells, d_ids16, n, kDistTile, d_lut, num_chunks, stride, d_M + tile_base, d_err);
  } else {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile_hist<uint32_t>), dim3((unsigned)tiles), dim3(kBlock), hist_lds, stream,
                       d_cells, d_cells, n, kDistTile, d_lut, num_chunks, stride, d_M + tile_base, d_err);
  }
  PCG_CHECK(hipGetLastError());
  return 0;
}

int chunker_place(const uint8_t* d_las, const uint32_t* d_ids32, const uint16_t* d_ids16, int64_t n,
                  int32_t num_chunks, const uint32_t* d_M, int64_t tile_base, int64_t stride, uint8_t* d_out,
                  hipStream_t stream) {
  const int64_t tiles = (n + kDistTile - 1) / kDistTile;
  if (tiles == 0) return 0;
  if (tile_base < 0 || tile_base + tiles > stride) {
    set_error("placement: tiles outside the matrix");
    return 1;
  }
  const uint32_t* M = d_M + tile_base;
  if (!d_ids16) {
    // the byte-wise transcode of any record layout (the layouts of point
    // formats 0-3 and 6-7 normally take the tile transcode)
    int key_bits = 1;
    while ((1 << key_bits) < num_chunks) key_bits++;
    const size_t cnt_lds = scatter_lds_bytes(num_chunks, g_bpp_in, g_bpp_out);
    if (cnt_lds > 150 * 1024) {
      set_error("too many chunks for the LDS scatter counters");
      return 1;
    }
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)tiles), dim3(kBlock), cnt_lds, stream, d_las, d_ids32, n, kDistTile,
                       num_chunks, key_bits, stride, M, d_out);
    PCG_CHECK(hipGetLastError());
    return 0;
  }
  if (!distribute_tail_fits(num_chunks)) {
    set_error("record layout or chunk count outside the tile transcode");
    return 1;
  }
  const size_t lds = distribute_lds_bytes(num_chunks, g_bpp_in, g_bpp_out);
  static bool attr = false;
  if (!attr) {
    PCG_CHECK(hipFuncSetAttribute((const void*)k_distribute<false, prefetch_vecs(32)>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    PCG_CHECK(hipFuncSetAttribute((const void*)k_distribute<false, prefetch_vecs(64)>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    PCG_CHECK(hipFuncSetAttribute((const void*)k_distribute<true, prefetch_vecs(32)>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    PCG_CHECK(hipFuncSetAttribute((const void*)k_distribute<true, prefetch_vecs(64)>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr = true;
  }
#define PCG_DIST(E, V)                                                                                    \
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_distribute<E, V>), dim3((unsigned)tiles), dim3(kBlock), lds, stream, d_las, \
                     d_ids16, n, num_chunks, stride, M, d_out)
  const bool short_rec = g_bpp_in <= 32;
  if (g_tail_layout == 0) {
    if (short_rec) PCG_DIST(false, prefetch_vecs(32));
    else PCG_DIST(false, prefetch_vecs(64));
  } else {
    if (short_rec) PCG_DIST(true, prefetch_vecs(32));
    else PCG_DIST(true, prefetch_vecs(64));
  }
#undef PCG_DIST
  PCG_CHECK(hipGetLastError());
  return 0;
}

int chunker_distribute(const uint8_t* d_las, uint32_t* d_ids, int64_t n, const int32_t* d_lut, int32_t num_chunks,
                       int32_t tile_size, uint32_t* d_M, void* d_scan_ws, uint8_t* d_out, int32_t* d_err,
                       hipStream_t stream) {
  const int64_t tiles = (n + kDistTile - 1) / kDistTile;
  if (tiles == 0) return 0;
  if (tile_size != kDistTile) {
    set_error("chunk id tiles are 4096 points");
    return 1;
  }
  if (num_chunks > 8192) PCG_CHECK(hipMemsetAsync(d_M, 0, sizeof(uint32_t) * tiles * num_chunks, stream));
  if (chunker_tile_hist(d_ids, nullptr, n, d_lut, num_chunks, d_M, 0, tiles, d_err, stream)) return 1;
  if (exclusive_scan_u32(d_M, tiles * num_chunks, d_scan_ws, nullptr, stream)) return 1;
  return chunker_place(d_las, d_ids, nullptr, n, num_chunks, d_M, 0, tiles, d_out, stream);
}

int chunker_distribute_tail(const uint8_t* d_las, const uint32_t* d_cells, uint16_t* d_ids, int64_t n,
                            const int32_t* d_lut, int32_t num_chunks, uint32_t* d_M, void* d_scan_ws,
                            uint8_t* d_out, int32_t* d_err, hipStream_t stream) {
  const int64_t tiles = (n + kDistTile - 1) / kDistTile;
  if (tiles == 0) return 0;
  if (!distribute_tail_fits(num_chunks)) {
    set_error("record layout or chunk count outside the tile transcode");
    return 1;
  }
  if (num_chunks > 8192) PCG_CHECK(hipMemsetAsync(d_M, 0, sizeof(uint32_t) * tiles * num_chunks, stream));
  if (chunker_tile_hist(const_cast<uint32_t*>(d_cells), d_ids, n, d_lut, num_chunks, d_M, 0, tiles, d_err, stream))
    return 1;
  if (exclusive_scan_u32(d_M, tiles * num_chunks, d_scan_ws, nullptr, stream)) return 1;
  return chunker_place(d_las, nullptr, d_ids, n, num_chunks, d_M, 0, tiles, d_out, stream);
}

int lasgen_device(const LasGenParams& p, int64_t first, int64_t count, uint8_t* d_out, int32_t* d_qminmax,
                  hipStream_t stream) {
  LasGenBuilding bld[64];
  int nb = p.num_buildings < 64 ? p.num_buildings : 64;
  for (int b = 0; b < nb; b++) bld[b] = lasgen_building(p, b);
  LasGenBuilding* d_bld = nullptr;
  PCG_CHECK(hipMalloc(&d_bld, sizeof(bld)));
  PCG_CHECK(hipMemcpyAsync(d_bld, bld, sizeof(bld), hipMemcpyHostToDevice, stream));
  int64_t blocks = (count + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_lasgen, dim3((unsigned)blocks), dim3(256), 0, stream, p, d_bld, first, count, d_out, d_qminmax,
                     d_qminmax + 3);
  PCG_CHECK(hipGetLastError());
  PCG_CHECK(hipStreamSynchronize(stream));
  PCG_CHECK(hipFree(d_bld));
  return 0;
}

}  // namespace pcg
