// Poisson, Poisson-average and random node sampling on gfx950: all nodes of one
// octree level per launch (a workgroup or a wavefront per node, or a
// level-wide grid over the candidates / output bytes).
//
// Restates Converter/include/sampler_poisson.h:14-304 and
// sampler_poisson_average.h:14-492 for one inner node:
//   k_prepare      gather the children's points in child-then-index order
//                  (sampler_poisson.h:82-108), squared distance to the node
//                  centre, stable radix sort by float(distance) with an
//                  exact (double, index) fix-up, detection of equal
//                  distances that can change a decision
//   k_exact_sort   libstdc++ std::sort order for those nodes (:182-203)
//   k_cells        cell table: candidates sorted by 2*spacing cell
//   k_neighbors    per candidate, the earlier candidates within spacing
//                  (FP32 filter, FP64 re-check near the threshold)
//   k_accept       greedy acceptance (:133-180 / :184-281) in windows of
//                  64 sorted candidates, one wavefront per node
//   k_avg_colors   Poisson-average colour sums (:357-416), int64 atomics
//   k_sequential   exact scans for nodes where the windowed form could
//                  differ (10 001-check cap, accepted points in each
//                  other's colour range)
//   k_route        accepted records to the node, the rest back to the
//                  children in their order (:247-291 / :418-482): ranks,
//                  selection lists and output ranges
//   k_copy_tiles   the record bytes: 64-record tiles of the children, each
//                  written as one accepted run and one remainder run
// SamplerRandom (sampler_random.h:14-234) shares k_route / k_copy_tiles:
//   k_shuffle_leaves  std::shuffle of every leaf (seeded variant)
//   k_random_prepare  first taker of each 128^3 grid cell
#include <algorithm>

#include "lstd_shuffle.h"
#include "lstd_sort.h"
#include "pcg_common.h"
#include "sampler.h"

namespace pcg {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kNil = 0xFFFFFFFFu;
constexpr int kMaxNbr = 16;  // earlier neighbours kept per candidate
constexpr uint8_t kNbrOverflow = 255;  // list too long for the pool: search again
constexpr uint8_t kNbrPooled = 254;    // list in the level's neighbour pool
constexpr int kAccWords = 2048;         // LDS accepted bits per wave in k_accept (65536 candidates)
constexpr int kAccDeepNodes = 2048;     // k_accept launches up to this many nodes prefetch three windows

__device__ inline int32_t ld_i32(const uint8_t* p) {
  return (int32_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
}
__device__ inline uint16_t ld_u16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

// X, Y, Z (3 x int32 at any alignment) from four aligned dword loads; the
// record buffers keep at least 3 bytes of their allocation after a record
__device__ inline void ld_xyz(const uint8_t* p, int32_t& X, int32_t& Y, int32_t& Z) {
  const uint32_t s = (uint32_t)((uintptr_t)p & 3);
  const uint32_t* w = (const uint32_t*)(p - s);
  const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
  X = (int32_t)__builtin_amdgcn_alignbyte(w1, w0, s);
  Y = (int32_t)__builtin_amdgcn_alignbyte(w2, w1, s);
  Z = (int32_t)__builtin_amdgcn_alignbyte(w3, w2, s);
}

__host__ __device__ inline uint64_t align16(uint64_t v) { return (v + 15) & ~15ull; }

// Pointers into the node scratch come from memory, so the compiler cannot
// tell their address space and emits flat instructions; a cast to a
// global-address pointer makes the loads and stores through it global ones.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ inline gptr<T> gbl(T* p) { return (gptr<T>)p; }
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
// 16 bytes through a global pointer (the HIP vector types cannot be assigned
// from a global-address reference)
__device__ inline uint4 ld_g16(gptr<const uint4> p) {
  const u32x4_t v = *(gptr<const u32x4_t>)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));
__device__ inline int4 ld_g16i(gptr<const int4> p) {
  const i32x4_t v = *(gptr<const i32x4_t>)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ inline void st_g16i(gptr<int4> p, int4 v) {
  i32x4_t t;
  t.x = v.x; t.y = v.y; t.z = v.z; t.w = v.w;
  *(gptr<i32x4_t>)p = t;
}

__host__ __device__ inline uint32_t hash_cap(int64_t n, int method) {
  uint32_t h = 64;
  while ((int64_t)h < 2 * n) h <<= 1;
  if (method == METHOD_POISSON_AVERAGE && h < 4096) h = 4096;
  return h;
}

struct Layout {
  int4* xyz;         // main order: X, Y, Z and 0, one 16-byte load per candidate
  double* add;       // main order
  uint32_t *keyA, *idxA, *keyB, *idxB;
  double* sadd;      // sorted order
  uint32_t* sidx;    // sorted -> main (points at idxA)
  uint8_t* acc;      // sorted order
  uint32_t* cell16;  // main order, packed 16^3 cell (average)
  uint16_t* rgb;     // main order, own colour (average)
  uint32_t H;
  uint4* htab;                     // cell table by hash slot: (2*spacing cell key, start, end, 0) in cell order
  uint32_t *hkey, *hhead, *next;   // lists of the sequential forms
  int4* ent;                       // cell order: (X, Y, Z, sorted position), positions ascending inside a cell
  double* eAdd;                    // cell order squared distance to the centre (Poisson)
  uint32_t* eC16;                  // cell order 16^3 cell (average)
  long long* sums;   // sorted order, 8 per candidate (average)
  uint32_t* nbr;     // sorted order, kMaxNbr earlier neighbours per candidate
  uint8_t* nbrCount; // kNbrOverflow: more than kMaxNbr
  uint32_t* tileAcc; // per 64-record tile of each child: accepted candidates of the node before it
  uint8_t* accMain;  // main order: accepted (written by k_route for the record copy)
};

// Carves the per-node scratch into the arrays above; with base == nullptr it
// only measures.  Returns the bytes used.
__host__ __device__ inline uint64_t layout_carve(uint8_t* base, int64_t n, int method, Layout* L) {
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) -> uint8_t* {
    uint8_t* r = base ? base + off : nullptr;
    off += align16(bytes);
    return r;
  };
  if (method == METHOD_RANDOM) {
    // candidates stay in main order; the cell table maps a grid cell to the
    // first candidate that may take it (hkey: cell, hhead: candidate)
    *L = Layout();
    L->idxA = (uint32_t*)take(4 * n);
    L->keyB = (uint32_t*)take(4 * n);
    L->acc = (uint8_t*)take(n);
    L->H = hash_cap(n, method);
    L->hkey = (uint32_t*)take(4 * (uint64_t)L->H);
    L->hhead = (uint32_t*)take(4 * (uint64_t)L->H);
    L->tileAcc = (uint32_t*)take(4 * (n / 64 + 16));
    L->accMain = (uint8_t*)take(n);
    L->sidx = L->idxA;
    return off + 64;
  }
  const bool avg = method == METHOD_POISSON_AVERAGE;
  L->xyz = (int4*)take(16 * n);
  L->add = (double*)take(8 * n);
  L->keyA = (uint32_t*)take(4 * n);
  L->idxA = (uint32_t*)take(4 * n);
  L->keyB = (uint32_t*)take(4 * n);
  L->idxB = (uint32_t*)take(4 * n);
  L->sadd = (double*)take(8 * n);
  L->acc = (uint8_t*)take(n);
  L->cell16 = (uint32_t*)take(4 * n);
  L->H = hash_cap(n, method);
  L->htab = (uint4*)take(16 * (uint64_t)L->H);
  L->hkey = (uint32_t*)take(4 * (uint64_t)L->H);
  L->hhead = (uint32_t*)take(4 * (uint64_t)L->H);
  L->next = (uint32_t*)take(4 * n);
  L->ent = (int4*)take(16 * n);
  L->eAdd = avg ? nullptr : (double*)take(8 * n);
  L->eC16 = avg ? (uint32_t*)take(4 * n) : nullptr;
  L->sums = avg ? (long long*)take(64 * n) : nullptr;
  L->rgb = avg ? (uint16_t*)take(6 * n) : nullptr;
  L->nbr = (uint32_t*)take(4 * kMaxNbr * n);
  L->nbrCount = (uint8_t*)take(n);
  L->tileAcc = (uint32_t*)take(4 * (n / 64 + 16));
  L->accMain = (uint8_t*)take(n);
  L->sidx = L->idxA;
  return off + 64;
}

__host__ __device__ inline uint64_t layout_bytes(int64_t n, int method) {
  Layout L;
  return layout_carve(nullptr, n, method, &L);
}

__device__ inline Layout make_layout(uint8_t* base, int64_t n, int method) {
  Layout L;
  layout_carve(base, n, method, &L);
  return L;
}

__device__ inline double coord(int32_t v, double scale, double offset) { return ((double)v * scale) + offset; }

struct NodeGeom {
  double cx, cy, cz;        // (min + max) * 0.5
  double spacing, s2;
  double min[3], size[3];
};

__device__ inline NodeGeom node_geom(const NodeInfo& info, const SampleParams& P) {
  NodeGeom g;
  g.cx = (info.min[0] + info.max[0]) * 0.5;
  g.cy = (info.min[1] + info.max[1]) * 0.5;
  g.cz = (info.min[2] + info.max[2]) * 0.5;
  // sampler_poisson.h:114 spacing = baseSpacing / pow(2, level) (exact: power of two)
  double p2 = 1.0;
  for (int i = 0; i < info.level; i++) p2 *= 2.0;
  g.spacing = P.base_spacing / p2;
  g.s2 = g.spacing * g.spacing;
  for (int a = 0; a < 3; a++) { g.min[a] = info.min[a]; g.size[a] = info.max[a] - info.min[a]; }
  return g;
}

// the 16^3 accept-grid cell of a position and the range searched around it
// (sampler_poisson_average.h:186-208)
__device__ inline int cell16_axis(double v, double mn, double size) {
  double d = 16.0 * (v - mn) / size;
  d = d < 15.0 ? d : 15.0;
  d = d > 0.0 ? d : 0.0;
  return (int)d;
}

__device__ inline uint32_t pack16(int x, int y, int z) { return (uint32_t)(x | (y << 4) | (z << 8)); }

__device__ inline void range16(const NodeGeom& g, double x, double y, double z, int lo[3], int hi[3]) {
  double v[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    lo[a] = cell16_axis(v[a] - g.spacing, g.min[a], g.size[a]);
    hi[a] = cell16_axis(v[a] + g.spacing, g.min[a], g.size[a]);
  }
}

__device__ inline bool in_range16(uint32_t c, const int lo[3], const int hi[3]) {
  int x = c & 15, y = (c >> 4) & 15, z = (c >> 8) & 15;
  return x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2];
}

// hash-grid cell (2 * spacing) of a coordinate, shifted by one so points on
// or just outside the box boundary stay in range
__device__ inline int gcell(double v, double mn, double cs) {
  double d = (v - mn) / cs;
  int64_t i = (int64_t)__builtin_floor(d) + 1;
  return (int)(i < 0 ? 0 : (i > 127 ? 127 : i));
}
__device__ inline uint32_t gkey(int x, int y, int z) { return (uint32_t)(x | (y << 7) | (z << 14)); }
__device__ inline uint32_t ghash(uint32_t k, uint32_t H) { return (k * 2654435761u) & (H - 1); }

__device__ inline uint64_t wave_match8(uint32_t key, uint64_t active) {
  uint64_t peers = active;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    uint64_t m = __ballot((key >> b) & 1);
    peers &= ((key >> b) & 1) ? m : ~m;
  }
  return peers;
}

// ------------------------------------------------------ block radix pass
constexpr int kRadixE = 4;  // keys per thread per tile

struct RadixShared {
  uint32_t hist[256];
  uint32_t base[256];
  uint32_t wc[kRadixE][kWaves][256];  // per sub-tile and wave digit counts (zero between tiles)
  uint32_t wp[kRadixE][kWaves][256];  // their exclusive prefix per digit, from the digit's base
  int skip;
};

// One stable counting pass of a block-wide LSD radix sort over the digit
// (key >> shift) & mask (mask < 256) of n (key, value) pairs.  Tiles of
// kRadixE * kBlock keys; sub-tile e of a tile precedes sub-tile e + 1.
// `pre` (LDS, 256 counts) is the digit's histogram when the caller counted
// it already; otherwise the pass counts.  Returns false and writes nothing
// when every key has the same digit.
template <bool kPrefetch = false>
__device__ __attribute__((always_inline)) inline bool block_radix_pass(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n, int shift,
                                 uint32_t mask, RadixShared& sh, const uint32_t* pre = nullptr) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  sh.hist[tid] = pre ? pre[tid] : 0u;
  for (int e = 0; e < kRadixE; e++)
    for (int w = 0; w < kWaves; w++) sh.wc[e][w][tid] = 0;
  __syncthreads();
  if (!pre) {
    for (int64_t j0 = 0; j0 < n; j0 += 4 * kBlock) {
      uint32_t kk[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = j0 + u * kBlock + tid;
        kk[u] = j < n ? kin[j] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (j0 + u * kBlock + tid < n) atomicAdd(&sh.hist[(kk[u] >> shift) & mask], 1u);
    }
    __syncthreads();
  }
  if (wave == 0) {
    uint32_t h0 = sh.hist[4 * lane], h1 = sh.hist[4 * lane + 1], h2 = sh.hist[4 * lane + 2],
             h3 = sh.hist[4 * lane + 3];
    uint32_t s = h0 + h1 + h2 + h3, incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    uint32_t ex = incl - s;
    sh.base[4 * lane] = ex;
    sh.base[4 * lane + 1] = ex + h0;
    sh.base[4 * lane + 2] = ex + h0 + h1;
    sh.base[4 * lane + 3] = ex + h0 + h1 + h2;
    const uint32_t un = (uint32_t)n;
    const bool full = h0 == un || h1 == un || h2 == un || h3 == un;
    const uint64_t any = __ballot(full);
    if (lane == 0) sh.skip = any != 0;
  }
  __syncthreads();
  if (sh.skip) return false;
  // a tile's keys and values (positions past the node read the last one);
  // kPrefetch: the next tile's are loaded while this one is ranked and
  // scattered (more registers), otherwise after the scatter
  uint32_t key[kRadixE], val[kRadixE];
  auto load_tile = [&](int64_t b0, uint32_t* kk, uint32_t* vv) {
#pragma unroll
    for (int e = 0; e < kRadixE; e++) {
      const int64_t j = b0 + e * kBlock + tid;
      const int64_t jc = j < n ? j : n - 1;
      kk[e] = kin[jc];
      vv[e] = vin[jc];
    }
  };
  if (n > 0) load_tile(0, key, val);
  for (int64_t b0 = 0; b0 < n; b0 += (int64_t)kRadixE * kBlock) {
    uint32_t d[kRadixE];
    uint64_t peers[kRadixE];
    bool valid[kRadixE];
#pragma unroll
    for (int e = 0; e < kRadixE; e++) valid[e] = b0 + e * kBlock + tid < n;
#pragma unroll
    for (int e = 0; e < kRadixE; e++) {
      d[e] = (key[e] >> shift) & mask;
      peers[e] = wave_match8(d[e], __ballot(valid[e]));
      if (valid[e] && (peers[e] & lt) == 0) sh.wc[e][wave][d[e]] = (uint32_t)__popcll(peers[e]);
    }
    uint32_t nkey[kRadixE], nval[kRadixE];
    if constexpr (kPrefetch) {
      const int64_t nb = b0 + (int64_t)kRadixE * kBlock;
      load_tile(nb < n ? nb : b0, nkey, nval);
    }
    __syncthreads();
    // per digit (one per thread): exclusive prefix over (sub-tile, wave)
    // from the digit's base; counts back to zero for the next tile
    {
      uint32_t run = sh.base[tid];
#pragma unroll
      for (int e = 0; e < kRadixE; e++)
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
          const uint32_t c = sh.wc[e][w][tid];
          sh.wc[e][w][tid] = 0;
          sh.wp[e][w][tid] = run;
          run += c;
        }
      sh.base[tid] = run;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kRadixE; e++) {
      if (!valid[e]) continue;
      const uint32_t pos = sh.wp[e][wave][d[e]] + (uint32_t)__popcll(peers[e] & lt);
      kout[pos] = key[e];
      vout[pos] = val[e];
    }
    if constexpr (kPrefetch) {
#pragma unroll
      for (int e = 0; e < kRadixE; e++) {
        key[e] = nkey[e];
        val[e] = nval[e];
      }
    } else {
      const int64_t nb = b0 + (int64_t)kRadixE * kBlock;
      if (nb < n) load_tile(nb, key, val);
    }
  }
  __syncthreads();
  return true;
}

// End of the run of keys equal to a[p] in the ascending array a[0, n):
// galloping from p, then a binary search, so a run of R members costs
// O(log R) dependent loads instead of R.
template <class T>
__device__ inline int64_t run_end(const T* a, int64_t p, int64_t n) {
  const T v = a[p];
  int64_t lo = p + 1;  // a[lo - 1] == v
  int64_t hi = n;
  for (int64_t step = 1;; step <<= 1) {
    const int64_t probe = lo + step - 1;
    if (probe >= n) break;
    if (!(a[probe] == v)) { hi = probe; break; }
    lo = probe + 1;
  }
  while (lo < hi) {  // a[lo - 1] == v, and hi == n or a[hi] != v
    const int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] == v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Runs of equal sort keys up to this length are ordered by one lane
// (insertion, at most kShortRun^2 / 2 steps); longer ones by block-wide
// passes, so that no node costs quadratic time in one lane.
constexpr int kShortRun = 32;

// Can two members of the tie group [p, e) (sorted positions of equal
// distance) exclude each other, so that their order changes a decision or
// a colour sum?  Groups above 64 members are taken as conflicting.
__device__ bool tie_group_conflicts(const Layout& L, const NodeGeom& g, const SampleParams& P, int64_t p, int64_t e) {
  if (e - p > 64) return true;
  for (int64_t a = p; a < e; a++) {
    for (int64_t b = a + 1; b < e; b++) {
      uint32_t ia = L.idxA[a], ib = L.idxA[b];
      const int4 pa = L.xyz[ia], pb = L.xyz[ib];
      double xa = coord(pa.x, P.scale[0], P.offset[0]), ya = coord(pa.y, P.scale[1], P.offset[1]),
             za = coord(pa.z, P.scale[2], P.offset[2]);
      double xb = coord(pb.x, P.scale[0], P.offset[0]), yb = coord(pb.y, P.scale[1], P.offset[1]),
             zb = coord(pb.z, P.scale[2], P.offset[2]);
      double dx = xa - xb, dy = ya - yb, dz = za - zb;
      double dd = dx * dx + dy * dy + dz * dz;
      if (!(dd < g.s2)) continue;
      if (P.method == METHOD_POISSON) {
        double cd = __builtin_sqrt(L.sadd[p]);
        double lim = cd - g.spacing;
        if (L.sadd[p] >= lim * lim) return true;
      } else {
        int lo[3], hi[3];
        range16(g, xa, ya, za, lo, hi);
        if (in_range16(L.cell16[ib], lo, hi)) return true;
        range16(g, xb, yb, zb, lo, hi);
        if (in_range16(L.cell16[ia], lo, hi)) return true;
      }
    }
  }
  return false;
}

// ---------------------------------------------------------------- prepare
// Bucket sort of a node's candidates: the squared distance to the centre is
// at most a quarter of the box diagonal squared, so floor(add * B / that)
// (clamped) is a bucket index that never decreases with the distance, and
// for points on a surface the buckets fill about evenly.  One counting pass
// during the gather, one scatter of (distance, index) and an insertion sort
// of each bucket by (distance, index) give the same order as the radix sort
// below, with the payload moved once.  Nodes with a bucket of more than
// kBucketMaxRun candidates (dense clusters) take the radix sort.
constexpr int kBuckets = 8192;      // most buckets of a node (LDS counters)
constexpr int kBucketMaxRun = 96;   // largest bucket sorted by one lane
constexpr int64_t kBucketMaxN = 65536;

// (distance, index) pairs [st, st + k) of sadd / keyB, k <= 8, sorted by an
// odd-even transposition network; indices to idxA
__device__ inline void bucket_sort8(const Layout& L, uint32_t st, uint32_t k) {
  if (k == 0) return;
  if (k == 1) {
    L.idxA[st] = L.keyB[st];
    return;
  }
  double d[8];
  uint32_t v[8];
#pragma unroll
  for (uint32_t t = 0; t < 8; t++) {
    d[t] = t < k ? L.sadd[st + t] : __builtin_inf();
    v[t] = t < k ? L.keyB[st + t] : 0xFFFFFFFFu;
  }
#pragma unroll
  for (int r = 0; r < 8; r++) {
#pragma unroll
    for (int t = r & 1; t + 1 < 8; t += 2) {
      const bool sw = d[t] > d[t + 1] || (d[t] == d[t + 1] && v[t] > v[t + 1]);
      const double dl = sw ? d[t + 1] : d[t], dh = sw ? d[t] : d[t + 1];
      const uint32_t vl = sw ? v[t + 1] : v[t], vh = sw ? v[t] : v[t + 1];
      d[t] = dl; d[t + 1] = dh;
      v[t] = vl; v[t + 1] = vh;
    }
  }
#pragma unroll
  for (uint32_t t = 0; t < 8; t++)
    if (t < k) {
      L.sadd[st + t] = d[t];
      L.idxA[st + t] = v[t];
    }
}

__device__ inline uint32_t dist_bucket(double add, double scale, uint32_t B) {
  const double t = add * scale;
  return t < (double)(B - 1) ? (uint32_t)t : B - 1;
}

__global__ __launch_bounds__(kBlock) void k_prepare(const int32_t* __restrict__ nodes, const NodeInfo* __restrict__ infos,
                                                     const NodeState* __restrict__ states, NodeWork* __restrict__ work,
                                                     SampleParams P) {
  __shared__ int64_t s_cs[9];
  __shared__ const uint8_t* s_cptr[8];
  __shared__ uint8_t* s_scr;
  __shared__ union {
    RadixShared rad;
    uint32_t bucket[kBuckets];
  } s_u;
  RadixShared& s_rad = s_u.rad;
  __shared__ uint32_t s_h4[4][256];  // histograms of the four key bytes, counted while gathering
  __shared__ uint32_t s_maxb, s_wsum[kWaves];
  const int k = blockIdx.x;
  const NodeInfo info = infos[nodes[k]];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < 4 * 256; i += kBlock) (&s_h4[0][0])[i] = 0;
  if (tid == 0) {
    s_cs[0] = 0;
    for (int c = 0; c < 8; c++) {
      int ch = info.children[c];
      int64_t cnt = ch >= 0 ? states[ch].count : 0;
      s_cptr[c] = ch >= 0 ? states[ch].ptr : nullptr;
      s_cs[c + 1] = s_cs[c] + cnt;
    }
    int64_t n = s_cs[8];
    uint64_t bytes = layout_bytes(n, P.method);
    uint64_t off = atomicAdd((unsigned long long*)P.scratch_top, (unsigned long long)bytes);
    NodeWork w;
    w.n = n;
    for (int c = 0; c < 9; c++) w.child_start[c] = s_cs[c];
    for (int c = 0; c < 8; c++) w.csrc[c] = s_cptr[c];
    w.flags = (P.method == METHOD_POISSON_AVERAGE || (int)gridDim.x < P.ring_min_nodes || n > P.ring_max_cand)
                  ? NW_CELL_PATH
                  : 0;
    w.pad = 0;
    if (off + bytes > P.scratch_cap) {
      w.flags = NW_OVERFLOW;
      w.scratch = nullptr;
      s_scr = nullptr;
      atomicAdd(&P.counters[6], 1ull);
    } else {
      w.scratch = P.scratch + off;
      s_scr = w.scratch;
    }
    work[k] = w;
  }
  __syncthreads();
  if (!s_scr) return;
  const int64_t n = s_cs[8];
  Layout L = make_layout(s_scr, n, P.method);
  NodeGeom g = node_geom(info, P);
  // bucket sort for nodes up to kBucketMaxN candidates (about two per
  // bucket up to 16 384 candidates), the radix sort otherwise
  const double maxAdd = 0.25 * (g.size[0] * g.size[0] + g.size[1] * g.size[1] + g.size[2] * g.size[2]);
  uint32_t B = 64;
  while ((int64_t)B * 2 < n && B < (uint32_t)kBuckets) B <<= 1;
  // (launches of few nodes, where one node's latency decides, keep the
  // radix sort: its passes are shorter for a single large node)
  bool bucketPath = n >= 2 && n <= kBucketMaxN && maxAdd > 0.0 && (int)gridDim.x >= P.bucket_min_nodes;
  const double bscale = bucketPath ? (double)B / maxAdd : 0.0;
  if (bucketPath) {
    for (uint32_t i = tid; i < B; i += kBlock) s_u.bucket[i] = 0;
    if (tid == 0) s_maxb = 0;
    __syncthreads();
  }
  // gather (sampler_poisson.h:82-108), four records per thread in flight;
  // the bucket counts or the byte histograms of the sort keys are counted
  // on the way
  for (int64_t j0 = 0; j0 < n; j0 += 4 * kBlock) {
    int32_t X4[4], Y4[4], Z4[4];
    const uint8_t* r4[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = j0 + u * kBlock + tid;
      r4[u] = nullptr;
      X4[u] = Y4[u] = Z4[u] = 0;
      if (j < n) {
        int c = 0;
        while (c < 7 && j >= s_cs[c + 1]) c++;
        r4[u] = s_cptr[c] + (j - s_cs[c]) * P.bpp;
        ld_xyz(r4[u], X4[u], Y4[u], Z4[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = j0 + u * kBlock + tid;
      if (j >= n) continue;
      const uint8_t* r = r4[u];
      const int32_t X = X4[u], Y = Y4[u], Z = Z4[u];
      double x = coord(X, P.scale[0], P.offset[0]), y = coord(Y, P.scale[1], P.offset[1]),
             z = coord(Z, P.scale[2], P.offset[2]);
      double ax = x - g.cx, ay = y - g.cy, az = z - g.cz;
      double add = ax * ax + ay * ay + az * az;
      L.xyz[j] = make_int4(X, Y, Z, 0);
      L.add[j] = add;
      if (bucketPath) {
        atomicAdd(&s_u.bucket[dist_bucket(add, bscale, B)], 1u);
      } else {
        float f = (float)add;
        uint32_t fb;
        __builtin_memcpy(&fb, &f, 4);
        L.keyA[j] = fb;
        L.idxA[j] = (uint32_t)j;
        atomicAdd(&s_h4[0][fb & 255u], 1u);
        atomicAdd(&s_h4[1][(fb >> 8) & 255u], 1u);
        atomicAdd(&s_h4[2][(fb >> 16) & 255u], 1u);
        atomicAdd(&s_h4[3][fb >> 24], 1u);
      }
      if (P.method == METHOD_POISSON_AVERAGE) {
        L.cell16[j] = pack16(cell16_axis(x, g.min[0], g.size[0]), cell16_axis(y, g.min[1], g.size[1]),
                             cell16_axis(z, g.min[2], g.size[2]));
        L.rgb[3 * j + 0] = ld_u16(r + P.rgb_offset);
        L.rgb[3 * j + 1] = ld_u16(r + P.rgb_offset + 2);
        L.rgb[3 * j + 2] = ld_u16(r + P.rgb_offset + 4);
        for (int q = 0; q < 8; q++) L.sums[8 * j + q] = 0;
      }
    }
  }
  __syncthreads();
  if (bucketPath) {
    // the largest bucket decides between the bucket sort and the radix sort
    uint32_t m = 0;
    for (uint32_t i = tid; i < B; i += kBlock) m = s_u.bucket[i] > m ? s_u.bucket[i] : m;
    if (m > (uint32_t)kBucketMaxRun) atomicMax(&s_maxb, m);
    __syncthreads();
    if (s_maxb) {
      // radix sort after all: its keys and indices, counted by its passes
      bucketPath = false;
      for (int64_t j = tid; j < n; j += kBlock) {
        float f = (float)L.add[j];
        uint32_t fb;
        __builtin_memcpy(&fb, &f, 4);
        L.keyA[j] = fb;
        L.idxA[j] = (uint32_t)j;
      }
      __syncthreads();
    }
  }
  if (bucketPath) {
    // exclusive prefix of the bucket counts: B / kBlock consecutive buckets
    // per thread (B is a power of two of at least 64)
    const uint32_t per = B >= (uint32_t)kBlock ? B / kBlock : 1u;
    const uint32_t b0 = (uint32_t)tid * per;
    uint32_t mine = 0;
    if (b0 < B)
      for (uint32_t i = 0; i < per; i++) mine += s_u.bucket[b0 + i];
    uint32_t incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d);
      if (lane >= d) incl += o;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t run = incl - mine;
    for (int w = 0; w < wave; w++) run += s_wsum[w];
    if (b0 < B)
      for (uint32_t i = 0; i < per; i++) {
        const uint32_t c = s_u.bucket[b0 + i];
        s_u.bucket[b0 + i] = run;
        run += c;
      }
    __syncthreads();
    // scatter (distance, index) to the buckets; each counter ends at its
    // bucket's end
    for (int64_t j0 = 0; j0 < n; j0 += 4 * kBlock) {
      double a4[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = j0 + u * kBlock + tid;
        a4[u] = j < n ? L.add[j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = j0 + u * kBlock + tid;
        if (j >= n) continue;
        const uint32_t pos = atomicAdd(&s_u.bucket[dist_bucket(a4[u], bscale, B)], 1u);
        L.sadd[pos] = a4[u];
        L.keyB[pos] = (uint32_t)j;
      }
    }
    __syncthreads();
    // each bucket in (distance, index) order by one lane: up to eight
    // entries loaded together and sorted in registers, longer buckets by
    // insertion in place
    for (uint32_t b = tid; b < B; b += kBlock) {
      const uint32_t e = s_u.bucket[b], st = b ? s_u.bucket[b - 1] : 0u;
      if (e - st <= 8) {
        bucket_sort8(L, st, e - st);
        continue;
      }
      for (uint32_t i = st + 1; i < e; i++) {
        const double dv = L.sadd[i];
        const uint32_t v = L.keyB[i];
        uint32_t q = i;
        while (q > st) {
          const double dq = L.sadd[q - 1];
          const uint32_t vq = L.keyB[q - 1];
          if (!(dq > dv || (dq == dv && vq > v))) break;
          L.sadd[q] = dq;
          L.keyB[q] = vq;
          q--;
        }
        L.sadd[q] = dv;
        L.keyB[q] = v;
      }
      for (uint32_t i = st; i < e; i++) L.idxA[i] = L.keyB[i];
    }
    __syncthreads();
  } else {
  // stable LSD radix sort on float(add) bits
  uint32_t *kc = L.keyA, *ic = L.idxA, *ko = L.keyB, *io = L.idxB;
  // the float key, 4 passes; runs of equal keys (distinct doubles that round
  // to the same float, or equal distances) are put in (double add, main
  // index) order below
  const bool precounted = n > kBucketMaxN || maxAdd <= 0.0 || n < 2 || (int)gridDim.x < P.bucket_min_nodes;
  for (int pass = 0; pass < 4; pass++) {
    if (!block_radix_pass(kc, ic, ko, io, n, 8 * pass, 255u, s_rad, precounted ? s_h4[pass] : nullptr)) continue;
    uint32_t* t = kc; kc = ko; ko = t;
    t = ic; ic = io; io = t;
  }
  // runs of equal keys: order by (double add, main index); run starts are
  // found four positions per thread at a time so that the key loads are in
  // flight together.  Runs longer than kShortRun (near-duplicate clusters)
  // are left to the full-width passes below.
  __shared__ int s_long;
  if (tid == 0) s_long = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    uint32_t kpv[4], kcur[4], knx[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      kcur[u] = p < n ? kc[p] : 0u;
      kpv[u] = (p > 0 && p < n) ? kc[p - 1] : ~0u;
      knx[u] = p + 1 < n ? kc[p + 1] : ~0u;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      if (p >= n || kpv[u] == kcur[u] || knx[u] != kcur[u]) continue;
      const int64_t e = run_end(kc, p, n);
      if (e - p > kShortRun) {
        s_long = 1;
        continue;
      }
      for (int64_t i = p + 1; i < e; i++) {
        uint32_t v = ic[i];
        double dv = L.add[v];
        int64_t q = i - 1;
        while (q >= p && (L.add[ic[q]] > dv || (L.add[ic[q]] == dv && ic[q] > v))) {
          ic[q + 1] = ic[q];
          q--;
        }
        ic[q + 1] = v;
      }
    }
  }
  __syncthreads();
  if (s_long) {
    // a long run of equal float keys: order the whole node by the 64 bits of
    // the double distance (non-negative, so its bit pattern ascends with
    // it), low word then high word, eight stable passes from the current
    // order, in which equal doubles are in main-index order
    for (int half = 0; half < 2; half++) {
      for (int64_t p = tid; p < n; p += kBlock) {
        uint64_t b;
        __builtin_memcpy(&b, &L.add[ic[p]], 8);
        kc[p] = half == 0 ? (uint32_t)b : (uint32_t)(b >> 32);
      }
      __syncthreads();
      for (int pass = 0; pass < 4; pass++) {
        if (!block_radix_pass(kc, ic, ko, io, n, 8 * pass, 255u, s_rad)) continue;
        uint32_t* t = kc; kc = ko; ko = t;
        t = ic; ic = io; io = t;
      }
    }
  }
  // sorted positions and distances
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    uint32_t v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      v[u] = p < n ? ic[p] : 0u;
    }
    double d[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      d[u] = p < n ? L.add[v[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      if (p < n) {
        L.sadd[p] = d[u];
        if (ic != L.idxA) L.idxA[p] = v[u];
      }
    }
  }
  __syncthreads();
  }
  // equal distances: the reference's unstable sort decides their order;
  // it matters only if two of them can exclude each other
  bool conflict = false;
  for (int64_t b0 = 0; b0 + 1 < n && !conflict; b0 += 4 * kBlock) {
    double dp[4], dc[4], dn[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      dc[u] = p < n ? L.sadd[p] : -1.0;
      dn[u] = p + 1 < n ? L.sadd[p + 1] : -2.0;
      dp[u] = (p > 0 && p < n) ? L.sadd[p - 1] : -3.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      if (conflict || p + 1 >= n || !(dc[u] == dn[u]) || dp[u] == dc[u]) continue;
      int64_t e = p + 1;
      while (e < n && L.sadd[e] == dc[u]) e++;
      conflict = tie_group_conflicts(L, g, P, p, e);
    }
  }
  if (conflict) atomicOr(&work[k].flags, (uint32_t)NW_EXACT_SORT);
}

// ------------------------------------------------------------- exact sort
// The reference orders candidates with libstdc++ std::sort, which places
// equal distances in an order fixed by its moves.  Introsort's final
// insertion sort is stable, so that order is the order of the equal keys
// after the partition phase.  The partition phase is replayed here: every
// median-of-three Hoare partition is evaluated by one wavefront from prefix
// counts (the k-th element >= pivot from the left swaps with the k-th
// element <= pivot from the right while the former lies left of the
// latter), segments are shared between the four waves through an LDS
// stack, and past the depth limit a segment is heap-sorted by one lane.

struct PairArr {
  double* K;
  uint32_t* H;
};

__device__ inline void pair_swap(PairArr a, int64_t i, int64_t j) {
  double k = a.K[i]; a.K[i] = a.K[j]; a.K[j] = k;
  uint32_t h = a.H[i]; a.H[i] = a.H[j]; a.H[j] = h;
}

__device__ void pair_push_heap(PairArr a, int64_t f, int64_t hole, int64_t top, double vk, uint32_t vh) {
  int64_t parent = (hole - 1) / 2;
  while (hole > top && a.K[f + parent] < vk) {
    a.K[f + hole] = a.K[f + parent];
    a.H[f + hole] = a.H[f + parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  a.K[f + hole] = vk;
  a.H[f + hole] = vh;
}

__device__ void pair_adjust_heap(PairArr a, int64_t f, int64_t hole, int64_t len, double vk, uint32_t vh) {
  const int64_t top = hole;
  int64_t second = hole;
  while (second < (len - 1) / 2) {
    second = 2 * (second + 1);
    if (a.K[f + second] < a.K[f + second - 1]) second--;
    a.K[f + hole] = a.K[f + second];
    a.H[f + hole] = a.H[f + second];
    hole = second;
  }
  if ((len & 1) == 0 && second == (len - 2) / 2) {
    second = 2 * (second + 1);
    a.K[f + hole] = a.K[f + second - 1];
    a.H[f + hole] = a.H[f + second - 1];
    hole = second - 1;
  }
  pair_push_heap(a, f, hole, top, vk, vh);
}

__device__ void pair_heap_sort(PairArr a, int64_t f, int64_t len) {
  if (len >= 2) {
    int64_t parent = (len - 2) / 2;
    while (true) {
      pair_adjust_heap(a, f, parent, len, a.K[f + parent], a.H[f + parent]);
      if (parent == 0) break;
      parent--;
    }
  }
  for (int64_t last = len; last > 1;) {
    --last;
    double vk = a.K[f + last];
    uint32_t vh = a.H[f + last];
    a.K[f + last] = a.K[f];
    a.H[f + last] = a.H[f];
    pair_adjust_heap(a, f, 0, last, vk, vh);
  }
}

// pair versions of the libstdc++ partition loop for one lane
__device__ void pair_partition_loop(PairArr a, int64_t first, int64_t last, int64_t depth) {
  struct Fr { int64_t f, l, d; };
  Fr st[64];
  int sp = 0;
  st[sp++] = Fr{first, last, depth};
  while (sp > 0) {
    Fr fr = st[--sp];
    int64_t f = fr.f, l = fr.l, d = fr.d;
    while (l - f > 16) {
      if (d == 0) {
        pair_heap_sort(a, f, l - f);
        break;
      }
      --d;
      int64_t mid = f + (l - f) / 2, x = f + 1, y = mid, z = l - 1;
      double kx = a.K[x], ky = a.K[y], kz = a.K[z];
      int64_t pick;
      if (kx < ky) pick = (ky < kz) ? y : ((kx < kz) ? z : x);
      else pick = (kx < kz) ? x : ((ky < kz) ? z : y);
      pair_swap(a, f, pick);
      const double pv = a.K[f];
      int64_t lo = f + 1, hi = l;
      while (true) {
        while (a.K[lo] < pv) ++lo;
        --hi;
        while (pv < a.K[hi]) --hi;
        if (!(lo < hi)) break;
        pair_swap(a, lo, hi);
        ++lo;
      }
      if (sp < 64) st[sp++] = Fr{lo, l, d};
      l = lo;
    }
  }
}

// The kc swaps of a partition (pair k: left position f + posGe[f + k] and
// right position f + posLe[f + leT - 1 - k]), k = k0, k0 + stride, ...:
// U pairs per thread with their loads in flight.  Every left position
// lies before every right one, so no element takes part in two swaps.
template <int U>
__device__ inline void swap_pairs(PairArr a, const uint32_t* posLe, const uint32_t* posGe, int64_t f, int64_t leT,
                                  int64_t kc, int64_t k0, int64_t stride) {
  const gptr<double> K = gbl(a.K);
  const gptr<uint32_t> H = gbl(a.H);
  for (int64_t kb = k0; kb < kc; kb += U * stride) {
    uint32_t I[U], J[U];  // positions (node-relative, below 2^31)
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t k = kb + u * stride;
      k = k < kc ? k : kc - 1;
      I[u] = (uint32_t)f + posGe[f + k];
      J[u] = (uint32_t)f + posLe[f + leT - 1 - k];
    }
    double KI[U], KJ[U];
    uint32_t HI[U], HJ[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      KI[u] = K[I[u]]; KJ[u] = K[J[u]];
      HI[u] = H[I[u]]; HJ[u] = H[J[u]];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (kb + u * stride < kc) {
        K[I[u]] = KJ[u]; K[J[u]] = KI[u];
        H[I[u]] = HJ[u]; H[J[u]] = HI[u];
      }
    }
  }
}

// Block-wide unguarded partition of [f+1, l) around the pivot at f: the
// k-th element >= pivot from the left swaps with the k-th element <= pivot
// from the right while the former lies left of the latter; the cut is the
// (K+1)-th such left element or the K-th right element, whichever is first.
// All control flow is uniform over the block.
struct PartShared {
  uint32_t wge[4][kWaves], wle[4][kWaves];  // per sub-chunk and wave
  unsigned long long total;
  long long fail;
  unsigned long long kc;
  unsigned long long markL, markR;  // marked elements left / right of the cut
};

__device__ int64_t block_partition(PairArr a, uint32_t* posLe, uint32_t* posGe, int64_t f, int64_t l, PartShared& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint64_t me = 1ull << lane;
  const double pv = a.K[f];
  if (tid == 0) { sh.total = 0; sh.fail = INT64_MAX; sh.kc = 0; sh.markL = 0; sh.markR = 0; }
  __syncthreads();
  unsigned long long myLe = 0;
#pragma unroll 4
  for (int64_t i = f + 1 + tid; i < l; i += kBlock) myLe += !(pv < a.K[i]);
  if (myLe) atomicAdd(&sh.total, myLe);
  __syncthreads();
  const int64_t leT = (int64_t)sh.total;
  // four consecutive sub-chunks of kBlock elements per step (their loads in
  // flight together, one barrier pair per step), each ranked after the
  // previous ones exactly as one sub-chunk per step would be
  constexpr int U = 4;
  int64_t cG = 0, cL = 0;
  for (int64_t b = f + 1; b < l; b += U * kBlock) {
    bool ge[U], le[U];
    uint64_t bg[U], bl[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = b + u * kBlock + tid;
      const bool in = i < l;
      const double v = in ? a.K[i] : pv;
      ge[u] = in && !(v < pv);
      le[u] = in && !(pv < v);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      bg[u] = __ballot(ge[u]);
      bl[u] = __ballot(le[u]);
      if (lane == 0) { sh.wge[u][wave] = (uint32_t)__popcll(bg[u]); sh.wle[u][wave] = (uint32_t)__popcll(bl[u]); }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = b + u * kBlock + tid;
      int64_t preG = cG, preL = cL, totG = 0, totL = 0;
      for (int w = 0; w < kWaves; w++) {
        if (w < wave) { preG += sh.wge[u][w]; preL += sh.wle[u][w]; }
        totG += sh.wge[u][w];
        totL += sh.wle[u][w];
      }
      if (le[u]) posLe[f + preL + __popcll(bl[u] & lt)] = (uint32_t)(i - f);
      bool ok = false, bad = false;
      if (ge[u]) {
        const int64_t k = preG + __popcll(bg[u] & lt) + 1;
        const int64_t leIncl = preL + __popcll(bl[u] & (lt | me));
        ok = (leT - leIncl) >= k;
        bad = !ok;
        if (ok) posGe[f + k - 1] = (uint32_t)(i - f);
      }
      // one LDS atomic per wave: the pairs it formed, and its first element
      // without a partner (elements rise with the lane)
      const uint64_t okm = __ballot(ok), badm = __ballot(bad);
      if (lane == 0) {
        if (okm) atomicAdd(&sh.kc, (unsigned long long)__popcll(okm));
        if (badm) atomicMin(&sh.fail, (long long)(b + u * kBlock + 64 * wave + __builtin_ctzll(badm)));
      }
      cG += totG;
      cL += totL;
    }
    __syncthreads();
  }
  const int64_t kc = (int64_t)sh.kc;
  swap_pairs<4>(a, posLe, posGe, f, leT, kc, tid, kBlock);
  __syncthreads();
  const int64_t fail = (int64_t)sh.fail;
  int64_t cut;
  if (kc == 0) {
    cut = fail;
  } else {
    int64_t rK = f + posLe[f + leT - kc];
    cut = fail < rK ? fail : rK;
  }
  __syncthreads();
  return cut;
}

// Wave-wide unguarded partition of [f+1, l) around the pivot at f: the
// rule of block_partition (the k-th element >= pivot from the left swaps with
// the k-th element <= pivot from the right while the former lies left of the
// latter; the cut is the first left element without a partner or the K-th
// right element, whichever is first) with one wavefront, 64 elements per
// step, and no workgroup barrier.  Segments of different wavefronts use
// disjoint ranges of posLe / posGe.
__device__ int64_t wave_partition(PairArr a, uint32_t* posLe, uint32_t* posGe, int64_t f, int64_t l) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint64_t me = 1ull << lane;
  const double pv = a.K[f];
  // four steps' loads in flight per iteration (the loop is bound by their
  // latency); out-of-range lanes read the pivot
  constexpr int U = 4;
  int64_t leT = 0;
  for (int64_t b = f + 1; b < l; b += 64 * U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = b + 64 * u + lane;
      v[u] = a.K[i < l ? i : f];
    }
#pragma unroll
    for (int u = 0; u < U; u++) leT += __popcll(__ballot(b + 64 * u + lane < l && !(pv < v[u])));
  }
  int64_t cG = 0, cL = 0, kc = 0, fail = INT64_MAX;
  for (int64_t b0 = f + 1; b0 < l; b0 += 64 * U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = b0 + 64 * u + lane;
      v[u] = a.K[i < l ? i : f];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t b = b0 + 64 * u;
      const int64_t i = b + lane;
      const bool in = i < l;
      const bool ge = in && !(v[u] < pv), le = in && !(pv < v[u]);
      const uint64_t bg = __ballot(ge), bl = __ballot(le);
      if (le) posLe[f + cL + __popcll(bl & lt)] = (uint32_t)(i - f);
      bool ok = false;
      if (ge) {
        const int64_t kk = cG + __popcll(bg & lt) + 1;
        const int64_t leIncl = cL + __popcll(bl & (lt | me));
        ok = (leT - leIncl) >= kk;
        if (ok) posGe[f + kk - 1] = (uint32_t)(i - f);
      }
      const uint64_t okm = __ballot(ok), badm = __ballot(ge && !ok);
      kc += __popcll(okm);
      // ok only falls with the rank, so the first element without a
      // partner is the lowest bad lane of the first step that has one
      if (badm && fail == INT64_MAX) fail = b + __builtin_ctzll(badm);
      cG += __popcll(bg);
      cL += __popcll(bl);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  swap_pairs<2>(a, posLe, posGe, f, leT, kc, lane, 64);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (kc == 0) return fail;
  const int64_t rK = f + posLe[f + leT - kc];
  return fail < rK ? fail : rK;
}

// median-of-three pivot of [f, l) moved to f (__move_median_to_first), by
// the calling lane
__device__ inline void pair_pivot_to_first(PairArr a, int64_t f, int64_t l) {
  const int64_t mid = f + (l - f) / 2, x = f + 1, y = mid, z = l - 1;
  const double kx = a.K[x], ky = a.K[y], kz = a.K[z];
  int64_t pick;
  if (kx < ky) pick = (ky < kz) ? y : ((kx < kz) ? z : x);
  else pick = (kx < kz) ? x : ((ky < kz) ? z : y);
  pair_swap(a, f, pick);
}

// A wavefront's segment staged in LDS: keys, main indices (bit 31: the
// element is marked), and the partition's position lists, 16 bytes per
// element, carved per wavefront from the launch's dynamic LDS.  The
// partition loop then runs on LDS latency instead of memory latency.
// Launches of few nodes (one workgroup per CU at most) stage segments of
// up to kStageNarrow elements, wider ones up to kStageWide (a small LDS
// footprint keeps several workgroups per CU).
constexpr int kStageNarrow = 2048;
constexpr int kStageWide = 128;
constexpr uint32_t kStageMark = 0x80000000u;
template <class T>
using lptr = __attribute__((address_space(3))) T*;
struct StageLds {
  lptr<double> K;
  lptr<uint32_t> H;
  lptr<uint16_t> le, ge;
};
__device__ inline StageLds stage_view(uint8_t* base, int cap) {
  return StageLds{(lptr<double>)base, (lptr<uint32_t>)(base + 8 * cap), (lptr<uint16_t>)(base + 12 * cap),
                  (lptr<uint16_t>)(base + 14 * cap)};
}

// wave_partition on a staged segment: [f+1, l) of K/H around the pivot at
// f, the same swaps (positions relative to f in le / ge)
__device__ __attribute__((always_inline)) inline int wave_partition_lds(const StageLds& S, int f, int l) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint64_t me = 1ull << lane;
  const double pv = S.K[f];
  int leT = 0;
  for (int b = f + 1; b < l; b += 64) {
    const int i = b + lane;
    const double v = S.K[i < l ? i : f];
    leT += __popcll(__ballot(i < l && !(pv < v)));
  }
  int cG = 0, cL = 0, kc = 0, fail = INT32_MAX;
  for (int b = f + 1; b < l; b += 64) {
    const int i = b + lane;
    const bool in = i < l;
    const double v = S.K[in ? i : f];
    const bool ge = in && !(v < pv), le = in && !(pv < v);
    const uint64_t bg = __ballot(ge), bl = __ballot(le);
    if (le) S.le[cL + __popcll(bl & lt)] = (uint16_t)(i - f);
    bool ok = false;
    if (ge) {
      const int kk = cG + __popcll(bg & lt) + 1;
      const int leIncl = cL + __popcll(bl & (lt | me));
      ok = (leT - leIncl) >= kk;
      if (ok) S.ge[kk - 1] = (uint16_t)(i - f);
    }
    const uint64_t okm = __ballot(ok), badm = __ballot(ge && !ok);
    kc += __popcll(okm);
    if (badm && fail == INT32_MAX) fail = b + __builtin_ctzll(badm);
    cG += __popcll(bg);
    cL += __popcll(bl);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int kk = lane; kk < kc; kk += 64) {
    const int i = f + S.ge[kk], j = f + S.le[leT - 1 - kk];
    const double ki = S.K[i], kj = S.K[j];
    const uint32_t hi = S.H[i], hj = S.H[j];
    S.K[i] = kj; S.K[j] = ki;
    S.H[i] = hj; S.H[j] = hi;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (kc == 0) return fail;
  const int rK = f + S.le[leT - kc];
  return fail < rK ? fail : rK;
}

__global__ __launch_bounds__(kBlock) void k_exact_sort(const int32_t* __restrict__ nodes,
                                                        const NodeInfo* __restrict__ infos, NodeWork* __restrict__ work,
                                                        SampleParams P, unsigned long long* counters, int stage_cap) {
  constexpr int kWaveSeg = 8192;   // segments up to this size: partitioned by one wavefront
  constexpr int kList = 256;       // larger segments per level (block-wide partitions)
  constexpr int kWaveList = 1024;  // segments handed to the wavefronts
  constexpr int kStack = 64;       // frames of a wavefront's partition loop
  __shared__ int32_t s_cf[kList], s_cl[kList], s_cd[kList];
  __shared__ int32_t s_nf[kList], s_nl[kList], s_nd[kList];
  __shared__ int32_t s_df[kWaveList], s_dl[kWaveList], s_dd[kWaveList];
  __shared__ int32_t s_sf[kWaves][kStack], s_sl[kWaves][kStack], s_sd[kWaves][kStack];
  __shared__ int s_nc, s_nn, s_nd_count, s_bad, s_any;
  __shared__ int32_t s_lf[kWaves][kStack], s_ll[kWaves][kStack], s_ld[kWaves][kStack];
  __shared__ PartShared sh;
  extern __shared__ __attribute__((aligned(16))) uint8_t s_stage[];  // kWaves x 16 x stage_cap bytes
  const int k = blockIdx.x;
  NodeWork w = work[k];
  if (!(w.flags & NW_EXACT_SORT) || (w.flags & NW_OVERFLOW)) return;
  const int64_t n = w.n;
  Layout L = make_layout(w.scratch, n, P.method);
  const NodeGeom g = node_geom(infos[nodes[k]], P);
  const int tid = threadIdx.x;
  // Only the order of marked elements is needed: the members of tie groups
  // that conflict (or of every tie group when the 10 001-check window of
  // the Poisson scan could bind).  Segments of the replay without a marked
  // element are not followed further.
  uint8_t* mark = L.acc;  // by main index
  for (int64_t i = tid; i < n; i += kBlock) mark[i] = 0;
  __shared__ int s_cap;
  if (tid == 0) { s_any = 0; s_cap = 0; }
  __syncthreads();
  // Can the 10 001-check window bind?  The test of k_accept, per 64-candidate
  // window: more than 10 000 candidates in [min limSq of the window, add).
  // Only then does the order of ties that cannot exclude each other matter
  // (the sequential form then scans the accepted points in sorted order).
  if (P.method == METHOD_POISSON && n > 10001) {
    // one window per wavefront (its candidates on the lanes), four windows'
    // loads in flight
    const int lane = tid & 63, wave = tid >> 6;
    bool cap = false;
    for (int64_t q0 = wave; 64 * q0 < n; q0 += 4 * kWaves) {
      double v[4], gate[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t w0 = 64 * (q0 + u * kWaves), p = w0 + lane;
        const int64_t hi = w0 + 64 < n ? w0 + 64 : n, m = hi - 10001;
        v[u] = w0 < n ? L.sadd[p < n ? p : n - 1] : 0.0;
        gate[u] = (w0 < n && w0 + 64 > 10001 && m >= 1) ? L.sadd[m - 1] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t w0 = 64 * (q0 + u * kWaves), p = w0 + lane;
        const int64_t hi = w0 + 64 < n ? w0 + 64 : n, m = hi - 10001;
        double ml = 1e308;
        if (p < n) {
          const double cd = __builtin_sqrt(v[u]);
          const double lim = cd - g.spacing;
          ml = lim * lim;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
          const double o = __shfl_xor(ml, d);
          ml = o < ml ? o : ml;
        }
        // more than 10 000 in [ml, add(hi - 1)): the first sadd >= ml lies
        // before hi - 10 001, i.e. sadd[hi - 10 002] >= ml (sadd ascending)
        if (w0 < n && w0 + 64 > 10001 && m >= 1 && !(gate[u] < ml)) cap = true;
      }
    }
    if (cap) s_cap = 1;
  }
  __syncthreads();
  const bool all_ties = s_cap != 0;
  // tie group heads, four positions per thread with their loads in flight
  for (int64_t b0 = 0; b0 + 1 < n; b0 += 4 * kBlock) {
    double dp[4], dc[4], dn[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      dc[u] = p < n ? L.sadd[p] : -1.0;
      dn[u] = p + 1 < n ? L.sadd[p + 1] : -2.0;
      dp[u] = (p > 0 && p < n) ? L.sadd[p - 1] : -3.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      if (p + 1 >= n || !(dc[u] == dn[u]) || dp[u] == dc[u]) continue;
      const int64_t e = run_end(L.sadd, p, n);
      if (all_ties || tie_group_conflicts(L, g, P, p, e)) {
        for (int64_t q = p; q < e; q++) mark[L.idxA[q]] = 1;
        s_any = 1;
      }
    }
  }
  __syncthreads();
  // the partition phase works on (distance, main index) pairs; bit 31 of
  // the main index marks the element, so that the count of marked elements
  // on either side of a cut reads the pairs only
  PairArr a{L.sadd, L.idxB};
  uint32_t* posLe = L.keyB;
  uint32_t* posGe = L.keyA;
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = b0 + u * kBlock + tid;
      v[u] = i < n ? L.add[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = b0 + u * kBlock + tid;
      if (i < n) { a.K[i] = v[u]; a.H[i] = (uint32_t)i | (mark[i] ? kStageMark : 0u); }
    }
  }
  if (tid == 0) {
    atomicAdd(&counters[3], 1ull);
    s_nc = 0; s_nd_count = 0; s_bad = 0;
    int64_t d = 0;
    for (int64_t m = n; m > 1; m >>= 1) d++;
    if (n > 16 && s_any) {
      if (n > kWaveSeg) { s_cf[0] = 0; s_cl[0] = (int32_t)n; s_cd[0] = (int32_t)(2 * d); s_nc = 1; }
      else { s_df[0] = 0; s_dl[0] = (int32_t)n; s_dd[0] = (int32_t)(2 * d); s_nd_count = 1; }
    }
  }
  __syncthreads();
  // levels of the recursion above kWaveSeg: every segment of a level is
  // partitioned by the whole block; both parts continue with the
  // decremented depth, the parts up to kWaveSeg on the wavefronts below
  for (int level = 0; level < 128 && s_nc > 0 && !s_bad; level++) {
    if (tid == 0) s_nn = 0;
    __syncthreads();
    const int nc = s_nc;
    for (int c = 0; c < nc; c++) {
      int64_t f = s_cf[c], l = s_cl[c];
      int depth = s_cd[c];
      if (depth == 0) {
        if (tid == 0) {
          if (s_nd_count < kWaveList) { s_df[s_nd_count] = (int32_t)f; s_dl[s_nd_count] = (int32_t)l; s_dd[s_nd_count] = 0; s_nd_count++; }
          else s_bad = 1;
        }
        __syncthreads();
        continue;
      }
      if (tid == 0) {
        int64_t mid = f + (l - f) / 2, x = f + 1, y = mid, z = l - 1;
        double kx = a.K[x], ky = a.K[y], kz = a.K[z];
        int64_t pick;
        if (kx < ky) pick = (ky < kz) ? y : ((kx < kz) ? z : x);
        else pick = (kx < kz) ? x : ((ky < kz) ? z : y);
        pair_swap(a, f, pick);
      }
      __syncthreads();
      int64_t cut = block_partition(a, posLe, posGe, f, l, sh);
      {
        unsigned long long mL = 0, mR = 0;
        for (int64_t b = f; b < l; b += 4 * kBlock) {
          uint32_t h[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int64_t i = b + u * kBlock + tid;
            h[u] = i < l ? a.H[i] : 0u;
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int64_t i = b + u * kBlock + tid;
            if (h[u] & kStageMark) { if (i < cut) mL++; else mR++; }
          }
        }
        if (mL) atomicAdd(&sh.markL, mL);
        if (mR) atomicAdd(&sh.markR, mR);
      }
      __syncthreads();
      if (tid == 0) {
        if (cut <= f || cut >= l) {
          s_bad = 1;  // cannot happen for a median-of-three pivot; stay bounded
        } else {
          int64_t pf[2] = {f, cut}, pl[2] = {cut, l};
          const unsigned long long marks[2] = {sh.markL, sh.markR};
          for (int side = 0; side < 2; side++) {
            int64_t sz = pl[side] - pf[side];
            if (sz <= 16 || marks[side] == 0) continue;
            if (sz > kWaveSeg && s_nn < kList) {
              s_nf[s_nn] = (int32_t)pf[side]; s_nl[s_nn] = (int32_t)pl[side]; s_nd[s_nn] = depth - 1; s_nn++;
            } else if (sz <= kWaveSeg && s_nd_count < kWaveList) {
              s_df[s_nd_count] = (int32_t)pf[side]; s_dl[s_nd_count] = (int32_t)pl[side]; s_dd[s_nd_count] = depth - 1;
              s_nd_count++;
            } else {
              s_bad = 1;
            }
          }
        }
      }
      __syncthreads();
      if (s_bad) break;
    }
    __syncthreads();
    if (tid == 0) {
      for (int c = 0; c < s_nn; c++) { s_cf[c] = s_nf[c]; s_cl[c] = s_nl[c]; s_cd[c] = s_nd[c]; }
      s_nc = s_nn;
    }
    __syncthreads();
  }
  if (s_nc > 0) s_bad = 1;
  __syncthreads();
  if (s_bad) {
    // lists exhausted: replay the whole sort on one lane (same result)
    if (tid == 0) {
      for (int64_t i = 0; i < n; i++) L.idxB[i] = (uint32_t)i;
      lstd::sort(L.idxB, n, lstd::KeyLess{L.add});
      for (int64_t i = 0; i < n; i++) L.sadd[i] = L.add[L.idxB[i]];
    }
    __syncthreads();
  } else {
    // segments up to kWaveSeg: the rest of the partition loop, one wavefront
    // each, with its own stack of the parts still to do (the right part is
    // pushed, the left one continues, as in pair_partition_loop; parts
    // without a marked element are dropped); parts up to stage_cap staged in
    // LDS and partitioned there by the same wavefront (LDS latency instead
    // of memory latency for every pass), parts at the depth limit
    // heap-sorted by one lane in place
    const int lane = tid & 63, wave = tid >> 6;
    bool wbad = false;
    for (int d = wave; d < s_nd_count && !wbad; d += kWaves) {
      int sp = 0;
      s_sf[wave][0] = s_df[d]; s_sl[wave][0] = s_dl[d]; s_sd[wave][0] = s_dd[d];
      sp = 1;
      while (sp > 0 && !wbad) {
        sp--;
        int64_t f = s_sf[wave][sp], l = s_sl[wave][sp];
        int depth = s_sd[wave][sp];
        while (l - f > 16) {
          if (depth == 0) {
            if (lane == 0) pair_heap_sort(a, f, l - f);
            break;
          }
          const int len = (int)(l - f);
          if (len <= stage_cap) {
            // the rest of this segment's partition loop on its LDS copy: the
            // same steps (wave partitions down to 16 elements, heap sort at
            // the depth limit, unmarked parts dropped)
            const StageLds S = stage_view(s_stage + (size_t)wave * 16 * stage_cap, stage_cap);
            for (int i0 = 0; i0 < len; i0 += 4 * 64) {
              double kv[4];
              uint32_t hv[4];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const int i = i0 + 64 * u + lane;
                const int ic = i < len ? i : len - 1;
                kv[u] = a.K[f + ic];
                hv[u] = a.H[f + ic];
              }
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const int i = i0 + 64 * u + lane;
                if (i < len) { S.K[i] = kv[u]; S.H[i] = hv[u]; }
              }
            }
            if (lane == 0) { s_lf[wave][0] = 0; s_ll[wave][0] = len; s_ld[wave][0] = depth; }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            int lsp = 1;
            while (lsp > 0 && !wbad) {
              lsp--;
              int lf = s_lf[wave][lsp], ll = s_ll[wave][lsp], ld = s_ld[wave][lsp];
              while (ll - lf > 16) {
                if (ld == 0) {
                  if (lane == 0) pair_heap_sort(PairArr{(double*)S.K, (uint32_t*)S.H}, lf, ll - lf);
                  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                  __builtin_amdgcn_wave_barrier();
                  break;
                }
                ld--;
                if (lane == 0) pair_pivot_to_first(PairArr{(double*)S.K, (uint32_t*)S.H}, lf, ll);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const int cut = wave_partition_lds(S, lf, ll);
                if (cut <= lf || cut >= ll) {  // cannot happen for a median-of-three pivot; stay bounded
                  wbad = true;
                  break;
                }
                uint64_t mL = 0, mR = 0;
                for (int b = lf; b < ll; b += 64) {
                  const int i = b + lane;
                  const bool m = i < ll && (S.H[i < ll ? i : lf] & kStageMark) != 0;
                  mL |= __ballot(m && i < cut);
                  mR |= __ballot(m && i >= cut);
                }
                if (ll - cut > 16 && mR) {
                  if (lsp >= kStack) { wbad = true; break; }
                  if (lane == 0) { s_lf[wave][lsp] = cut; s_ll[wave][lsp] = ll; s_ld[wave][lsp] = ld; }
                  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                  __builtin_amdgcn_wave_barrier();
                  lsp++;
                }
                if (!mL) break;
                ll = cut;
              }
            }
            for (int i = lane; i < len; i += 64) {
              a.K[f + i] = S.K[i];
              a.H[f + i] = S.H[i];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            break;
          }
          depth--;
          if (lane == 0) pair_pivot_to_first(a, f, l);
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          const int64_t cut = wave_partition(a, posLe, posGe, f, l);
          if (cut <= f || cut >= l) {  // cannot happen for a median-of-three pivot; stay bounded
            wbad = true;
            break;
          }
          uint64_t mL = 0, mR = 0;
          for (int64_t b = f; b < l; b += 256) {
            uint32_t h[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const int64_t i = b + 64 * u + lane;
              h[u] = a.H[i < l ? i : f];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const int64_t i = b + 64 * u + lane;
              const bool m = i < l && (h[u] & kStageMark) != 0;
              mL += __popcll(__ballot(m && i < cut));
              mR += __popcll(__ballot(m && i >= cut));
            }
          }
          if (l - cut > 16 && mR) {
            if (sp >= kStack) { wbad = true; break; }
            if (lane == 0) { s_sf[wave][sp] = (int32_t)cut; s_sl[wave][sp] = (int32_t)l; s_sd[wave][sp] = depth; }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            sp++;
          }
          if (!mL) break;
          l = cut;
        }
      }
    }
    if (wbad && lane == 0) s_bad = 1;
    __syncthreads();
    if (s_bad) {
      // a wavefront's stack or partition failed: replay the whole sort on
      // one lane (same result)
      if (tid == 0) {
        for (int64_t i = 0; i < n; i++) L.idxB[i] = (uint32_t)i;
        lstd::sort(L.idxB, n, lstd::KeyLess{L.add});
        for (int64_t i = 0; i < n; i++) L.sadd[i] = L.add[L.idxB[i]];
      }
      __syncthreads();
    }
  }
  // Positions after the partition phase, then each marked run of equal
  // distances [p, e) (radix order in idxA) is put in the order of those
  // positions.  Every marked element lies in a final segment of at most 16
  // elements (or a heap-sorted one, where it already holds its final
  // place), and the final insertion sort keeps it inside that segment, so
  // the run's members sit at post-partition positions [p - 15, e + 15).
  // Short runs: one lane, insertion by position.  Long runs: one wave walks
  // that window in order and compacts the run's members by ballot.
  // Unmarked runs keep the radix order: no decision depends on it.
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    uint32_t h[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = b0 + u * kBlock + tid;
      h[u] = i < n ? a.H[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = b0 + u * kBlock + tid;
      if (i < n) posLe[h[u] & ~kStageMark] = (uint32_t)i;
    }
  }
  __syncthreads();
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    uint32_t v[4];
    double d[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      v[u] = p < n ? L.idxA[p] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      d[u] = p < n ? L.add[v[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = b0 + u * kBlock + tid;
      if (p < n) L.sadd[p] = d[u];
    }
  }
  __syncthreads();
  if (s_any) {
    const int lane = tid & 63, wave = tid >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int64_t base = 64 * (int64_t)wave; base < n; base += kBlock) {
      const int64_t p = base + lane;
      bool head = p + 1 < n && L.sadd[p] == L.sadd[p + 1] && !(p > 0 && L.sadd[p - 1] == L.sadd[p]);
      if (head) head = mark[L.idxA[p]] != 0;
      const int64_t e = head ? run_end(L.sadd, p, n) : p;
      const bool longrun = head && e - p > kShortRun;
      if (head && !longrun) {
        for (int64_t i = p + 1; i < e; i++) {
          uint32_t v = L.idxA[i];
          uint32_t pv = posLe[v];
          int64_t q = i - 1;
          while (q >= p && posLe[L.idxA[q]] > pv) {
            L.idxA[q + 1] = L.idxA[q];
            q--;
          }
          L.idxA[q + 1] = v;
        }
      }
      for (uint64_t lm = __ballot(longrun); lm; lm &= lm - 1) {
        const int src = __builtin_ctzll(lm);
        const int64_t rp = __shfl((int)p, src), re = __shfl((int)e, src);
        const double key = L.sadd[rp];
        const int64_t w0 = rp > 16 ? rp - 16 : 0, w1 = re + 16 < n ? re + 16 : n;
        int64_t r = rp;
        for (int64_t j0 = w0; j0 < w1; j0 += 64) {
          const int64_t j = j0 + lane;
          uint32_t h = 0;
          bool hit = false;
          if (j < w1) {
            h = a.H[j] & ~kStageMark;
            hit = L.add[h] == key;
          }
          const uint64_t hm = __ballot(hit);
          if (hit && r + __popcll(hm & lt) < re) L.idxA[r + __popcll(hm & lt)] = h;
          r += __popcll(hm);
        }
        // the window held exactly the run: anything else is a broken
        // invariant, reported instead of a wrong order
        if (lane == 0 && r != re) atomicAdd(&counters[8], 1ull);
      }
    }
  }
}

// ------------------------------------------------------------- acceptance
__device__ inline double dd64_xyz(int32_t ax, int32_t ay, int32_t az, int32_t bx, int32_t by, int32_t bz,
                                  const SampleParams& P) {
  double dx = coord(ax, P.scale[0], P.offset[0]) - coord(bx, P.scale[0], P.offset[0]);
  double dy = coord(ay, P.scale[1], P.offset[1]) - coord(by, P.scale[1], P.offset[1]);
  double dz = coord(az, P.scale[2], P.offset[2]) - coord(bz, P.scale[2], P.offset[2]);
  return dx * dx + dy * dy + dz * dz;
}

__device__ inline double dd64_pair(const Layout& L, uint32_t a, uint32_t b, const SampleParams& P) {
  const int4 pa = L.xyz[a], pb = L.xyz[b];
  return dd64_xyz(pa.x, pa.y, pa.z, pb.x, pb.y, pb.z, P);
}

// exact "dd < spacing^2" through an FP32 filter on integer deltas; the
// FP64 reference expression is evaluated only inside a +-1e-3 band
struct Filter {
  float sc[3];
  float s2f;
  double s2;
  unsigned long long rechecks = 0, tests = 0;
  float maxerr = 0.f;
};

__device__ inline bool within_xyz(int32_t ax, int32_t ay, int32_t az, int32_t bx, int32_t by, int32_t bz,
                                  const SampleParams& P, Filter& F) {
  F.tests++;
  float dx = (float)((int64_t)ax - (int64_t)bx) * F.sc[0];
  float dy = (float)((int64_t)ay - (int64_t)by) * F.sc[1];
  float dz = (float)((int64_t)az - (int64_t)bz) * F.sc[2];
  float d32 = dx * dx + dy * dy + dz * dz;
  if (d32 > F.s2f * 1.001f) return false;
  if (d32 < F.s2f * 0.999f) return true;
  F.rechecks++;
  double d64 = dd64_xyz(ax, ay, az, bx, by, bz, P);
  float e = (float)(__builtin_fabs((double)d32 - d64) / F.s2);
  F.maxerr = e > F.maxerr ? e : F.maxerr;
  return d64 < F.s2;
}

__device__ inline Filter make_filter(const NodeGeom& g, const SampleParams& P) {
  Filter F;
  for (int a = 0; a < 3; a++) F.sc[a] = (float)P.scale[a];
  F.s2f = (float)g.s2;
  F.s2 = g.s2;
  return F;
}

// one atomic per wave for the filter statistics
__device__ inline void flush_filter(const Filter& F, const SampleParams& P) {
  unsigned long long r = F.rechecks, t = F.tests;
  float m = F.maxerr;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    r += __shfl_xor(r, d);
    t += __shfl_xor(t, d);
    float o = __shfl_xor(m, d);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) {
    if (r) atomicAdd(&P.counters[0], r);
    if (t) atomicAdd(&P.counters[1], t);
    if (m > 0.f) atomicMax((unsigned int*)&P.counters[2], __float_as_uint(m));
  }
}

// slot of a cell in the cell table, kNil when the cell holds no candidate
__device__ inline uint32_t table_find(const Layout& L, uint32_t key) {
  uint32_t slot = ghash(key, L.H);
  for (uint32_t probe = 0; probe < L.H; probe++) {
    uint32_t k = L.htab[slot].x;
    if (k == key) return slot;
    if (k == kNil) return kNil;
    slot = (slot + 1) & (L.H - 1);
  }
  return kNil;
}

struct CellBox {
  int lx, hx, ly, hy, lz, hz;
};

__device__ inline CellBox cell_box(const NodeGeom& g, double x, double y, double z) {
  const double cs = 2.0 * g.spacing, sq = g.spacing * (1.0 + 1e-9);
  CellBox b;
  b.lx = gcell(x - sq, g.min[0], cs); b.hx = gcell(x + sq, g.min[0], cs);
  b.ly = gcell(y - sq, g.min[1], cs); b.hy = gcell(y + sq, g.min[1], cs);
  b.lz = gcell(z - sq, g.min[2], cs); b.hz = gcell(z + sq, g.min[2], cs);
  return b;
}

// Cell table of a node: candidates sorted by their 2*spacing cell (stable,
// so ascending sorted position inside a cell) with their coordinates
// copied alongside, and a hash from cell to its range.
constexpr int kCellsU = 8;  // candidates per thread in flight in k_cells' gathers
__global__ __launch_bounds__(kBlock) void k_cells(const int32_t* __restrict__ nodes, const NodeInfo* __restrict__ infos,
                                                   NodeWork* __restrict__ work, SampleParams P) {
  __shared__ RadixShared s_rad;
  __shared__ uint32_t s_h3[3][256];  // histograms of the three 7-bit key digits, counted with the keys
  const int k = blockIdx.x;
  const NodeWork w = work[k];
  if ((w.flags & NW_OVERFLOW) || !(w.flags & NW_CELL_PATH)) return;
  const int64_t n = w.n;
  const int tid = threadIdx.x;
  Layout L = make_layout(w.scratch, n, P.method);
  const NodeGeom g = node_geom(infos[nodes[k]], P);
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  const double cs = 2.0 * g.spacing;
  for (int i = tid; i < 3 * 256; i += kBlock) (&s_h3[0][0])[i] = 0;
  __syncthreads();
  // kCellsU candidates per thread in flight (the loop is bound by the
  // latency of its two dependent loads); positions past the node read
  // position 0 and are not used
  const gptr<const uint32_t> gIdxA = gbl((const uint32_t*)L.idxA);
  const gptr<const int4> gXyz = gbl((const int4*)L.xyz);
  for (int64_t p0 = 0; p0 < n; p0 += kCellsU * kBlock) {
    uint32_t j4[kCellsU];
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int64_t p = p0 + u * kBlock + tid;
      j4[u] = gIdxA[p < n ? p : 0];
    }
    int32_t X4[kCellsU], Y4[kCellsU], Z4[kCellsU];
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int64_t p = p0 + u * kBlock + tid;
      const int4 q = ld_g16i(gXyz + j4[u]);
      X4[u] = q.x; Y4[u] = q.y; Z4[u] = q.z;
    }
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int64_t p = p0 + u * kBlock + tid;
      if (p >= n) continue;
      double x = coord(X4[u], P.scale[0], P.offset[0]), y = coord(Y4[u], P.scale[1], P.offset[1]),
             z = coord(Z4[u], P.scale[2], P.offset[2]);
      const uint32_t key = gkey(gcell(x, g.min[0], cs), gcell(y, g.min[1], cs), gcell(z, g.min[2], cs));
      gbl(L.keyA)[p] = key;
      gbl(L.idxB)[p] = (uint32_t)p;
      atomicAdd(&s_h3[0][key & 127u], 1u);
      atomicAdd(&s_h3[1][(key >> 7) & 127u], 1u);
      atomicAdd(&s_h3[2][(key >> 14) & 127u], 1u);
    }
  }
  for (uint32_t i = tid; i < L.H; i += kBlock) L.htab[i] = make_uint4(kNil, 0u, 0u, 0u);
  __syncthreads();
  uint32_t *kc = L.keyA, *vc = L.idxB, *ko = L.keyB, *vo = L.next;
  for (int pass = 0; pass < 3; pass++) {
    if (!block_radix_pass<true>(kc, vc, ko, vo, n, 7 * pass, 127u, s_rad, s_h3[pass])) continue;
    uint32_t* t = kc; kc = ko; ko = t;
    t = vc; vc = vo; vo = t;
  }
  // cell-ordered entries, kCellsU per thread with each level of the index
  // chain (sorted position -> main index -> coordinates) loaded together
  // (positions past the node read position 0 and are not used)
  const gptr<const uint32_t> gVc = gbl((const uint32_t*)vc), gKc = gbl((const uint32_t*)kc);
  for (int64_t e0 = tid; e0 < n; e0 += kCellsU * kBlock) {
    uint32_t p4[kCellsU], k4[kCellsU], kp4[kCellsU], j4[kCellsU];
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int64_t e = e0 + u * kBlock;
      const int64_t ec = e < n ? e : 0;
      p4[u] = gVc[ec];
      k4[u] = gKc[ec];
      kp4[u] = gKc[ec > 0 ? ec - 1 : 0];
    }
#pragma unroll
    for (int u = 0; u < kCellsU; u++) j4[u] = gIdxA[p4[u]];
    int4 v4[kCellsU];
    double a4[kCellsU];
    uint32_t c4[kCellsU];
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int4 q = ld_g16i(gXyz + j4[u]);
      v4[u] = make_int4(q.x, q.y, q.z, (int32_t)p4[u]);
      if (avg) c4[u] = gbl((const uint32_t*)L.cell16)[j4[u]];
      else a4[u] = gbl((const double*)L.sadd)[p4[u]];
    }
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int64_t e = e0 + u * kBlock;
      if (e >= n) continue;
      st_g16i(gbl(L.ent) + e, v4[u]);
      if (avg) gbl(L.eC16)[e] = c4[u];
      else gbl(L.eAdd)[e] = a4[u];
    }
    // each cell's first entry claims the cell's hash slot and writes its
    // range: the claims of all kCellsU entries issued together (a key is
    // claimed once: keys ascend in kc), probing on only for collisions; the
    // cell's end found by galloping over kc meanwhile
    uint32_t slot[kCellsU];
    uint64_t pend = 0;
#pragma unroll
    for (int u = 0; u < kCellsU; u++) {
      const int64_t e = e0 + u * kBlock;
      slot[u] = ghash(k4[u], L.H);
      if (e < n && (e == 0 || kp4[u] != k4[u])) pend |= 1ull << u;
    }
    uint32_t end4[kCellsU];
#pragma unroll
    for (int u = 0; u < kCellsU; u++)
      end4[u] = ((pend >> u) & 1) ? (uint32_t)run_end(kc, e0 + u * kBlock, n) : 0u;
    while (pend) {
      uint32_t old[kCellsU];
#pragma unroll
      for (int u = 0; u < kCellsU; u++) old[u] = ((pend >> u) & 1) ? atomicCAS(&L.htab[slot[u]].x, kNil, k4[u]) : 0u;
#pragma unroll
      for (int u = 0; u < kCellsU; u++) {
        if (!((pend >> u) & 1)) continue;
        if (old[u] == kNil) {
          L.htab[slot[u]].y = (uint32_t)(e0 + u * kBlock);
          L.htab[slot[u]].z = end4[u];
          pend &= ~(1ull << u);
        } else {
          slot[u] = (slot[u] + 1) & (L.H - 1);
        }
      }
    }
  }
}

// Work items of a level: 256 candidates of one node each.  item_start[k]
// is the first item of node k, item_start[num_nodes] the total.
__global__ __launch_bounds__(1024) void k_plan(const NodeWork* __restrict__ work, int32_t num_nodes,
                                                uint32_t* __restrict__ item_start) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_run;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_run = 0;
  __syncthreads();
  for (int b = 0; b < num_nodes; b += 1024) {
    const int k = b + tid;
    uint32_t v = 0;
    if (k < num_nodes && (work[k].flags & (NW_OVERFLOW | NW_CELL_PATH)) == NW_CELL_PATH)
      v = (uint32_t)((work[k].n + kBlock - 1) / kBlock);
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < 16; q++) {
      if (q < wave) woff += s_w[q];
      tot += s_w[q];
    }
    const uint32_t run = s_run;
    if (k < num_nodes) item_start[k] = run + woff + incl - v;
    __syncthreads();
    if (tid == 0) s_run = run + tot;
    __syncthreads();
  }
  if (tid == 0) item_start[num_nodes] = s_run;
}

// Per candidate, the earlier candidates that can reject it: within spacing
// and (Poisson) not behind the scan's early exit (sampler_poisson.h:160) or
// (Poisson-average) inside its 16^3 search range.  Candidates are taken in
// cell order so that the lanes of a wave read the same cell ranges; all
// nodes of the level share one grid.
__global__ __launch_bounds__(kBlock) void k_neighbors(const int32_t* __restrict__ nodes,
                                                       const NodeInfo* __restrict__ infos,
                                                       NodeWork* __restrict__ work, int32_t num_nodes,
                                                       const uint32_t* __restrict__ item_start, SampleParams P) {
  const uint32_t total = item_start[num_nodes];
  const int tid = threadIdx.x;
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  Filter stats;
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    int lo = 0, hi = num_nodes - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (item_start[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int k = __builtin_amdgcn_readfirstlane(lo);
    const NodeWork& w = work[k];
    const int64_t n = w.n;
    const int64_t e = (int64_t)(item - item_start[k]) * kBlock + tid;
    if (e >= n) continue;
    const Layout L = make_layout(w.scratch, n, P.method);
    const NodeGeom g = node_geom(infos[nodes[k]], P);
    Filter F = make_filter(g, P);
    const int4 me = L.ent[e];
    const uint32_t p = (uint32_t)me.w;
    const int32_t X = me.x, Y = me.y, Z = me.z;
    const double x = coord(X, P.scale[0], P.offset[0]), y = coord(Y, P.scale[1], P.offset[1]),
                 z = coord(Z, P.scale[2], P.offset[2]);
    int lo16[3], hi16[3];
    double limSq = 0.0;
    if (!avg) {
      double cd = __builtin_sqrt(L.eAdd[e]);
      double lim = cd - g.spacing;
      limSq = lim * lim;
    } else {
      range16(g, x, y, z, lo16, hi16);
    }
    const CellBox cb = cell_box(g, x, y, z);
    // every earlier candidate within spacing that passes the filter, in cell
    // order; emit(q, i) receives the i-th one
    auto scan = [&](auto&& emit) {
      int cnt = 0;
      for (int cz = cb.lz; cz <= cb.hz; cz++)
        for (int cy = cb.ly; cy <= cb.hy; cy++)
          for (int cx = cb.lx; cx <= cb.hx; cx++) {
            const uint32_t slot = table_find(L, gkey(cx, cy, cz));
            if (slot == kNil) continue;
            const uint4 t = L.htab[slot];
            // four entries loaded per step, then taken in order (entries of
            // a cell rise in sorted position: the first later one ends it)
            bool stop = false;
            for (uint32_t f0 = t.y; f0 < t.z && !stop; f0 += 4) {
              int4 o4[4];
              double a4[4];
              uint32_t c4[4];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const uint32_t f = f0 + u;
                o4[u] = f < t.z ? L.ent[f] : make_int4(0, 0, 0, -1);
                if (!avg) a4[u] = f < t.z ? L.eAdd[f] : 0.0;
                else c4[u] = f < t.z ? L.eC16[f] : 0u;
              }
#pragma unroll
              for (int u = 0; u < 4; u++) {
                if (stop || f0 + u >= t.z) continue;
                const int4 o = o4[u];
                const uint32_t q = (uint32_t)o.w;
                if (q >= p) { stop = true; continue; }
                if (!within_xyz(o.x, o.y, o.z, X, Y, Z, P, F)) continue;
                if (!avg) {
                  if (a4[u] < limSq) continue;
                } else {
                  if (!in_range16(c4[u], lo16, hi16)) continue;
                }
                emit(q, cnt);
                cnt++;
              }
            }
          }
      return cnt;
    };
    uint32_t* out = L.nbr + (uint64_t)p * kMaxNbr;
    const int cnt = scan([&](uint32_t q, int i) {
      if (i < kMaxNbr) out[i] = q;
    });
    if (cnt <= kMaxNbr) {
      L.nbrCount[p] = (uint8_t)cnt;
    } else {
      // the full list goes to the level's pool (a second identical scan);
      // the row keeps its offset and length
      atomicOr(&work[k].flags, (uint32_t)NW_MANY_NBR);
      atomicAdd(&P.counters[7], 1ull);
      const unsigned long long off = atomicAdd(P.nbr_pool_top, (unsigned long long)cnt);
      if (off + (unsigned long long)cnt <= P.nbr_pool_cap) {
        uint32_t* dst = P.nbr_pool + off;
        Filter F2 = make_filter(g, P);
        Filter keep = F;
        F = F2;
        scan([&](uint32_t q, int i) {
          if (i < cnt) dst[i] = q;
        });
        F = keep;
        out[0] = (uint32_t)off;
        out[1] = (uint32_t)(off >> 32);
        out[2] = (uint32_t)cnt;
        L.nbrCount[p] = kNbrPooled;
      } else {
        L.nbrCount[p] = kNbrOverflow;  // searched again by k_accept
      }
    }
    stats.rechecks += F.rechecks;
    stats.tests += F.tests;
    stats.maxerr = F.maxerr > stats.maxerr ? F.maxerr : stats.maxerr;
  }
  flush_filter(stats, P);
}

// Greedy acceptance in sorted order, one wavefront per node, 64 candidates
// per window: a candidate is rejected by an accepted neighbour of an earlier
// window; inside the window the dependencies are resolved by ballots.  The
// accepted flags of the node live in LDS (global memory past 65536
// candidates).
// The rare forms of one candidate's test in a window: a neighbour list
// longer than a row (walked from the level's pool, four entries in flight per
// step) or one that overflowed the pool (the candidate's cells searched in the
// cell table).  Sets `rej` when an accepted earlier window's neighbour is
// found, and the in-window neighbours in `in_mask`.
template <class BitsPtr>
__device__ inline void lane_slow(const SampleParams& P, const Layout& L, const NodeGeom& g, Filter& F, bool avg,
                                 int64_t p, int64_t w0, int cnt, uint32_t r0, uint32_t r1, uint32_t r2, double limSq,
                                 BitsPtr bits, uint64_t& in_mask, bool& rej) {
  if (cnt == kNbrPooled) {
    const uint32_t* lst = P.nbr_pool + ((uint64_t)r0 | ((uint64_t)r1 << 32));
    const int len = (int)r2;
    for (int i0 = 0; i0 < len && !rej; i0 += 4) {
      uint32_t q4[4];
#pragma unroll
      for (int u = 0; u < 4; u++) q4[u] = i0 + u < len ? lst[i0 + u] : 0xFFFFFFFFu;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t q = q4[u];
        if (q == 0xFFFFFFFFu || rej) continue;
        if ((int64_t)q >= w0) {
          in_mask |= 1ull << (q - w0);
        } else if ((bits[q >> 5] >> (q & 31)) & 1u) {
          rej = true;
        }
      }
    }
    return;
  }
  const uint32_t j = L.idxA[p];
  const int4 q = L.xyz[j];
  const int32_t X = q.x, Y = q.y, Z = q.z;
  const double x = coord(X, P.scale[0], P.offset[0]), y = coord(Y, P.scale[1], P.offset[1]),
               z = coord(Z, P.scale[2], P.offset[2]);
  int lo16[3], hi16[3];
  if (avg) range16(g, x, y, z, lo16, hi16);
  const CellBox cb = cell_box(g, x, y, z);
  for (int cz = cb.lz; cz <= cb.hz && !rej; cz++)
    for (int cy = cb.ly; cy <= cb.hy && !rej; cy++)
      for (int cx = cb.lx; cx <= cb.hx && !rej; cx++) {
        const uint32_t slot = table_find(L, gkey(cx, cy, cz));
        if (slot == kNil) continue;
        const uint4 t = L.htab[slot];
        for (uint32_t f = t.y; f < t.z; f++) {
          const int4 o = L.ent[f];
          const uint32_t qp = (uint32_t)o.w;
          if ((int64_t)qp >= p) break;
          if ((int64_t)qp < w0 && !((bits[qp >> 5] >> (qp & 31)) & 1u)) continue;
          if (!avg) {
            if (L.eAdd[f] < limSq) continue;
          } else {
            if (!in_range16(L.eC16[f], lo16, hi16)) continue;
          }
          if (!within_xyz(o.x, o.y, o.z, X, Y, Z, P, F)) continue;
          if ((int64_t)qp < w0) { rej = true; break; }
          in_mask |= 1ull << (qp - w0);
        }
      }
}

template <int kAccAhead, class BitsPtr>
__device__ void accept_windows(NodeWork* __restrict__ work, int k, const SampleParams& P, const Layout& L,
                               const NodeGeom& g, Filter& F, const int64_t n, BitsPtr bits);

// Greedy acceptance of node k by one wavefront; the accepted bits live in
// `lds_bits` (lds_words words) when they fit, in global memory otherwise.
template <int kAccAhead>
__device__ void accept_node(const int32_t* __restrict__ nodes, const NodeInfo* __restrict__ infos,
                            NodeWork* __restrict__ work, int k, const SampleParams& P, uint32_t* lds_bits,
                            int64_t lds_words) {
  const int lane = threadIdx.x & 63;
  const NodeWork w = work[k];
  if ((w.flags & NW_OVERFLOW) || !(w.flags & NW_CELL_PATH)) return;
  const int64_t n = w.n;
  const Layout L = make_layout(w.scratch, n, P.method);
  const NodeGeom g = node_geom(infos[nodes[k]], P);
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  Filter F = make_filter(g, P);
  const bool lds = n <= lds_words * 32;
  if (lds)
    accept_windows<kAccAhead>(work, k, P, L, g, F, n, lds_bits);
  else
    accept_windows<kAccAhead>(work, k, P, L, g, F, n, gbl(L.keyA));
  flush_filter(F, P);
}

// The window loop of accept_node with the accepted bits at `bits`, an LDS
// or a global pointer, so that each copy of the loop reads them with LDS or
// global instructions.  The node's arrays are read through global-address
// pointers: the flat instructions that a generic pointer compiles to also
// count as LDS operations, and a wait for the accepted-bit reads would then
// wait for every prefetch in flight too.
template <int kAccAhead, class BitsPtr>
__device__ void accept_windows(NodeWork* __restrict__ work, int k, const SampleParams& P, const Layout& L,
                               const NodeGeom& g, Filter& F, const int64_t n, BitsPtr bits) {
  const int lane = threadIdx.x & 63;
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  const gptr<const uint8_t> gCnt = gbl((const uint8_t*)L.nbrCount);
  const gptr<const double> gSadd = gbl((const double*)L.sadd);
  const gptr<const uint4> gNbr = gbl((const uint4*)L.nbr);
  const gptr<uint8_t> gAcc = gbl(L.acc);
  if (n <= 0) return;
  const int64_t words = (n + 31) / 32;
  for (int64_t i = lane; i < words; i += 64) bits[i] = 0;
  // the bits are written and read by this wavefront only: ordering its own
  // accesses needs no cache maintenance
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  // The window loop is bound by memory latency (one wavefront per node), so
  // the neighbour rows, counts and distances of the next kAccAhead windows
  // are in flight while a window is decided; the slots are unrolled so that
  // no register of a load in flight is moved.  The value the 10 001-check
  // test of a window reads is loaded with its window.
  struct Slot {
    uint4 r0, r1, r2, r3;
    int cnt;
    double sadd, capv;
  };
  Slot sl[kAccAhead];
#pragma unroll
  for (int s = 0; s < kAccAhead; s++) sl[s].r0 = sl[s].r1 = sl[s].r2 = sl[s].r3 = make_uint4(0, 0, 0, 0);
  // Every lane loads on every window, from a clamped position when it lies
  // past the node (its values are then unused): with no branch around the
  // loads, the number of loads issued after a slot's is fixed, and the wait
  // for a slot leaves the younger slots' loads in flight instead of waiting
  // for all of them.
  auto prefetch = [&](Slot& s, int64_t w0) {
    const int64_t p = w0 + lane < n ? w0 + lane : n - 1;
    s.cnt = gCnt[p];
    s.sadd = gSadd[p];
    const gptr<const uint4> rw = gNbr + (uint64_t)p * (kMaxNbr / 4);
    s.r0 = ld_g16(rw); s.r1 = ld_g16(rw + 1); s.r2 = ld_g16(rw + 2); s.r3 = ld_g16(rw + 3);
    if (kAccAhead > 1) {
      const int64_t hi = (w0 + 64 < n ? w0 + 64 : n);
      const int64_t c = hi - 10002;
      s.capv = gSadd[c < 0 ? 0 : (c < n ? c : n - 1)];
    }
  };
#pragma unroll
  for (int s = 0; s < kAccAhead; s++) prefetch(sl[s], 64 * s);
  bool cap_may_bind = false;
  for (int64_t wb = 0; wb < n; wb += 64 * kAccAhead) {
#pragma unroll
    for (int s = 0; s < kAccAhead; s++) {
      const int64_t w0 = wb + 64 * s;
      if (w0 >= n) break;
      const int64_t p = w0 + lane;
      const bool valid = p < n;
      uint64_t in_mask = 0;
      bool rej = false;
      double limSq = 1e308;
      const int cnt = sl[s].cnt;
      const double sa = sl[s].sadd, capv = sl[s].capv;
      const uint32_t row[16] = {sl[s].r0.x, sl[s].r0.y, sl[s].r0.z, sl[s].r0.w, sl[s].r1.x, sl[s].r1.y,
                                sl[s].r1.z, sl[s].r1.w, sl[s].r2.x, sl[s].r2.y, sl[s].r2.z, sl[s].r2.w,
                                sl[s].r3.x, sl[s].r3.y, sl[s].r3.z, sl[s].r3.w};
      prefetch(sl[s], w0 + 64 * kAccAhead);
      if (valid) {
        if (!avg) {
          double cd = __builtin_sqrt(sa);
          double lim = cd - g.spacing;
          limSq = lim * lim;
        }
        if (cnt != kNbrOverflow && cnt != kNbrPooled) {
          // every entry tested (no early exit), so that the accepted-bit loads
          // are issued together; a rejected candidate's in-window mask is not
          // used
          uint32_t hit = 0;
#pragma unroll
          for (int i = 0; i < kMaxNbr; i++) {
            const uint32_t q = row[i];
            const bool use = i < cnt;
            if (use && (int64_t)q >= w0) in_mask |= 1ull << (q - w0);
            const bool earlier = use && (int64_t)q < w0;
            const uint32_t word = earlier ? bits[q >> 5] : 0u;
            hit |= (word >> (q & 31)) & 1u;
          }
          rej = hit != 0;
        } else {
          lane_slow(P, L, g, F, avg, p, w0, cnt, row[0], row[1], row[2], limSq, bits, in_mask, rej);
        }
      }
      // in-window dependencies: the lowest undecided lane always resolves
      bool decided = !valid || rej;
      bool accepted = false;
      uint64_t dec = __ballot(decided), accm = 0;
      while (dec != ~0ull) {
        if (!decided) {
          if (in_mask & accm) {
            decided = true;
          } else if ((in_mask & ~dec) == 0) {
            decided = true;
            accepted = true;
          }
        }
        dec = __ballot(decided);
        accm = __ballot(accepted);
      }
      if (lane == 0) {
        bits[w0 >> 5] = (uint32_t)accm;
        if ((w0 >> 5) + 1 < words) bits[(w0 >> 5) + 1] = (uint32_t)(accm >> 32);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      if (!avg && w0 + 64 > 10001) {
        // the 10 001-check window of the reference scan may bind when more
        // than 10 000 candidates lie in [min limSq of the window, add)
        double ml = limSq;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
          double o = __shfl_xor(ml, d);
          ml = o < ml ? o : ml;
        }
        // the first sadd >= ml lies more than 10 000 positions before hi - 1
        // exactly when sadd[hi - 10 002] >= ml (sadd ascending): one value,
        // loaded with the window when several windows are in flight
        const int64_t hi = (w0 + 64 < n ? w0 + 64 : n);
        const int64_t m = hi - 10001;
        if (m >= 1 && lane == 0 && !((kAccAhead > 1 ? capv : gSadd[m - 1]) < ml)) cap_may_bind = true;
      }
    }
  }
  // no store in the window loop (a store there would count among the loads
  // that the slot waits see): the cap flag and the accepted bytes afterwards
  if (cap_may_bind) atomicOr(&work[k].flags, (uint32_t)NW_SEQ_ACCEPT);
  for (int64_t i = lane; i < n; i += 64) gAcc[i] = (uint8_t)((bits[i >> 5] >> (i & 31)) & 1u);
}

// one wavefront per node (accepted bits past kAccWords * 32 candidates in
// global memory: a kernel holding a large LDS share would wait for CUs that
// the other sampling streams keep busy)
// ------------------------------------------------- pipelined acceptance
// In a launch of few nodes each k_accept wavefront sits alone on its SIMD and
// its window loop is bound by the latency of each step: the neighbour rows'
// loads, the masks built from them and the accepted-bit lookups.  Here one
// workgroup takes a node: kPipeProducers producer wavefronts prepare the
// windows ahead into a ring of LDS slots, and one consumer wavefront decides
// the windows in order.  For each candidate a producer writes its in-window
// neighbours as a mask and its earlier neighbours as a list, except those in
// windows the consumer has already decided, which it looks up itself: an
// accepted one rejects the candidate at once.  The consumer looks up the
// rest, resolves the window by ballots (as accept_windows) and publishes the
// window's accepted bits.  The 10 001-check test needs no decision and is
// made by the producers.  Lists kept in the pool and overflowed lists are
// walked by the consumer (lane_slow).  Two forms: 16-bit positions and 16
// slots for nodes of at most 65 536 candidates, 32-bit positions and 8 slots
// (and 32 KB of bits) up to 262 144; larger nodes keep their accepted bits in
// global memory and take accept_windows.
constexpr int kPipeProducers = 7;
constexpr uint32_t kPipeRejected = 0x100u;  // slot kind: rejected by a decided window
constexpr uint32_t kPipeMask = 0x200u;      // slot kind: earlier neighbours as a mask
constexpr int kPipeSpan = 512;              // positions a mask covers: the window's 512 before it
constexpr uint32_t kPipeSpin = 1u << 22;   // bound of every wait (an error when reached)
template <bool kWide>
struct PipeCfg {
  // the windows not yet decided when a slot is filled lie within the
  // kPipeSpan positions before it
  static constexpr int kSlots = 8;
  static_assert(64 * (kSlots - 1) <= kPipeSpan, "a mask covers every window not yet decided");
  static constexpr int kWords = kWide ? 8192 : 2048;  // LDS bits: the most candidates / 32
  static constexpr int kVecs = kWide ? 4 : 2;         // 16-byte words of a lane's 16 positions
  // Empty entry of a slot's list: the last position of the bits.  It is
  // never an earlier neighbour (a node has at most kWords * 32 candidates),
  // and its bit is zero whenever it is looked up: every word of the bits is
  // cleared, and a node that has the position decides it in its last
  // window, after every lookup.
  static constexpr uint32_t kNone = (uint32_t)kWords * 32u - 1u;
};
// A slot holds, per lane, the in-window mask and the kind of its list in one
// 16-byte word, and the earlier neighbours left to look up as 16 positions
// (kNone where there is none): the consumer reads a lane's whole slot with
// three or five 16-byte LDS reads, then looks up all 16 bits at once.
template <bool kWide>
struct PipeLds {
  using C = PipeCfg<kWide>;
  uint4 head[C::kSlots][64];  // mask low, mask high, kind (0, kPipeRejected, kPipeMask, kNbrPooled, kNbrOverflow), 0
  // 16 positions (kind 0), or for a list longer than 16 entries (dense
  // clusters) its entries in windows not yet decided as a mask of the
  // kPipeSpan positions before the window (kPipeMask); an overflowed list's
  // row words 0-2 in the first three dwords
  uint4 early[C::kSlots][64][kPipeSpan / 128];
  uint32_t ready[C::kSlots];             // window index + 1 of the slot's contents
  uint32_t consumed;                     // windows decided
  uint32_t abort;
  uint32_t bits[C::kWords];
};

__device__ inline uint32_t lds_acquire(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until *p >= want (wave-uniform); false when the workgroup gave up.
__device__ inline bool pipe_wait(uint32_t* abort, uint32_t* p, uint32_t want) {
  for (uint32_t spin = 0;; spin++) {
    if (lds_acquire(p) >= want) return true;
    if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
    if (spin >= kPipeSpin) {
      __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// entry i of a lane's 16 positions in its dwords
template <bool kWide>
__device__ inline uint32_t pipe_entry(const uint32_t* e, int i) {
  if constexpr (kWide) return e[i];
  else return (i & 1) ? (e[i >> 1] >> 16) : (e[i >> 1] & 0xFFFFu);
}
template <bool kWide>
__device__ inline void pipe_set_entry(uint32_t* e, int i, uint32_t v) {
  if constexpr (kWide) e[i] = v;
  else e[i >> 1] = (i & 1) ? ((e[i >> 1] & 0xFFFFu) | (v << 16)) : ((e[i >> 1] & 0xFFFF0000u) | v);
}

template <bool kWide>
__device__ void pipe_produce(NodeWork* __restrict__ work, int k, const SampleParams& P, const Layout& L,
                             const NodeGeom& g, const int64_t n, PipeLds<kWide>& S, const int pw) {
  using C = PipeCfg<kWide>;
  const int lane = threadIdx.x & 63;
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  const gptr<const uint8_t> gCnt = gbl((const uint8_t*)L.nbrCount);
  const gptr<const double> gSadd = gbl((const double*)L.sadd);
  const gptr<const uint4> gNbr = gbl((const uint4*)L.nbr);
  const int64_t nw = (n + 63) / 64;
  bool cap_may_bind = false;
  for (int64_t j = pw; j < nw; j += kPipeProducers) {
    const int sl = (int)(j % C::kSlots);
    const int64_t w0 = 64 * j;
    const int64_t p = w0 + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : n - 1;
    const int cnt = gCnt[pc];
    const double sa = gSadd[pc];
    const gptr<const uint4> rw = gNbr + (uint64_t)pc * (kMaxNbr / 4);
    const uint4 r0 = ld_g16(rw), r1 = ld_g16(rw + 1), r2 = ld_g16(rw + 2), r3 = ld_g16(rw + 3);
    if (!avg && w0 + 64 > 10001) {
      // the 10 001-check window of the reference scan may bind (the test of
      // accept_windows)
      double limSq = 1e308;
      if (valid) {
        const double lim = __builtin_sqrt(sa) - g.spacing;
        limSq = lim * lim;
      }
      double ml = limSq;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        const double o = __shfl_xor(ml, d);
        ml = o < ml ? o : ml;
      }
      const int64_t hi = (w0 + 64 < n ? w0 + 64 : n);
      const int64_t m = hi - 10001;
      if (m >= 1 && lane == 0 && !(gSadd[m - 1] < ml)) cap_may_bind = true;
    }
    // the slot is free once the window kSlots before is decided; the
    // windows before that one are decided too, and their bits are final
    const int64_t freeAt = j - C::kSlots + 1;
    if (!pipe_wait(&S.abort, &S.consumed, freeAt > 0 ? (uint32_t)freeAt : 0u)) break;
    const int64_t decidedEnd = 64 * (int64_t)lds_acquire(&S.consumed);  // candidates of decided windows
    uint32_t info = 0;
    uint64_t in_mask = 0;
    uint32_t e[4 * C::kVecs];
#pragma unroll
    for (int i = 0; i < 4 * C::kVecs; i++) e[i] = kWide ? C::kNone : (C::kNone | (C::kNone << 16));
    if (valid) {
      if (cnt != kNbrOverflow && cnt != kNbrPooled) {
        const uint32_t row[16] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w,
                                  r2.x, r2.y, r2.z, r2.w, r3.x, r3.y, r3.z, r3.w};
        uint32_t hit = 0;
#pragma unroll
        for (int i = 0; i < kMaxNbr; i++) {
          const uint32_t q = row[i];
          const bool use = i < cnt;
          if (use && (int64_t)q >= w0) in_mask |= 1ull << (q - w0);
          const bool old = use && (int64_t)q < decidedEnd;
          // every lookup is made (kNone's bit is zero) so that none waits
          // alone in a branch
          const uint32_t qo = old ? q : C::kNone;
          hit |= (S.bits[qo >> 5] >> (qo & 31)) & 1u;
          pipe_set_entry<kWide>(e, i, (use && !old && (int64_t)q < w0) ? q : C::kNone);
        }
        info = hit ? kPipeRejected : 0u;
      } else if (cnt == kNbrPooled) {
        // a long list (dense clusters): walked here, four entries in flight
        // per step, like a row.  Its entries in windows not yet decided go
        // to the lane's mask (they all lie in the kPipeSpan positions before
        // the window: the windows before kSlots - 1 ago are decided).
        const gptr<const uint32_t> lst = gbl((const uint32_t*)P.nbr_pool) + ((uint64_t)r0.x | ((uint64_t)r0.y << 32));
        const int len = (int)r0.z;
        const int64_t base = w0 - kPipeSpan > 0 ? w0 - kPipeSpan : 0;
        uint4* const sp = S.early[sl][lane];
#pragma unroll
        for (int v = 0; v < kPipeSpan / 128; v++) sp[v] = make_uint4(0u, 0u, 0u, 0u);
        uint32_t hit = 0;
        for (int i0 = 0; i0 < len && !hit; i0 += 4) {
          uint32_t q4[4];
#pragma unroll
          for (int u = 0; u < 4; u++) q4[u] = lst[i0 + u < len ? i0 + u : len - 1];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const uint32_t q = q4[u];
            const bool use = i0 + u < len;
            if (use && (int64_t)q >= w0) in_mask |= 1ull << (q - w0);
            const bool old = use && (int64_t)q < decidedEnd;
            const uint32_t qo = old ? q : C::kNone;
            hit |= (S.bits[qo >> 5] >> (qo & 31)) & 1u;
            if (use && !old && (int64_t)q < w0) {
              const uint32_t r = (uint32_t)(q - base);
              atomicOr(&((uint32_t*)sp)[r >> 5], 1u << (r & 31));
            }
          }
        }
        info = hit ? kPipeRejected : kPipeMask;
      } else {
        info = (uint32_t)cnt;
        e[0] = r0.x;
        e[1] = r0.y;
        e[2] = r0.z;
      }
    }
    S.head[sl][lane] = make_uint4((uint32_t)in_mask, (uint32_t)(in_mask >> 32), info, 0u);
    if (info != kPipeMask)
#pragma unroll
      for (int v = 0; v < C::kVecs; v++) S.early[sl][lane][v] = make_uint4(e[4 * v], e[4 * v + 1], e[4 * v + 2], e[4 * v + 3]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) lds_release(&S.ready[sl], (uint32_t)(j + 1));
  }
  if (cap_may_bind) atomicOr(&work[k].flags, (uint32_t)NW_SEQ_ACCEPT);
}

template <bool kWide>
__device__ void pipe_consume(const SampleParams& P, const Layout& L, const NodeGeom& g, Filter& F,
                             const int64_t n, PipeLds<kWide>& S) {
  using C = PipeCfg<kWide>;
  const int lane = threadIdx.x & 63;
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  const int64_t nw = (n + 63) / 64;
  uint32_t* bits = S.bits;
  for (int64_t j = 0; j < nw; j++) {
    const int sl = (int)(j % C::kSlots);
    if (!pipe_wait(&S.abort, &S.ready[sl], (uint32_t)(j + 1))) break;
    const int64_t w0 = 64 * j;
    const int64_t p = w0 + lane;
    const bool valid = p < n;
    const uint4 h = S.head[sl][lane];
    uint32_t ew[4 * C::kVecs];
#pragma unroll
    for (int v = 0; v < C::kVecs; v++) {
      const uint4 t = S.early[sl][lane][v];
      ew[4 * v] = t.x; ew[4 * v + 1] = t.y; ew[4 * v + 2] = t.z; ew[4 * v + 3] = t.w;
    }
    const uint32_t info = h.z;
    uint64_t in_mask = ((uint64_t)h.y << 32) | h.x;
    bool rej = false;
    if (valid) {
      if (info == kPipeRejected) {
        rej = true;
      } else if (info == kPipeMask) {
        const uint4* sp = S.early[sl][lane];
        const uint32_t* bw = bits + ((w0 - kPipeSpan > 0 ? w0 - kPipeSpan : 0) >> 5);
        uint32_t hit = 0;
#pragma unroll
        for (int v = 0; v < kPipeSpan / 128; v++) {
          const uint4 m = sp[v];
          hit |= (m.x & bw[4 * v]) | (m.y & bw[4 * v + 1]) | (m.z & bw[4 * v + 2]) | (m.w & bw[4 * v + 3]);
        }
        rej = hit != 0;
      } else if (info == 0u) {
        // all 16 lookups issued together: an empty entry looks up kNone's
        // bit, which is zero
        uint32_t hit = 0;
#pragma unroll
        for (int i = 0; i < kMaxNbr; i++) {
          const uint32_t q = pipe_entry<kWide>(ew, i);
          hit |= (bits[q >> 5] >> (q & 31)) & 1u;
        }
        rej = hit != 0;
      } else {
        double limSq = 1e308;
        if (!avg) {
          const double lim = __builtin_sqrt(L.sadd[p]) - g.spacing;
          limSq = lim * lim;
        }
        lane_slow(P, L, g, F, avg, p, w0, (int)info, ew[0], ew[1], ew[2], limSq, bits, in_mask, rej);
      }
    }
    // in-window dependencies: the lowest undecided lane always resolves
    bool decided = !valid || rej;
    bool accepted = false;
    uint64_t dec = __ballot(decided), accm = 0;
    while (dec != ~0ull) {
      if (!decided) {
        if (in_mask & accm) {
          decided = true;
        } else if ((in_mask & ~dec) == 0) {
          decided = true;
          accepted = true;
        }
      }
      dec = __ballot(decided);
      accm = __ballot(accepted);
    }
    if (lane == 0) {
      bits[2 * j] = (uint32_t)accm;
      bits[2 * j + 1] = (uint32_t)(accm >> 32);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) lds_release(&S.consumed, (uint32_t)(j + 1));
  }
}

// The cell-path nodes of a launch for the pipelined acceptance, in two
// lists in `list` (num_nodes entries): the nodes of at most 65 536
// candidates from the front, the larger ones from the back; counts[0] and
// counts[1] their lengths.  The pipelined kernels then take a grid sized to
// the GPU over their list instead of a workgroup (and its LDS) per node of
// the launch: in a launch the ring decides, most nodes are not theirs.
__global__ __launch_bounds__(1024) void k_pipe_select(const NodeWork* __restrict__ work, int32_t num_nodes,
                                                      int32_t* __restrict__ list, uint32_t* __restrict__ counts,
                                                      int32_t form) {
  __shared__ uint32_t s_w[2][16];
  __shared__ uint32_t s_run[2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  constexpr int64_t kNarrowMax = (int64_t)PipeCfg<false>::kWords * 32;
  if (tid < 2) s_run[tid] = 0;
  __syncthreads();
  for (int b = 0; b < num_nodes; b += 1024) {
    const int k = b + tid;
    bool narrow = false, wide = false;
    if (k < num_nodes) {
      const uint32_t f = work[k].flags;
      if ((f & (NW_OVERFLOW | NW_CELL_PATH)) == NW_CELL_PATH) {
        narrow = form == 0 && work[k].n <= kNarrowMax;
        wide = !narrow;
      }
    }
    const uint64_t bn = __ballot(narrow), bw = __ballot(wide);
    if (lane == 0) { s_w[0][wave] = (uint32_t)__popcll(bn); s_w[1][wave] = (uint32_t)__popcll(bw); }
    __syncthreads();
    uint32_t on = s_run[0], ow = s_run[1], tn = 0, tw = 0;
    for (int q = 0; q < 16; q++) {
      if (q < wave) { on += s_w[0][q]; ow += s_w[1][q]; }
      tn += s_w[0][q];
      tw += s_w[1][q];
    }
    if (narrow) list[on + __popcll(bn & lt)] = k;
    if (wide) list[num_nodes - 1 - (int)(ow + __popcll(bw & lt))] = k;
    __syncthreads();
    if (tid == 0) { s_run[0] += tn; s_run[1] += tw; }
    __syncthreads();
  }
  if (tid == 0) { counts[0] = s_run[0]; counts[1] = s_run[1]; }
}

// kWide == false: the nodes of k_pipe_select's first list (at most 65 536
// candidates); kWide: its second list, with accept_windows past 262 144.
// Each workgroup takes the list's entries blockIdx.x, + gridDim.x, ...
template <bool kWide>
__global__ __launch_bounds__(64 * (1 + kPipeProducers)) void k_accept_pipe(const int32_t* __restrict__ nodes,
                                                                          const NodeInfo* __restrict__ infos,
                                                                          NodeWork* __restrict__ work,
                                                                          const int32_t* __restrict__ list,
                                                                          const uint32_t* __restrict__ counts,
                                                                          int32_t num_nodes, SampleParams P) {
  using C = PipeCfg<kWide>;
  __shared__ PipeLds<kWide> S;
  const uint32_t count = counts[kWide ? 1 : 0];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const int k = kWide ? list[num_nodes - 1 - (int)e] : list[e];
    const NodeWork w = work[k];
    const int64_t n = w.n;
    const Layout L = make_layout(w.scratch, n, P.method);
    const NodeGeom g = node_geom(infos[nodes[k]], P);
    if (n > (int64_t)C::kWords * 32 || P.pipe_form == 2) {
      if (wave == 0) {
        Filter F = make_filter(g, P);
        accept_windows<3>(work, k, P, L, g, F, n, gbl(L.keyA));
        flush_filter(F, P);
      }
      continue;
    }
    // every word, also past the node's: kNone's bit must read zero
    for (int i = threadIdx.x; i < C::kWords; i += blockDim.x) S.bits[i] = 0;
    if (threadIdx.x < C::kSlots) S.ready[threadIdx.x] = 0;
    if (threadIdx.x == 0) { S.consumed = 0; S.abort = 0; }
    __syncthreads();
    Filter F = make_filter(g, P);
    if (wave == 0) pipe_consume<kWide>(P, L, g, F, n, S);
    else pipe_produce<kWide>(work, k, P, L, g, n, S, wave - 1);
    __syncthreads();
    if (S.abort) {
      if (threadIdx.x == 0) atomicAdd(&P.counters[9], 1ull);
    } else {
      const gptr<uint8_t> gAcc = gbl(L.acc);
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        gAcc[i] = (uint8_t)((S.bits[i >> 5] >> (i & 31)) & 1u);
      if (wave == 0) flush_filter(F, P);
    }
    __syncthreads();  // the next node reuses S
  }
}

// kAccAhead windows of neighbour rows in flight: 3 for launches of few
// nodes (each wavefront alone on its SIMD, bound by latency), 1 for wide
// launches (more wavefronts per SIMD hide the latency instead)
template <int kAccAhead>
__global__ __launch_bounds__(kBlock) void k_accept(const int32_t* __restrict__ nodes, const NodeInfo* __restrict__ infos,
                                                    NodeWork* __restrict__ work, int32_t num_nodes, SampleParams P) {
  __shared__ uint32_t s_bits[kWaves][kAccWords];
  const int wave = threadIdx.x >> 6;
  const int k = __builtin_amdgcn_readfirstlane(blockIdx.x * kWaves + wave);
  if (k >= num_nodes) return;
  accept_node<kAccAhead>(nodes, infos, work, k, P, s_bits[wave], kAccWords);
}

// ------------------------------------------------------ accepted shell ring
// Poisson acceptance in one pass, one wavefront per node, 64 sorted
// candidates per window.  A candidate c can only be rejected by an accepted
// point q with add(q) >= limSq(c) = (sqrt(add(c)) - spacing)^2
// (sampler_poisson.h:147-170).  Candidates come in ascending distance, so
// once the window's first candidate lies beyond spacing, accepted points
// below that window's (cd - spacing)^2 can never reject again: the live
// accepted points form a shell of width ~spacing that a ring buffer in LDS
// holds, chained per 2*spacing cell (newest first).  The 10 001-check limit
// of the reference scan cannot bind while the shell fits the ring.  A node
// whose shell outgrows the ring is handed to the cell-table path.
// Distances are filtered in FP32 on coordinates relative to the node
// centre (error below 1e-4 * spacing^2); pairs within the +-1e-3 band are
// re-checked in FP64 with the reference expression on the integer
// coordinates.
constexpr int kRingCap = 512;     // accepted points in the shell (power of two)
constexpr int kRingSlots = 1024;  // cell chains (direct-mapped, power of two)
constexpr int kRingWaves = 4;     // wavefronts per k_ring block
constexpr int16_t kRingEnd = INT16_MIN;

// 13 KB per wavefront, so that three blocks of four fit the 160 KB of a CU.
// `add` holds the squared distance rounded up to FP32: it only decides that
// an entry is too close to the centre to reject (add < limit), and rounding
// up keeps every such decision true of the exact value.  `next` is the
// entry's index minus the next one in its chain (a chain older than 32 767
// entries is dead: the ring holds 512).  Each entry's sorted position, read
// only for the FP64 re-check, is kept in the node's scratch.
struct RingLds {
  float x[kRingCap], y[kRingCap], z[kRingCap];
  float add[kRingCap];
  int16_t next[kRingCap];
  int32_t head[kRingSlots];
};
static_assert(3 * kRingWaves * sizeof(RingLds) <= 160 * 1024, "three k_ring blocks per CU");

__device__ inline float round_up_f32(double a) {
  float f = (float)a;
  if ((double)f < a) f = __uint_as_float(__float_as_uint(f) + 1u);  // a >= 0
  return f;
}

struct RingGeom {
  float off, inv_cs;  // cell = floor((v + off) * inv_cs), v relative to the centre
  float s2f, band_lo, band_hi;
  double s2;
};

__device__ inline int ring_cell(float v, const RingGeom& G) {
  int c = (int)__builtin_floorf((v + G.off) * G.inv_cs);
  return c < 0 ? 0 : (c > 127 ? 127 : c);
}

__device__ inline uint32_t ring_slot(int cx, int cy, int cz) {
  return ((uint32_t)(cx | (cy << 7) | (cz << 14)) * 2654435761u) >> (32 - 10);  // 10 bits: kRingSlots
}
static_assert(kRingSlots == 1024, "ring_slot hashes cells to 10 bits");

__device__ void ring_node(const int32_t* __restrict__ nodes, const NodeInfo* __restrict__ infos,
                          NodeWork* __restrict__ work, const int k, const SampleParams& P, RingLds& R) {
  const int lane = threadIdx.x & 63;
  const NodeWork& w = work[k];
  if (w.flags & (NW_OVERFLOW | NW_CELL_PATH)) return;
  const int64_t n = w.n;
  const Layout L = make_layout(w.scratch, n, P.method);
  const NodeGeom g = node_geom(infos[nodes[k]], P);
  RingGeom G;
  G.inv_cs = (float)(1.0 / (2.0 * g.spacing));
  G.off = (float)(66.0 * g.spacing);
  G.s2 = g.s2;
  G.s2f = (float)g.s2;
  G.band_lo = G.s2f * 0.999f;
  G.band_hi = G.s2f * 1.001f;
  const float sq = (float)(g.spacing * (1.0 + 1e-4));
  unsigned long long rechecks = 0, tests = 0;
  float maxerr = 0.f;
  const gptr<int32_t> rpos = gbl((int32_t*)L.keyB);  // sorted position of each ring entry
  for (int i = lane; i < kRingSlots; i += 64) R.head[i] = -1;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int32_t total = 0, tail = 0;  // accepted points appended / first live one (wave-uniform)
  bool overflow = false;
  // exact "dd < spacing^2" for a pair within the FP32 band
  auto exact = [&](int32_t qX, int32_t qY, int32_t qZ, int32_t X, int32_t Y, int32_t Z, float d32) {
    rechecks++;
    const double d64 = dd64_xyz(qX, qY, qZ, X, Y, Z, P);
    const float e = (float)(__builtin_fabs((double)d32 - d64) / G.s2);
    maxerr = e > maxerr ? e : maxerr;
    return d64 < G.s2;
  };
  if (n <= 0) return;
  // the node's arrays through global-address pointers (a flat load would
  // also count as an LDS operation, and every wait of the chain walks would
  // wait for the next windows' loads too); every lane loads, from a clamped
  // position past the node (values unused), so the loads take no branch
  const gptr<const uint32_t> gIdx = gbl((const uint32_t*)L.idxA);
  const gptr<const int4> gXyz = gbl((const int4*)L.xyz);
  const gptr<const double> gSadd = gbl((const double*)L.sadd);
  auto clampn = [n](int64_t q) { return q < n ? q : n - 1; };
  int32_t X, Y, Z;
  double a;
  {
    const uint32_t j = gIdx[clampn(lane)];
    const int4 q = ld_g16i(gXyz + j);
    X = q.x; Y = q.y; Z = q.z;
    a = gSadd[clampn(lane)];
  }
  // the next window's candidates are loaded while this one is decided; the
  // window after it has its main indices loaded already, so no load waits
  // on another load of the same window
  uint32_t jn = gIdx[clampn(lane + 64)];
  for (int64_t w0 = 0; w0 < n; w0 += 64) {
    const int64_t p = w0 + lane;
    const bool valid = p < n;
    int32_t nX, nY, nZ;
    double na;
    {
      const int4 q = ld_g16i(gXyz + jn);
      nX = q.x; nY = q.y; nZ = q.z;
      na = gSadd[clampn(p + 64)];
    }
    jn = gIdx[clampn(p + 128)];
    const double cd = __builtin_sqrt(a);
    const double lim = cd - g.spacing;
    const double limSq = lim * lim;
    const float xf = (float)(coord(X, P.scale[0], P.offset[0]) - g.cx);
    const float yf = (float)(coord(Y, P.scale[1], P.offset[1]) - g.cy);
    const float zf = (float)(coord(Z, P.scale[2], P.offset[2]) - g.cz);
    // evict accepted points that no current or later candidate can see
    const double lim0 = __shfl(lim, 0);
    const double thr = lim0 >= 0.0 ? lim0 * lim0 : 0.0;
    while (tail < total) {
      const int32_t idx = tail + lane;
      const bool dead = idx < total && (double)R.add[idx & (kRingCap - 1)] < thr;
      const int d = __popcll(__ballot(dead));
      tail += d;
      if (d < 64) break;
    }
    bool rej = false;
    uint64_t in_mask = 0;
    if (valid) {
      // the cells around the candidate, their chains walked side by side
      const int lx = ring_cell(xf - sq, G), hx = ring_cell(xf + sq, G);
      const int ly = ring_cell(yf - sq, G), hy = ring_cell(yf + sq, G);
      const int lz = ring_cell(zf - sq, G), hz = ring_cell(zf + sq, G);
      const int nx = hx - lx + 1, nxy = nx * (hy - ly + 1);
      const int ncell = nxy * (hz - lz + 1);
      for (int base = 0; base < ncell && !rej; base += 8) {
        int32_t cur[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
          const int ci = base + t;
          cur[t] = -1;
          if (ci < ncell) {
            const int cz = lz + ci / nxy, rem = ci % nxy;
            cur[t] = R.head[ring_slot(lx + rem % nx, ly + rem / nx, cz)];
          }
        }
        bool any = true;
        while (any && !rej) {
          any = false;
#pragma unroll
          for (int t = 0; t < 8; t++) {
            const int32_t idx = cur[t];
            if (idx < 0 || idx < tail - 64) continue;
            any = true;
            const int e = idx & (kRingCap - 1);
            const int16_t nx = R.next[e];
            cur[t] = nx == kRingEnd ? -1 : idx - nx;
            // chains are newest first up to reordering inside one window,
            // so nothing live follows an entry older than tail - 64
            if (idx < tail) continue;
            tests++;
            const float dx = R.x[e] - xf, dy = R.y[e] - yf, dz = R.z[e] - zf;
            const float d32 = dx * dx + dy * dy + dz * dz;
            if (d32 > G.band_hi) continue;
            // an entry nearer the centre than limSq cannot reject (checked
            // only for the few entries inside the band's upper edge)
            if ((double)R.add[e] < limSq) continue;
            bool within = d32 < G.band_lo;
            if (!within) {
              const uint32_t jq = gIdx[rpos[e]];
              const int4 q = ld_g16i(gXyz + jq);
              within = exact(q.x, q.y, q.z, X, Y, Z, d32);
            }
            if (within) rej = true;
          }
        }
      }
    }
    // earlier candidates of the window, broadcast from their lanes: a pass
    // sorts the pairs into clearly within spacing and inside the FP32 band
    // (the band pairs get the FP64 check afterwards).  Only candidates that
    // the ring did not reject are broadcast: a rejected one is never
    // accepted, so it cannot exclude a later one.
    uint64_t m_band = 0;
    const uint64_t surv = __ballot(valid && !rej) & ~(1ull << 63);
    for (uint64_t mm = surv; mm; mm &= mm - 1) {
      const int q = __builtin_amdgcn_readfirstlane(__builtin_ctzll(mm));
      const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xf), q));
      const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(yf), q));
      const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(zf), q));
      const float dx = qx - xf, dy = qy - yf, dz = qz - zf;
      const float d32 = dx * dx + dy * dy + dz * dz;
      // most earlier candidates of a window lie far away: the rest of the
      // test only where some later lane has q within the band's upper edge
      const bool nearq = lane > q && !(d32 > G.band_hi);
      if (__ballot(nearq) == 0) continue;
      const double qa = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(a), q),
                                         __builtin_amdgcn_readlane(__double2loint(a), q));
      const bool cand = nearq && !(qa < limSq);
      const uint64_t bit = 1ull << q;
      in_mask |= (cand && d32 < G.band_lo) ? bit : 0ull;
      m_band |= (cand && !(d32 < G.band_lo)) ? bit : 0ull;
    }
    if (!valid || rej) { in_mask = 0; m_band = 0; }
    tests += (uint64_t)__popcll(surv & lt);
    uint64_t band_any = 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const uint64_t o = (uint64_t)__shfl_xor((unsigned long long)(m_band | band_any), d);
      band_any |= o | m_band;
    }
    while (band_any) {
      const int q = __builtin_ctzll(band_any);
      band_any &= band_any - 1;
      const int32_t qX = __builtin_amdgcn_readlane(X, q), qY = __builtin_amdgcn_readlane(Y, q),
                    qZ = __builtin_amdgcn_readlane(Z, q);
      const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xf), q));
      const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(yf), q));
      const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(zf), q));
      if ((m_band >> q) & 1ull) {
        const float dx = qx - xf, dy = qy - yf, dz = qz - zf;
        if (exact(qX, qY, qZ, X, Y, Z, dx * dx + dy * dy + dz * dz)) in_mask |= 1ull << q;
      }
    }
    bool decided = !valid || rej;
    bool accepted = false;
    uint64_t dec = __ballot(decided), accm = 0;
    while (dec != ~0ull) {
      if (!decided) {
        if (in_mask & accm) {
          decided = true;
        } else if ((in_mask & ~dec) == 0) {
          decided = true;
          accepted = true;
        }
      }
      dec = __ballot(decided);
      accm = __ballot(accepted);
    }
    if (valid) gbl(L.acc)[p] = accepted ? 1 : 0;
    // append the window's accepted points; entries back to tail - 64 must
    // stay intact for the chain walks
    const int32_t kacc = __popcll(accm);
    const int32_t keep = tail - 64 > 0 ? tail - 64 : 0;
    if (total + kacc - keep > kRingCap) {
      overflow = true;
      break;
    }
    if (accepted) {
      const int32_t idx = total + __popcll(accm & lt);
      const int e = idx & (kRingCap - 1);
      R.x[e] = xf; R.y[e] = yf; R.z[e] = zf; R.add[e] = round_up_f32(a);
      rpos[e] = (int32_t)p;
      const int32_t prev = atomicExch(&R.head[ring_slot(ring_cell(xf, G), ring_cell(yf, G), ring_cell(zf, G))], idx);
      // inside one window a chain may point at a newer entry (at most 63 on)
      R.next[e] = (prev < 0 || idx - prev > 32767) ? kRingEnd : (int16_t)(idx - prev);
    }
    total += kacc;
    X = nX; Y = nY; Z = nZ; a = na;
  }
  if (overflow && lane == 0) atomicOr(&work[k].flags, (uint32_t)NW_CELL_PATH);
  Filter F;
  F.rechecks = rechecks;
  F.tests = tests;
  F.maxerr = maxerr;
  flush_filter(F, P);
}

// Work queue of the ring: the launch's nodes that the ring decides, largest
// first by power-of-two classes of their candidate counts, so that the
// longest nodes start first.  ctl[0] = queue length, ctl[1] = next ticket.
__global__ __launch_bounds__(1024) void k_ring_queue(const NodeWork* __restrict__ work, int32_t num_nodes,
                                                     int32_t* __restrict__ queue, uint32_t* __restrict__ ctl) {
  __shared__ uint32_t s_cnt[33], s_base[33];
  const int tid = threadIdx.x;
  if (tid < 33) s_cnt[tid] = 0;
  __syncthreads();
  for (int k = tid; k < num_nodes; k += 1024) {
    const NodeWork& w = work[k];
    if (w.flags & (NW_OVERFLOW | NW_CELL_PATH)) continue;
    const int c = w.n > 0 ? 64 - __builtin_clzll((unsigned long long)w.n) : 0;
    atomicAdd(&s_cnt[c < 32 ? c : 32], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int c = 32; c >= 0; c--) {
      s_base[c] = run;
      run += s_cnt[c];
    }
    ctl[0] = run;
    ctl[1] = 0;
  }
  __syncthreads();
  for (int k = tid; k < num_nodes; k += 1024) {
    const NodeWork& w = work[k];
    if (w.flags & (NW_OVERFLOW | NW_CELL_PATH)) continue;
    const int c = w.n > 0 ? 64 - __builtin_clzll((unsigned long long)w.n) : 0;
    queue[atomicAdd(&s_base[c < 32 ? c : 32], 1u)] = k;
  }
}

// Persistent form: every wavefront takes the next node of the queue until
// the queue is drained, so a wavefront that finishes a small node starts
// another while a large one is still running next to it.
__global__ __launch_bounds__(64 * kRingWaves) void k_ring(const int32_t* __restrict__ nodes,
                                                           const NodeInfo* __restrict__ infos,
                                                           NodeWork* __restrict__ work,
                                                           const int32_t* __restrict__ queue,
                                                           uint32_t* __restrict__ ctl, SampleParams P) {
  __shared__ RingLds s_ring[kRingWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t len = ctl[0];
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&ctl[1], 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= len) break;
    ring_node(nodes, infos, work, __builtin_amdgcn_readfirstlane(queue[t]), P, s_ring[wave]);
  }
}

// Accepted entries of a Poisson-average node in cell order, with pre[e] =
// accepted entries before entry e (n + 1 values), so that the colour pass
// scans only the accepted points of a cell.  The arrays take the neighbour
// list space, which nothing reads after k_accept.
__device__ inline void avg_accepted_arrays(const Layout& L, int64_t n, int4*& aent, uint32_t*& aC16,
                                           uint32_t*& pre) {
  aent = (int4*)L.nbr;
  aC16 = (uint32_t*)(aent + n);
  pre = aC16 + n;
}

__global__ __launch_bounds__(kBlock) void k_avg_compact(NodeWork* __restrict__ work, SampleParams P) {
  __shared__ uint32_t s_w[4][kWaves];
  const int k = blockIdx.x;
  const NodeWork w = work[k];
  if (w.flags & NW_OVERFLOW) return;
  const int64_t n = w.n;
  const Layout L = make_layout(w.scratch, n, P.method);
  int4* aent;
  uint32_t *aC16, *pre;
  avg_accepted_arrays(L, n, aent, aC16, pre);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t run = 0;
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    int4 o[4];
    bool a[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t e = b0 + u * kBlock + tid;
      o[u] = e < n ? L.ent[e] : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) a[u] = b0 + u * kBlock + tid < n && L.acc[(uint32_t)o[u].w];
    uint64_t bal[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      bal[u] = __ballot(a[u]);
      if (lane == 0) s_w[u][wave] = (uint32_t)__popcll(bal[u]);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t e = b0 + u * kBlock + tid;
      uint32_t woff = 0, tot = 0;
      for (int v = 0; v < kWaves; v++) { if (v < wave) woff += s_w[u][v]; tot += s_w[u][v]; }
      const uint32_t off = run + woff + (uint32_t)__popcll(bal[u] & lt);
      if (e < n) {
        pre[e] = off;
        if (a[u]) {
          aent[off] = o[u];
          aC16[off] = L.eC16[e];
        }
      }
      run += tot;
    }
    __syncthreads();
  }
  if (tid == 0) pre[n] = run;
}

// Poisson-average colour sums: every rejected candidate C adds its own colour
// (weight 1) to each accepted P with P inside C's 16^3 range and dd < s^2,
// into P's "before" or "after" slots by sorted rank.  With accepted points
// pairwise outside each other's range the reference's sequential pass
// (sampler_poisson_average.h:369-415) then yields
//   P = 2 * (own + before) + after.
__global__ __launch_bounds__(kBlock) void k_avg_colors(const int32_t* __restrict__ nodes,
                                                        const NodeInfo* __restrict__ infos,
                                                        NodeWork* __restrict__ work, SampleParams P) {
  const int k = blockIdx.x;
  const NodeWork w = work[k];
  if (w.flags & NW_OVERFLOW) return;
  const int64_t n = w.n;
  const Layout L = make_layout(w.scratch, n, P.method);
  const NodeGeom g = node_geom(infos[nodes[k]], P);
  bool violation = false;
  // a node's candidates are split over gridDim.y workgroups (the sums are
  // atomic), so that narrow levels fill the GPU
  const int64_t per = (n + gridDim.y - 1) / gridDim.y;
  const int64_t e0 = (int64_t)blockIdx.y * per, e1 = e0 + per < n ? e0 + per : n;
  int4* aent;
  uint32_t *aC16, *pre;
  avg_accepted_arrays(L, n, aent, aC16, pre);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
    const int4 me = L.ent[e];
    const uint32_t p = (uint32_t)me.w;
    const int32_t X = me.x, Y = me.y, Z = me.z;
    const double x = coord(X, P.scale[0], P.offset[0]), y = coord(Y, P.scale[1], P.offset[1]),
                 z = coord(Z, P.scale[2], P.offset[2]);
    int lo16[3], hi16[3];
    range16(g, x, y, z, lo16, hi16);
    const CellBox cb = cell_box(g, x, y, z);
    const bool self_acc = L.acc[p];
    const uint32_t j = L.idxA[p];
    const long long r = L.rgb[3 * j], gg = L.rgb[3 * j + 1], b = L.rgb[3 * j + 2];
    for (int cz = cb.lz; cz <= cb.hz; cz++)
      for (int cy = cb.ly; cy <= cb.hy; cy++)
        for (int cx = cb.lx; cx <= cb.hx; cx++) {
          const uint32_t slot = table_find(L, gkey(cx, cy, cz));
          if (slot == kNil) continue;
          const uint4 t = L.htab[slot];
          const uint32_t f0 = pre[t.y], f1 = pre[t.z];
          for (uint32_t fb = f0; fb < f1; fb += 4) {
            int4 o4[4];
            uint32_t c4[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              o4[u] = fb + u < f1 ? aent[fb + u] : make_int4(0, 0, 0, (int32_t)p);
              c4[u] = fb + u < f1 ? aC16[fb + u] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const int4 o = o4[u];
              const uint32_t q = (uint32_t)o.w;
              if (q == p) continue;
              if (!in_range16(c4[u], lo16, hi16)) continue;
              if (!(dd64_xyz(o.x, o.y, o.z, X, Y, Z, P) < g.s2)) continue;
              if (self_acc) { violation = true; continue; }
              unsigned long long* d = (unsigned long long*)(L.sums + 8 * (uint64_t)q + (p < q ? 0 : 4));
              atomicAdd(&d[0], (unsigned long long)r);
              atomicAdd(&d[1], (unsigned long long)gg);
              atomicAdd(&d[2], (unsigned long long)b);
              atomicAdd(&d[3], 1ull);
            }
          }
        }
  }
  if (violation) atomicOr(&work[k].flags, (uint32_t)NW_SEQ_AVERAGE);
}

// Poisson acceptance with the 10 001-check limit of the reference scan
// (sampler_poisson.h:147-176), for nodes where it can bind; one wavefront.
// The scan visits the accepted points newest first and stops at the first
// one closer to the centre than (cd - spacing)^2.  Accepted points come in
// ascending distance, so an accepted neighbour q within spacing (the
// neighbour lists hold exactly those not behind that stop) rejects
// candidate p iff the scan reaches it inside its 10 001 checks:
//   na(p) - 1 - accIndex(q) <= 10 000,
// where na(p) is the number accepted before p and accIndex(q) the number
// accepted before q.  accIndex comes from the accepted bits and a prefix
// count per 32-candidate word; in-window neighbours are always inside the
// limit; a candidate whose latest earlier neighbour sits at the edge of
// the limit waits until every lower lane of its window is decided.
__device__ void accept_node_capped(const Layout& L, const NodeGeom& g, int64_t n, const SampleParams& P) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t* bits = L.keyA;  // accepted flags by sorted position
  uint32_t* pre = L.keyB;   // accepted candidates before each word of `bits`
  const int64_t words = (n + 31) / 32;
  for (int64_t i = lane; i < words; i += 64) { bits[i] = 0; pre[i] = 0; }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
  Filter F = make_filter(g, P);
  int64_t A0 = 0;  // accepted before the window
  for (int64_t w0 = 0; w0 < n; w0 += 64) {
    const int64_t p = w0 + lane;
    const bool valid = p < n;
    uint64_t in_mask = 0;
    int64_t amax = -1;  // largest accIndex of an accepted earlier-window neighbour
    auto consider = [&](uint32_t q) {
      if ((int64_t)q >= w0) {
        in_mask |= 1ull << (q - w0);
        return;
      }
      const uint32_t wd = bits[q >> 5];
      if (!((wd >> (q & 31)) & 1u)) return;
      const int64_t ai = (int64_t)pre[q >> 5] + __popc(wd & ((1u << (q & 31)) - 1u));
      amax = ai > amax ? ai : amax;
    };
    if (valid) {
      const int cnt = L.nbrCount[p];
      if (cnt != kNbrOverflow && cnt != kNbrPooled) {
        const uint32_t* row = L.nbr + (uint64_t)p * kMaxNbr;
        for (int i = 0; i < cnt; i++) consider(row[i]);
      } else if (cnt == kNbrPooled) {
        const uint32_t* rw = L.nbr + (uint64_t)p * kMaxNbr;
        const uint32_t* lst = P.nbr_pool + ((uint64_t)rw[0] | ((uint64_t)rw[1] << 32));
        const uint32_t len = rw[2];
        for (uint32_t i = 0; i < len; i++) consider(lst[i]);
      } else {
        const double cd = __builtin_sqrt(L.sadd[p]);
        const double lim = cd - g.spacing;
        const double limSq = lim * lim;
        const uint32_t j = L.idxA[p];
        const int4 q = L.xyz[j];
        const int32_t X = q.x, Y = q.y, Z = q.z;
        const double x = coord(X, P.scale[0], P.offset[0]), y = coord(Y, P.scale[1], P.offset[1]),
                     z = coord(Z, P.scale[2], P.offset[2]);
        const CellBox cb = cell_box(g, x, y, z);
        for (int cz = cb.lz; cz <= cb.hz; cz++)
          for (int cy = cb.ly; cy <= cb.hy; cy++)
            for (int cx = cb.lx; cx <= cb.hx; cx++) {
              const uint32_t slot = table_find(L, gkey(cx, cy, cz));
              if (slot == kNil) continue;
              const uint4 t = L.htab[slot];
              for (uint32_t f = t.y; f < t.z; f++) {
                const int4 o = L.ent[f];
                const uint32_t q = (uint32_t)o.w;
                if ((int64_t)q >= p) break;
                if (L.eAdd[f] < limSq) continue;
                if (!within_xyz(o.x, o.y, o.z, X, Y, Z, P, F)) continue;
                consider(q);
              }
            }
      }
    }
    // na(p) lies in [A0, A0 + lane]
    const bool rej = valid && amax >= 0 && amax >= A0 + lane - 10001;
    const bool edge = valid && !rej && amax >= 0 && amax >= A0 - 10001;
    bool decided = !valid || rej;
    bool accepted = false;
    uint64_t dec = __ballot(decided), accm = 0;
    while (dec != ~0ull) {
      if (!decided) {
        if (in_mask & accm) {
          decided = true;
        } else if ((in_mask & ~dec) == 0) {
          if (!edge) {
            decided = true;
            accepted = true;
          } else if ((~dec & lt) == 0) {
            const int64_t na = A0 + __popcll(accm & lt);
            decided = true;
            accepted = !(amax >= na - 10001);
          }
        }
      }
      dec = __ballot(decided);
      accm = __ballot(accepted);
    }
    if (valid) L.acc[p] = accepted ? 1 : 0;
    if (lane == 0) {
      const int64_t wd = w0 >> 5;
      bits[wd] = (uint32_t)accm;
      pre[wd] = (uint32_t)A0;
      if (wd + 1 < words) {
        bits[wd + 1] = (uint32_t)(accm >> 32);
        pre[wd + 1] = (uint32_t)(A0 + __popc((uint32_t)accm));
      }
    }
    A0 += __popcll(accm);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
  }
  flush_filter(F, P);
}

// Exact forms for the nodes flagged above: the capped Poisson scan by one
// wavefront, the Poisson-average colour pass by one lane.
__global__ void k_sequential(const int32_t* __restrict__ nodes, const NodeInfo* __restrict__ infos,
                             NodeWork* __restrict__ work, SampleParams P) {
  const int k = blockIdx.x;
  NodeWork w = work[k];
  if ((w.flags & NW_OVERFLOW) || !(w.flags & (NW_SEQ_ACCEPT | NW_SEQ_AVERAGE))) return;
  if (threadIdx.x == 0) atomicAdd(&P.counters[4], 1ull);
  const int64_t n = w.n;
  Layout L = make_layout(w.scratch, n, P.method);
  NodeGeom g = node_geom(infos[nodes[k]], P);
  if (P.method == METHOD_POISSON) {
    accept_node_capped(L, g, n, P);
    return;
  }
  if (threadIdx.x != 0) return;
  // sampler_poisson_average.h:184-230 and :357-416 with per-cell lists in
  // push order (head in hkey, tail in hhead)
  for (int c = 0; c < 4096; c++) { L.hkey[c] = kNil; L.hhead[c] = kNil; }
  for (int64_t p = 0; p < n; p++) {
    uint32_t j = L.idxA[p];
    const int4 pj = L.xyz[j];
    double x = coord(pj.x, P.scale[0], P.offset[0]), y = coord(pj.y, P.scale[1], P.offset[1]),
           z = coord(pj.z, P.scale[2], P.offset[2]);
    int lo[3], hi[3];
    range16(g, x, y, z, lo, hi);
    bool accept = true;
    for (int cx = lo[0]; cx <= hi[0] && accept; cx++)
      for (int cy = lo[1]; cy <= hi[1] && accept; cy++)
        for (int cz = lo[2]; cz <= hi[2] && accept; cz++)
          for (uint32_t q = L.hkey[cx + cy * 16 + cz * 256]; q != kNil; q = L.next[q])
            if (dd64_pair(L, L.idxA[q], j, P) < g.s2) { accept = false; break; }
    L.acc[p] = accept ? 1 : 0;
    L.next[p] = kNil;
    if (accept) {
      uint32_t c = L.cell16[j];
      int cell = (c & 15) + ((c >> 4) & 15) * 16 + ((c >> 8) & 15) * 256;
      if (L.hkey[cell] == kNil) L.hkey[cell] = (uint32_t)p; else L.next[L.hhead[cell]] = (uint32_t)p;
      L.hhead[cell] = (uint32_t)p;
    }
  }
  for (int64_t p = 0; p < n; p++) {
    uint32_t j = L.idxA[p];
    L.sums[8 * p + 0] = L.rgb[3 * j]; L.sums[8 * p + 1] = L.rgb[3 * j + 1];
    L.sums[8 * p + 2] = L.rgb[3 * j + 2]; L.sums[8 * p + 3] = 1;
  }
  for (int64_t p = 0; p < n; p++) {
    uint32_t j = L.idxA[p];
    const int4 pj = L.xyz[j];
    double x = coord(pj.x, P.scale[0], P.offset[0]), y = coord(pj.y, P.scale[1], P.offset[1]),
           z = coord(pj.z, P.scale[2], P.offset[2]);
    int lo[3], hi[3];
    range16(g, x, y, z, lo, hi);
    for (int cx = lo[0]; cx <= hi[0]; cx++)
      for (int cy = lo[1]; cy <= hi[1]; cy++)
        for (int cz = lo[2]; cz <= hi[2]; cz++)
          for (uint32_t q = L.hkey[cx + cy * 16 + cz * 256]; q != kNil; q = L.next[q])
            if (dd64_pair(L, L.idxA[q], j, P) < g.s2)
              for (int c = 0; c < 4; c++) L.sums[8 * q + c] += L.sums[8 * p + c];
  }
}

// Routes the node's records (sampler_poisson.h:247-294,
// sampler_poisson_average.h:418-482): the accepted records, in
// child-then-index order, become the node's points; the rejected ones stay
// in their child in order (Poisson-average children keep every record, the
// accepted ones with the averaged colour).  Allocates the output ranges,
// writes the selection lists and the output segment table, and updates the
// node states; k_copy_tiles moves the bytes.
__global__ __launch_bounds__(kBlock) void k_route(const int32_t* __restrict__ nodes,
                                                   const NodeInfo* __restrict__ infos,
                                                   NodeState* __restrict__ states, NodeWork* __restrict__ work,
                                                   SampleParams P) {
  __shared__ int64_t s_acc[8];
  __shared__ uint8_t* s_out[9];
  __shared__ int s_fail;
  __shared__ uint32_t s_wsum[4][kWaves];
  const int k = blockIdx.x;
  const int node = nodes[k];
  const NodeWork& w = work[k];
  OutSeg* segs = P.segs + 9 * (int64_t)k;
  if (w.flags & NW_OVERFLOW) {
    if (threadIdx.x < 9) segs[threadIdx.x].count = 0;
    return;
  }
  const NodeInfo info = infos[node];
  const int64_t n = w.n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  const bool seq_avg = (w.flags & NW_SEQ_AVERAGE) != 0;
  Layout L = make_layout(w.scratch, n, P.method);
  // accepted flags in main order: scattered from sorted order (Poisson), or
  // looked up through the inverse permutation, which Poisson-average needs
  // for its colour sums
  uint32_t* inv = L.keyB;  // main index -> sorted position (Poisson-average)
  for (int64_t p0 = 0; p0 < n; p0 += 4 * kBlock) {
    uint32_t j4[4];
    uint8_t a4[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = p0 + u * kBlock + tid;
      j4[u] = p < n ? L.idxA[p] : 0u;
      a4[u] = (!avg && p < n) ? L.acc[p] : 0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = p0 + u * kBlock + tid;
      if (p >= n) continue;
      if (avg) inv[j4[u]] = (uint32_t)p;
      else L.accMain[j4[u]] = a4[u];
    }
  }
  if (tid < 8) s_acc[tid] = 0;
  __syncthreads();
  for (int c = 0; c < 8; c++) {
    unsigned long long cnt = 0;
    for (int64_t j = w.child_start[c] + tid; j < w.child_start[c + 1]; j += kBlock)
      cnt += avg ? L.acc[inv[j]] : L.accMain[j];
    if (cnt) atomicAdd((unsigned long long*)&s_acc[c], cnt);
  }
  __syncthreads();
  int64_t total = 0;
  for (int c = 0; c < 8; c++) total += s_acc[c];
  if (tid == 0) {
    const int bpp = P.bpp;
    auto seg_bytes = [&](int64_t recs) { return ((uint64_t)recs * bpp + 15) & ~15ull; };
    uint64_t bytes = seg_bytes(total);
    for (int c = 0; c < 8; c++) {
      int64_t cnt = w.child_start[c + 1] - w.child_start[c];
      int64_t rej = cnt - s_acc[c];
      if (info.children[c] >= 0 && rej > 0) bytes += seg_bytes(avg ? cnt : rej);
    }
    uint64_t off = atomicAdd((unsigned long long*)P.arena_top, (unsigned long long)bytes);
    s_fail = off + bytes > P.arena_cap;
    if (s_fail) {
      atomicOr(&work[k].flags, (uint32_t)NW_OVERFLOW);
      atomicAdd(&P.counters[6], 1ull);
      for (int q = 0; q < 9; q++) segs[q].count = 0;
    } else {
      uint8_t* o = P.arena + off;
      s_out[8] = o;
      segs[0] = OutSeg{o, nullptr, total, k, -1};
      o += seg_bytes(total);
      for (int c = 0; c < 8; c++) {
        int64_t cnt = w.child_start[c + 1] - w.child_start[c];
        int64_t rej = cnt - s_acc[c];
        s_out[c] = nullptr;
        segs[1 + c] = OutSeg{nullptr, nullptr, 0, k, c};
        if (info.children[c] >= 0 && rej > 0) {
          s_out[c] = o;
          segs[1 + c] = OutSeg{o, nullptr, avg ? cnt : rej, k, c};
          o += seg_bytes(avg ? cnt : rej);
        }
      }
    }
  }
  __syncthreads();
  if (s_fail) return;
  // ranks in main order: accepted -> sel[rank]; rejected (Poisson) -> the
  // child's remainder list; Poisson-average colours of accepted records
  __shared__ int64_t tile_base[8];  // first tile of each child in tileAcc
  if (tid == 0) {
    int64_t tb = 0;
    for (int c = 0; c < 8; c++) {
      tile_base[c] = tb;
      tb += (w.child_start[c + 1] - w.child_start[c] + 63) / 64;
    }
  }
  __syncthreads();
  int64_t run = 0;  // accepted before this step
  // four consecutive sub-chunks of kBlock candidates per step: their flag
  // loads in flight together, one barrier pair per step
  for (int64_t b0 = 0; b0 < n; b0 += 4 * kBlock) {
    uint32_t q4[4];
    bool a4[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = b0 + u * kBlock + tid;
      q4[u] = (avg && j < n) ? inv[j] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = b0 + u * kBlock + tid;
      a4[u] = j < n && (avg ? L.acc[q4[u]] : L.accMain[j]);
    }
    uint64_t bal[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      bal[u] = __ballot(a4[u]);
      if (lane == 0) s_wsum[u][wave] = (uint32_t)__popcll(bal[u]);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = b0 + u * kBlock + tid;
      const uint32_t q = q4[u];
      const bool a = a4[u];
      const uint32_t before_in_wave = (uint32_t)__popcll(bal[u] & (lane == 0 ? 0ull : (~0ull >> (64 - lane))));
      uint32_t woff = 0, tot = 0;
      for (int v = 0; v < kWaves; v++) { if (v < wave) woff += s_wsum[u][v]; tot += s_wsum[u][v]; }
      if (j < n) {
        int c = 0;
#pragma unroll
        for (int x = 1; x < 8; x++) c += j >= w.child_start[x] ? 1 : 0;
        const int64_t local = j - w.child_start[c];
        const int64_t apos = run + woff + before_in_wave;
        if ((local & 63) == 0) L.tileAcc[tile_base[c] + (local >> 6)] = (uint32_t)apos;
        if (avg) L.accMain[j] = a ? 1 : 0;
        if (a) {
          if (avg) {
            long long rs[3], ws;
            if (seq_avg) {
              for (int ch = 0; ch < 3; ch++) rs[ch] = L.sums[8 * q + ch];
              ws = L.sums[8 * q + 3];
            } else {
              for (int ch = 0; ch < 3; ch++)
                rs[ch] = 2 * ((long long)L.rgb[3 * j + ch] + L.sums[8 * q + ch]) + L.sums[8 * q + 4 + ch];
              ws = 2 * (1 + L.sums[8 * q + 3]) + L.sums[8 * q + 7];
            }
            for (int ch = 0; ch < 3; ch++) L.rgb[3 * j + ch] = (uint16_t)(rs[ch] / ws);
          }
        }
      }
      run += tot;
    }
    __syncthreads();
  }
  if (tid == 0) {
    NodeState& me = states[node];
    me.ptr = s_out[8];
    me.count = total;
    me.accepted = total;
    me.flags |= NS_SAMPLED;
    for (int c = 0; c < 8; c++) {
      int ch = info.children[c];
      if (ch < 0) continue;
      int64_t cnt = w.child_start[c + 1] - w.child_start[c];
      int64_t rej = cnt - s_acc[c];
      NodeState& cs = states[ch];
      if (avg) {
        if (rej == 0) {
          cs.flags |= NS_DISCARDED;
        } else {
          cs.ptr = s_out[c];
          cs.flags |= NS_COMPLETED;
        }
      } else {
        cs.ptr = rej > 0 ? s_out[c] : nullptr;
        cs.count = rej;
        cs.flags |= NS_COMPLETED;
      }
    }
    atomicAdd(&P.counters[5], (unsigned long long)n);
  }
}

// Copy tiles of the level: 64 consecutive records of one child each.
// tile_start[8 k + c] is the first tile of child c of node k.
__global__ __launch_bounds__(1024) void k_tile_plan(const NodeWork* __restrict__ work, int32_t num_nodes,
                                                     uint32_t* __restrict__ tile_start) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_run;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_run = 0;
  __syncthreads();
  const int num = 8 * num_nodes;
  for (int b = 0; b < num; b += 1024) {
    const int q = b + tid;
    uint32_t v = 0;
    if (q < num) {
      const NodeWork& w = work[q >> 3];
      if (!(w.flags & NW_OVERFLOW)) v = (uint32_t)((w.child_start[(q & 7) + 1] - w.child_start[q & 7] + 63) / 64);
    }
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int x = 0; x < 16; x++) {
      if (x < wave) woff += s_w[x];
      tot += s_w[x];
    }
    const uint32_t run = s_run;
    if (q < num) tile_start[q] = run + woff + incl - v;
    __syncthreads();
    if (tid == 0) s_run = run + tot;
    __syncthreads();
  }
  if (tid == 0) tile_start[num] = s_run;
}

// 4 bytes at any address inside an LDS staging area (two aligned dword
// reads and a byte align; the area is padded so the second read stays in
// it; the aligned pointer comes from pointer arithmetic so that the compiler
// keeps the LDS address space)
__device__ inline uint32_t lds_u32(const uint8_t* p) {
  const uint32_t sh = (uint32_t)((uintptr_t)p & 3);
  const uint32_t* w = (const uint32_t*)(p - sh);
  return __builtin_amdgcn_alignbyte(w[1], w[0], sh);
}

// Writes `count` records from the wave's staged tile (records `list[r]` of
// the stage, in order) to dst + first * bpp: aligned dwords inside the range,
// bytes at its edges; averaged colours replace RGB of accepted records when
// `col` is set (Poisson-average).  An inner dword holds bytes of at most two
// consecutive records (bpp >= 12): one or two unaligned 4-byte LDS reads;
// the record of a byte offset comes from a multiply by `rmul` =
// 2^21 / bpp + 1 (exact for offsets below 64 * 130 + 64 and 12 <= bpp <= 130).
__device__ inline void write_run(uint8_t* __restrict__ dst, uint64_t first, uint32_t count, const uint8_t* stage,
                                 const uint8_t* list, int bpp, uint32_t rmul, const uint16_t* col,
                                 const int64_t* colj, uint64_t colmask, int rgb0) {
  const int lane = threadIdx.x & 63;
  const uint64_t g0 = first * (uint64_t)bpp;
  const uint32_t nbytes = count * (uint32_t)bpp;
  if (nbytes == 0) return;
  const uint64_t wfirst = g0 >> 2, wlast = (g0 + nbytes - 1) >> 2;
  for (uint64_t wi = wfirst + lane; wi <= wlast; wi += 64) {
    const uint64_t b0 = wi << 2;
    const bool full = b0 >= g0 && b0 + 4 <= g0 + nbytes;
    bool bytewise = !full;
    uint32_t word = 0;
    if (full) {
      const uint32_t ob = (uint32_t)(b0 - g0);
      const uint32_t r = (ob * rmul) >> 21;
      const uint32_t bb = ob - r * (uint32_t)bpp;
      const int l = list[r];
      word = lds_u32(stage + l * bpp + bb);
      if (bb + 4 > (uint32_t)bpp) {
        const uint32_t kb = 8 * ((uint32_t)bpp - bb);  // bits from record r
        const uint32_t w2 = lds_u32(stage + list[r + 1] * bpp);
        word = (word & ((1u << kb) - 1u)) | (w2 << kb);
      }
      // an accepted record's RGB bytes in this dword take the slow path
      if (col && ((colmask >> l) & 1ull) && bb < (uint32_t)rgb0 + 6 && bb + 4 > (uint32_t)rgb0) bytewise = true;
    }
    if (bytewise) {
      word = 0;
      for (int t = 0; t < 4; t++) {
        const uint64_t gb = b0 + t;
        if (gb < g0 || gb >= g0 + nbytes) continue;
        const uint32_t ob = (uint32_t)(gb - g0);
        const uint32_t r = (ob * rmul) >> 21;
        const int bb = (int)(ob - r * (uint32_t)bpp);
        const int l = list[r];
        uint32_t byte = stage[l * bpp + bb];
        if (col && ((colmask >> l) & 1ull) && bb >= rgb0 && bb < rgb0 + 6) {
          const uint16_t v = col[3 * colj[l] + (bb - rgb0) / 2];
          byte = ((bb - rgb0) & 1) ? (v >> 8) : (v & 0xFF);
        }
        word |= byte << (8 * t);
      }
    }
    if (full) {
      *(uint32_t*)(dst + b0) = word;
    } else {
      for (int t = 0; t < 4; t++) {
        const uint64_t gb = b0 + t;
        if (gb >= g0 && gb < g0 + nbytes) dst[gb] = (uint8_t)(word >> (8 * t));
      }
    }
  }
}

// Moves the records of every sampled node of the level: each wavefront takes
// a tile of 64 consecutive records of one child (16-byte loads into LDS) and
// writes its accepted records as one run of the node's range and its
// rejected records as one run of the child's remainder (Poisson-average: the
// whole tile to the remainder, accepted records with the averaged colour,
// and the accepted ones to the node).  Ranks come from the accepted counts
// k_route recorded per tile.
constexpr int kTileStage = 64 * 130 + 32;  // record bytes of a staged tile (bpp <= 130)
constexpr int kTileVecs = (64 * 130 + 15 + 15) / 16 / 64 + 1;  // 16-byte loads per lane for one tile

__global__ __launch_bounds__(kBlock) void k_copy_tiles(const NodeWork* __restrict__ work, int32_t num_nodes,
                                                        const uint32_t* __restrict__ tile_start, SampleParams P) {
  __shared__ __attribute__((aligned(16))) uint8_t s_stage[kWaves][kTileStage];
  __shared__ uint8_t s_list[kWaves][3][64];  // accepted lanes, rejected lanes, all lanes
  __shared__ int64_t s_j[kWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int num = 8 * num_nodes;
  const uint32_t total = tile_start[num];
  const int bpp = P.bpp;
  const uint32_t rmul = (1u << 21) / (uint32_t)bpp + 1u;
  const bool avg = P.method == METHOD_POISSON_AVERAGE;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint8_t* stage = s_stage[wave];
  uint8_t* accList = s_list[wave][0];
  uint8_t* rejList = s_list[wave][1];
  uint8_t* allList = s_list[wave][2];
  allList[lane] = (uint8_t)lane;
  // each wave takes a contiguous range of tiles: one search for its first
  // tile's (node, child), then a forward walk
  const uint32_t waves_total = gridDim.x * kWaves;
  const uint32_t per = (total + waves_total - 1) / waves_total;
  const uint32_t i0 = (blockIdx.x * kWaves + wave) * per;
  const uint32_t i1 = (i0 + per) < total ? (i0 + per) : total;
  int q = 0;
  if (i0 < i1) {
    int lo = 0, hi = num - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (tile_start[mid] <= i0) lo = mid; else hi = mid - 1;
    }
    q = __builtin_amdgcn_readfirstlane(lo);
  }
  // values of the current (node, child), reloaded when the walk moves on;
  // a wave's tiles are consecutive, so most tiles reuse them
  int cq = -1;
  uint32_t qbeg = 0, qend = 0;
  const uint8_t* csrc = nullptr;
  int64_t cstart = 0, ccnt = 0;
  const uint8_t* accMain = nullptr;
  const uint32_t* tacc = nullptr;
  const uint16_t* rgb = nullptr;
  uint32_t accChild = 0;
  uint8_t* dAcc = nullptr;
  uint8_t* dRej = nullptr;
  for (uint32_t item = i0; item < i1; item++) {
    if (cq < 0 || item >= qend) {
      while (tile_start[q + 1] <= item) q++;  // tile_start[num] == total ends the walk
      cq = q;
      qbeg = tile_start[q];
      qend = tile_start[q + 1];
      const int k = q >> 3, c = q & 7;
      const NodeWork& w = work[k];
      const Layout L = make_layout(w.scratch, w.n, P.method);
      csrc = w.csrc[c];
      cstart = w.child_start[c];
      ccnt = w.child_start[c + 1] - cstart;
      accMain = L.accMain;
      rgb = L.rgb;
      // the node's accepted records before the child
      int64_t tb = 0;
      for (int x = 0; x < c; x++) tb += (w.child_start[x + 1] - w.child_start[x] + 63) / 64;
      tacc = L.tileAcc + tb;
      accChild = tacc[0];
      const OutSeg* sg = P.segs + 9 * (int64_t)k;
      dAcc = sg[0].dst;
      dRej = sg[1 + c].dst;
    }
    const uint32_t t = item - qbeg;
    const int cnt = (int)((ccnt - (int64_t)t * 64) < 64 ? (ccnt - (int64_t)t * 64) : 64);
    // the tile's records (16-byte loads), its flags and the accepted count
    // before it, all in flight together
    const uint8_t* src = csrc + (uint64_t)t * 64 * bpp;
    const uint32_t b0 = (uint32_t)((uintptr_t)src & 15);
    const uint8_t* a0 = src - b0;
    const int nvec = (int)((b0 + (uint32_t)cnt * bpp + 15) >> 4);
    uint4 rv[kTileVecs];
#pragma unroll
    for (int u = 0; u < kTileVecs; u++) {
      const int v = lane + 64 * u;
      if (v < nvec) rv[u] = *(const uint4*)(a0 + 16 * (uint64_t)v);
    }
    const bool valid = lane < cnt;
    const int64_t j = cstart + (int64_t)t * 64 + lane;
    const bool a = valid && accMain[j];
    const uint32_t accBefore = tacc[t];
#pragma unroll
    for (int u = 0; u < kTileVecs; u++) {
      const int v = lane + 64 * u;
      if (v < nvec) *(uint4*)(stage + 16 * v) = rv[u];
    }
    const uint64_t accm = __ballot(a), rejm = __ballot(valid && !a);
    if (a) accList[__popcll(accm & lt)] = (uint8_t)lane;
    if (valid && !a) rejList[__popcll(rejm & lt)] = (uint8_t)lane;
    s_j[wave][lane] = j;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint8_t* stg = stage + b0;
    write_run(dAcc, accBefore, (uint32_t)__popcll(accm), stg, accList, bpp, rmul, avg ? rgb : nullptr, s_j[wave],
              accm, P.rgb_offset);
    if (dRej) {
      if (!avg)
        write_run(dRej, (uint64_t)t * 64 - (accBefore - accChild), (uint32_t)__popcll(rejm), stg, rejList, bpp, rmul,
                  nullptr, nullptr, 0ull, 0);
      else
        write_run(dRej, (uint64_t)t * 64, (uint32_t)cnt, stg, allList, bpp, rmul, rgb, s_j[wave], accm,
                  P.rgb_offset);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
}

// ------------------------------------------------------ SamplerRandom
// Restates Converter/include/sampler_random.h:14-234 for one inner node.
// Candidates are the children's points in child-then-index order (:133-146);
// each falls into a cell of a 128^3 grid over the node box and is taken
// when its child has fewer than 100 points, or when it lies within
// 0.7 * sqrt(3) (in half-cell units) of its cell centre and the cell is
// still free (:147-178).  The grid is filled in candidate order, so the
// free-cell test passes exactly for the first candidate of its cell among
// those that take a cell (all points of small children, and the points
// near their cell centre).  That first candidate is an atomicMin per cell;
// the hand-back to the children (:184-228) is k_route + k_copy_tiles, as
// for SamplerPoisson.

// fmod(v, 1.0) up to the sign of a zero result: v - trunc(v) is exact (the
// fractional part of v is representable in v's precision), and the
// reference only uses the value as 2 * f - 1
__device__ inline double frac_toward_zero(double v) { return v - __builtin_trunc(v); }

__device__ inline uint32_t random_slot(const Layout& L, uint32_t cell) {
  uint32_t h = (cell * 2654435761u) & (L.H - 1u);
  while (true) {
    const uint32_t kk = L.hkey[h];
    if (kk == cell || kk == kNil) return h;
    h = (h + 1u) & (L.H - 1u);
  }
}

__global__ __launch_bounds__(kBlock) void k_random_prepare(const int32_t* __restrict__ nodes,
                                                            const NodeInfo* __restrict__ infos,
                                                            const NodeState* __restrict__ states,
                                                            NodeWork* __restrict__ work, SampleParams P) {
  __shared__ int64_t s_cs[9];
  __shared__ const uint8_t* s_cptr[8];
  __shared__ uint8_t* s_scr;
  const int k = blockIdx.x;
  const NodeInfo info = infos[nodes[k]];
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_cs[0] = 0;
    for (int c = 0; c < 8; c++) {
      const int ch = info.children[c];
      s_cptr[c] = ch >= 0 ? states[ch].ptr : nullptr;
      s_cs[c + 1] = s_cs[c] + (ch >= 0 ? states[ch].count : 0);
    }
    const int64_t n = s_cs[8];
    const uint64_t bytes = layout_bytes(n, P.method);
    const uint64_t off = atomicAdd((unsigned long long*)P.scratch_top, (unsigned long long)bytes);
    NodeWork w;
    w.n = n;
    for (int c = 0; c < 9; c++) w.child_start[c] = s_cs[c];
    for (int c = 0; c < 8; c++) w.csrc[c] = s_cptr[c];
    w.flags = 0;
    w.pad = 0;
    if (off + bytes > P.scratch_cap) {
      w.flags = NW_OVERFLOW;
      w.scratch = nullptr;
      atomicAdd(&P.counters[6], 1ull);
    } else {
      w.scratch = P.scratch + off;
    }
    s_scr = w.scratch;
    work[k] = w;
  }
  __syncthreads();
  if (!s_scr) return;
  const int64_t n = s_cs[8];
  Layout L = make_layout(s_scr, n, P.method);
  for (uint32_t h = tid; h < L.H; h += kBlock) {
    L.hkey[h] = kNil;
    L.hhead[h] = kNil;
  }
  __syncthreads();
  // toCellIndex (:64-87) on the node box, FP64 in the reference's order
  const double mnx = info.min[0], mny = info.min[1], mnz = info.min[2];
  const double sx = info.max[0] - info.min[0], sy = info.max[1] - info.min[1], sz = info.max[2] - info.min[2];
  for (int64_t j = tid; j < n; j += kBlock) {
    int c = 0;
    while (c < 7 && j >= s_cs[c + 1]) c++;
    int32_t X, Y, Z;
    ld_xyz(s_cptr[c] + (j - s_cs[c]) * P.bpp, X, Y, Z);
    const double x = coord(X, P.scale[0], P.offset[0]), y = coord(Y, P.scale[1], P.offset[1]),
                 z = coord(Z, P.scale[2], P.offset[2]);
    const double nx = (x - mnx) / sx, ny = (y - mny) / sy, nz = (z - mnz) / sz;
    const double lx = 2.0 * frac_toward_zero(128.0 * nx) - 1.0;
    const double ly = 2.0 * frac_toward_zero(128.0 * ny) - 1.0;
    const double lz = 2.0 * frac_toward_zero(128.0 * nz) - 1.0;
    const double dist = __builtin_sqrt(lx * lx + ly * ly + lz * lz);
    int64_t ix = (int64_t)(128.0 * nx), iy = (int64_t)(128.0 * ny), iz = (int64_t)(128.0 * nz);
    ix = ix < 0 ? 0 : (ix > 127 ? 127 : ix);
    iy = iy < 0 ? 0 : (iy > 127 ? 127 : iy);
    iz = iz < 0 ? 0 : (iz > 127 ? 127 : iz);
    const uint32_t cell = (uint32_t)(ix + iy * 128 + iz * 128 * 128);
    const bool small = s_cs[c + 1] - s_cs[c] < 100;
    const bool near = dist < P.random_limit;
    L.idxA[j] = (uint32_t)j;
    L.keyB[j] = cell;
    L.acc[j] = small ? 2 : (near ? 1 : 0);
    if (small || near) {
      uint32_t h = (cell * 2654435761u) & (L.H - 1u);
      while (true) {
        const uint32_t prev = atomicCAS(&L.hkey[h], kNil, cell);
        if (prev == kNil || prev == cell) break;
        h = (h + 1u) & (L.H - 1u);
      }
      atomicMin(&L.hhead[h], (uint32_t)j);
    }
  }
  __syncthreads();
  for (int64_t j = tid; j < n; j += kBlock) {
    const uint8_t a = L.acc[j];
    if (a == 1) L.acc[j] = L.hhead[random_slot(L, L.keyB[j])] == (uint32_t)j ? 1 : 0;
    else if (a == 2) L.acc[j] = 1;
  }
}

// One wavefront per leaf: lane 0 replays std::shuffle's swaps on the slot
// table (LDS, or global scratch for large leaves), then the wave moves each
// record to its slot.
__global__ __launch_bounds__(64) void k_shuffle_leaves(const ShuffleLeaf* __restrict__ leaves, int32_t bpp,
                                                       uint32_t seed) {
  __shared__ uint32_t s_idx[kShuffleLdsPoints];
  const ShuffleLeaf lf = leaves[blockIdx.x];
  const int64_t n = lf.count;
  uint32_t* idx = n <= kShuffleLdsPoints ? s_idx : lf.idx;
  for (int64_t i = threadIdx.x; i < n; i += 64) idx[i] = (uint32_t)i;
  __syncthreads();
  if (threadIdx.x == 0)
    lstd::shuffle_swaps(n, seed, [idx](int64_t i, int64_t j) {
      const uint32_t t = idx[i];
      idx[i] = idx[j];
      idx[j] = t;
    });
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += 64) {
    const uint8_t* src = lf.src + i * bpp;
    uint8_t* dst = lf.dst + (int64_t)idx[i] * bpp;
    for (int b = 0; b < bpp; b++) dst[b] = src[b];
  }
}

// ------------------------------------------------------ BROTLI layout
// converter_utils.h:89-104: the low 21 bits of each coordinate interleaved,
// x at bit 0 of each triple
__device__ inline uint64_t spread21(uint32_t v) {
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < 21; i++) r |= (uint64_t)((v >> i) & 1u) << (3 * i);
  return r;
}
__device__ inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread21(x) | spread21(y) << 1 | spread21(z) << 2;
}

// (upper, lower) code order through the codes' ranks
struct RankLess {
  const uint32_t* rk;
  __device__ bool operator()(uint32_t a, uint32_t b) const { return rk[a] < rk[b]; }
};

// One block per node: position codes relative to the node's minimum
// coordinate (whose start value is int64 max converted to int32, i.e. -1:
// indexer.cpp:1164-1167), a stable LSD radix sort of the 96-bit (upper,
// lower) code in three 32-bit words, then libstdc++'s std::sort on one lane
// when two points share a code (their order is the sort's), and the arrays.
__global__ __launch_bounds__(kBlock) void k_soa_layout(const EncNode* __restrict__ nodes, EncParams E,
                                                        uint8_t* __restrict__ scratch, uint8_t* __restrict__ out) {
  __shared__ RadixShared s_rad;
  __shared__ int32_t s_min[3];
  __shared__ int s_dup;
  __shared__ int64_t s_start[kBlock];
  const EncNode nd = nodes[blockIdx.x];
  const int64_t n = nd.count;
  const int tid = threadIdx.x;
  if (n <= 0) return;
  const int bpp = E.bpp;
  uint8_t* base = scratch + nd.scr_off;
  uint64_t* up = (uint64_t*)base;
  uint64_t* lo = (uint64_t*)(base + align16(8 * (uint64_t)n));
  uint32_t* kA = (uint32_t*)(base + 2 * align16(8 * (uint64_t)n));
  uint32_t* kB = (uint32_t*)((uint8_t*)kA + align16(4 * (uint64_t)n));
  uint32_t* iA = (uint32_t*)((uint8_t*)kB + align16(4 * (uint64_t)n));
  uint32_t* iB = (uint32_t*)((uint8_t*)iA + align16(4 * (uint64_t)n));
  if (tid < 3) s_min[tid] = (int32_t)-1;
  if (tid == 0) s_dup = 0;
  __syncthreads();
  int32_t mn[3] = {-1, -1, -1};
  for (int64_t i = tid; i < n; i += kBlock) {
    const uint8_t* r = nd.src + i * bpp + E.pos_off;
    for (int a = 0; a < 3; a++) {
      const int32_t v = ld_i32(r + 4 * a);
      mn[a] = v < mn[a] ? v : mn[a];
    }
  }
  for (int a = 0; a < 3; a++) atomicMin(&s_min[a], mn[a]);
  __syncthreads();
  for (int64_t i = tid; i < n; i += kBlock) {
    const uint8_t* r = nd.src + i * bpp + E.pos_off;
    uint32_t m[3];
    for (int a = 0; a < 3; a++) m[a] = (uint32_t)(ld_i32(r + 4 * a) - s_min[a]);
    lo[i] = morton3(m[0] & 0xffffu, m[1] & 0xffffu, m[2] & 0xffffu);
    up[i] = morton3(m[0] >> 16, m[1] >> 16, m[2] >> 16);
    kA[i] = (uint32_t)lo[i];
    iA[i] = (uint32_t)i;
  }
  __syncthreads();
  uint32_t *kc = kA, *ic = iA, *ko = kB, *io = iB;
  for (int word = 0; word < 3; word++) {
    if (word > 0) {
      for (int64_t p = tid; p < n; p += kBlock) {
        const uint32_t j = ic[p];
        kc[p] = word == 1 ? ((uint32_t)(lo[j] >> 32) | ((uint32_t)up[j] << 16)) : (uint32_t)(up[j] >> 16);
      }
      __syncthreads();
    }
    for (int pass = 0; pass < 4; pass++) {
      if (!block_radix_pass(kc, ic, ko, io, n, 8 * pass, 255u, s_rad)) continue;
      uint32_t* t = kc; kc = ko; ko = t;
      t = ic; ic = io; io = t;
    }
  }
  for (int64_t p = 1 + tid; p < n; p += kBlock)
    if (up[ic[p]] == up[ic[p - 1]] && lo[ic[p]] == lo[ic[p - 1]]) s_dup = 1;
  __syncthreads();
  if (s_dup) {
    // rank of every point: the sorted position of the first point with its
    // code (a run-start max-scan over one contiguous slice per thread)
    uint32_t* rk = ko;
    const int64_t per = (n + kBlock - 1) / kBlock;
    const int64_t b = (int64_t)tid * per, e = b + per < n ? b + per : n;
    int64_t last = -1;
    for (int64_t p = b; p < e; p++)
      if (p == 0 || up[ic[p]] != up[ic[p - 1]] || lo[ic[p]] != lo[ic[p - 1]]) last = p;
    s_start[tid] = last;
    __syncthreads();
    if (tid == 0) {
      int64_t run = 0;
      for (int t = 0; t < kBlock; t++) {
        const int64_t v = s_start[t];
        s_start[t] = run;
        if (v >= 0) run = v;
      }
    }
    __syncthreads();
    int64_t cur = s_start[tid];
    for (int64_t p = b; p < e; p++) {
      if (p == 0 || up[ic[p]] != up[ic[p - 1]] || lo[ic[p]] != lo[ic[p - 1]]) cur = p;
      rk[ic[p]] = (uint32_t)cur;
    }
    for (int64_t i = tid; i < n; i += kBlock) io[i] = (uint32_t)i;
    __syncthreads();
    if (tid == 0) lstd::sort(io, n, RankLess{rk});
    __syncthreads();
    ic = io;
  }
  uint8_t* o = out + nd.out_off;
  uint64_t colbase = 0;
  for (int c = 0; c < E.num_cols; c++) {
    const EncColumn col = E.cols[c];
    for (int64_t p = tid; p < n; p += kBlock) {
      const uint32_t j = ic[p];
      uint8_t* d = o + colbase + (uint64_t)p * col.size;
      const uint8_t* r = nd.src + (uint64_t)j * bpp;
      if (col.kind == EC_POSITION_MORTON) {
        const uint64_t u = up[j], l = lo[j];
        for (int b = 0; b < 8; b++) { d[b] = (uint8_t)(u >> (8 * b)); d[8 + b] = (uint8_t)(l >> (8 * b)); }
      } else if (col.kind == EC_RGB_MORTON) {
        const uint32_t cr = (uint32_t)r[col.src] | ((uint32_t)r[col.src + 1] << 8);
        const uint32_t cg = (uint32_t)r[col.src + 2] | ((uint32_t)r[col.src + 3] << 8);
        const uint32_t cb = (uint32_t)r[col.src + 4] | ((uint32_t)r[col.src + 5] << 8);
        const uint64_t mc = morton3(cr, cg, cb);
        for (int b = 0; b < 8; b++) d[b] = (uint8_t)(mc >> (8 * b));
      } else {
        for (int b = 0; b < col.size; b++) d[b] = r[col.src + b];
      }
    }
    colbase += (uint64_t)n * col.size;
  }
}

}  // namespace

}  // namespace pcg

namespace pcg {

int sample_level(const int32_t* d_nodes, int32_t num_nodes, const NodeInfo* d_info, NodeState* d_state,
                 NodeWork* d_work, const SampleParams& p, hipStream_t s) {
  if (num_nodes == 0) return 0;
  if (p.bpp > 130 || p.bpp < 12) {
    set_error("record size outside 12..130 bytes is not supported by the sampler");
    return 1;
  }
  dim3 grid((unsigned)num_nodes);
  if (p.method == METHOD_RANDOM) {
    hipLaunchKernelGGL(k_random_prepare, grid, dim3(kBlock), 0, s, d_nodes, d_info, d_state, d_work, p);
    PCG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_route, grid, dim3(kBlock), 0, s, d_nodes, d_info, d_state, d_work, p);
    PCG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_tile_plan, dim3(1), dim3(1024), 0, s, d_work, num_nodes, p.seg_start);
    PCG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_copy_tiles, dim3((unsigned)p.copy_blocks), dim3(kBlock), 0, s, d_work, num_nodes, p.seg_start, p);
    PCG_CHECK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_prepare, grid, dim3(kBlock), 0, s, d_nodes, d_info, d_state, d_work, p);
  PCG_CHECK(hipGetLastError());
  {
    const int cus = p.num_cus > 0 ? p.num_cus : 256;
    const int cap = num_nodes <= cus ? kStageNarrow : kStageWide;
    const size_t lds = (size_t)kWaves * 16 * cap;
    PCG_CHECK(hipFuncSetAttribute((const void*)k_exact_sort, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)((size_t)kWaves * 16 * kStageNarrow)));
    hipLaunchKernelGGL(k_exact_sort, grid, dim3(kBlock), lds, s, d_nodes, d_info, d_work, p, p.counters, cap);
  }
  PCG_CHECK(hipGetLastError());
  if (p.method == METHOD_POISSON && num_nodes >= p.ring_min_nodes) {
    // the queue takes the item table, which k_plan fills only afterwards
    hipLaunchKernelGGL(k_ring_queue, dim3(1), dim3(1024), 0, s, d_work, num_nodes, (int32_t*)p.item_start,
                       p.ring_ctl);
    PCG_CHECK(hipGetLastError());
    const int blocks = std::max(1, std::min((num_nodes + kRingWaves - 1) / kRingWaves, p.ring_blocks));
    hipLaunchKernelGGL(k_ring, dim3((unsigned)blocks), dim3(64 * kRingWaves), 0, s, d_nodes, d_info, d_work,
                       (const int32_t*)p.item_start, p.ring_ctl, p);
    PCG_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_cells, grid, dim3(kBlock), 0, s, d_nodes, d_info, d_work, p);
  PCG_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, s, d_work, num_nodes, p.item_start);
  PCG_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_neighbors, dim3((unsigned)p.neighbor_blocks), dim3(kBlock), 0, s, d_nodes, d_info, d_work, num_nodes,
                     p.item_start, p);
  PCG_CHECK(hipGetLastError());
  // pipelined: launches of few nodes, and launches the ring decides (their
  // cell-path nodes are the few whose shell outgrew the ring, often the
  // level's largest)
  const bool ring_level = p.method == METHOD_POISSON && num_nodes >= p.ring_min_nodes;
  if (p.accept_pipe_nodes > 0 && (num_nodes <= p.accept_pipe_nodes || ring_level)) {
    // the lists in the item table (k_neighbors has read it), their counts
    // in the ring's control words (the ring is done)
    int32_t* list = (int32_t*)p.item_start;
    uint32_t* counts = p.ring_ctl;
    hipLaunchKernelGGL(k_pipe_select, dim3(1), dim3(1024), 0, s, d_work, num_nodes, list, counts, p.pipe_form);
    PCG_CHECK(hipGetLastError());
    const int cus = p.num_cus > 0 ? p.num_cus : 256;
    hipLaunchKernelGGL(k_accept_pipe<false>, dim3((unsigned)std::min(num_nodes, 3 * cus)), dim3(64 * (1 + kPipeProducers)),
                       0, s, d_nodes, d_info, d_work, (const int32_t*)list, (const uint32_t*)counts, num_nodes, p);
    PCG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_accept_pipe<true>, dim3((unsigned)std::min(num_nodes, cus)), dim3(64 * (1 + kPipeProducers)), 0,
                       s, d_nodes, d_info, d_work, (const int32_t*)list, (const uint32_t*)counts, num_nodes, p);
  } else if (num_nodes <= kAccDeepNodes)
    hipLaunchKernelGGL(k_accept<3>, dim3((unsigned)((num_nodes + kWaves - 1) / kWaves)), dim3(kBlock), 0, s, d_nodes,
                       d_info, d_work, num_nodes, p);
  else
    hipLaunchKernelGGL(k_accept<1>, dim3((unsigned)((num_nodes + kWaves - 1) / kWaves)), dim3(kBlock), 0, s, d_nodes,
                       d_info, d_work, num_nodes, p);
  PCG_CHECK(hipGetLastError());

  if (p.method == METHOD_POISSON_AVERAGE) {
    // about avg_blocks workgroups in all, at most 64 per node
    const unsigned slices = (unsigned)std::max(1, std::min(64, p.avg_blocks / std::max(1, (int)num_nodes)));
    hipLaunchKernelGGL(k_avg_compact, grid, dim3(kBlock), 0, s, d_work, p);
    PCG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_avg_colors, dim3((unsigned)num_nodes, slices), dim3(kBlock), 0, s, d_nodes, d_info, d_work,
                       p);
    PCG_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sequential, grid, dim3(64), 0, s, d_nodes, d_info, d_work, p);
  PCG_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_route, grid, dim3(kBlock), 0, s, d_nodes, d_info, d_state, d_work, p);
  PCG_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_tile_plan, dim3(1), dim3(1024), 0, s, d_work, num_nodes, p.seg_start);
  PCG_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_copy_tiles, dim3((unsigned)p.copy_blocks), dim3(kBlock), 0, s, d_work, num_nodes, p.seg_start, p);
  PCG_CHECK(hipGetLastError());
  return 0;
}

namespace {
__global__ void k_zero_u64(uint64_t* __restrict__ p, int32_t n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}
}  // namespace

int sample_zero(uint64_t* d, int32_t n, hipStream_t s) {
  if (n < 0 || n > 64) {
    set_error("sample_zero: at most 64 words");
    return 1;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, d, n);
  PCG_CHECK(hipGetLastError());
  return 0;
}

uint64_t sample_scratch_bytes(int64_t n, int method) { return layout_bytes(n, method); }

uint64_t sample_scratch_bound(int64_t candidates, int64_t max_nodes, int method) {
  // A node of n candidates takes its per-candidate arrays (A bytes per
  // candidate, each array rounded up to 16 bytes), at most 24 bytes per hash
  // slot (htab, hkey, hhead; the random method keeps hkey and hhead only)
  // for H(n) <= max(minimum, 4n) slots, and 64 bytes.  A is measured at
  // n = 2^20 with the hash taken out; layout_bytes(1) covers the minimum
  // hash, the 64 bytes and 16 bytes per array, which bounds the rounding of
  // every array.  So layout_bytes(n) <= (A + 96) n + layout_bytes(1) for
  // every n.
  const int64_t m = int64_t(1) << 20;
  const uint64_t slotBytes = method == METHOD_RANDOM ? 8 : 24;
  const uint64_t lin =
      (layout_bytes(m, method) - 64 - slotBytes * (uint64_t)hash_cap(m, method)) / (uint64_t)m + 1;
  return (uint64_t)candidates * (lin + 96) + (uint64_t)max_nodes * (layout_bytes(1, method) + 512);
}

namespace {
// Byte ranges are arbitrary (records of any size): aligned dword stores
// into the destination, each assembled from the two aligned source dwords
// that cover it (byte align by the source/destination offset difference);
// bytes only at the edges of a range.  Sources are 16-byte aligned arena
// ranges, so the dword reads around a range stay inside its allocation.
__global__ __launch_bounds__(256) void k_gather_ranges(const CopyDesc* __restrict__ desc, uint8_t* __restrict__ dst) {
  const CopyDesc d = desc[blockIdx.x];
  const uint8_t* s = d.src;
  uint8_t* o = dst + d.dst_offset;
  const uint64_t n = d.bytes;
  if (n == 0) return;
  const uint64_t head = (4 - ((uintptr_t)o & 3)) & 3;  // bytes before the first aligned destination word
  const uint64_t h = head < n ? head : n;
  if (threadIdx.x < h) o[threadIdx.x] = s[threadIdx.x];
  const uint64_t words = (n - h) / 4;
  const uint8_t* sw = s + h;  // source of the first aligned destination word
  const uint32_t shift = (uint32_t)((uintptr_t)sw & 3);
  const uint32_t* sa = (const uint32_t*)(sw - shift);
  uint32_t* o4 = (uint32_t*)(o + h);
  for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) {
    const uint32_t lo = sa[i];
    const uint32_t hi = shift ? sa[i + 1] : 0u;
    o4[i] = __builtin_amdgcn_alignbyte(hi, lo, shift);
  }
  for (uint64_t i = h + words * 4 + threadIdx.x; i < n; i += blockDim.x) o[i] = s[i];
}
}  // namespace

int gather_ranges(const CopyDesc* d_desc, int64_t num, uint8_t* d_dst, hipStream_t s) {
  if (num == 0) return 0;
  hipLaunchKernelGGL(k_gather_ranges, dim3((unsigned)num), dim3(256), 0, s, d_desc, d_dst);
  PCG_CHECK(hipGetLastError());
  return 0;
}

int shuffle_leaves(const ShuffleLeaf* d_leaves, int64_t num, int32_t bpp, uint32_t seed, hipStream_t s) {
  if (num == 0) return 0;
  hipLaunchKernelGGL(k_shuffle_leaves, dim3((unsigned)num), dim3(64), 0, s, d_leaves, bpp, seed);
  PCG_CHECK(hipGetLastError());
  return 0;
}

uint64_t enc_scratch_bytes(int64_t count) {
  return 2 * align16(8 * (uint64_t)count) + 4 * align16(4 * (uint64_t)count);
}

int encode_layout(const EncNode* d_nodes, int64_t num_nodes, const EncParams& p, uint8_t* d_scratch, uint8_t* d_out,
                  hipStream_t s) {
  if (num_nodes == 0) return 0;
  hipLaunchKernelGGL(k_soa_layout, dim3((unsigned)num_nodes), dim3(kBlock), 0, s, d_nodes, p, d_scratch, d_out);
  PCG_CHECK(hipGetLastError());
  return 0;
}

}  // namespace pcg
