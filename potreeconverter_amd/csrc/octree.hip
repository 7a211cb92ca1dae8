// Octree build on gfx950: one refinement round of buildHierarchy
// (Converter/src/indexer.cpp:823-1109) for many segments at once.
//
// A segment is a contiguous range of output records plus the double box of
// the node being refined.  Per round:
//   k_seg_keys      15-bit Morton cell over 32^3 (indexer.cpp:851-871)
//   k_seg_pyramid   32^3 histogram -> sum pyramid -> realized nodes
//                   (createSumPyramid :691-730, createNodes :732-817)
//   k_seg_hist /
//   k_seg_scatter   stable LSD radix pass (8 then 7 bits) over the keys,
//                   moving whole records; two passes equal the reference's
//                   stable counting sort by the 15-bit key (:874-909)
#include <algorithm>

#include "octree.h"
#include "pcg_common.h"
#include "scan.h"

namespace pcg {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kLeafCap = 10000;  // indexer.h:47
constexpr int kMaxBpp = 128;

__device__ inline int32_t ld_i32(const uint8_t* p) {
  return (int32_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
}

// 5-bit spread for the 15-bit key: bit 3k+2 = x_k, 3k+1 = y_k, 3k = z_k
// (mortonEncode_magicbits(iz, iy, ix), converter_utils.h:89-104)
__device__ inline uint32_t spread5(uint32_t v) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 5; k++) r |= ((v >> k) & 1u) << (3 * k);
  return r;
}

__device__ inline int64_t cell32(int32_t X, double scale, double offset, double mn, double size) {
  double x = ((double)X * scale) + offset;
  int64_t i = (int64_t)(32.0 * (x - mn) / size);
  return i < 0 ? 0 : (i > 31 ? 31 : i);
}

__global__ __launch_bounds__(kBlock) void k_seg_keys(const uint8_t* __restrict__ rec, int32_t bpp,
                                                      const SegDesc* __restrict__ segs,
                                                      const TileDesc* __restrict__ tiles, OctreeParams p,
                                                      uint16_t* __restrict__ keys) {
  const TileDesc t = tiles[blockIdx.x];
  const SegDesc s = segs[t.seg];
  for (int64_t k = threadIdx.x; k < t.count; k += kBlock) {
    int64_t i = t.start + k;
    const uint8_t* r = rec + i * bpp;
    int64_t ix = cell32(ld_i32(r + 0), p.scale[0], p.offset[0], s.min[0], s.size[0]);
    int64_t iy = cell32(ld_i32(r + 4), p.scale[1], p.offset[1], s.min[1], s.size[1]);
    int64_t iz = cell32(ld_i32(r + 8), p.scale[2], p.offset[2], s.min[2], s.size[2]);
    keys[i] = (uint16_t)(spread5((uint32_t)iz) | (spread5((uint32_t)iy) << 1) | (spread5((uint32_t)ix) << 2));
  }
}

// One workgroup per segment: histogram (32768 bins in LDS), pyramid, and
// the list of realized nodes (level, morton index, start, count).
constexpr int kPyrBlock = 1024;
// Per-segment 32^3 cell counts: gridDim.y workgroups per segment each count
// a slice of its keys in LDS and add their non-zero bins to the segment's
// global histogram (zeroed by the caller), so that every CU takes part.
__global__ __launch_bounds__(kPyrBlock) void k_seg_count(const uint16_t* __restrict__ keys,
                                                          const SegDesc* __restrict__ segs,
                                                          uint32_t* __restrict__ h5g) {
  __shared__ uint32_t h5[32768];
  const SegDesc s = segs[blockIdx.x];
  for (int i = threadIdx.x; i < 32768; i += kPyrBlock) h5[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t per = ((s.count + gridDim.y - 1) / gridDim.y + kPyrBlock - 1) / kPyrBlock * kPyrBlock;
  const int64_t a = (int64_t)blockIdx.y * per, b = a + per < s.count ? a + per : s.count;
  for (int64_t k0 = a; k0 < b; k0 += kPyrBlock) {
    int64_t k = k0 + threadIdx.x;
    bool valid = k < b;
    uint32_t key = valid ? keys[s.start + k] : 0xFFFFu;
    // merge runs of equal keys inside the wave before the LDS atomics
    uint32_t prev = __shfl_up(key, 1, 64);
    bool head = valid && (lane == 0 || prev != key);
    uint64_t heads = __ballot(head);
    uint64_t valids = __ballot(valid);
    if (head) {
      uint64_t after = (lane == 63) ? 0ull : (heads >> (lane + 1));
      int len = after ? (__ffsll((long long)after)) : (64 - lane);
      int last_valid = 63 - __clzll((long long)valids);
      if (lane + len - 1 > last_valid) len = last_valid - lane + 1;
      atomicAdd(&h5[key], (uint32_t)len);
    }
  }
  __syncthreads();
  uint32_t* g = h5g + (size_t)blockIdx.x * 32768;
  for (int i = threadIdx.x; i < 32768; i += kPyrBlock)
    if (h5[i]) atomicAdd(&g[i], h5[i]);
}

__global__ __launch_bounds__(kPyrBlock) void k_seg_pyramid(const uint32_t* __restrict__ h5g,
                                                            const SegDesc* __restrict__ segs,
                                                            NodeOut* __restrict__ out, uint32_t* __restrict__ counter,
                                                            uint32_t max_out) {
  __shared__ uint32_t h5[32768];
  __shared__ uint32_t up[4096 + 512 + 64 + 8 + 1];
  __shared__ uint32_t s_wsum[kPyrBlock / 64];
  const SegDesc s = segs[blockIdx.x];
  const uint32_t* g = h5g + (size_t)blockIdx.x * 32768;
  for (int i = threadIdx.x; i < 32768; i += kPyrBlock) h5[i] = g[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // pyramid in morton layout: level l cell m = sum of level l+1 cells 8m..8m+7
  uint32_t* lv4 = up;
  uint32_t* lv3 = up + 4096;
  uint32_t* lv2 = lv3 + 512;
  uint32_t* lv1 = lv2 + 64;
  uint32_t* lv0 = lv1 + 8;
  for (int m = threadIdx.x; m < 4096; m += kPyrBlock) {
    uint32_t a = 0;
    for (int i = 0; i < 8; i++) a += h5[8 * m + i];
    lv4[m] = a;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < 512; m += kPyrBlock) { uint32_t a = 0; for (int i = 0; i < 8; i++) a += lv4[8 * m + i]; lv3[m] = a; }
  __syncthreads();
  if (threadIdx.x < 64) { uint32_t a = 0; for (int i = 0; i < 8; i++) a += lv3[8 * threadIdx.x + i]; lv2[threadIdx.x] = a; }
  __syncthreads();
  if (threadIdx.x < 8) { uint32_t a = 0; for (int i = 0; i < 8; i++) a += lv2[8 * threadIdx.x + i]; lv1[threadIdx.x] = a; }
  __syncthreads();
  if (threadIdx.x == 0) { uint32_t a = 0; for (int i = 0; i < 8; i++) a += lv1[i]; lv0[0] = a; }
  __syncthreads();
  // realized nodes of levels 0..4 (their start needs the level-5 prefix, so
  // emit them as (level, m) and compute starts from the scanned histogram)
  // first: exclusive prefix of h5 kept beside the counts
  // (32 items per thread, then a block scan over the 1024 partial sums)
  uint32_t cnt5[32];
  uint32_t local = 0;
  const int base = threadIdx.x * 32;
#pragma unroll
  for (int i = 0; i < 32; i++) { cnt5[i] = h5[base + i]; local += cnt5[i]; }
  uint32_t incl = local;
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) s_wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t woff = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) woff += s_wsum[w];
  __syncthreads();
  uint32_t run = woff + incl - local;
#pragma unroll
  for (int i = 0; i < 32; i++) { h5[base + i] = run; run += cnt5[i]; }
  __syncthreads();
  // level 5 nodes: count > 0 and parent (level 4) > cap
  const uint32_t cap = kLeafCap;
  for (int i = 0; i < 32; i++) {
    int m = base + i;
    uint32_t c = cnt5[i];
    if (c == 0) continue;
    if (lv4[m >> 3] <= cap || lv3[m >> 6] <= cap || lv2[m >> 9] <= cap || lv1[m >> 12] <= cap || lv0[0] <= cap) continue;
    uint32_t slot = atomicAdd(counter, 1u);
    if (slot < max_out) out[slot] = NodeOut{(int32_t)blockIdx.x, 5, (uint32_t)m, h5[m], c};
  }
  // levels 0..4
  for (int l = 0; l <= 4; l++) {
    const int cells = 1 << (3 * l);
    const uint32_t* lv = l == 0 ? lv0 : (l == 1 ? lv1 : (l == 2 ? lv2 : (l == 3 ? lv3 : lv4)));
    for (int m = threadIdx.x; m < cells; m += kPyrBlock) {
      uint32_t c = lv[m];
      if (c == 0 || c > cap) continue;
      bool ok = true;
      for (int a = l - 1; a >= 0; a--) {
        const uint32_t* pa = a == 0 ? lv0 : (a == 1 ? lv1 : (a == 2 ? lv2 : (a == 3 ? lv3 : lv4)));
        if (pa[m >> (3 * (l - a))] <= cap) { ok = false; break; }
      }
      if (!ok) continue;
      uint32_t slot = atomicAdd(counter, 1u);
      if (slot < max_out) out[slot] = NodeOut{(int32_t)blockIdx.x, l, (uint32_t)m, h5[m << (3 * (5 - l))], c};
    }
  }
}

// per-tile digit histogram; M layout per segment: [bin][tile-in-segment]
__global__ __launch_bounds__(kBlock) void k_seg_hist(const uint16_t* __restrict__ keys,
                                                      const TileDesc* __restrict__ tiles, int shift, int bins,
                                                      uint32_t* __restrict__ M) {
  __shared__ uint32_t h[256];
  const TileDesc t = tiles[blockIdx.x];
  for (int i = threadIdx.x; i < 256; i += kBlock) h[i] = 0;
  __syncthreads();
  for (int64_t k = threadIdx.x; k < t.count; k += kBlock) {
    uint32_t d = (keys[t.start + k] >> shift) & (bins - 1);
    atomicAdd(&h[d], 1u);
  }
  __syncthreads();
  // all 256 bins are written so a 7-bit pass leaves zeros in the upper half
  for (int b = threadIdx.x; b < 256; b += kBlock) M[t.m_base + (int64_t)b * t.seg_tiles + t.tile_in_seg] = h[b];
}

__device__ inline uint64_t wave_match_bits(uint32_t key, int bits, uint64_t active) {
  uint64_t peers = active;
  for (int b = 0; b < bits; b++) {
    uint64_t m = __ballot((key >> b) & 1);
    peers &= ((key >> b) & 1) ? m : ~m;
  }
  return peers;
}

// stable scatter of records + keys for one digit; each wave owns a
// contiguous quarter of the tile (same scheme as the chunker scatter)
// bytes of one wave's record staging area and output staging area in LDS
__host__ __device__ inline int seg_area_bytes(int bpp) { return (64 * bpp + 48 + 15) & ~15; }

// 4 bytes at any address inside an LDS staging area (two aligned dword
// reads and a byte align; the areas are padded so the second read stays in
// them)
__device__ inline uint32_t lds_u32(const uint8_t* p) {
  const uint32_t sh = (uint32_t)((uintptr_t)p & 3);
  const uint32_t* w = (const uint32_t*)(p - sh);
  return __builtin_amdgcn_alignbyte(w[1], w[0], sh);
}

// The 64 records of a step are assembled in LDS grouped into runs of one
// digit (a run is contiguous in the output: the stable order of equal
// digits is the lane order), and every run is written with aligned dword
// stores, bytes only where a run starts or ends inside a word.
__global__ __launch_bounds__(kBlock) void k_seg_scatter(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                         const uint16_t* __restrict__ ksrc,
                                                         uint16_t* __restrict__ kdst, int32_t bpp,
                                                         const TileDesc* __restrict__ tiles,
                                                         const SegDesc* __restrict__ segs, int shift, int bins,
                                                         const uint32_t* __restrict__ M) {
  __shared__ uint32_t s_cnt[kWaves][256];
  extern __shared__ __attribute__((aligned(16))) uint8_t s_dyn[];  // per wave: record area, output area
  const TileDesc t = tiles[blockIdx.x];
  const int64_t delta = segs[t.seg].delta;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint8_t* rec_area = s_dyn + (size_t)wave * 2 * seg_area_bytes(bpp);
  uint8_t* out_area = rec_area + seg_area_bytes(bpp);
  const int bits = bins == 256 ? 8 : 7;
  for (int i = threadIdx.x; i < kWaves * 256; i += kBlock) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t per_wave = (t.count + kWaves - 1) / kWaves;
  const int64_t w0 = wave * per_wave < t.count ? wave * per_wave : t.count;
  const int64_t w1 = (w0 + per_wave) < t.count ? (w0 + per_wave) : t.count;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int64_t f = w0; f < w1; f += 64) {
    int64_t k = f + lane;
    bool valid = k < w1;
    uint32_t d = valid ? ((ksrc[t.start + k] >> shift) & (bins - 1)) : 0;
    uint64_t peers = wave_match_bits(d, bits, __ballot(valid));
    if (valid && (peers & lt_mask) == 0) s_cnt[wave][d] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  for (int b = threadIdx.x; b < bins; b += kBlock) {
    uint32_t run = M[t.m_base + (int64_t)b * t.seg_tiles + t.tile_in_seg];
    for (int w = 0; w < kWaves; w++) {
      uint32_t v = s_cnt[w][b];
      s_cnt[w][b] = run;
      run += v;
    }
  }
  __syncthreads();
  const int out_vecs = (64 * bpp + 15) / 16;
  const int ow = (bpp + 3) / 4;     // dwords of a record
  const int inner = (bpp - 4) / 4;  // dwords 1 .. inner lie inside the record at any alignment
  const uint32_t last_mask = (bpp & 3) ? ((1u << (8 * (bpp & 3))) - 1u) : 0xFFFFFFFFu;
  for (int64_t f = w0; f < w1; f += 64) {
    int cnt = (int)((w1 - f) < 64 ? (w1 - f) : 64);
    // stage the wave's records with 16-byte loads
    {
      uint64_t b0 = (uint64_t)(t.start + f) * bpp;
      uint64_t a0 = b0 & ~15ull;
      uint64_t end = b0 + (uint64_t)cnt * bpp;
      int nvec = (int)((end - a0 + 15) >> 4);
      for (int v = lane; v < nvec; v += 64) *(uint4*)(rec_area + 16 * v) = *(const uint4*)(src + a0 + 16 * (uint64_t)v);
    }
    for (int i = lane; i < out_vecs; i += 64) ((uint4*)out_area)[i] = make_uint4(0u, 0u, 0u, 0u);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint32_t shift_b = (uint32_t)(((uint64_t)(t.start + f) * bpp) & 15ull);
    int64_t k = f + lane;
    bool valid = lane < cnt;
    uint16_t key = valid ? ksrc[t.start + k] : 0;
    uint32_t d = (key >> shift) & (bins - 1);
    uint64_t peers = wave_match_bits(d, bits, __ballot(valid));
    uint32_t base = valid ? s_cnt[wave][d] : 0;
    __builtin_amdgcn_wave_barrier();
    const bool leader = valid && (peers & lt_mask) == 0;
    const uint32_t len = leader ? (uint32_t)__popcll(peers) : 0u;
    if (leader) s_cnt[wave][d] = base + len;
    // runs in the output area in leader-lane order
    uint32_t incl = len;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
      const uint32_t o = __shfl_up(incl, dd);
      if (lane >= dd) incl += o;
    }
    const uint32_t run_start = incl - len;  // valid on leader lanes
    const int my_leader = valid ? __builtin_ctzll(peers) : lane;
    const uint32_t my_run = __shfl(run_start, my_leader);
    if (valid) {
      // the record as dwords: one byte align per dword from the staged
      // record, one byte permute to its place in the output area, plain
      // stores inside it, ORs (the area is zeroed per step) for the two
      // dwords shared with the neighbouring records
      const uint8_t* r = rec_area + shift_b + lane * bpp;
      const uint32_t sh = (uint32_t)((uintptr_t)r & 3);
      const uint32_t* ra = (const uint32_t*)(r - sh);
      const uint32_t off = (my_run + (uint32_t)__popcll(peers & lt_mask)) * (uint32_t)bpp;
      const uint32_t sel = 0x07060504u - (off & 3) * 0x01010101u;
      uint32_t* dw = (uint32_t*)out_area + (off >> 2);
      uint32_t prev = 0u, lo = ra[0];
      for (int q = 0; q < ow; q++) {
        const uint32_t hi = ra[q + 1];
        uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, sh);
        lo = hi;
        if (q == ow - 1) v &= last_mask;
        const uint32_t x = __builtin_amdgcn_perm(v, prev, sel);
        if (q >= 1 && q <= inner) dw[q] = x;
        else atomicOr(dw + q, x);
        prev = v;
      }
      atomicOr(dw + ow, __builtin_amdgcn_perm(0u, prev, sel));
      kdst[(int64_t)base + __popcll(peers & lt_mask) + delta] = key;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    uint64_t leaders = __ballot(leader);
    while (leaders) {
      const int L = __builtin_ctzll(leaders);
      leaders &= leaders - 1;
      const uint64_t g0 = ((uint64_t)__builtin_amdgcn_readlane(base, L) + (uint64_t)delta) * (uint64_t)bpp;
      const uint32_t nbytes = (uint32_t)__builtin_amdgcn_readlane(len, L) * (uint32_t)bpp;
      const uint8_t* sr = out_area + (size_t)__builtin_amdgcn_readlane(run_start, L) * bpp;
      const uint64_t wfirst = g0 >> 2, wlast = (g0 + nbytes - 1) >> 2;
      for (uint64_t wi = wfirst + lane; wi <= wlast; wi += 64) {
        const uint64_t b0 = wi << 2;
        if (b0 >= g0 && b0 + 4 <= g0 + nbytes) {
          *(uint32_t*)(dst + b0) = lds_u32(sr + (b0 - g0));
        } else {
          for (uint64_t bb = b0; bb < b0 + 4; bb++)
            if (bb >= g0 && bb < g0 + nbytes) dst[bb] = sr[bb - g0];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
}

}  // namespace

int octree_round(uint8_t* d_a, uint8_t* d_b, uint16_t* d_ka, uint16_t* d_kb, int32_t bpp, const OctreeParams& p,
                 const SegDesc* d_segs, int32_t num_segs, const TileDesc* d_tiles, int64_t num_tiles,
                 int64_t m_total, uint32_t* d_M, void* d_scan_ws, NodeOut* d_nodes, uint32_t* d_node_counter,
                 uint32_t max_nodes, uint32_t* d_h5, hipStream_t s) {
  if (num_segs == 0) return 0;
  if (bpp > kMaxBpp) {
    set_error("record size above 128 bytes is not supported by the octree build");
    return 1;
  }
  hipLaunchKernelGGL(k_seg_keys, dim3((unsigned)num_tiles), dim3(kBlock), 0, s, d_a, bpp, d_segs, d_tiles, p, d_ka);
  PCG_CHECK(hipGetLastError());
  // about 512 counting workgroups in all, at most 16 per segment
  const unsigned slices = (unsigned)std::max(1, std::min(16, 512 / std::max(1, (int)num_segs)));
  PCG_CHECK(hipMemsetAsync(d_h5, 0, (size_t)num_segs * 32768 * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_seg_count, dim3((unsigned)num_segs, slices), dim3(kPyrBlock), 0, s, d_ka, d_segs, d_h5);
  PCG_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_seg_pyramid, dim3((unsigned)num_segs), dim3(kPyrBlock), 0, s, d_h5, d_segs, d_nodes,
                     d_node_counter, max_nodes);
  PCG_CHECK(hipGetLastError());
  // pass 1: low 8 bits A -> B, pass 2: high 7 bits B -> A
  for (int pass = 0; pass < 2; pass++) {
    const int shift = pass == 0 ? 0 : 8;
    const int bins = pass == 0 ? 256 : 128;
    uint8_t* src = pass == 0 ? d_a : d_b;
    uint8_t* dst = pass == 0 ? d_b : d_a;
    uint16_t* ks = pass == 0 ? d_ka : d_kb;
    uint16_t* kd = pass == 0 ? d_kb : d_ka;
    hipLaunchKernelGGL(k_seg_hist, dim3((unsigned)num_tiles), dim3(kBlock), 0, s, ks, d_tiles, shift, bins, d_M);
    PCG_CHECK(hipGetLastError());
    if (exclusive_scan_u32(d_M, m_total, d_scan_ws, nullptr, s)) return 1;
    static bool lds_set = false;  // staging above 64 KB for records near kMaxBpp
    if (!lds_set) {
      PCG_CHECK(hipFuncSetAttribute((const void*)k_seg_scatter, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    kWaves * 2 * seg_area_bytes(kMaxBpp)));
      lds_set = true;
    }
    hipLaunchKernelGGL(k_seg_scatter, dim3((unsigned)num_tiles), dim3(kBlock),
                       (size_t)kWaves * 2 * seg_area_bytes(bpp), s, src, dst, ks, kd, bpp, d_tiles, d_segs, shift,
                       bins, d_M);
    PCG_CHECK(hipGetLastError());
  }
  return 0;
}

}  // namespace pcg
