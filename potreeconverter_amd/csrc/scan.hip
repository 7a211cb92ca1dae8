// Single-pass exclusive prefix sum with decoupled look-back (Merrill &
// Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back",
// 2016), written for 64-wide wavefronts.  Used to turn per-(bucket, tile)
// counts into scatter offsets for the chunker and the octree passes.
#include "pcg_common.h"
#include "scan.h"

namespace pcg {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;                       // per thread
constexpr int kScanTile = kScanBlock * kScanItems;  // 4096 elements per tile
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

__device__ inline uint64_t wave_incl_scan(uint64_t v) {
  int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_u32(uint32_t* __restrict__ data, int64_t n,
                                                          uint64_t* __restrict__ status,
                                                          uint32_t* __restrict__ tile_counter,
                                                          uint64_t* __restrict__ total_out) {
  __shared__ uint64_t s_wave[kScanBlock / 64];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const int64_t base = (int64_t)tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t local = 0;
  if (base + kScanItems <= n) {
    const uint4* src = (const uint4*)(data + base);  // 64-byte aligned
#pragma unroll
    for (int k = 0; k < kScanItems / 4; k++) {
      const uint4 q = src[k];
      v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      int64_t i = base + k;
      v[k] = i < n ? data[i] : 0u;
    }
  }
#pragma unroll
  for (int k = 0; k < kScanItems; k++) local += v[k];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t incl = wave_incl_scan(local);
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint64_t wave_off = 0, agg = 0;
  for (int w = 0; w < kScanBlock / 64; w++) {
    if (w < wave) wave_off += s_wave[w];
    agg += s_wave[w];
  }
  // publish the aggregate, then wave 0 looks back over 64 predecessors at a
  // time: the nearest one with an inclusive prefix ends the walk, the ones
  // after it contribute their aggregates; a predecessor that has published
  // nothing yet (tiles are taken in order, so it is running) is re-read
  if (wave == 0) {
    if (tile == 0) {
      if (lane == 0) {
        __hip_atomic_store(&status[0], kFlagPre | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        s_prefix = 0;
      }
    } else {
      if (lane == 0) __hip_atomic_store(&status[tile], kFlagAgg | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      uint64_t prefix = 0;
      int64_t j = (int64_t)tile - 1;  // newest predecessor of the window
      while (true) {
        const int64_t idx = j - lane;
        uint64_t st = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                               : kFlagPre;  // before tile 0: an empty inclusive prefix
        const uint64_t flag = st & ~kValMask;
        const uint64_t pre = __ballot(flag == kFlagPre);
        const uint64_t none = __ballot(flag == 0);
        const int stop = pre ? __builtin_ctzll(pre) : 64;  // nearest inclusive prefix
        const uint64_t upto = stop == 63 ? ~0ull : ((stop == 64) ? ~0ull : ((1ull << (stop + 1)) - 1));
        if (none & upto) continue;  // a predecessor in the window is not published yet
        uint64_t v = ((uint64_t)lane <= (uint64_t)stop) ? (st & kValMask) : 0ull;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        prefix += v;
        if (stop < 64) break;
        j -= 64;
      }
      if (lane == 0) {
        __hip_atomic_store(&status[tile], kFlagPre | (prefix + agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        s_prefix = prefix;
      }
    }
  }
  __syncthreads();
  uint64_t run = s_prefix + wave_off + incl - local;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int64_t i = base + k;
    if (i < n) data[i] = (uint32_t)run;
    run += v[k];
  }
  // the thread holding the last element publishes the grand total
  if (total_out && base <= n - 1 && n - 1 < base + kScanItems) *total_out = run;
}

}  // namespace

size_t scan_workspace_bytes(int64_t n) {
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  return (size_t)(tiles + 1) * sizeof(uint64_t) + 256;
}

int exclusive_scan_u32(uint32_t* d_data, int64_t n, void* d_workspace, uint64_t* d_total, hipStream_t stream) {
  if (n <= 0) return 0;
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  uint64_t* status = (uint64_t*)d_workspace;
  uint32_t* counter = (uint32_t*)((uint8_t*)d_workspace + (size_t)tiles * sizeof(uint64_t) + 64);
  PCG_CHECK(hipMemsetAsync(d_workspace, 0, scan_workspace_bytes(n), stream));
  hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)tiles), dim3(kScanBlock), 0, stream, d_data, n, status, counter,
                     d_total);
  PCG_CHECK(hipGetLastError());
  return 0;
}

}  // namespace pcg
