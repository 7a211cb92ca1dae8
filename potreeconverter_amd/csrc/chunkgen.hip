// GPU side of the synthetic pre-chunked point sets (chunkgen.h): records of
// every chunk and the statistics of chunks/metadata.json in one launch.
#include <hip/hip_runtime.h>

#include <vector>

#include "chunkgen.h"
#include "converter.h"
#include "pcg_common.h"

namespace pcg {

namespace {

constexpr int kGenBlock = 256;
constexpr int kGenTile = 4096;  // records per workgroup

struct GenTile {
  int32_t chunk;
  int32_t tile;  // kGenTile records from tile * kGenTile
};

// One workgroup per tile of one chunk, kGenTile / kGenBlock records per
// lane.  The minima and maxima are folded per lane in registers, then per
// workgroup in LDS, then into `stats` with one atomic per component; the
// classification histogram is counted in LDS.
__global__ __launch_bounds__(kGenBlock) void k_chunkgen(ChunkGenParams p, const GenTile* __restrict__ tiles,
                                                         const int64_t* __restrict__ offsets, uint8_t* __restrict__ out,
                                                         unsigned long long* __restrict__ stats,
                                                         unsigned long long* __restrict__ hist) {
  __shared__ unsigned long long s_min[kChunkGenStats], s_max[kChunkGenStats];
  __shared__ uint32_t s_hist[256];
  const GenTile t = tiles[blockIdx.x];
  const int64_t first = offsets[t.chunk];
  const int64_t count = offsets[t.chunk + 1] - first;
  const int32_t m = chunkgen_morton(p, t.chunk);
  if (threadIdx.x < kChunkGenStats) {
    s_min[threadIdx.x] = ~0ull;
    s_max[threadIdx.x] = 0ull;
  }
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long lmin[kChunkGenStats], lmax[kChunkGenStats];
  for (int k = 0; k < kChunkGenStats; k++) {
    lmin[k] = ~0ull;
    lmax[k] = 0ull;
  }
  const int64_t base = (int64_t)t.tile * kGenTile;
  for (int j = threadIdx.x; j < kGenTile; j += kGenBlock) {
    const int64_t i = base + j;
    if (i >= count) break;
    uint8_t rec[kChunkGenBytes];
    chunkgen_point(p, m, i, rec);
    uint8_t* dst = out + (size_t)(first + i) * kChunkGenBytes;
    for (int b = 0; b < kChunkGenBytes; b++) dst[b] = rec[b];
    double v[kChunkGenStats];
    chunkgen_values(rec, v);
    for (int k = 0; k < kChunkGenStats; k++) {
      const unsigned long long o = ordered_bits(v[k]);
      lmin[k] = o < lmin[k] ? o : lmin[k];
      lmax[k] = o > lmax[k] ? o : lmax[k];
    }
    atomicAdd(&s_hist[rec[16]], 1u);
  }
  for (int k = 0; k < kChunkGenStats; k++) {
    if (lmin[k] != ~0ull) atomicMin(&s_min[k], lmin[k]);
    if (lmax[k] != 0ull) atomicMax(&s_max[k], lmax[k]);
  }
  __syncthreads();
  if (threadIdx.x < kChunkGenStats) {
    if (s_min[threadIdx.x] != ~0ull) atomicMin(&stats[threadIdx.x], s_min[threadIdx.x]);
    if (s_max[threadIdx.x] != 0ull) atomicMax(&stats[kChunkGenStats + threadIdx.x], s_max[threadIdx.x]);
  }
  if (s_hist[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)s_hist[threadIdx.x]);
}

__global__ void k_chunkgen_init(unsigned long long* stats, unsigned long long* hist) {
  const int i = threadIdx.x;
  if (i < kChunkGenStats) stats[i] = ~0ull;
  if (i >= kChunkGenStats && i < 2 * kChunkGenStats) stats[i] = 0ull;
  hist[i] = 0ull;
}

}  // namespace

// Records of every chunk (counts[i] each, chunk i from offset
// sum(counts[0..i)) into d_out, and the ordered-bit minima / maxima and the
// classification histogram of all of them.
int chunkgen_device(const ChunkGenParams& p, const std::vector<int64_t>& counts, uint8_t* d_out,
                    std::vector<uint64_t>& stats, std::vector<int64_t>& hist) {
  hipStream_t s = device().stream;
  std::vector<int64_t> offsets(counts.size() + 1, 0);
  std::vector<GenTile> tiles;
  for (size_t i = 0; i < counts.size(); i++) {
    offsets[i + 1] = offsets[i] + counts[i];
    const int64_t nt = (counts[i] + kGenTile - 1) / kGenTile;
    if (nt > (1ll << 31)) {
      set_error("chunk too large for the generator");
      return 1;
    }
    for (int64_t k = 0; k < nt; k++) tiles.push_back(GenTile{(int32_t)i, (int32_t)k});
  }
  GenTile* dtiles = (GenTile*)workspace(WS_GEN_TILES, tiles.size() * sizeof(GenTile) + 16);
  int64_t* doff = (int64_t*)workspace(WS_GEN_OFFSETS, offsets.size() * 8);
  unsigned long long* st = (unsigned long long*)workspace(WS_GEN_STATS, (2 * kChunkGenStats + 256) * 8);
  if (!dtiles || !doff || !st) return 1;
  PCG_CHECK(hipMemcpyAsync(dtiles, tiles.data(), tiles.size() * sizeof(GenTile), hipMemcpyHostToDevice, s));
  PCG_CHECK(hipMemcpyAsync(doff, offsets.data(), offsets.size() * 8, hipMemcpyHostToDevice, s));
  unsigned long long* hi = st + 2 * kChunkGenStats;
  hipLaunchKernelGGL(k_chunkgen_init, dim3(1), dim3(256), 0, s, st, hi);
  PCG_CHECK(hipGetLastError());
  if (!tiles.empty()) {
    hipLaunchKernelGGL(k_chunkgen, dim3((unsigned)tiles.size()), dim3(kGenBlock), 0, s, p, dtiles, doff, d_out, st, hi);
    PCG_CHECK(hipGetLastError());
  }
  std::vector<unsigned long long> h(2 * kChunkGenStats + 256);
  PCG_CHECK(stream_copy(h.data(), st, h.size() * 8, hipMemcpyDeviceToHost));
  stats.assign(h.begin(), h.begin() + 2 * kChunkGenStats);
  hist.assign(256, 0);
  for (int i = 0; i < 256; i++) hist[i] = (int64_t)h[2 * kChunkGenStats + i];
  return 0;
}

}  // namespace pcg
