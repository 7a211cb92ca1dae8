// Host side of the synthetic pre-chunked point sets (chunkgen.h): counts,
// chunks/metadata.json, the host writer of a chunks/ directory and the GPU
// generator's entry point.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <filesystem>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/potree_gpu.h"
#include "attributes.h"
#include "chunkgen.h"
#include "converter.h"

namespace fs = std::filesystem;

namespace pcg {

int chunkgen_device(const ChunkGenParams& p, const std::vector<int64_t>& counts, uint8_t* d_out,
                    std::vector<uint64_t>& stats, std::vector<int64_t>& hist);

namespace {

bool valid(const ChunkGenParams& p) {
  if (p.num_chunks < 1 || p.num_chunks > kChunkGenCells || (p.num_chunks & (p.num_chunks - 1))) {
    set_error("num_chunks must be a power of two in [1, 4096]");
    return false;
  }
  if (!(p.median > 0) || !(p.sigma >= 0) || p.min_count < 0 || p.max_count < p.min_count ||
      p.max_count > (1ll << 40)) {
    set_error("chunk count parameters out of range");
    return false;
  }
  if (!(p.cube_size > 0) || !(p.scale > 0) || p.cube_size / p.scale >= 2147483647.0) {
    set_error("cube size / scale out of the int32 range");
    return false;
  }
  return true;
}

std::vector<int64_t> counts_of(const ChunkGenParams& p) {
  std::vector<int64_t> c(p.num_chunks);
  const uint64_t s = p.seed ^ 0xC0C0000000000000ull;
  for (int32_t i = 0; i < p.num_chunks; i++) {
    double u1 = lasgen_u01(lasgen_hash(s, (uint64_t)i, 100));
    const double u2 = lasgen_u01(lasgen_hash(s, (uint64_t)i, 101));
    if (u1 < 1e-300) u1 = 1e-300;
    const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double v = p.median * exp(p.sigma * g);
    v = std::min(std::max(v, (double)p.min_count), (double)p.max_count);
    c[i] = (int64_t)llround(v);
  }
  return c;
}

// output attributes of the generated records (LAS point format 2 + the three
// extra-bytes attributes) with the folded statistics
bool attributes_of(const ChunkGenParams& p, const std::vector<uint64_t>& stats, const std::vector<int64_t>& hist,
                   int64_t total, Attributes& A) {
  lasfile::Header h;
  h.point_format = 2;
  h.record_length = 26 + 7;
  for (int a = 0; a < 3; a++) {
    h.scale[a] = p.scale;
    h.offset[a] = 0.0;
    h.min[a] = 0.0;
    h.max[a] = p.cube_size;
  }
  lasfile::ExtraBytesDescriptor g, u, b;
  g.data_type = 9; g.name = "gauss"; g.description = "float32 N(0,1)";
  u.data_type = 3; u.name = "uniform16"; u.description = "uint16 uniform";
  b.data_type = 2; b.name = "uniform8"; b.description = "int8 uniform";
  h.extra = {g, u, b};
  ChunkerParams P;
  if (!output_attributes_for(h, A, P)) return false;
  if (A.bytes != kChunkGenBytes) {
    set_error("generated record layout differs from the output attributes");
    return false;
  }
  if (total <= 0) return true;
  auto mn = [&](int k) { return from_ordered_bits(stats[k]); };
  auto mx = [&](int k) { return from_ordered_bits(stats[kChunkGenStats + k]); };
  Attribute& pos = A.list[0];
  for (int a = 0; a < 3; a++) {
    pos.min[a] = p.scale * mn(a) + 0.0;
    pos.max[a] = p.scale * mx(a) + 0.0;
  }
  int k = 3;
  for (size_t i = 1; i < A.list.size(); i++) {
    Attribute& at = A.list[i];
    for (int e = 0; e < at.numElements && e < 3; e++, k++) {
      at.min[e] = mn(k);
      at.max[e] = mx(k);
    }
  }
  Attribute* cls = A.get("classification");
  for (int i = 0; i < 256; i++) cls->histogram[i] = hist[i];
  return true;
}

std::string metadata_of(const ChunkGenParams& p, const Attributes& A) {
  return chunk_metadata_json(Vec3(0, 0, 0), Vec3(p.cube_size, p.cube_size, p.cube_size), A);
}

}  // namespace
}  // namespace pcg

using namespace pcg;

static_assert(sizeof(ChunkGenParams) <= 128, "ChunkGenParams must fit the caller's 128-byte buffer");

extern "C" {

int pcg_chunkgen_params(int32_t num_chunks, double median, double sigma, int64_t min_count, int64_t max_count,
                        uint64_t seed, void* params) {
  ChunkGenParams p;
  memset(&p, 0, sizeof(p));
  p.seed = seed;
  p.cube_size = 2048.0;
  p.scale = 0.001;
  p.median = median;
  p.sigma = sigma;
  p.min_count = min_count;
  p.max_count = max_count;
  p.num_chunks = num_chunks;
  if (!valid(p)) return 1;
  memcpy(params, &p, sizeof(p));
  return 0;
}

int pcg_chunkgen_counts(const void* params, int64_t* counts, char* names) {
  const ChunkGenParams& p = *(const ChunkGenParams*)params;
  if (!valid(p)) return 1;
  const std::vector<int64_t> c = counts_of(p);
  for (int32_t i = 0; i < p.num_chunks; i++) {
    if (counts) counts[i] = c[i];
    if (names) chunkgen_name(chunkgen_morton(p, i), names + 6 * i);
  }
  return 0;
}

void pcg_chunkgen_fill(const void* params, int32_t chunk, int64_t first, int64_t count, uint8_t* out) {
  const ChunkGenParams& p = *(const ChunkGenParams*)params;
  const int32_t m = chunkgen_morton(p, chunk);
  for (int64_t i = 0; i < count; i++) chunkgen_point(p, m, first + i, out + (size_t)i * kChunkGenBytes);
}

int pcg_chunkgen_write(const void* params, const char* outdir) {
  const ChunkGenParams& p = *(const ChunkGenParams*)params;
  if (!valid(p)) return 1;
  const std::vector<int64_t> counts = counts_of(p);
  const std::string dir = std::string(outdir) + "/chunks";
  std::error_code ec;
  fs::create_directories(dir, ec);
  if (ec) {
    set_error("cannot create " + dir);
    return 1;
  }
  // per worker: statistics and histogram of the chunks it wrote
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<uint64_t>> wstats(nt, std::vector<uint64_t>(2 * kChunkGenStats));
  std::vector<std::vector<int64_t>> whist(nt, std::vector<int64_t>(256, 0));
  std::atomic<int32_t> next(0);
  std::atomic<bool> failed(false);
  auto work = [&](int w) {
    std::vector<uint64_t>& st = wstats[w];
    for (int k = 0; k < kChunkGenStats; k++) { st[k] = ~0ull; st[kChunkGenStats + k] = 0; }
    std::vector<uint8_t> buf;
    for (int32_t i; (i = next.fetch_add(1)) < p.num_chunks;) {
      const int32_t m = chunkgen_morton(p, i);
      buf.resize((size_t)counts[i] * kChunkGenBytes);
      for (int64_t j = 0; j < counts[i]; j++) {
        uint8_t* r = buf.data() + (size_t)j * kChunkGenBytes;
        chunkgen_point(p, m, j, r);
        double v[kChunkGenStats];
        chunkgen_values(r, v);
        for (int k = 0; k < kChunkGenStats; k++) {
          const uint64_t o = ordered_bits(v[k]);
          st[k] = std::min(st[k], o);
          st[kChunkGenStats + k] = std::max(st[kChunkGenStats + k], o);
        }
        whist[w][r[16]]++;
      }
      char name[6];
      chunkgen_name(m, name);
      std::ofstream f(dir + "/" + name + ".bin", std::ios::binary);
      f.write((const char*)buf.data(), (std::streamsize)buf.size());
      if (!f) failed = true;
    }
  };
  std::vector<std::thread> th;
  for (int w = 0; w < nt; w++) th.emplace_back(work, w);
  for (auto& t : th) t.join();
  if (failed) {
    set_error("short write under " + dir);
    return 1;
  }
  std::vector<uint64_t> stats(2 * kChunkGenStats);
  std::vector<int64_t> hist(256, 0);
  for (int k = 0; k < kChunkGenStats; k++) {
    stats[k] = ~0ull;
    stats[kChunkGenStats + k] = 0;
    for (int w = 0; w < nt; w++) {
      stats[k] = std::min(stats[k], wstats[w][k]);
      stats[kChunkGenStats + k] = std::max(stats[kChunkGenStats + k], wstats[w][kChunkGenStats + k]);
    }
  }
  int64_t total = 0;
  for (int64_t c : counts) total += c;
  for (int w = 0; w < nt; w++)
    for (int i = 0; i < 256; i++) hist[i] += whist[w][i];
  Attributes A;
  if (!attributes_of(p, stats, hist, total, A)) return 1;
  std::ofstream mf(dir + "/metadata.json", std::ios::binary);
  mf << metadata_of(p, A);
  if (!mf) {
    set_error("cannot write " + dir + "/metadata.json");
    return 1;
  }
  return 0;
}

int pcg_chunkgen_device(const void* params, void* d_records, char* metadata, int64_t metadata_capacity) {
  if (ensure_device()) return 1;
  const ChunkGenParams& p = *(const ChunkGenParams*)params;
  if (!valid(p)) return 1;
  const std::vector<int64_t> counts = counts_of(p);
  std::vector<uint64_t> stats;
  std::vector<int64_t> hist;
  if (chunkgen_device(p, counts, (uint8_t*)d_records, stats, hist)) return 1;
  int64_t total = 0;
  for (int64_t c : counts) total += c;
  Attributes A;
  if (!attributes_of(p, stats, hist, total, A)) return 1;
  const std::string text = metadata_of(p, A);
  if ((int64_t)text.size() + 1 > metadata_capacity) {
    set_error("metadata buffer too small: " + std::to_string(text.size() + 1) + " bytes needed");
    return 1;
  }
  memcpy(metadata, text.c_str(), text.size() + 1);
  return 0;
}

}  // extern "C"
