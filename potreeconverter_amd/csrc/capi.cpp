// extern "C" entry points (include/potree_gpu.h).
#include <string.h>

#include <algorithm>
#include <chrono>

#include <filesystem>
#include <string>
#include <vector>

#include "../../include/potree_gpu.h"
#include "chunker.h"
#include "converter.h"
#include "indexing.h"
#include "lasgen.h"
#include "sampler.h"

namespace fs = std::filesystem;
using namespace pcg;

static_assert(sizeof(LasGenParams) <= 256, "LasGenParams must fit the caller's 256-byte buffer");

extern "C" {

int pcg_init(void) {
  if (ensure_device()) return 1;
  uint8_t* scr = nullptr;  // the no-allocation scratch's address on this device
  return noalloc_scratch(&scr);
}

const char* pcg_last_error(void) { return last_error().c_str(); }

static int load_las_to_device(const char* las_path, lasfile::Header& h, uint8_t** d_las) {
  std::string err;
  if (!lasfile::read_header(las_path, h, &err)) {
    set_error(err);
    return 1;
  }
  std::vector<uint8_t> host((size_t)h.num_points * h.record_length);
  if (!host.empty() && !lasfile::read_points(las_path, h, 0, h.num_points, host.data())) {
    set_error(std::string("short read of point records: ") + las_path);
    return 1;
  }
  PCG_CHECK(hipMalloc(d_las, host.size() + 64));
  if (!host.empty()) PCG_CHECK(stream_copy(*d_las, host.data(), host.size(), hipMemcpyHostToDevice));
  return 0;
}

int pcg_do_chunking(const char* las_path, const char* outdir, double* timings) {
  if (ensure_device()) return 1;
  lasfile::Header h;
  uint8_t* d_las = nullptr;
  int rc = load_las_to_device(las_path, h, &d_las);
  if (rc) { if (d_las) (void)hipFree(d_las); return rc; }
  DeviceChunks dc;
  rc = gpu_chunking(d_las, h, dc);
  (void)hipFree(d_las);
  if (rc) { dc.release(); return rc; }
  rc = write_chunk_files(dc, outdir);
  if (timings) { timings[0] = dc.t_count; timings[1] = dc.t_distribute; timings[2] = dc.t_total; }
  dc.release();
  return rc;
}

static void fill_timings(double* t, const DeviceChunks& dc, const IndexResult& r, double tWrite);

// every source file's records in HBM (freed by the caller)
static int load_sources(const char* const* las_paths, int32_t num_sources, std::vector<ChunkSource>& src,
                        std::vector<uint8_t*>& bufs) {
  if (num_sources < 1 || !las_paths) {
    set_error("no source files");
    return 1;
  }
  for (int32_t i = 0; i < num_sources; i++) {
    lasfile::Header h;
    uint8_t* d = nullptr;
    int rc = load_las_to_device(las_paths[i], h, &d);
    if (d) bufs.push_back(d);
    if (rc) return rc;
    src.push_back(ChunkSource{d, h});
  }
  return 0;
}

static void free_sources(std::vector<uint8_t*>& bufs) {
  for (uint8_t* d : bufs) (void)hipFree(d);
  bufs.clear();
}

static std::vector<std::string> name_list(const char* const* names, int32_t n) {
  std::vector<std::string> v;
  for (int32_t i = 0; names && i < n; i++) v.push_back(names[i] ? names[i] : "");
  return v;
}

int pcg_do_chunking_sources(const char* const* las_paths, int32_t num_sources, const char* outdir, const double* min,
                            const double* max, const char* const* attributes, int32_t num_attributes,
                            double* timings) {
  if (ensure_device()) return 1;
  if ((min == nullptr) != (max == nullptr)) {
    set_error("pcg_do_chunking_sources: give both min and max, or neither");
    return 1;
  }
  std::vector<ChunkSource> src;
  std::vector<uint8_t*> bufs;
  int rc = load_sources(las_paths, num_sources, src, bufs);
  DeviceChunks dc;
  if (!rc) rc = gpu_chunking_sources(src, min, max, name_list(attributes, num_attributes), dc);
  free_sources(bufs);
  if (!rc) rc = write_chunk_files(dc, outdir);
  if (timings) { timings[0] = dc.t_count; timings[1] = dc.t_distribute; timings[2] = dc.t_total; }
  dc.release();
  return rc;
}

int pcg_convert_sources(const char* const* las_paths, int32_t num_sources, const char* outdir, const char* method,
                        const char* encoding, const char* const* attributes, int32_t num_attributes, int32_t flags,
                        double* timings) {
  if (ensure_device()) return 1;
  std::vector<ChunkSource> src;
  std::vector<uint8_t*> bufs;
  int rc = load_sources(las_paths, num_sources, src, bufs);
  DeviceChunks dc;
  if (!rc) rc = gpu_chunking_sources(src, nullptr, nullptr, name_list(attributes, num_attributes), dc);
  free_sources(bufs);
  if (rc) { dc.release(); return rc; }
  fs::create_directories(outdir);
  if (flags & 1) {
    rc = write_chunk_files(dc, outdir);
    if (rc) { dc.release(); return rc; }
  }
  IndexResult r;
  // the name is the first source's file stem (main.cpp:151-203)
  const std::string name = fs::path(las_paths[0]).stem().string();
  rc = gpu_indexing(dc, method, name, dc.pointsTotal, r, encoding ? encoding : "DEFAULT");
  if (!rc) rc = write_index_files(r, outdir);
  fill_timings(timings, dc, r, 0.0);
  dc.release();
  r.release();
  return rc;
}

static void fill_timings(double* t, const DeviceChunks& dc, const IndexResult& r, double tWrite) {
  (void)tWrite;
  if (!t) return;
  t[0] = dc.t_count;
  t[1] = dc.t_distribute;
  t[2] = dc.t_total;
  t[3] = r.t_total;
  t[4] = r.t_build;
  t[5] = r.t_sample;
  t[6] = (double)r.counters[3];
  t[7] = (double)r.counters[4];
}

int pcg_release(void) {
  if (ensure_device()) return 1;
  PCG_CHECK(hipStreamSynchronize(device().stream));
  release_workspace();
  return 0;
}

int pcg_set_option(const char* key, int64_t value) {
  const std::string k = key ? key : "";
  if (k == "plain_sampling") {
    g_plain_sampling = value != 0;
    return 0;
  }
  if (k == "ring_min_nodes") {
    if (value < 1 || value > (1 << 30)) {
      set_error("ring_min_nodes must be in [1, 2^30]");
      return 1;
    }
    g_ring_min_nodes = (int32_t)value;
    return 0;
  }
  if (k == "random_seed") {
    if (value < -1 || value > 0xFFFFFFFFll) {
      set_error("random_seed must be -1 (clock) or in [0, 2^32 - 1]");
      return 1;
    }
    g_random_seed = value;
    return 0;
  }
  if (k == "sample_groups") {
    if (value < 0 || value > kMaxSampleGroups) {
      set_error("sample_groups must be in [0, " + std::to_string(kMaxSampleGroups) + "]");
      return 1;
    }
    g_sample_groups = (int32_t)value;
    return 0;
  }
  if (k == "sample_sets") {
    if (value < 0 || value > kMaxSampleSets) {
      set_error("sample_sets must be in [0, " + std::to_string(kMaxSampleSets) + "]");
      return 1;
    }
    g_sample_sets = (int32_t)value;
    return 0;
  }
  if (k == "sample_order") {
    if (value != 0 && value != 1) {
      set_error("sample_order must be 0 (octree levels) or 1 (heights)");
      return 1;
    }
    g_sample_order = (int32_t)value;
    return 0;
  }
  struct Shape {
    const char* name;
    int32_t* value;
    int64_t lo, hi;
  };
  const Shape shapes[] = {{"ring_blocks_per_cu", &g_ring_blocks_per_cu, 1, 3},
                          {"copy_blocks", &g_copy_blocks, 256, 1 << 20},
                          {"neighbor_blocks", &g_neighbor_blocks, 256, 1 << 20},
                          {"avg_blocks", &g_avg_blocks, 64, 1 << 20},
                          {"bucket_min_nodes", &g_bucket_min_nodes, 1, 1 << 30},
                          {"accept_pipe_nodes", &g_accept_pipe_nodes, 0, 1 << 30},
                          {"ring_max_cand", &g_ring_max_cand, 64, 1 << 30},
                          {"pipe_form", &g_pipe_form, 0, 2}};
  for (const Shape& sh : shapes)
    if (k == sh.name) {
      if (value < sh.lo || value > sh.hi) {
        set_error(k + " must be in [" + std::to_string(sh.lo) + ", " + std::to_string(sh.hi) + "]");
        return 1;
      }
      *sh.value = (int32_t)value;
      return 0;
    }
  if (k == "set_policy") {
    if (value != 0 && value != 1) {
      set_error("set_policy must be 0 (points) or 1 (deep subtrees apart)");
      return 1;
    }
    g_set_policy = (int32_t)value;
    return 0;
  }
  if (k == "sample_graph") {
    g_sample_graph = value != 0;
    return 0;
  }
  if (k == "distribute_tail") {
    g_distribute_tail = value != 0;
    return 0;
  }
  if (k == "chunk_grid") {
    if (value != 0 && value != 128 && value != 256 && value != 512) {
      set_error("chunk_grid must be 0 (the reference's rule), 128, 256 or 512");
      return 1;
    }
    g_chunk_grid = (int32_t)value;
    return 0;
  }
  if (k == "chunk_cap") {
    if (value < 0) {
      set_error("chunk_cap must be 0 (the reference's rule) or positive");
      return 1;
    }
    g_chunk_cap = value;
    return 0;
  }
  set_error("unknown option: " + k);
  return 1;
}

int pcg_get_option(const char* key, int64_t* value) {
  const std::string k = key ? key : "";
  if (!value) {
    set_error("pcg_get_option: null value pointer");
    return 1;
  }
  const std::pair<const char*, int64_t> all[] = {
      {"plain_sampling", g_plain_sampling ? 1 : 0}, {"ring_min_nodes", g_ring_min_nodes},
      {"random_seed", g_random_seed},               {"sample_groups", g_sample_groups},
      {"sample_sets", g_sample_sets},               {"sample_order", g_sample_order},
      {"sample_graph", g_sample_graph ? 1 : 0},      {"set_policy", g_set_policy},
      {"ring_blocks_per_cu", g_ring_blocks_per_cu}, {"copy_blocks", g_copy_blocks},
      {"neighbor_blocks", g_neighbor_blocks},       {"avg_blocks", g_avg_blocks},
      {"bucket_min_nodes", g_bucket_min_nodes},     {"accept_pipe_nodes", g_accept_pipe_nodes},
      {"ring_max_cand", g_ring_max_cand},           {"pipe_form", g_pipe_form},
      {"distribute_tail", g_distribute_tail ? 1 : 0}, {"chunk_grid", g_chunk_grid},
      {"chunk_cap", g_chunk_cap}};
  for (const auto& kv : all)
    if (k == kv.first) {
      *value = kv.second;
      return 0;
    }
  set_error("unknown option: " + k);
  return 1;
}

static std::string encoding_or_default(const char* encoding) { return encoding ? encoding : "DEFAULT"; }

// The runtime options the autotuner searches, in the order of its output,
// with the values it tries (every value gives the same output bytes).
namespace {
struct TuneParam {
  const char* name;
  std::vector<int64_t> values;
};
const std::vector<TuneParam>& tune_params() {
  static const std::vector<TuneParam> p = {
      {"sample_groups", {1, 2, 3, 4}},
      {"sample_order", {0, 1}},
      {"sample_graph", {0, 1}},
      {"set_policy", {0, 1}},
      {"ring_min_nodes", {64, 128, 256, 512, 1024, 2048, 4096}},
      {"ring_blocks_per_cu", {1, 2, 3}},
      {"copy_blocks", {2048, 4096, 8192, 16384}},
      {"neighbor_blocks", {1024, 2048, 4096, 8192}},
      {"avg_blocks", {512, 1024, 2048, 4096}},
      {"bucket_min_nodes", {128, 512, 2048, 1 << 30}},
      {"accept_pipe_nodes", {0, 256, 1024, 4096}},
      {"ring_max_cand", {32768, 65536, 1 << 30}},
  };
  return p;
}
}  // namespace

int32_t pcg_autotune_params(void) { return (int32_t)tune_params().size(); }

const char* pcg_autotune_param_name(int32_t i) {
  const auto& p = tune_params();
  return i >= 0 && i < (int32_t)p.size() ? p[i].name : nullptr;
}

int pcg_autotune_device(const void* d_las, int64_t n, const uint8_t* las_header, const char* method,
                        int32_t passes, int32_t repeats, int64_t* chosen, double* best_s, double* default_s) {
  if (ensure_device()) return 1;
  if (!d_las || n <= 0 || !las_header || !method || passes < 1 || passes > 8 || repeats < 1 || repeats > 8) {
    set_error("pcg_autotune_device: needs device records, n > 0, a header, a method, 1-8 passes and repeats");
    return 1;
  }
  const auto& params = tune_params();
  std::vector<int64_t> cur(params.size());
  for (size_t i = 0; i < params.size(); i++)
    if (pcg_get_option(params[i].name, &cur[i])) return 1;
  // one configuration: the fastest of `repeats` conversions (chunking +
  // indexing in HBM, no files), wall time to the end of the job
  auto trial = [&](double& t) -> int {
    t = 1e30;
    for (int r = 0; r < repeats; r++) {
      const auto t0 = std::chrono::steady_clock::now();
      if (pcg_convert_device(d_las, n, las_header, method, nullptr, nullptr)) return 1;
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      t = std::min(t, dt);
    }
    return 0;
  };
  auto apply = [&](size_t i, int64_t v) { return pcg_set_option(params[i].name, v); };
  double best = 0;
  if (trial(best)) return 1;  // the first conversion also sizes the workspace
  if (trial(best)) return 1;
  if (default_s) *default_s = best;
  // coordinate descent: each parameter over its values with the others
  // fixed, a change kept when it is at least 1 % faster
  for (int pass = 0; pass < passes; pass++) {
    bool changed = false;
    for (size_t i = 0; i < params.size(); i++) {
      for (int64_t v : params[i].values) {
        if (v == cur[i]) continue;
        if (apply(i, v)) return 1;
        double t = 0;
        if (trial(t)) {
          apply(i, cur[i]);
          return 1;
        }
        if (t < 0.99 * best) {
          best = t;
          cur[i] = v;
          changed = true;
        }
      }
      if (apply(i, cur[i])) return 1;
    }
    if (!changed) break;
  }
  if (chosen)
    for (size_t i = 0; i < params.size(); i++) chosen[i] = cur[i];
  if (best_s) *best_s = best;
  return 0;
}

int pcg_convert(const char* las_path, const char* outdir, const char* method, int32_t flags, double* timings) {
  return pcg_convert_encoded(las_path, outdir, method, nullptr, flags, timings);
}

int pcg_convert_encoded(const char* las_path, const char* outdir, const char* method, const char* encoding,
                        int32_t flags, double* timings) {
  if (ensure_device()) return 1;
  lasfile::Header h;
  uint8_t* d_las = nullptr;
  int rc = load_las_to_device(las_path, h, &d_las);
  if (rc) { if (d_las) (void)hipFree(d_las); return rc; }
  DeviceChunks dc;
  rc = gpu_chunking(d_las, h, dc);
  (void)hipFree(d_las);
  if (rc) { dc.release(); return rc; }
  fs::create_directories(outdir);
  if (flags & 1) {
    rc = write_chunk_files(dc, outdir);
    if (rc) { dc.release(); return rc; }
  }
  IndexResult r;
  std::string name = fs::path(las_path).stem().string();
  rc = gpu_indexing(dc, method, name, (int64_t)h.num_points, r, encoding_or_default(encoding));
  if (!rc) rc = write_index_files(r, outdir);
  fill_timings(timings, dc, r, 0.0);
  dc.release();
  r.release();
  return rc;
}

int pcg_do_indexing(const char* outdir, const char* method, const char* name, int64_t points_total,
                    double* timings) {
  return pcg_do_indexing_encoded(outdir, method, nullptr, name, points_total, timings);
}

int pcg_do_indexing_encoded(const char* outdir, const char* method, const char* encoding, const char* name,
                            int64_t points_total, double* timings) {
  if (ensure_device()) return 1;
  DeviceChunks dc;
  int rc = read_chunk_files(outdir, dc);
  if (rc) { dc.release(); return rc; }
  IndexResult r;
  rc = gpu_indexing(dc, method, name ? name : "", points_total > 0 ? points_total : dc.numPoints, r,
                    encoding_or_default(encoding));
  if (!rc) rc = write_index_files(r, outdir);
  fill_timings(timings, dc, r, 0.0);
  dc.release();
  r.release();
  return rc;
}

static void fill_stats(double* stats, const DeviceChunks& dc, const IndexResult& r) {
  if (!stats) return;
  stats[0] = dc.t_count;
  stats[1] = dc.t_distribute;
  stats[2] = dc.t_total;
  stats[3] = r.t_total;
  stats[4] = r.t_build;
  stats[5] = r.t_sample;
  stats[6] = r.t_assemble;
  stats[7] = (double)r.octreeBytes;
  stats[8] = (double)r.numNodes;
  stats[9] = (double)r.numSampled;
  stats[10] = (double)r.counters[5];  // candidates
  stats[11] = (double)r.counters[0];  // fp64 re-checks
  stats[12] = (double)r.counters[1];  // fp32 pair tests
  uint32_t eb = (uint32_t)r.counters[2];
  float e;
  memcpy(&e, &eb, 4);
  stats[13] = (double)e;              // max |dd32 - dd64| / spacing^2 on re-checked pairs
  stats[14] = (double)r.counters[3];  // nodes sorted in libstdc++ order
  stats[15] = (double)r.counters[4];  // nodes resolved sequentially
  stats[16] = (double)dc.chunks.size();
  stats[17] = (double)r.buildRounds;
  stats[18] = (double)r.octreeDepth;
  stats[19] = (double)dc.ms_count_kernel;
  stats[20] = (double)dc.ms_distribute_kernels;
  stats[21] = r.t_setup;
  stats[22] = r.t_order;
  stats[23] = r.t_hierarchy;
}

int pcg_convert_device(const void* d_las, int64_t n, const uint8_t* las_header, const char* method,
                       const char* outdir, double* stats) {
  return pcg_convert_device_encoded(d_las, n, las_header, method, nullptr, outdir, stats);
}

int pcg_convert_device_encoded(const void* d_las, int64_t n, const uint8_t* las_header, const char* method,
                               const char* encoding, const char* outdir, double* stats) {
  if (ensure_device()) return 1;
  lasfile::Header h;
  h.point_format = las_header[104] & 0x3F;
  h.record_length = lasfile::rd<uint16_t>(las_header + 105);
  h.num_points = (uint64_t)n;
  for (int i = 0; i < 3; i++) h.scale[i] = lasfile::rd<double>(las_header + 131 + 8 * i);
  for (int i = 0; i < 3; i++) h.offset[i] = lasfile::rd<double>(las_header + 155 + 8 * i);
  h.max[0] = lasfile::rd<double>(las_header + 179); h.min[0] = lasfile::rd<double>(las_header + 187);
  h.max[1] = lasfile::rd<double>(las_header + 195); h.min[1] = lasfile::rd<double>(las_header + 203);
  h.max[2] = lasfile::rd<double>(las_header + 211); h.min[2] = lasfile::rd<double>(las_header + 219);
  DeviceChunks dc;
  int rc = gpu_chunking((const uint8_t*)d_las, h, dc);
  if (rc) { dc.release(); return rc; }
  IndexResult r;
  rc = gpu_indexing(dc, method, "device", n, r, encoding_or_default(encoding));
  if (!rc && outdir) rc = write_index_files(r, outdir);
  fill_stats(stats, dc, r);
  dc.release();
  r.release();
  return rc;
}

int pcg_do_indexing_device(const void* d_records, int32_t num_chunks, const char* const* chunk_names,
                           const int64_t* chunk_counts, const char* chunks_metadata, const char* method,
                           const char* encoding, const char* name, const char* outdir, double* stats) {
  if (ensure_device()) return 1;
  if (num_chunks < 0 || (num_chunks > 0 && (!chunk_names || !chunk_counts || !d_records)) || !chunks_metadata) {
    set_error("pcg_do_indexing_device: missing chunk list, records or metadata");
    return 1;
  }
  DeviceChunks dc;
  if (!parse_chunk_metadata(chunks_metadata, dc.min, dc.max, dc.attributes)) return 1;
  dc.bpp = dc.attributes.bytes;
  int64_t off = 0;
  for (int32_t i = 0; i < num_chunks; i++) {
    ChunkDesc d;
    d.id = chunk_names[i] ? chunk_names[i] : "";
    bool ok = d.id.size() >= 1 && d.id.size() <= 21 && d.id[0] == 'r' && chunk_counts[i] >= 0;
    for (size_t c = 1; ok && c < d.id.size(); c++) ok = d.id[c] >= '0' && d.id[c] <= '7';
    if (!ok) {
      set_error("bad chunk " + std::to_string(i) + ": name \"" + d.id + "\", count " + std::to_string(chunk_counts[i]));
      return 1;
    }
    // chunks in ascending name, the canonical order in which the library
    // reads a chunks/ directory (getChunks, Converter/src/indexer.cpp:53-179,
    // takes the directory listing's order)
    if (i > 0 && !(dc.chunks.back().id < d.id)) {
      set_error("chunk names must be distinct and in ascending order: " + dc.chunks.back().id + ", " + d.id);
      return 1;
    }
    d.count = chunk_counts[i];
    d.offset = off;
    off += d.count;
    d.level = (int32_t)d.id.size() - 1;
    chunk_box(d.id, dc.min, dc.max, d.min, d.max);
    dc.chunks.push_back(d);
  }
  dc.numPoints = off;
  dc.pointsTotal = off;
  // the octree build reorders the records in place: work on a copy
  dc.d_records = (uint8_t*)workspace(WS_RECORDS, (size_t)off * dc.bpp + 64);
  dc.owned = false;
  if (!dc.d_records) return 1;
  hipStream_t s = device().stream;
  if (off > 0) PCG_CHECK(hipMemcpyAsync(dc.d_records, d_records, (size_t)off * dc.bpp, hipMemcpyDeviceToDevice, s));
  IndexResult r;
  int rc = gpu_indexing(dc, method, name ? name : "", off, r, encoding_or_default(encoding));
  if (!rc && outdir) {
    fs::create_directories(outdir);
    rc = write_index_files(r, outdir);
  }
  fill_stats(stats, dc, r);
  dc.release();
  r.release();
  return rc;
}

int pcg_sample_subtree(const char* method, int32_t num_nodes, const int32_t* children, const double* bounds,
                       const int32_t* levels, const uint8_t* sampled, const int64_t* counts, const void* records,
                       int32_t bpp, int32_t rgb_offset, const double scale[3], const double offset[3],
                       double base_spacing, int64_t* out_counts, void* out_records, int64_t out_capacity,
                       int32_t* out_children, int32_t* events, int32_t* num_events) {
  if (ensure_device()) return 1;
  const std::string m = method ? method : "";
  if (m != "poisson" && m != "poisson_average" && m != "random") {
    set_error("unknown sampling method: " + m);
    return 1;
  }
  if (num_nodes < 1 || !children || !bounds || !levels || !counts || !scale || !offset || !out_counts ||
      !out_children || !events || !num_events) {
    set_error("pcg_sample_subtree: missing argument");
    return 1;
  }
  if (bpp < 12 || bpp > 4096 || (m == "poisson_average" && (rgb_offset < 12 || rgb_offset + 6 > bpp))) {
    set_error("pcg_sample_subtree: bytes per point " + std::to_string(bpp) + " / rgb offset " +
              std::to_string(rgb_offset) + " out of range");
    return 1;
  }
  // a tree rooted at node 0: one parent per node, levels one apart, records
  // only at the nodes without children
  std::vector<int32_t> parent(num_nodes, -1);
  std::vector<int64_t> offsets(num_nodes, 0);
  int64_t total = 0;
  for (int32_t i = 0; i < num_nodes; i++) {
    bool inner = false;
    for (int c = 0; c < 8; c++) {
      const int32_t ch = children[8 * i + c];
      if (ch == -1) continue;
      if (ch <= 0 || ch >= num_nodes || parent[ch] != -1 || levels[ch] != levels[i] + 1) {
        set_error("pcg_sample_subtree: bad child " + std::to_string(ch) + " of node " + std::to_string(i));
        return 1;
      }
      parent[ch] = i;
      inner = true;
    }
    if (counts[i] < 0 || (inner && counts[i] != 0) || levels[i] < 0 || levels[i] > 60) {
      set_error("pcg_sample_subtree: node " + std::to_string(i) + " has " + std::to_string(counts[i]) +
                " records at level " + std::to_string(levels[i]) + " (only nodes without children hold records)");
      return 1;
    }
    offsets[i] = total;
    total += counts[i];
  }
  for (int32_t i = 1; i < num_nodes; i++) {
    int32_t a = i, steps = 0;
    while (a > 0 && steps <= num_nodes) { a = parent[a]; steps++; }
    if (a != 0) {
      set_error("pcg_sample_subtree: node " + std::to_string(i) + " is not below node 0");
      return 1;
    }
  }
  if (total > 0 && !records) {
    set_error("pcg_sample_subtree: missing records");
    return 1;
  }
  DeviceChunks dc;
  dc.bpp = bpp;
  dc.numPoints = total;
  dc.pointsTotal = total;
  for (int a = 0; a < 3; a++) {
    dc.attributes.posScale[a] = scale[a];
    dc.attributes.posOffset[a] = offset[a];
  }
  dc.attributes.bytes = bpp;
  dc.d_records = (uint8_t*)workspace(WS_RECORDS, (size_t)total * bpp + 64);
  if (!dc.d_records) return 1;
  if (total > 0) PCG_CHECK(stream_copy(dc.d_records, records, (size_t)total * bpp, hipMemcpyHostToDevice));
  SubtreeSpec spec;
  spec.numNodes = num_nodes;
  spec.children = children;
  spec.bounds = bounds;
  spec.levels = levels;
  spec.sampled = sampled;
  spec.counts = counts;
  spec.offsets = offsets.data();
  spec.baseSpacing = base_spacing;
  spec.rgbOffset = rgb_offset;
  IndexResult r;
  SubtreeResult res;
  int rc = gpu_indexing(dc, m, "", total, r, "DEFAULT", &spec, &res);
  if (rc) return rc;
  // every node's records back to back in node order, gathered on the GPU
  int64_t outTotal = 0;
  std::vector<CopyDesc> copies;
  for (int32_t i = 0; i < num_nodes; i++) {
    if (res.count[i] > 0)
      copies.push_back(CopyDesc{res.ptr[i], (uint64_t)outTotal * bpp, (uint64_t)res.count[i] * bpp});
    outTotal += res.count[i];
  }
  if (outTotal > out_capacity || (outTotal > 0 && !out_records)) {
    set_error("pcg_sample_subtree: the results hold " + std::to_string(outTotal) + " records, the output buffer " +
              std::to_string(out_capacity));
    return 1;
  }
  if (outTotal > 0) {
    hipStream_t s = device().stream;
    uint8_t* d_out = (uint8_t*)workspace(WS_OCTREE, (size_t)outTotal * bpp + 64);
    CopyDesc* d_desc = (CopyDesc*)workspace(WS_COPYDESC, copies.size() * sizeof(CopyDesc) + 16);
    if (!d_out || !d_desc) return 1;
    PCG_CHECK(hipMemcpyAsync(d_desc, copies.data(), copies.size() * sizeof(CopyDesc), hipMemcpyHostToDevice, s));
    if (gather_ranges(d_desc, (int64_t)copies.size(), d_out, s)) return 1;
    PCG_CHECK(stream_copy(out_records, d_out, (size_t)outTotal * bpp, hipMemcpyDeviceToHost));
  }
  for (int32_t i = 0; i < num_nodes; i++) out_counts[i] = res.count[i];
  memcpy(out_children, res.children.data(), sizeof(int32_t) * 8 * (size_t)num_nodes);
  for (size_t e = 0; e < res.events.size(); e++) {
    events[2 * e] = res.events[e].first;
    events[2 * e + 1] = res.events[e].second;
  }
  *num_events = (int32_t)res.events.size();
  return 0;
}

int pcg_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
  if (ensure_device()) return 1;
  size_t f = 0, t = 0;
  PCG_CHECK(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return 0;
}

uint64_t pcg_workspace_bytes(void) { return workspace_total(); }

int pcg_device_alloc(void** ptr, uint64_t bytes) {
  if (ensure_device()) return 1;
  PCG_CHECK(hipMalloc(ptr, bytes ? bytes : 16));
  return 0;
}

int pcg_device_free(void* ptr) {
  if (ptr) PCG_CHECK(hipFree(ptr));
  return 0;
}

int pcg_copy(void* dst, const void* src, uint64_t bytes, int32_t kind) {
  if (ensure_device()) return 1;
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  PCG_CHECK(stream_copy(dst, src, bytes, k));
  return 0;
}

int pcg_synchronize(void) {
  if (ensure_device()) return 1;
  PCG_CHECK(hipDeviceSynchronize());
  return 0;
}

// chunker parameters of a counting-only pass (no attribute statistics)
static int count_params(int64_t n, int32_t bpp, const double las_scale[3], const double las_offset[3],
                        const double pos_scale[3], const double pos_offset[3], const double min[3],
                        const double max[3], int32_t grid_size, hipStream_t s) {
  if (n < 0 || bpp < 12 || bpp > 128 || grid_size < 1 || grid_size > 512) {
    set_error("count_cells: need n >= 0, 12 <= bpp <= 128 and 1 <= grid_size <= 512");
    return 1;
  }
  ChunkerParams P;
  memset(&P, 0, sizeof(P));
  Vec3 mn(min[0], min[1], min[2]), mx(max[0], max[1], max[2]);
  for (int a = 0; a < 3; a++) {
    P.las_scale[a] = las_scale[a]; P.las_offset[a] = las_offset[a];
    P.pos_scale[a] = pos_scale[a]; P.pos_offset[a] = pos_offset[a];
    P.min[a] = min[a];
  }
  P.cube = (mx - mn).max();
  P.grid_size = grid_size;
  P.bpp_in = bpp;
  P.class_src = -1;
  P.point_format = -1;  // counts only: no attribute statistics
  return chunker_set_params(P, s);
}

int pcg_count_cells_noalloc(const void* d_las, int64_t n, int32_t bpp, const double las_scale[3],
                            const double las_offset[3], const double pos_scale[3], const double pos_offset[3],
                            const double min[3], const double max[3], int32_t grid_size, int32_t* d_grid,
                            void* stream) {
  if (ensure_device()) return 1;
  hipStream_t s = stream ? (hipStream_t)stream : device().stream;
  if (count_params(n, bpp, las_scale, las_offset, pos_scale, pos_offset, min, max, grid_size, s)) return 1;
  const int64_t cells = (int64_t)grid_size * grid_size * grid_size;
  int32_t herr = 0;
  if (chunker_count_noalloc((const uint8_t*)d_las, n, d_grid, cells, device().num_cus, s, &herr)) return 1;
  if (herr & 1) {
    set_error("encountered point outside bounding box.");
    return 123;
  }
  return 0;
}

int pcg_count_cells(const void* d_las, int64_t n, int32_t bpp, const double las_scale[3], const double las_offset[3],
                    const double pos_scale[3], const double pos_offset[3], const double min[3], const double max[3],
                    int32_t grid_size, int32_t* d_grid, void* stream) {
  if (ensure_device()) return 1;
  hipStream_t s = stream ? (hipStream_t)stream : device().stream;
  ChunkerParams P;
  memset(&P, 0, sizeof(P));
  Vec3 mn(min[0], min[1], min[2]), mx(max[0], max[1], max[2]);
  for (int a = 0; a < 3; a++) {
    P.las_scale[a] = las_scale[a]; P.las_offset[a] = las_offset[a];
    P.pos_scale[a] = pos_scale[a]; P.pos_offset[a] = pos_offset[a];
    P.min[a] = min[a];
  }
  P.cube = (mx - mn).max();
  P.grid_size = grid_size;
  P.bpp_in = bpp;
  P.class_src = -1;
  P.point_format = -1;  // counts only: no attribute statistics
  if (chunker_set_params(P, s)) return 1;
  int64_t cells = (int64_t)grid_size * grid_size * grid_size;
  void *stats = nullptr, *hist = nullptr, *err = nullptr;
  PCG_CHECK(hipMalloc(&stats, sizeof(uint64_t) * 2 * kMaxStats));
  PCG_CHECK(hipMalloc(&hist, 256 * sizeof(unsigned long long)));
  PCG_CHECK(hipMalloc(&err, 16));
  PCG_CHECK(hipMemsetAsync(d_grid, 0, sizeof(int32_t) * cells, s));
  PCG_CHECK(hipMemsetAsync(err, 0, 16, s));
  int rc = chunker_count((const uint8_t*)d_las, n, d_grid, nullptr, (uint64_t*)stats, (uint64_t*)stats + kMaxStats,
                         (unsigned long long*)hist, (int32_t*)err, device().num_cus, s);
  int32_t herr = 0;
  if (!rc) {
    PCG_CHECK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    PCG_CHECK(hipStreamSynchronize(s));
  }
  (void)hipFree(stats); (void)hipFree(hist); (void)hipFree(err);
  if (rc) return rc;
  if (herr & 1) {
    set_error("encountered point outside bounding box.");
    return 123;
  }
  return 0;
}

int pcg_lasgen_device(const void* params, int64_t first, int64_t count, uint8_t* d_out, int32_t* d_qminmax) {
  if (ensure_device()) return 1;
  return lasgen_device(*(const LasGenParams*)params, first, count, d_out, d_qminmax, device().stream);
}

void pcg_lasgen_bounds(const LasGenParams* p, const int32_t qmin[3], const int32_t qmax[3], double mn[3], double mx[3]);

int pcg_lasgen_header(const void* params, const int32_t qminmax[6], uint8_t* header375) {
  const LasGenParams* p = (const LasGenParams*)params;
  lasfile::Header h;
  h.point_format = (uint8_t)p->point_format;
  h.record_length = (uint16_t)lasgen_record_length(*p);
  h.num_points = (uint64_t)p->n;
  for (int i = 0; i < 3; i++) {
    h.scale[i] = 1.0 / p->inv_scale[i];
    h.offset[i] = p->offset[i];
  }
  pcg_lasgen_bounds(p, qminmax, qminmax + 3, h.min, h.max);
  memset(header375, 0, 375);
  memcpy(header375, "LASF", 4);
  header375[24] = 1;
  header375[25] = 4;
  lasfile::wr<uint16_t>(header375 + 94, 375);
  lasfile::wr<uint32_t>(header375 + 96, 375);
  header375[104] = h.point_format;
  lasfile::wr<uint16_t>(header375 + 105, h.record_length);
  // legacy point count: 0 for formats 6-10 and above 2^32 - 1 points (LAS 1.4)
  lasfile::wr<uint32_t>(header375 + 107,
                        h.num_points <= 0xFFFFFFFFull && h.point_format < 6 ? (uint32_t)h.num_points : 0u);
  for (int i = 0; i < 3; i++) lasfile::wr<double>(header375 + 131 + 8 * i, h.scale[i]);
  for (int i = 0; i < 3; i++) lasfile::wr<double>(header375 + 155 + 8 * i, h.offset[i]);
  lasfile::wr<double>(header375 + 179, h.max[0]); lasfile::wr<double>(header375 + 187, h.min[0]);
  lasfile::wr<double>(header375 + 195, h.max[1]); lasfile::wr<double>(header375 + 203, h.min[1]);
  lasfile::wr<double>(header375 + 211, h.max[2]); lasfile::wr<double>(header375 + 219, h.min[2]);
  lasfile::wr<uint64_t>(header375 + 247, h.num_points);
  return 0;
}

}  // extern "C"
