// Device-resident doChunking: count -> pyramid/LUT -> stable scatter.
// Restates Converter/src/chunker_countsort_laszip.cpp:1206-1261 with the
// per-point work on gfx950 (chunker.hip) and the chunk list on the host.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <filesystem>
#include <fstream>

#include "chunker.h"
#include "converter.h"
#include "scan.h"

namespace fs = std::filesystem;

namespace pcg {

static std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const std::string& last_error() { return g_error; }

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Device& device() {
  static Device d;
  return d;
}

int32_t g_chunk_grid = 0;
int64_t g_chunk_cap = 0;
bool g_distribute_tail = true;

int ensure_device() {
  static bool ready = false;
  if (ready) return 0;
  Device& d = device();
  int count = 0;
  PCG_CHECK(hipGetDeviceCount(&count));
  if (count < 1) {
    set_error("no HIP device available");
    return 1;
  }
  PCG_CHECK(hipSetDevice(d.id));
  hipDeviceProp_t prop;
  PCG_CHECK(hipGetDeviceProperties(&prop, d.id));
  d.num_cus = prop.multiProcessorCount;
  PCG_CHECK(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
  ready = true;
  return 0;
}

void DeviceChunks::release() {
  if (d_records && owned) (void)hipFree(d_records);
  d_records = nullptr;
  owned = false;
}

namespace {
struct WsEntry {
  void* p = nullptr;
  size_t bytes = 0;
};
WsEntry g_ws[WS_COUNT];
}  // namespace

void* workspace(int slot, size_t bytes) {
  if (slot < 0 || slot >= WS_COUNT) {
    set_error("bad workspace slot");
    return nullptr;
  }
  WsEntry& e = g_ws[slot];
  if (e.p && e.bytes >= bytes) return e.p;
  if (e.p) (void)hipFree(e.p);
  e.p = nullptr;
  e.bytes = 0;
  const size_t want = (bytes ? bytes : 16) + bytes / 8;
  hipError_t err = hipMalloc(&e.p, want);
  if (err != hipSuccess) {
    e.p = nullptr;
    set_error(std::string("hipMalloc of ") + std::to_string(want) + " bytes failed: " + hipGetErrorString(err));
    return nullptr;
  }
  e.bytes = want;
  return e.p;
}

size_t workspace_bytes(int slot) { return slot >= 0 && slot < WS_COUNT ? g_ws[slot].bytes : 0; }

size_t workspace_total() {
  size_t t = 0;
  for (const WsEntry& e : g_ws) t += e.bytes;
  return t;
}

void release_workspace_slot(int slot) {
  if (slot < 0 || slot >= WS_COUNT || !g_ws[slot].p) return;
  (void)hipFree(g_ws[slot].p);
  g_ws[slot] = WsEntry();
}

void release_workspace() {
  for (WsEntry& e : g_ws) {
    if (e.p) (void)hipFree(e.p);
    e = WsEntry();
  }
}

void child_box(const Vec3& min, const Vec3& max, int index, Vec3& cmin, Vec3& cmax) {
  // converter_utils.h:106-136
  Vec3 size = max - min;
  Vec3 center = min + (size * 0.5);
  cmin.x = (index & 4) ? center.x : min.x; cmax.x = (index & 4) ? max.x : center.x;
  cmin.y = (index & 2) ? center.y : min.y; cmax.y = (index & 2) ? max.y : center.y;
  cmin.z = (index & 1) ? center.z : min.z; cmax.z = (index & 1) ? max.z : center.z;
}

void chunk_box(const std::string& id, const Vec3& min, const Vec3& max, Vec3& cmin, Vec3& cmax) {
  cmin = min;
  cmax = max;
  for (size_t i = 1; i < id.size(); i++) {
    Vec3 a, b;
    child_box(cmin, cmax, id[i] - '0', a, b);
    cmin = a;
    cmax = b;
  }
}

// chunker_countsort_laszip.cpp:82-115
static std::string node_id(int level, int64_t gridSize, int64_t x, int64_t y, int64_t z) {
  std::string id = "r";
  int64_t cur = gridSize;
  for (int i = 0; i < level; i++) {
    int index = 0;
    if (x >= cur / 2) { index += 4; x -= cur / 2; }
    if (y >= cur / 2) { index += 2; y -= cur / 2; }
    if (z >= cur / 2) { index += 1; z -= cur / 2; }
    id += char('0' + index);
    cur /= 2;
  }
  return id;
}

// a workspace buffer of the job
struct DevBuf {
  void* p = nullptr;
  int alloc(int slot, size_t bytes) {
    p = workspace(slot, bytes);
    return p ? 0 : 1;
  }
  template <class T>
  T* as() { return (T*)p; }
};

int gpu_chunking(const uint8_t* d_las, const lasfile::Header& h, DeviceChunks& out) {
  return gpu_chunking_sources(std::vector<ChunkSource>{ChunkSource{d_las, h}}, nullptr, nullptr, {}, out);
}

int gpu_chunking_sources(const std::vector<ChunkSource>& src, const double* cubeMin, const double* cubeMax,
                         const std::vector<std::string>& requested, DeviceChunks& out) {
  if (ensure_device()) return 1;
  Device& dev = device();
  hipStream_t s = dev.stream;
  double t0 = now_s();
  out.release();
  const int S = (int)src.size();
  if (S < 1) {
    set_error("no sources");
    return 1;
  }
  // output attributes of all sources and each source's transcode and
  // statistics tables (PotreeConverter.h:195-315)
  std::vector<lasfile::Header> hs;
  for (const ChunkSource& c : src) hs.push_back(c.header);
  std::vector<ChunkerParams> Ps;
  std::vector<std::vector<std::pair<std::string, int>>> statTargets;
  if (!output_attributes_for_sources(hs, requested, out.attributes, Ps, statTargets)) return 1;
  Attributes& A = out.attributes;

  // main.cpp:215-284 (computeStats over the sources' headers, or the
  // caller's box) then chunker_countsort_laszip.cpp:205-207
  Vec3 hmin, hmax;
  if (cubeMin && cubeMax) {
    hmin = Vec3(cubeMin[0], cubeMin[1], cubeMin[2]);
    hmax = Vec3(cubeMax[0], cubeMax[1], cubeMax[2]);
  } else {
    hmin = Vec3(INFINITY, INFINITY, INFINITY);
    hmax = Vec3(-INFINITY, -INFINITY, -INFINITY);
    for (const auto& h : hs)
      for (int k = 0; k < 3; k++) {
        hmin[k] = std::min(hmin[k], h.min[k]);
        hmax[k] = std::max(hmax[k], h.max[k]);
      }
    const double statsCube = (hmax - hmin).max();
    if (statsCube == 0.0) {
      set_error("invalid bounding box. at least one axis has a size of zero.");
      return 123;
    }
    hmax = hmin + statsCube;
  }
  double cube = (hmax - hmin).max();
  if (!(cube > 0.0)) {
    set_error("invalid bounding box. at least one axis has a size of zero.");
    return 123;
  }
  int64_t N = 0;
  std::vector<int64_t> firstPoint(S + 1, 0), idOff(S + 1, 0);
  for (int i = 0; i < S; i++) {
    N += (int64_t)hs[i].num_points;
    firstPoint[i + 1] = N;
    // each source's id slice starts 16-byte aligned (4 ids)
    idOff[i + 1] = idOff[i] + (((int64_t)hs[i].num_points + 3) & ~3ll);
  }
  if (N >= (int64_t)4000000000LL) {
    set_error("point counts of 4e9 and above are not supported");
    return 1;
  }
  out.pointsTotal = N;
  out.numPoints = N;
  out.bpp = A.bytes;
  out.min = hmin;
  out.max = hmin + cube;
  // :1210-1220 (a test option may force both, as on a larger input)
  int64_t cap = g_chunk_cap > 0 ? g_chunk_cap : std::min(N / 20, (int64_t)10000000);
  int32_t G = g_chunk_grid > 0 ? g_chunk_grid : (N < 100000000 ? 128 : (N < 500000000 ? 256 : 512));
  int32_t Lmax = G == 128 ? 7 : (G == 256 ? 8 : 9);
  out.gridSize = G;
  out.maxPointsPerChunk = cap;
  for (int i = 0; i < S; i++) {
    ChunkerParams& P = Ps[i];
    for (int a = 0; a < 3; a++) {
      P.las_scale[a] = hs[i].scale[a];
      P.las_offset[a] = hs[i].offset[a];
      P.pos_scale[a] = A.posScale[a];
      P.pos_offset[a] = A.posOffset[a];
      P.min[a] = hmin[a];
    }
    P.cube = cube;
    P.grid_size = G;
  }

  const int64_t cells = (int64_t)G * G * G;
  DevBuf grid, stats, hist, err, ids;
  if (grid.alloc(WS_GRID, sizeof(int32_t) * cells) ||
      stats.alloc(WS_STATS, sizeof(uint64_t) * 2 * kMaxStats * S) ||
      ids.alloc(WS_POINT_IDS, sizeof(uint32_t) * std::max<int64_t>(idOff[S], 1) + 64) ||
      hist.alloc(WS_HIST, sizeof(unsigned long long) * 256) || err.alloc(WS_ERR, sizeof(int32_t) * 4))
    return 1;
  PCG_CHECK(hipMemsetAsync(grid.p, 0, sizeof(int32_t) * cells, s));
  for (int i = 0; i < S; i++) {
    PCG_CHECK(hipMemsetAsync(stats.as<uint64_t>() + 2 * kMaxStats * i, 0xFF, sizeof(uint64_t) * kMaxStats, s));
    PCG_CHECK(hipMemsetAsync(stats.as<uint64_t>() + 2 * kMaxStats * i + kMaxStats, 0, sizeof(uint64_t) * kMaxStats, s));
  }
  PCG_CHECK(hipMemsetAsync(hist.p, 0, sizeof(unsigned long long) * 256, s));
  PCG_CHECK(hipMemsetAsync(err.p, 0, sizeof(int32_t) * 4, s));
  static hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  if (!ev[0])
    for (int i = 0; i < 4; i++) PCG_CHECK(hipEventCreate(&ev[i]));
  // countPointsInCells (:131-361): every source's points into one grid
  for (int i = 0; i < S; i++) {
    if (chunker_set_params(Ps[i], s)) return 1;
    if (i == 0) PCG_CHECK(hipEventRecord(ev[0], s));
    uint64_t* st = stats.as<uint64_t>() + 2 * kMaxStats * i;
    if (chunker_count(src[i].d_las, (int64_t)hs[i].num_points, grid.as<int32_t>(), ids.as<uint32_t>() + idOff[i], st,
                      st + kMaxStats, hist.as<unsigned long long>(), err.as<int32_t>(), dev.num_cus, s))
      return 1;
  }
  PCG_CHECK(hipEventRecord(ev[1], s));
  int32_t herr[4];
  std::vector<uint64_t> hstats((size_t)2 * kMaxStats * S);
  unsigned long long hhist[256];
  PCG_CHECK(hipMemcpyAsync(herr, err.p, sizeof(herr), hipMemcpyDeviceToHost, s));
  PCG_CHECK(hipMemcpyAsync(hstats.data(), stats.p, sizeof(uint64_t) * hstats.size(), hipMemcpyDeviceToHost, s));
  PCG_CHECK(hipMemcpyAsync(hhist, hist.p, sizeof(hhist), hipMemcpyDeviceToHost, s));
  PCG_CHECK(hipStreamSynchronize(s));
  if (herr[0] & 1) {
    set_error("encountered point outside bounding box.");
    return 123;
  }
  double t1 = now_s();
  out.t_count = t1 - t0;

  // attribute min/max (chunker_countsort_laszip.cpp:385-389, :914-932),
  // merged over the sources; position uses the decoded coordinate, monotone
  // in the stored integer
  for (int i = 0; i < S; i++) {
    if (hs[i].num_points == 0) continue;
    const uint64_t* hsi = hstats.data() + 2 * kMaxStats * i;
    for (size_t k = 0; k < statTargets[i].size(); k++) {
      const auto& tg = statTargets[i][k];
      Attribute* at = tg.first.empty() ? nullptr : A.get(tg.first);
      if (!at) continue;
      double mn = from_ordered_bits(hsi[k]), mx = from_ordered_bits(hsi[kMaxStats + k]);
      if (tg.first == "position") {
        mn = hs[i].scale[tg.second] * mn + hs[i].offset[tg.second];
        mx = hs[i].scale[tg.second] * mx + hs[i].offset[tg.second];
      }
      at->min[tg.second] = std::min(at->min[tg.second], mn);
      at->max[tg.second] = std::max(at->max[tg.second], mx);
    }
  }
  if (Attribute* cls = A.get("classification"))
    for (int i = 0; i < 256; i++) cls->histogram[i] = (int64_t)hhist[i];

  // pyramid + chunk cells (:1074-1183)
  std::vector<DevBuf> levels(Lmax);
  std::vector<int32_t*> lv(Lmax + 1);
  for (int l = 0; l < Lmax; l++) {
    int64_t g = (int64_t)1 << l;
    if (levels[l].alloc(WS_LEVEL0 + l, sizeof(int32_t) * g * g * g)) return 1;
    lv[l] = levels[l].as<int32_t>();
  }
  lv[Lmax] = grid.as<int32_t>();
  if (chunker_pyramid(lv.data(), Lmax, cap, s)) return 1;
  uint32_t maxOut = 1 << 16;
  std::vector<int64_t> raw;
  uint32_t numChunks = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    DevBuf list, counter;
    if (list.alloc(WS_CHUNK_LIST, sizeof(int64_t) * 5 * maxOut) || counter.alloc(WS_CHUNK_COUNTER, sizeof(uint32_t)))
      return 1;
    PCG_CHECK(hipMemsetAsync(counter.p, 0, sizeof(uint32_t), s));
    if (chunker_chunk_cells(lv.data(), Lmax, list.as<int64_t>(), counter.as<uint32_t>(), maxOut, s)) return 1;
    PCG_CHECK(hipMemcpyAsync(&numChunks, counter.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PCG_CHECK(hipStreamSynchronize(s));
    if (numChunks <= maxOut) {
      raw.resize((size_t)5 * numChunks);
      PCG_CHECK(stream_copy(raw.data(), list.p, sizeof(int64_t) * raw.size(), hipMemcpyDeviceToHost));
      break;
    }
    maxOut = numChunks;
  }
  out.chunks.clear();
  for (uint32_t c = 0; c < numChunks; c++) {
    ChunkDesc d;
    d.level = (int32_t)raw[5 * c + 0];
    d.x = raw[5 * c + 1]; d.y = raw[5 * c + 2]; d.z = raw[5 * c + 3];
    d.count = raw[5 * c + 4];
    d.id = node_id(d.level, (int64_t)1 << d.level, d.x, d.y, d.z);
    out.chunks.push_back(d);
  }
  std::sort(out.chunks.begin(), out.chunks.end(), [](const ChunkDesc& a, const ChunkDesc& b) { return a.id < b.id; });
  std::vector<int64_t> sorted(4 * out.chunks.size());
  int64_t off = 0;
  for (size_t c = 0; c < out.chunks.size(); c++) {
    ChunkDesc& d = out.chunks[c];
    sorted[4 * c + 0] = d.level; sorted[4 * c + 1] = d.x; sorted[4 * c + 2] = d.y; sorted[4 * c + 3] = d.z;
    d.offset = off;
    off += d.count;
    chunk_box(d.id, out.min, out.max, d.min, d.max);
  }
  if (off != N) {
    set_error("chunk counts do not add up to the point count");
    return 1;
  }
  DevBuf dchunks, lut;
  if (dchunks.alloc(WS_CHUNKS, sizeof(int64_t) * (sorted.size() + 4))) return 1;
  PCG_CHECK(hipMemcpyAsync(dchunks.p, sorted.data(), sizeof(int64_t) * sorted.size(), hipMemcpyHostToDevice, s));
  if (lut.alloc(WS_LUT, sizeof(int32_t) * cells)) return 1;
  if (chunker_lut_fill(dchunks.as<int64_t>(), (int64_t)out.chunks.size(), Lmax, lut.as<int32_t>(), s)) return 1;
  levels.clear();

  // distribute (:660-1003): the tiles of every source side by side in one
  // chunk x tile matrix, so that the scan puts each chunk's points in source
  // order, then input order
  const int32_t C = (int32_t)out.chunks.size();
  std::vector<int64_t> tileBase(S + 1, 0);
  for (int i = 0; i < S; i++) tileBase[i + 1] = tileBase[i] + ((int64_t)hs[i].num_points + 4095) / 4096;
  const int64_t tiles = tileBase[S];
  out.d_records = (uint8_t*)workspace(WS_RECORDS, (size_t)N * out.bpp + 64);
  out.owned = false;
  if (!out.d_records) return 1;
  // the distribute kernels write every byte of every output record (unset bytes 0)
  DevBuf M, ws, ids16;
  if (M.alloc(WS_TILE_M, sizeof(uint32_t) * std::max<int64_t>(1, tiles * C)) ||
      ws.alloc(WS_SCAN, scan_workspace_bytes(tiles * C)))
    return 1;
  std::vector<char> tail(S, 0);
  bool anyTail = false;
  for (int i = 0; i < S; i++) {
    if (chunker_set_params(Ps[i], s)) return 1;
    tail[i] = g_distribute_tail && distribute_tail_fits(C);
    anyTail = anyTail || tail[i];
  }
  if (anyTail && ids16.alloc(WS_CHUNK_IDS16, sizeof(uint16_t) * std::max<int64_t>(idOff[S], 1) + 64)) return 1;
  PCG_CHECK(hipEventRecord(ev[2], s));
  if (C > 8192 && tiles > 0) PCG_CHECK(hipMemsetAsync(M.p, 0, sizeof(uint32_t) * tiles * C, s));
  for (int i = 0; i < S; i++) {
    if (S > 1 && chunker_set_params(Ps[i], s)) return 1;
    if (chunker_tile_hist(ids.as<uint32_t>() + idOff[i], tail[i] ? ids16.as<uint16_t>() + idOff[i] : nullptr,
                          (int64_t)hs[i].num_points, lut.as<int32_t>(), C, M.as<uint32_t>(), tileBase[i], tiles,
                          err.as<int32_t>(), s))
      return 1;
  }
  if (tiles > 0 && exclusive_scan_u32(M.as<uint32_t>(), tiles * C, ws.p, nullptr, s)) return 1;
  for (int i = 0; i < S; i++) {
    if (S > 1 && chunker_set_params(Ps[i], s)) return 1;
    if (chunker_place(src[i].d_las, ids.as<uint32_t>() + idOff[i], tail[i] ? ids16.as<uint16_t>() + idOff[i] : nullptr,
                      (int64_t)hs[i].num_points, C, M.as<uint32_t>(), tileBase[i], tiles, out.d_records, s))
      return 1;
  }
  PCG_CHECK(hipEventRecord(ev[3], s));
  PCG_CHECK(hipMemcpyAsync(herr, err.p, sizeof(herr), hipMemcpyDeviceToHost, s));
  PCG_CHECK(hipStreamSynchronize(s));
  if (herr[0] & 2) {
    set_error("point to node lookup failed, no node found.");
    return 123;
  }
  PCG_CHECK(hipEventElapsedTime(&out.ms_count_kernel, ev[0], ev[1]));
  PCG_CHECK(hipEventElapsedTime(&out.ms_distribute_kernels, ev[2], ev[3]));
  double t2 = now_s();
  out.t_distribute = t2 - t1;
  out.t_total = t2 - t0;
  return 0;
}

int write_chunk_files(const DeviceChunks& dc, const std::string& targetDir) {
  std::string dir = targetDir + "/chunks";
  fs::create_directories(dir);
  for (const auto& e : fs::directory_iterator(dir)) fs::remove(e);
  std::vector<uint8_t> host((size_t)dc.numPoints * dc.bpp);
  if (!host.empty()) PCG_CHECK(stream_copy(host.data(), dc.d_records, host.size(), hipMemcpyDeviceToHost));
  for (const auto& c : dc.chunks) {
    if (c.count == 0) continue;
    std::ofstream f(dir + "/" + c.id + ".bin", std::ios::binary);
    f.write((const char*)host.data() + c.offset * dc.bpp, (std::streamsize)(c.count * dc.bpp));
  }
  std::ofstream m(dir + "/metadata.json", std::ios::binary);
  m << chunk_metadata_json(dc.min, dc.max, dc.attributes);
  return 0;
}

int read_chunk_files(const std::string& targetDir, DeviceChunks& out) {
  if (ensure_device()) return 1;
  std::string dir = targetDir + "/chunks";
  std::ifstream mf(dir + "/metadata.json", std::ios::binary);
  if (!mf) {
    set_error("missing " + dir + "/metadata.json");
    return 1;
  }
  std::string text((std::istreambuf_iterator<char>(mf)), std::istreambuf_iterator<char>());
  out.release();
  if (!parse_chunk_metadata(text, out.min, out.max, out.attributes)) return 1;
  out.bpp = out.attributes.bytes;
  out.chunks.clear();
  for (const auto& e : fs::directory_iterator(dir)) {
    std::string fn = e.path().filename().string();
    if (fn.size() < 4 || fn.substr(fn.size() - 4) != ".bin") continue;
    ChunkDesc d;
    d.id = fn.substr(0, fn.size() - 4);
    d.count = (int64_t)fs::file_size(e.path()) / out.bpp;
    out.chunks.push_back(d);
  }
  std::sort(out.chunks.begin(), out.chunks.end(), [](const ChunkDesc& a, const ChunkDesc& b) { return a.id < b.id; });
  int64_t off = 0;
  for (auto& d : out.chunks) {
    d.offset = off;
    off += d.count;
    d.level = (int32_t)d.id.size() - 1;
    chunk_box(d.id, out.min, out.max, d.min, d.max);
  }
  out.numPoints = off;
  out.pointsTotal = off;
  std::vector<uint8_t> host((size_t)off * out.bpp);
  for (auto& d : out.chunks) {
    std::ifstream f(dir + "/" + d.id + ".bin", std::ios::binary);
    f.read((char*)host.data() + d.offset * out.bpp, (std::streamsize)(d.count * out.bpp));
  }
  out.d_records = (uint8_t*)workspace(WS_RECORDS, host.size() + 64);
  out.owned = false;
  if (!out.d_records) return 1;
  if (!host.empty()) PCG_CHECK(stream_copy(out.d_records, host.data(), host.size(), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace pcg
