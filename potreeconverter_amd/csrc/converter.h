// Host orchestration of the device-resident chunker and indexer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "attributes.h"
#include "lasfile.h"

namespace pcg {

struct ChunkDesc {
  std::string id;     // "r" + one octal digit per level (x -> 4, y -> 2, z -> 1)
  int32_t level = 0;
  int64_t x = 0, y = 0, z = 0;
  int64_t count = 0;
  int64_t offset = 0;  // first record of the chunk in DeviceChunks::d_records
  Vec3 min, max;       // childBoundingBoxOf chain from the cubic box
};

// Device buffers that stay allocated across calls.  The library runs one
// job at a time on its stream, so a job reuses the memory of the previous
// one instead of allocating and freeing tens of GB (hipMalloc / hipFree of
// that size cost tens of milliseconds per job).  workspace() returns a
// buffer of at least `bytes` for `slot`, grown (with 1/8 headroom) when
// smaller, or nullptr with the error set.  The memory is held until
// release_workspace() or the end of the process.
enum WorkspaceSlot : int {
  WS_GRID, WS_STATS, WS_HIST, WS_ERR, WS_CHUNK_LIST, WS_CHUNK_COUNTER, WS_CHUNKS, WS_LUT, WS_TILE_M, WS_SCAN,
  WS_RECORDS, WS_OCT_B, WS_OCT_KEYS_A, WS_OCT_KEYS_B, WS_OCT_M, WS_OCT_SCAN, WS_OCT_SEGS, WS_OCT_TILES,
  WS_OCT_NODES, WS_OCT_COUNTER, WS_INFO, WS_STATE, WS_LIST, WS_ARENA, WS_TOPS, WS_COUNTERS, WS_OCTREE, WS_COPYDESC, WS_CHUNK_IDS16,
  WS_POINT_IDS, WS_ENC_NODES, WS_ENC_SCRATCH, WS_SHUF_DESC, WS_SHUF_OUT, WS_SHUF_IDX, WS_OCT_H5,
  WS_GEN_TILES, WS_GEN_OFFSETS, WS_GEN_STATS,
  WS_LEVEL0,  // pyramid levels: WS_LEVEL0 + level
  WS_GROUP0 = WS_LEVEL0 + 16,  // sampling group g, buffer k: WS_GROUP0 + 8 * g + k
  WS_COUNT = WS_GROUP0 + 8 * 8
};
constexpr int kMaxSampleGroups = 8;  // streams of the chunk-subtree sampling
constexpr int kMaxSampleSets = 64;   // sets of chunk subtrees sampled on those streams
inline int ws_group_slot(int g, int k) { return WS_GROUP0 + 8 * g + k; }
void* workspace(int slot, size_t bytes);
void release_workspace();
// bytes held by one slot / by all slots; frees one slot (its next use
// allocates again)
size_t workspace_bytes(int slot);
size_t workspace_total();
void release_workspace_slot(int slot);

struct DeviceChunks {
  uint8_t* d_records = nullptr;  // numPoints * bpp bytes, grouped by chunk (ascending ID)
  bool owned = false;            // d_records is the caller's to free (else a workspace buffer)
  int64_t numPoints = 0;
  int32_t bpp = 0;
  std::vector<ChunkDesc> chunks;
  Attributes attributes;  // posScale/posOffset, min/max, classification histogram
  Vec3 min, max;          // cubic bounding box written to chunks/metadata.json
  int64_t pointsTotal = 0;
  int32_t gridSize = 0;
  int64_t maxPointsPerChunk = 0;
  double t_count = 0, t_distribute = 0, t_total = 0;
  float ms_count_kernel = 0, ms_distribute_kernels = 0;  // HIP events on the library stream
  void release();
};

struct Device {
  int id = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
};

Device& device();
int ensure_device();

// Blocking copy ordered after everything queued on the library stream.  The
// stream is non-blocking, so a plain hipMemcpy (null stream) would not wait
// for it.
inline hipError_t stream_copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipStream_t s = device().stream;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

// Chunk grid size and chunk point cap forced by pcg_set_option ("chunk_grid",
// "chunk_cap"; 0 = the reference's rule from the point count,
// chunker_countsort_laszip.cpp:1210-1220): lets a small input take the
// G = 256 / 512 paths of a large one.
extern int32_t g_chunk_grid;
extern int64_t g_chunk_cap;
// The tile transcode (16-bit chunk ids, aligned 16-byte run stores) where the
// record layout allows, else the byte-wise scatter ("distribute_tail" option).
extern bool g_distribute_tail;

// doChunking on LAS records already resident in HBM (header describes them).
int gpu_chunking(const uint8_t* d_las, const lasfile::Header& h, DeviceChunks& out);

// One source of a multi-source doChunking: its records in HBM and its header.
struct ChunkSource {
  const uint8_t* d_las;
  lasfile::Header header;
};
// doChunking(sources, targetDir, min, max, state, outputAttributes, monitor)
// (Converter/include/chunker_countsort_laszip.h:19) on sources resident in
// HBM: cubeMin / cubeMax (both or neither) are the box main.cpp's
// computeStats hands over (else computed from the headers the same way),
// `requested` the --attributes filter (PotreeConverter.h:260-283; empty:
// every attribute of every source).  Chunks hold each source's points in
// source order, then input order.
int gpu_chunking_sources(const std::vector<ChunkSource>& src, const double* cubeMin, const double* cubeMax,
                         const std::vector<std::string>& requested, DeviceChunks& out);

// Writes chunks/<id>.bin + chunks/metadata.json like the reference chunker.
int write_chunk_files(const DeviceChunks& dc, const std::string& targetDir);

// Reads chunks/ written by any chunker (the reference's or ours) into HBM.
int read_chunk_files(const std::string& targetDir, DeviceChunks& out);

// Box of a chunk from its ID digits (indexer.cpp:161-170).
void chunk_box(const std::string& id, const Vec3& min, const Vec3& max, Vec3& cmin, Vec3& cmax);
void child_box(const Vec3& min, const Vec3& max, int index, Vec3& cmin, Vec3& cmax);

}  // namespace pcg
