// Launchers for the octree build kernels (octree.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcg {

struct OctreeParams {
  double scale[3];   // attributes.posScale
  double offset[3];  // attributes.posOffset
};

struct SegDesc {
  int64_t start;    // first record (absolute index)
  int64_t count;
  double min[3];
  double size[3];   // node max - min (indexer.cpp:843)
  int64_t delta;    // start - (records of earlier segments in this round)
};

struct TileDesc {
  int32_t seg;
  int32_t tile_in_seg;
  int64_t start;      // absolute first record
  int64_t count;
  int64_t m_base;     // 256 * (tiles of earlier segments)
  int64_t seg_tiles;  // tiles of this segment
};

struct NodeOut {
  int32_t seg;
  int32_t level;    // 0..5 below the segment node
  uint32_t morton;  // cell index at that level
  uint32_t start;   // first record relative to the segment
  uint32_t count;
};

// One refinement round: keys, realized nodes, stable sort of every segment
// by its 15-bit key (result in d_a / d_ka).
int octree_round(uint8_t* d_a, uint8_t* d_b, uint16_t* d_ka, uint16_t* d_kb, int32_t bpp, const OctreeParams& p,
                 const SegDesc* d_segs, int32_t num_segs, const TileDesc* d_tiles, int64_t num_tiles,
                 int64_t m_total, uint32_t* d_M, void* d_scan_ws, NodeOut* d_nodes, uint32_t* d_node_counter,
                 uint32_t max_nodes, uint32_t* d_h5, hipStream_t s);

}  // namespace pcg
