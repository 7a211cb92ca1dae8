// Bottom-up Poisson / Poisson-average / random node sampling on gfx950 (sampler.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcg {

enum SampleMethod : int32_t { METHOD_POISSON = 0, METHOD_POISSON_AVERAGE = 1, METHOD_RANDOM = 2 };

// node state flags
enum : int32_t {
  NS_COMPLETED = 1,   // onNodeCompleted was called for the node
  NS_DISCARDED = 2,   // removed from its parent without completion (Poisson-average)
  NS_SAMPLED = 4,
};

// per-node work flags (set by the kernels)
enum : uint32_t {
  NW_EXACT_SORT = 1,   // equal-distance candidates within spacing: libstdc++ sort order needed
  NW_SEQ_ACCEPT = 2,   // the 10 001-check window of the Poisson scan may bind
  NW_SEQ_AVERAGE = 4,  // two accepted points inside each other's colour range
  NW_OVERFLOW = 8,     // scratch or arena exhausted
  NW_MANY_NBR = 16,    // a candidate has more earlier neighbours than the list holds
  NW_CELL_PATH = 32,   // decided by the cell table + neighbour lists (Poisson-average, or the
                       // accepted shell outgrew the LDS ring of k_ring)
};

struct NodeInfo {
  int32_t children[8];  // node ids, -1 = none
  double min[3];
  double max[3];
  int32_t level;        // name.size() - 1
  int32_t pad;
};

struct NodeState {
  uint8_t* ptr;      // current records
  int64_t count;     // current numPoints
  int64_t accepted;  // numPoints right after the node's own sampling
  int32_t flags;     // NS_*
  int32_t pad;
};

struct NodeWork {
  uint8_t* scratch;
  int64_t n;              // candidates
  int64_t child_start[9]; // candidate index of each child's first point
  const uint8_t* csrc[8]; // records of each child before the node is sampled
  uint32_t flags;         // NW_*
  uint32_t pad;
};

// One output record range of a sampled node: its accepted records
// (child == -1) or the remainder of child `child`.  `sel` lists the source
// candidates in output order (main index for the node's range, index in the
// child for a remainder); nullptr means every record of the child, in order.
struct OutSeg {
  uint8_t* dst;          // 16-byte aligned
  const uint32_t* sel;
  int64_t count;         // records
  int32_t k;             // node index in the level
  int32_t child;
};

struct SampleParams {
  double scale[3];
  double offset[3];
  double base_spacing;
  int32_t bpp;
  int32_t rgb_offset;
  int32_t method;
  int32_t num_cus;         // compute units of the device (grids sized to the GPU)
  uint8_t* arena;          // record arena
  uint64_t arena_cap;      // bytes
  uint64_t* arena_top;     // bytes used (device counter)
  uint8_t* scratch;        // per-level scratch
  uint64_t scratch_cap;
  uint64_t* scratch_top;
  int32_t ring_min_nodes;  // levels with fewer nodes decide acceptance on the cell-table path
  int32_t ring_blocks;     // most k_ring blocks of a launch (the blocks the GPU holds at once)
  uint32_t* item_start;    // num_nodes + 1 entries: work items of the level
  OutSeg* segs;            // 9 per node of the level
  uint32_t* seg_start;     // 9 * num_nodes + 1 entries: copy items per segment
  uint32_t* nbr_pool;      // neighbour lists longer than a row
  unsigned long long nbr_pool_cap;   // entries
  unsigned long long* nbr_pool_top;  // entries used in the level (device counter)
  unsigned long long* counters;  // [0] fp64 re-checks, [1] pair tests, [2] max |dd32-dd64|/s^2 bits,
                                 // [3] exact-sort nodes, [4] sequential nodes, [5] candidates,
                                 // [6] nodes that ran out of scratch or arena,
                                 // [7] candidates with more than 16 earlier neighbours,
                                 // [8] runs of equal distances whose libstdc++ order failed
                                 // its window check (an error), [9] pipelined acceptances that
                                 // gave up waiting (an error); 16 entries
  double random_limit;     // SamplerRandom: 0.7 * sqrt(3), the cell-centre distance bound
  uint32_t* ring_ctl;      // ring work queue: [0] length, [1] next ticket (device)
  int32_t copy_blocks;     // grid of the level-wide record copy
  int32_t neighbor_blocks; // grid of the level-wide neighbour search
  int32_t avg_blocks;      // workgroups of the Poisson-average colour sums per launch
  int32_t bucket_min_nodes;  // launches of at least this many nodes bucket-sort their candidates
  int32_t accept_pipe_nodes; // launches of at most this many nodes, and the ring's launches (Poisson),
                             // take the pipelined acceptance; 0: never
  int32_t ring_max_cand;     // nodes above this many candidates take the cell-table path
  int32_t pipe_form;         // pipelined acceptance: 0 by node size, 1 every node in the 32-bit form,
                             // 2 every node by one wavefront (test forms, same results)
  int32_t pad3;
};

// Samples every node of `nodes` (all of one level, children already done).
int sample_level(const int32_t* d_nodes, int32_t num_nodes, const NodeInfo* d_info, NodeState* d_state,
                 NodeWork* d_work, const SampleParams& p, hipStream_t s);

// Scratch bytes for one node of n candidates.
uint64_t sample_scratch_bytes(int64_t n, int method);

// Clears n <= 64 words with a kernel (no memset node, so that the sampling
// launch sequence can be captured in a graph).
int sample_zero(uint64_t* d, int32_t n, hipStream_t s);
// Scratch bytes that cover any level of `max_nodes` nodes holding at most
// `candidates` candidates in total.
uint64_t sample_scratch_bound(int64_t candidates, int64_t max_nodes, int method);

// SamplerRandom's leaf shuffle (sampler_random.h:89-121): record i of a
// leaf moves to slot perm[i] of std::shuffle over the identity with
// std::default_random_engine(seed).  `idx` is 4 * count bytes of scratch
// for leaves too large for the kernel's LDS table (may be null otherwise).
struct ShuffleLeaf {
  const uint8_t* src;
  uint8_t* dst;
  uint32_t* idx;
  int64_t count;
};
constexpr int64_t kShuffleLdsPoints = 10240;
int shuffle_leaves(const ShuffleLeaf* d_leaves, int64_t num, int32_t bpp, uint32_t seed, hipStream_t s);

// Copies each (src, bytes) range to dst + offset (one workgroup per range).
struct CopyDesc {
  const uint8_t* src;
  uint64_t dst_offset;
  uint64_t bytes;
};
int gather_ranges(const CopyDesc* d_desc, int64_t num, uint8_t* d_dst, hipStream_t s);

// BROTLI node layout (indexer.cpp:1123-1354): every node's records as one
// array per attribute (position as two 48-bit Morton codes, rgb as one),
// points in ascending (upper, lower) position code, std::sort order for equal
// codes.  One EncNode per node with points; `out_off` / `scr_off` are byte
// offsets into the output and scratch buffers (scratch: enc_scratch_bytes).
struct EncNode {
  const uint8_t* src;
  int64_t count;
  uint64_t out_off;
  uint64_t scr_off;
};
enum EncColumnKind : int32_t { EC_RAW = 0, EC_POSITION_MORTON = 1, EC_RGB_MORTON = 2 };
struct EncColumn {
  int32_t kind;
  int32_t src;   // byte offset of the attribute in the record
  int32_t size;  // bytes per point in the layout
  int32_t pad;
};
constexpr int kMaxEncColumns = 40;
struct EncParams {
  int32_t bpp;
  int32_t num_cols;
  int32_t pos_off;
  int32_t pad;
  EncColumn cols[kMaxEncColumns];
};
uint64_t enc_scratch_bytes(int64_t count);
int encode_layout(const EncNode* d_nodes, int64_t num_nodes, const EncParams& p, uint8_t* d_scratch, uint8_t* d_out,
                  hipStream_t s);

}  // namespace pcg
