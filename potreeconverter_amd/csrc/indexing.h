// Device-resident doIndexing (octree build + bottom-up sampling + writers).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "converter.h"

namespace pcg {

// One onNodeCompleted call (indexer.cpp:1605-1608), in completion order.
struct CompletedNode {
  // "r" + d child digits as the key (1 << 3d) | digits (base 8) for d <= 20;
  // key 0: the node is named by `name` instead
  uint64_t key = 0;
  std::string name;
  int64_t numPoints = 0;
  int64_t byteOffset = 0;
  int64_t byteSize = 0;
};

struct IndexResult {
  uint8_t* d_octree = nullptr;   // octree.bin image in HBM (encoding DEFAULT)
  std::vector<uint8_t> h_octree; // octree.bin image in host memory (encoding BROTLI)
  bool octreeOnHost = false;
  uint64_t octreeBytes = 0;
  std::vector<CompletedNode> completed;  // the completion records when hierarchy.bin was built from them
  std::vector<uint8_t> hierarchy;  // hierarchy.bin
  std::string metadata;            // metadata.json
  int64_t octreeDepth = 0;
  int64_t firstChunkSize = 0;
  double t_build = 0, t_sample = 0, t_assemble = 0, t_total = 0;
  double t_setup = 0, t_order = 0, t_hierarchy = 0;  // host: tree, completion order + copy list, hierarchy + metadata
  uint64_t counters[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t numNodes = 0, numSampled = 0, buildRounds = 0;
  int64_t duplicateNodes = 0;  // nodes whose duplicate positions were dropped (indexer.cpp:1002-1098)
  void release();
};

// method: "poisson" | "poisson_average" | "random".  `name`/`pointsTotal` go to
// metadata.json as in createMetadata (indexer.cpp:354-524).
// Sampling launch mode: false = every node of an octree level in one launch
// per kernel (batched), true = one launch per node (plain).
extern bool g_plain_sampling;
// Poisson levels with at least this many nodes decide acceptance with the
// LDS shell ring (one wavefront per node); narrower levels, where one
// wavefront per node cannot hide latency, use the cell-table path.
extern int32_t g_ring_min_nodes;
// Chunk subtrees are sampled in this many sets, each on a stream of its own
// (1 .. kMaxSampleGroups; 1: one level-synchronous pass over the whole tree;
// 0: one set per hardware queue besides the library stream's).
extern int32_t g_sample_groups;
// Sets of chunk subtrees (sample_groups streams take them in turn, each
// with buffers for its largest set): 0 = as many as streams unless the job
// needs more sets to fit in device memory; n = max(streams, n), at most
// kMaxSampleSets and the chunk count.
extern int32_t g_sample_sets;
// Launch rounds of the node sampling: 0 = octree levels, deepest first;
// 1 = heights (a round holds every node whose children were all sampled in
// earlier rounds: leaves' parents first, whatever their depth).
extern int32_t g_sample_order;
// Capture the chunk-subtree sampling (every set's launch sequence on its
// stream) in a graph and launch that instead of the launches one by one.
extern bool g_sample_graph;
// How chunk subtrees are split into sets: 0 = longest processing time first
// on their points over all sets; 1 = subtrees deeper than where most points
// end get sets of their own (in proportion to their points), the others the
// rest.
extern int32_t g_set_policy;
// Launch shapes of the node sampling (same results for every value): k_ring
// blocks per CU (1-3), grids of the level-wide record copy and neighbour
// search, workgroups of a Poisson-average colour-sum launch.
extern int32_t g_ring_blocks_per_cu;
extern int32_t g_copy_blocks;
extern int32_t g_neighbor_blocks;
extern int32_t g_avg_blocks;
// Sampling launches of at least this many nodes sort each node's candidates
// by distance buckets, smaller ones by radix passes (same order).
extern int32_t g_bucket_min_nodes;
extern int32_t g_accept_pipe_nodes;
extern int32_t g_ring_max_cand;
extern int32_t g_pipe_form;
// Seed of every SamplerRandom leaf shuffle (0 .. 2^32-1), or -1 for the
// reference's choice, the system clock at the start of indexing.
extern int64_t g_random_seed;

// One Sampler::sample call (Converter/include/structures.h:152-164) on a
// caller-built subtree instead of a tree built from chunks: node 0 is the
// node handed to sample(); nodes without children hold records (leaves, and
// nodes sampled before, which the traversal of sampler_poisson.h:27-36 does
// not enter).  Node i holds counts[i] records from dc.d_records +
// offsets[i] * bpp.
struct SubtreeSpec {
  int32_t numNodes = 0;
  const int32_t* children = nullptr;  // 8 per node, -1 = none
  const double* bounds = nullptr;     // 6 per node: min x, y, z, max x, y, z
  const int32_t* levels = nullptr;    // name.size() - 1 of every node
  const uint8_t* sampled = nullptr;   // nullable: 1 = sampled before (not visited)
  const int64_t* counts = nullptr;
  const int64_t* offsets = nullptr;
  double baseSpacing = 0;
  int32_t rgbOffset = -1;
};
// Results: every node's records after the call (device pointers into the
// library's buffers, valid until the next call), the children that stay
// linked, and the callbacks in the reference's order: kind 0 =
// onNodeCompleted, 1 = onNodeDiscarded.
struct SubtreeResult {
  std::vector<const uint8_t*> ptr;
  std::vector<int64_t> count;
  std::vector<int32_t> children;
  std::vector<std::pair<int32_t, int32_t>> events;
};

// encoding: "BROTLI" compresses every node (indexer.cpp:1110-1401); any
// other value writes the records unchanged and is echoed in metadata.json
// (indexer.cpp:516, 1462).  With `sub` the tree is the given subtree and the
// call ends after sampling with `subOut` filled (no octree image).
int gpu_indexing(DeviceChunks& dc, const std::string& method, const std::string& name, int64_t pointsTotal,
                 IndexResult& out, const std::string& encoding = "DEFAULT", const SubtreeSpec* sub = nullptr,
                 SubtreeResult* subOut = nullptr);

int write_index_files(const IndexResult& r, const std::string& targetDir);

// HierarchyFlusher + HierarchyBuilder over the completion records
// (indexer.h:120-245, HierarchyBuilder.h:26-354); batches in ascending name.
std::vector<uint8_t> build_hierarchy_bin(const std::vector<CompletedNode>& recs, int64_t& firstChunkSize);
// The layout of hierarchy.bin depends on the set of records only, except
// for each node's NORMAL / LEAF type (whether a child is recorded after it)
// and the records' point counts and byte ranges.  plan_hierarchy fixes the
// layout for the records' node keys (items 0..n-1, any order; false when a
// key is missing or deeper than 20 levels); fill_hierarchy writes the file
// for one record order (rank[item]) and the records' values.
struct HierarchyPlan {
  enum : uint8_t { S_PLAIN = 0, S_NORMAL = 1, S_LEAF = 2, S_PROXY = 3 };
  struct Slot {
    int32_t item;
    uint8_t kind;
    uint8_t mask;
    uint64_t offset, size;  // S_PROXY
  };
  std::vector<Slot> slots;      // one per 22-byte record, in file order
  std::vector<int32_t> parent;  // parent item or -1
  int64_t firstChunkSize = 0;
  int32_t items = 0;
};
bool plan_hierarchy(const std::vector<uint64_t>& keys, HierarchyPlan& plan);
std::vector<uint8_t> fill_hierarchy(const HierarchyPlan& plan, const int32_t* rank, const int64_t* numPoints,
                                    const int64_t* byteOffset, const int64_t* byteSize);
// the same from node names only (the batch / chunk maps of HierarchyBuilder)
std::vector<uint8_t> build_hierarchy_bin_maps(const std::vector<CompletedNode>& recs, int64_t& firstChunkSize);
// node key <-> name ("r" + digits), keys for names of at most 21 characters
uint64_t node_key(const std::string& name);
std::string node_name(uint64_t key);

std::string create_metadata_json(const Attributes& attributes, const Vec3& min, const Vec3& max, double spacing,
                                 const std::string& name, int64_t pointsTotal, int64_t firstChunkSize,
                                 int64_t octreeDepth, const std::string& encoding = "DEFAULT");

}  // namespace pcg
