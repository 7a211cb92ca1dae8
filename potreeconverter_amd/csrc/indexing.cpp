// Device-resident doIndexing (Converter/src/indexer.cpp:1581-1795).
//
// 1. Octree build: every chunk with >= 10 000 points is refined in rounds
//    of the 32^3 counting sort (octree.hip); the host turns the realized
//    cells into the node tree (expandTo, indexer.cpp:915-947).
// 2. Sampling: the whole tree (root, nodes above the chunk roots, chunk
//    subtrees) is sampled bottom-up one round per launch sequence, a round
//    being an octree level or a height (sampler.hip).
//    Per-node results depend only on the children's states, so this equals
//    the reference's per-chunk post-order followed by the chunk-root phase.
// 3. Canonical completion order (per chunk in ascending ID, post-order,
//    children 0..7; then the chunk-root tasks of processChunkRoots in
//    pre-order; then the root), octree.bin assembled in HBM by one gather,
//    hierarchy.bin and metadata.json on the host.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <functional>
#include <map>
#include <unordered_map>

#include "encoding.h"
#include "indexing.h"
#include "octree.h"
#include "sampler.h"
#include "scan.h"

namespace pcg {

namespace {

constexpr int64_t kLeafCap = 10000;  // indexer.h:47

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct TNode {
  uint64_t key = 0;      // node_key of the name for levels up to 20
  std::string name;      // the name, kept only for levels above 20 (key 0)
  int32_t level = 0;     // name.size() - 1
  Vec3 min, max;
  int32_t children[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int32_t parent = -1;
  int64_t off = -1;      // leaf records: first record in the chunk buffer
  int64_t count = 0;     // leaf records: count
  bool chunkRoot = false;
  bool upper = false;    // root or intermediate node above the chunk roots
  bool presampled = false;  // holds records sampled before (a subtree's holder not visited)
  bool hasChildren() const {
    for (int c = 0; c < 8; c++) if (children[c] >= 0) return true;
    return false;
  }
  std::string full_name() const { return key ? node_name(key) : name; }
};

struct Tree {
  std::vector<TNode> nodes;
  int32_t push(uint64_t key, std::string&& name, int32_t level, const Vec3& mn, const Vec3& mx, int32_t parent,
               int digit) {
    nodes.emplace_back();
    TNode& t = nodes.back();
    t.key = key;
    t.name = std::move(name);
    t.level = level;
    t.min = mn;
    t.max = mx;
    t.parent = parent;
    const int32_t id = (int32_t)nodes.size() - 1;
    if (parent >= 0) nodes[parent].children[digit] = id;
    return id;
  }
  int32_t add(const std::string& name, const Vec3& mn, const Vec3& mx, int32_t parent) {
    const bool keyed = name.size() <= 21;
    return push(keyed ? node_key(name) : 0, keyed ? std::string() : std::string(name), (int32_t)name.size() - 1, mn,
                mx, parent, name.back() - '0');
  }
  // child of `id` with digit c, created with childBoundingBoxOf if missing
  int32_t child(int32_t id, int c) {
    if (nodes[id].children[c] >= 0) return nodes[id].children[c];
    Vec3 a, b;
    child_box(nodes[id].min, nodes[id].max, c, a, b);
    const TNode& p = nodes[id];
    const int32_t level = p.level + 1;
    if (level <= 20) return push((p.key << 3) | (uint64_t)c, std::string(), level, a, b, id, c);
    return push(0, p.full_name() + char('0' + c), level, a, b, id, c);
  }
};

// Host arrays of a job keep their memory for the next job (the library runs
// one job at a time): the vector takes over the storage of `keep` and hands
// it back, emptied, when the job ends.  Tens of MB of fresh pages per job
// would otherwise cost milliseconds of page faults on the critical path.
template <class T>
struct Reused {
  std::vector<T>& keep;
  std::vector<T> v;
  explicit Reused(std::vector<T>& k) : keep(k) {
    v.swap(keep);
    v.clear();
  }
  ~Reused() {
    v.clear();
    v.swap(keep);
  }
};

// a workspace buffer of the job
struct DevMem {
  void* p = nullptr;
  int slot;
  explicit DevMem(int s) : slot(s) {}
  int alloc(size_t b) {
    p = workspace(slot, b);
    return p ? 0 : 1;
  }
  template <class T>
  T* as() const { return (T*)p; }
};

// streams and events of the sampling groups (default priority, created once)
int group_streams(int G, std::vector<hipStream_t>& out) {
  static std::vector<hipStream_t> streams;
  while ((int)streams.size() < G) {
    hipStream_t st = nullptr;
    PCG_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    streams.push_back(st);
  }
  out.assign(streams.begin(), streams.begin() + G);
  return 0;
}

hipEvent_t group_event(int i) {
  static std::vector<hipEvent_t> events;
  while ((int)events.size() <= i) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      set_error("hipEventCreateWithFlags failed");
      return nullptr;
    }
    events.push_back(e);
  }
  return events[i];
}

// Distinct (X, Y, Z) positions among n records of bpp bytes in HBM: the key
// of the reference's duplicate test (indexer.cpp:1010-1027, "X, Y, Z"
// strings).  Only nodes that a refinement round cannot split come here.
int count_distinct_positions(const uint8_t* d, int64_t n, int32_t bpp, int64_t& distinct) {
  std::vector<uint8_t> h((size_t)n * bpp);
  PCG_CHECK(stream_copy(h.data(), d, h.size(), hipMemcpyDeviceToHost));
  std::vector<std::array<int32_t, 3>> k((size_t)n);
  for (int64_t i = 0; i < n; i++) memcpy(k[i].data(), &h[(size_t)i * bpp], 12);
  std::sort(k.begin(), k.end());
  distinct = std::unique(k.begin(), k.end()) - k.begin();
  return 0;
}

}  // namespace

bool g_plain_sampling = false;
int64_t g_random_seed = -1;
int32_t g_ring_min_nodes = 512;
int32_t g_sample_groups = 0;
int32_t g_sample_sets = 0;
int32_t g_sample_order = 0;
bool g_sample_graph = false;
int32_t g_set_policy = 0;
int32_t g_ring_blocks_per_cu = 3;
int32_t g_accept_pipe_nodes = 1024;
int32_t g_ring_max_cand = 1 << 30;
int32_t g_pipe_form = 0;
int32_t g_copy_blocks = 8192;
int32_t g_neighbor_blocks = 4096;
int32_t g_avg_blocks = 2048;
int32_t g_bucket_min_nodes = 512;

void IndexResult::release() {
  d_octree = nullptr;  // a workspace buffer
  std::vector<uint8_t>().swap(h_octree);
}

int gpu_indexing(DeviceChunks& dc, const std::string& method, const std::string& name, int64_t pointsTotal,
                 IndexResult& out, const std::string& encoding, const SubtreeSpec* sub, SubtreeResult* subOut) {
  if (ensure_device()) return 1;
  hipStream_t s = device().stream;
  const double t0 = now_s();
  out.release();
  out = IndexResult();
  const bool avg = method == "poisson_average";
  const bool rnd = method == "random";
  if (!avg && !rnd && method != "poisson") {
    set_error("unknown sampling method: " + method);
    return 1;
  }
  // any other encoding name writes the records unchanged (indexer.cpp:1462)
  const bool brotli = encoding == "BROTLI";
  if (brotli && !brotli_available()) return 1;
  Attributes attributes = dc.attributes;
  // chunks/metadata.json round trip: non-finite min/max read back as +inf
  for (auto& a : attributes.list)
    for (int i = 0; i < 3; i++) {
      if (!std::isfinite(a.min[i])) a.min[i] = INFINITY;
      if (!std::isfinite(a.max[i])) a.max[i] = INFINITY;
    }
  const int32_t bpp = dc.bpp;
  const int rgbOffset = sub ? sub->rgbOffset : attributes.getOffset("rgb");
  if (avg && rgbOffset < 0) {
    set_error("poisson_average needs an rgb attribute");
    return 1;
  }
  const int64_t N = dc.numPoints;

  // ---------------------------------------------------------------- tree
  static std::vector<TNode> keepNodes;
  Reused<TNode> nodesMem(keepNodes);
  Tree tree;
  tree.nodes.swap(nodesMem.v);
  struct GiveBack {
    Tree& t;
    std::vector<TNode>& v;
    ~GiveBack() { t.nodes.swap(v); }
  } giveBack{tree, nodesMem.v};
  std::vector<int32_t> chunkRoots;
  std::vector<int64_t> chunkRootPoints;
  if (sub) {
    // the caller's subtree: node 0 takes the place of the single chunk root
    // (nothing above it), the nodes without children hold records
    tree.nodes.resize(sub->numNodes);
    for (int32_t i = 0; i < sub->numNodes; i++) {
      TNode& t = tree.nodes[i];
      t = TNode();
      t.level = sub->levels[i];
      for (int a = 0; a < 3; a++) {
        t.min[a] = sub->bounds[6 * i + a];
        t.max[a] = sub->bounds[6 * i + 3 + a];
      }
      for (int c = 0; c < 8; c++) t.children[c] = sub->children[8 * i + c];
      if (!t.hasChildren()) {
        t.off = sub->offsets[i];
        t.count = sub->counts[i];
        t.presampled = sub->sampled && sub->sampled[i];
      }
    }
    // keys from the child digits (node 0 is "r"), for hierarchy planning
    std::vector<int32_t> stack{0};
    tree.nodes[0].key = 1;
    while (!stack.empty()) {
      const int32_t id = stack.back();
      stack.pop_back();
      for (int c = 0; c < 8; c++) {
        const int32_t ch = tree.nodes[id].children[c];
        if (ch < 0) continue;
        tree.nodes[ch].parent = id;
        const uint64_t k = tree.nodes[id].key;
        tree.nodes[ch].key = (k && k < (1ull << 60)) ? ((k << 3) | (uint64_t)c) : 0;
        if (!tree.nodes[ch].key) tree.nodes[ch].name = "r";
        stack.push_back(ch);
      }
    }
    tree.nodes[0].chunkRoot = true;
    chunkRoots.push_back(0);
    chunkRootPoints.push_back(N);
  } else {
    tree.add("r", dc.min, dc.max, -1);
    tree.nodes[0].upper = true;
  }
  static const std::vector<ChunkDesc> kNoChunks;
  for (const auto& c : sub ? kNoChunks : dc.chunks) {
    int32_t id;
    if (c.id == "r") {
      id = 0;
    } else {
      int32_t cur = 0;
      for (size_t l = 1; l + 1 < c.id.size(); l++) {
        cur = tree.child(cur, c.id[l] - '0');
        tree.nodes[cur].upper = true;
      }
      // the chunk box comes from the ID chain of the cubic box, as getChunks
      Vec3 mn, mx;
      chunk_box(c.id, dc.min, dc.max, mn, mx);
      id = tree.add(c.id, mn, mx, cur);
    }
    tree.nodes[id].chunkRoot = true;
    tree.nodes[id].upper = false;
    tree.nodes[id].off = c.offset;
    tree.nodes[id].count = c.count;
    chunkRoots.push_back(id);
    chunkRootPoints.push_back(c.count);
  }

  // ------------------------------------------------------ octree build
  const double tb0 = now_s();
  DevMem bufB(WS_OCT_B), keysA(WS_OCT_KEYS_A), keysB(WS_OCT_KEYS_B), dM(WS_OCT_M), ws(WS_OCT_SCAN), dsegs(WS_OCT_SEGS),
      dtiles(WS_OCT_TILES), dnodes(WS_OCT_NODES), dcounter(WS_OCT_COUNTER), dh5(WS_OCT_H5);
  struct Seg { int32_t node; int64_t start, count; Vec3 mn, mx; };
  std::vector<Seg> segs;
  for (int32_t id : chunkRoots)
    if (!sub && tree.nodes[id].count >= kLeafCap)
      segs.push_back({id, tree.nodes[id].off, tree.nodes[id].count, tree.nodes[id].min, tree.nodes[id].max});
  if (!segs.empty()) {
    // the previous job's sampling buffers and octree image are dead: give
    // them back first when the build buffers would not fit beside them
    const size_t buildBytes = ((size_t)N * bpp + (size_t)N * 4) * 9 / 8 + ((size_t)1 << 30);
    if (workspace_bytes(WS_OCT_B) < (size_t)N * bpp) {
      size_t freeB = 0, totalB = 0;
      PCG_CHECK(hipMemGetInfo(&freeB, &totalB));
      if (freeB < buildBytes) {
        release_workspace_slot(WS_ARENA);
        release_workspace_slot(WS_OCTREE);
        for (int k = 0; k < kMaxSampleGroups * 8; k++) release_workspace_slot(WS_GROUP0 + k);
      }
    }
    if (bufB.alloc((size_t)N * bpp + 64) || keysA.alloc((size_t)N * 2 + 64) || keysB.alloc((size_t)N * 2 + 64) ||
        dcounter.alloc(16))
      return 1;
  }
  OctreeParams op;
  for (int a = 0; a < 3; a++) { op.scale[a] = attributes.posScale[a]; op.offset[a] = attributes.posOffset[a]; }
  const int64_t tileSize = 4096;
  int64_t octreeDepth = 0;
  // One round per level of refinement.  The next round's segments (nodes
  // above the leaf cap, with their boxes from the childBoundingBoxOf chain)
  // follow from a round's output alone, so that round is launched before
  // the host expands the tree from the output.
  std::vector<SegDesc> sd;  // descriptors of the round in flight (kept until it is done)
  std::vector<TileDesc> td;
  uint32_t maxNodes = 0;
  auto launch_round = [&](const std::vector<Seg>& sg) -> int {
    sd.assign(sg.size(), SegDesc());
    td.clear();
    int64_t before = 0, tilesBefore = 0;
    for (size_t i = 0; i < sg.size(); i++) {
      SegDesc& d = sd[i];
      d.start = sg[i].start;
      d.count = sg[i].count;
      for (int a = 0; a < 3; a++) { d.min[a] = sg[i].mn[a]; d.size[a] = sg[i].mx[a] - sg[i].mn[a]; }
      d.delta = d.start - before;
      int64_t nt = (d.count + tileSize - 1) / tileSize;
      for (int64_t k = 0; k < nt; k++) {
        TileDesc x;
        x.seg = (int32_t)i;
        x.tile_in_seg = (int32_t)k;
        x.start = d.start + k * tileSize;
        x.count = std::min(tileSize, d.count - k * tileSize);
        x.m_base = 256 * tilesBefore;
        x.seg_tiles = nt;
        td.push_back(x);
      }
      tilesBefore += nt;
      before += d.count;
    }
    const int64_t mTotal = 256 * tilesBefore;
    maxNodes = (uint32_t)std::min<int64_t>(sg.size() * 32768ll, 1ll << 26);
    if (dsegs.alloc(sd.size() * sizeof(SegDesc)) || dtiles.alloc(td.size() * sizeof(TileDesc)) ||
        dM.alloc((size_t)mTotal * 4 + 64) || ws.alloc(scan_workspace_bytes(mTotal)) ||
        dnodes.alloc((size_t)maxNodes * sizeof(NodeOut)) || dh5.alloc(sd.size() * 32768 * sizeof(uint32_t)))
      return 1;
    PCG_CHECK(hipMemcpyAsync(dsegs.p, sd.data(), sd.size() * sizeof(SegDesc), hipMemcpyHostToDevice, s));
    PCG_CHECK(hipMemcpyAsync(dtiles.p, td.data(), td.size() * sizeof(TileDesc), hipMemcpyHostToDevice, s));
    PCG_CHECK(hipMemsetAsync(dcounter.p, 0, 16, s));
    return octree_round(dc.d_records, bufB.as<uint8_t>(), keysA.as<uint16_t>(), keysB.as<uint16_t>(), bpp, op,
                        dsegs.as<SegDesc>(), (int32_t)sd.size(), dtiles.as<TileDesc>(), (int64_t)td.size(), mTotal,
                        dM.as<uint32_t>(), ws.p, dnodes.as<NodeOut>(), dcounter.as<uint32_t>(), maxNodes,
                        dh5.as<uint32_t>(), s);
  };
  if (!segs.empty() && launch_round(segs)) return 1;
  while (!segs.empty()) {
    out.buildRounds++;
    uint32_t numOut = 0;
    PCG_CHECK(hipMemcpyAsync(&numOut, dcounter.p, 4, hipMemcpyDeviceToHost, s));
    PCG_CHECK(hipStreamSynchronize(s));
    if (numOut > maxNodes) {
      set_error("octree round produced more nodes than its buffer holds");
      return 1;
    }
    std::vector<NodeOut> no(numOut);
    if (numOut) PCG_CHECK(stream_copy(no.data(), dnodes.p, numOut * sizeof(NodeOut), hipMemcpyDeviceToHost));
    std::vector<Seg> next;
    std::vector<size_t> nextFrom;  // the output record of each next segment
    std::vector<int64_t> kept(no.size(), -1);  // records kept by the duplicate path
    for (size_t k = 0; k < no.size(); k++) {
      const NodeOut& o = no[k];
      if ((int64_t)o.count <= kLeafCap) continue;
      const Seg& sg = segs[o.seg];
      int64_t count = o.count;
      if (count == sg.count) {
        // indexer.cpp:1002-1098: every point of the segment in one cell five
        // levels down.  With at least maxPointsPerChunk / 2 duplicate
        // positions the reference keeps #distinct records of the node, and
        // they are its FIRST #distinct records in the current order (the
        // copy at :1088-1090 indexes the buffer by the loop counter, not by
        // the distinct list); then it tries again (:1093-1106): a leaf
        // below the leaf cap, otherwise another round.  With fewer
        // duplicates it only warns and refines the node as it is.
        int64_t distinct = 0;
        if (count_distinct_positions(dc.d_records + (size_t)(sg.start + (int64_t)o.start) * bpp, count, bpp,
                                     distinct))
          return 1;
        if (count - distinct >= kLeafCap / 2) {
          count = distinct;
          kept[k] = distinct;
          out.duplicateNodes++;
          if (count < kLeafCap) continue;
        }
      }
      Vec3 mn = sg.mn, mx = sg.mx;
      for (int i = 0; i < o.level; i++) {
        Vec3 a, b;
        child_box(mn, mx, (int)((o.morton >> (3 * (o.level - 1 - i))) & 7), a, b);
        mn = a;
        mx = b;
      }
      next.push_back({-1, sg.start + (int64_t)o.start, count, mn, mx});
      nextFrom.push_back(k);
    }
    if (!next.empty() && launch_round(next)) return 1;
    // the tree from this round's output, while the GPU refines the next
    std::vector<int32_t> outId(no.size());
    tree.nodes.reserve(tree.nodes.size() + 2 * no.size() + 64);
    for (size_t k = 0; k < no.size(); k++) {
      const NodeOut& o = no[k];
      const Seg& sg = segs[o.seg];
      int32_t id = sg.node;
      for (int i = 0; i < o.level; i++) id = tree.child(id, (o.morton >> (3 * (o.level - 1 - i))) & 7);
      TNode& t = tree.nodes[id];
      t.off = sg.start + o.start;
      t.count = kept[k] >= 0 ? kept[k] : (int64_t)o.count;
      octreeDepth = std::max<int64_t>(octreeDepth, (int64_t)t.level);
      outId[k] = id;
    }
    // refined nodes hold no records themselves
    for (size_t j = 0; j < next.size(); j++) {
      next[j].node = outId[nextFrom[j]];
      tree.nodes[next[j].node].off = -1;
      tree.nodes[next[j].node].count = 0;
    }
    segs.swap(next);
  }
  // segment roots that were split keep no records of their own
  for (int32_t id : chunkRoots) {
    TNode& t = tree.nodes[id];
    if (t.count >= kLeafCap && t.hasChildren()) { t.off = -1; t.count = 0; }
  }
  PCG_CHECK(hipStreamSynchronize(s));
  const double tb1 = now_s();
  out.t_build = tb1 - tb0;
  out.octreeDepth = octreeDepth;

  // ------------------------------------------------------------ sampling
  const int32_t numNodes = (int32_t)tree.nodes.size();
  out.numNodes = numNodes;
  static std::vector<NodeInfo> keepInfo;
  static std::vector<NodeState> keepState;
  Reused<NodeInfo> infoMem(keepInfo);
  Reused<NodeState> stateMem(keepState);
  std::vector<NodeInfo>& info = infoMem.v;
  std::vector<NodeState>& state = stateMem.v;
  info.resize(numNodes);
  state.resize(numNodes);
  int maxLevel = 0;
  for (int32_t i = 0; i < numNodes; i++) {
    const TNode& t = tree.nodes[i];
    NodeInfo& f = info[i];
    for (int c = 0; c < 8; c++) f.children[c] = t.children[c];
    for (int a = 0; a < 3; a++) { f.min[a] = t.min[a]; f.max[a] = t.max[a]; }
    f.level = t.level;
    f.pad = 0;
    maxLevel = std::max(maxLevel, f.level);
    NodeState& st = state[i];
    st.ptr = (t.off >= 0 && t.count > 0) ? dc.d_records + (size_t)t.off * bpp : nullptr;
    st.count = t.off >= 0 ? t.count : 0;
    st.accepted = st.count;
    st.flags = 0;
    st.pad = 0;
  }
  const bool singleChunk = chunkRoots.size() == 1;
  const char* envStats = getenv("PCG_LEVEL_STATS");
  const bool levelStats = envStats && envStats[0] == '1';
  // Sampling groups: the chunk subtrees are split into G sets (longest
  // processing time first on their point counts); each set samples its own
  // levels on a stream of its own, so that the narrow, latency-bound levels
  // of one set overlap the wide levels of another.  The nodes above the
  // chunk roots follow on the library stream once every set is done.  With
  // one group every sampled node is in the level lists of that group.
  int S = 1;  // streams of the chunk-subtree sampling
  if (!g_plain_sampling && !levelStats && chunkRoots.size() > 1) {
    int want = g_sample_groups;
    if (want <= 0) {
      // one group per hardware queue besides the library stream's: streams
      // that share a queue run one after the other
      const char* q = getenv("GPU_MAX_HW_QUEUES");
      const int queues = q && atoi(q) > 0 ? atoi(q) : 4;
      want = std::max(1, std::min(queues - 1, kMaxSampleGroups));
    }
    S = (int)std::min<size_t>((size_t)want, chunkRoots.size());
  }
  // every output range is rounded up to 16 bytes (9 ranges per node at most)
  const uint64_t arenaCap = (uint64_t)N * bpp * (avg ? 4 : 3) + (1u << 20) + (uint64_t)numNodes * 9 * 16;
  // Sets: the chunk subtrees are split into G >= S sets, set g sampled on
  // stream g % S with that stream's buffers.  A stream's buffers hold one
  // level of one set at a time (the scratch and neighbour pool restart at
  // every level), so the sets of a stream interleave level by level and the
  // buffers are sized by the stream's largest set.  More sets than streams
  // bound those buffers when the job would not fit in device memory
  // otherwise (waves of chunk subtrees); "sample_sets" forces a count.
  int G = S;
  if (!g_plain_sampling && !levelStats && chunkRoots.size() > 1) {
    const int maxSets = (int)std::min<size_t>(chunkRoots.size(), (size_t)kMaxSampleSets);
    if (g_sample_sets > 0) {
      G = std::max(S, std::min(g_sample_sets, maxSets));
    } else {
      size_t freeB = 0, totalB = 0;
      PCG_CHECK(hipMemGetInfo(&freeB, &totalB));
      // memory the sampling reuses or may release: the group buffers, the
      // arena and the octree image of the previous job, and the octree-build
      // and chunking buffers of this one
      size_t held = workspace_bytes(WS_ARENA) + workspace_bytes(WS_OCTREE);
      for (int k = 0; k < kMaxSampleGroups * 8; k++) held += workspace_bytes(WS_GROUP0 + k);
      static const int kRelease[] = {WS_OCT_B, WS_OCT_KEYS_A, WS_OCT_KEYS_B, WS_OCT_M, WS_OCT_SCAN,
                                     WS_POINT_IDS, WS_CHUNK_IDS16, WS_TILE_M, WS_SCAN};
      size_t releasable = 0;
      for (int k : kRelease) releasable += workspace_bytes(k);
      const double margin = 2e9 + 0.03 * (double)totalB;
      // per-candidate buffer bytes of a stream (scratch bound + neighbour pool)
      const int mId = avg ? METHOD_POISSON_AVERAGE : (rnd ? METHOD_RANDOM : METHOD_POISSON);
      const double perCand = (double)sample_scratch_bound(1 << 20, 1, mId) / (1 << 20) + 16.0;
      const double image = (double)N * bpp * (avg ? 2.0 : 1.0);
      auto need = [&](int sets) {
        return 1.125 * ((double)arenaCap + image + S * perCand * (double)((N + sets - 1) / sets));
      };
      const double avail = (double)freeB + (double)held - margin;
      if (need(S) > avail && releasable > 0) {
        for (int k : kRelease) release_workspace_slot(k);
      }
      const double avail2 = avail + (need(S) > avail ? (double)releasable : 0.0);
      while (G < maxSets && need(G) > avail2) G = std::min(maxSets, G * 2);
    }
  }
  std::vector<int32_t> groupOf(numNodes, G > 1 ? -1 : 0);  // -1: above the chunk roots
  std::vector<int64_t> groupPoints(G, 0);
  if (G > 1) {
    std::vector<size_t> crs(chunkRoots.size());
    for (size_t i = 0; i < crs.size(); i++) crs[i] = i;
    std::stable_sort(crs.begin(), crs.end(), [&](size_t x, size_t y) { return chunkRootPoints[x] > chunkRootPoints[y]; });
    // Set policy 1: the chunk subtrees that reach below the level where most
    // points end (dense spots: long chains of narrow, latency-bound levels)
    // take sets of their own, in proportion to their points, so that their
    // chains run beside the wide levels of the others instead of every set
    // starting with narrow levels at once.
    std::vector<int> deepest(chunkRoots.size(), 0);
    int setsDeep = 0;
    if (g_set_policy == 1) {
      std::map<int, int64_t> pointsAt;
      for (size_t ci = 0; ci < chunkRoots.size(); ci++) {
        std::vector<int32_t> stack{chunkRoots[ci]};
        while (!stack.empty()) {
          const int32_t id = stack.back();
          stack.pop_back();
          deepest[ci] = std::max(deepest[ci], tree.nodes[id].level);
          for (int c = 0; c < 8; c++)
            if (tree.nodes[id].children[c] >= 0) stack.push_back(tree.nodes[id].children[c]);
        }
        pointsAt[deepest[ci]] += chunkRootPoints[ci];
      }
      int mode = 0;
      int64_t best = -1;
      for (auto& kv : pointsAt)
        if (kv.second > best) { best = kv.second; mode = kv.first; }
      int64_t deepPoints = 0;
      for (size_t ci = 0; ci < chunkRoots.size(); ci++)
        if (deepest[ci] > mode) deepPoints += chunkRootPoints[ci];
      if (deepPoints > 0 && deepPoints < N)
        setsDeep = std::max(1, std::min(G - 1, (int)std::lround((double)G * (double)deepPoints / (double)N)));
      for (size_t ci = 0; ci < chunkRoots.size(); ci++) deepest[ci] = deepest[ci] > mode ? 1 : 0;
    }
    for (size_t ci : crs) {
      // longest processing time first within the chunk's share of the sets
      const bool deep = setsDeep > 0 && deepest[ci];
      const int lo = setsDeep == 0 ? 0 : (deep ? 0 : setsDeep), hi = setsDeep == 0 ? G : (deep ? setsDeep : G);
      int g = (int)(std::min_element(groupPoints.begin() + lo, groupPoints.begin() + hi) - groupPoints.begin());
      groupPoints[g] += chunkRootPoints[ci];
      std::vector<int32_t> stack{chunkRoots[ci]};
      while (!stack.empty()) {
        int32_t id = stack.back();
        stack.pop_back();
        groupOf[id] = g;
        for (int c = 0; c < 8; c++)
          if (tree.nodes[id].children[c] >= 0) stack.push_back(tree.nodes[id].children[c]);
      }
    }
  } else {
    groupPoints[0] = N;
  }
  // Launch rounds.  A node samples the records its children kept, so any
  // order that puts every node after its children gives the same records;
  // the output is assembled in the canonical order afterwards.  By octree
  // level (deepest first) a round also waits for unrelated deeper subtrees
  // (dense spots reach many levels below the surrounding terrain); by height
  // (1 + the largest height of its children, 0 for a leaf) each node runs in
  // the first round after its children's, so the leaves' parents of every
  // depth start at once.  Nodes of one height are never ancestor and
  // descendant, so a round's candidates are distinct points, as a level's.
  // Rounds are indexed like levels: round r runs when the loop reaches
  // index maxLevel + 1 - r (heights run from 1 to at most maxLevel).
  std::vector<int32_t> roundOf(numNodes, 0);
  if (g_sample_order == 1) {
    std::vector<std::vector<int32_t>> byLevel(maxLevel + 1);
    for (int32_t i = 0; i < numNodes; i++) byLevel[info[i].level].push_back(i);
    std::vector<int32_t> height(numNodes, 0);
    for (int l = maxLevel; l >= 0; l--)
      for (int32_t i : byLevel[l]) {
        int32_t h = 0;
        for (int c = 0; c < 8; c++) {
          const int32_t ch = tree.nodes[i].children[c];
          if (ch >= 0) h = std::max(h, height[ch] + 1);
        }
        height[i] = h;
        roundOf[i] = maxLevel + 1 - h;
      }
  } else {
    for (int32_t i = 0; i < numNodes; i++) roundOf[i] = info[i].level;
  }
  // round lists: group g in levels[g], the nodes above the chunk roots in
  // levels[G] (empty with one group)
  std::vector<std::vector<std::vector<int32_t>>> levels(G + 1, std::vector<std::vector<int32_t>>(maxLevel + 1));
  for (int32_t i = 0; i < numNodes; i++) {
    if (!tree.nodes[i].hasChildren() || (singleChunk && tree.nodes[i].upper)) continue;
    const int g = groupOf[i] >= 0 ? groupOf[i] : G;
    levels[g][roundOf[i]].push_back(i);
  }
  std::vector<int64_t> maxLevelNodes(G + 1, 1);
  for (int g = 0; g <= G; g++)
    for (auto& l : levels[g]) maxLevelNodes[g] = std::max<int64_t>(maxLevelNodes[g], (int64_t)l.size());
  DevMem dinfo(WS_INFO), dstate(WS_STATE), dlist(WS_LIST), arena(WS_ARENA), counters(WS_COUNTERS),
      gtop(WS_TOPS);
  const int methodId = avg ? METHOD_POISSON_AVERAGE : (rnd ? METHOD_RANDOM : METHOD_POISSON);
  // SamplerRandom shuffles every leaf before its parent samples it
  // (sampler_random.h:89-121): the shuffled copies replace the leaves'
  // records, so the state table points at them from the start
  std::vector<ShuffleLeaf> shuffle;
  DevMem shufDesc(WS_SHUF_DESC), shufOut(WS_SHUF_OUT), shufIdx(WS_SHUF_IDX);
  if (rnd) {
    uint64_t outBytes = 0, idxBytes = 0;
    for (int32_t i = 0; i < numNodes; i++) {
      if (tree.nodes[i].hasChildren() || tree.nodes[i].presampled || state[i].count <= 0) continue;
      const int64_t c = state[i].count;
      shuffle.push_back(ShuffleLeaf{state[i].ptr, (uint8_t*)(uintptr_t)outBytes,
                                    (uint32_t*)(uintptr_t)(c > kShuffleLdsPoints ? idxBytes + 1 : 0), c});
      outBytes += ((uint64_t)c * bpp + 15) & ~15ull;
      if (c > kShuffleLdsPoints) idxBytes += (4 * (uint64_t)c + 15) & ~15ull;
    }
    if (shufOut.alloc(outBytes + 64) || shufIdx.alloc(idxBytes + 64) ||
        shufDesc.alloc(sizeof(ShuffleLeaf) * shuffle.size() + 16))
      return 1;
    size_t q = 0;
    for (int32_t i = 0; i < numNodes; i++) {
      if (tree.nodes[i].hasChildren() || tree.nodes[i].presampled || state[i].count <= 0) continue;
      ShuffleLeaf& sl = shuffle[q++];
      sl.dst = shufOut.as<uint8_t>() + (uintptr_t)sl.dst;
      sl.idx = sl.idx ? (uint32_t*)(shufIdx.as<uint8_t>() + ((uintptr_t)sl.idx - 1)) : nullptr;
      state[i].ptr = sl.dst;
    }
    if (!shuffle.empty())
      PCG_CHECK(hipMemcpyAsync(shufDesc.p, shuffle.data(), sizeof(ShuffleLeaf) * shuffle.size(),
                               hipMemcpyHostToDevice, s));
    uint32_t seed = (uint32_t)g_random_seed;
    if (g_random_seed < 0)  // the reference's seed: the clock (sampler_random.h:102)
      seed = (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
    if (shuffle_leaves(shufDesc.as<ShuffleLeaf>(), (int64_t)shuffle.size(), bpp, seed, s)) return 1;
  }
  // all level lists in one upload: per group, deepest level first
  std::vector<int32_t> allLevels;
  std::vector<std::vector<size_t>> levelOffset(G + 1, std::vector<size_t>(maxLevel + 1, 0));
  for (int g = 0; g <= G; g++)
    for (int l = maxLevel; l >= 0; l--) {
      levelOffset[g][l] = allLevels.size();
      allLevels.insert(allLevels.end(), levels[g][l].begin(), levels[g][l].end());
    }
  if (dinfo.alloc(sizeof(NodeInfo) * numNodes) || dstate.alloc(sizeof(NodeState) * numNodes) ||
      dlist.alloc(sizeof(int32_t) * (allLevels.size() + 1)) || arena.alloc(arenaCap) || counters.alloc(128) ||
      gtop.alloc(64))
    return 1;
  // per-group buffers: node work, item and segment tables, neighbour pool,
  // scratch, and the group's scratch / pool tops
  struct GroupMem {
    DevMem work, items, outsegs, segstart, pool, scratch, tops;
    explicit GroupMem(int g)
        : work(ws_group_slot(g, 0)), items(ws_group_slot(g, 1)), outsegs(ws_group_slot(g, 2)),
          segstart(ws_group_slot(g, 3)), pool(ws_group_slot(g, 4)), scratch(ws_group_slot(g, 5)),
          tops(ws_group_slot(g, 6)) {}
  };
  std::vector<GroupMem> gm;
  for (int k = 0; k < S; k++) gm.emplace_back(k);
  SampleParams sp;
  memset(&sp, 0, sizeof(sp));
  for (int a = 0; a < 3; a++) { sp.scale[a] = attributes.posScale[a]; sp.offset[a] = attributes.posOffset[a]; }
  const double spacing = sub ? sub->baseSpacing : (dc.max - dc.min).x / 128.0;  // indexer.cpp:1603
  sp.base_spacing = spacing;
  sp.bpp = bpp;
  sp.rgb_offset = rgbOffset;
  sp.method = methodId;
  sp.random_limit = 0.7 * std::sqrt(3.0);  // sampler_random.h:159-164
  sp.arena = arena.as<uint8_t>();
  sp.arena_cap = arenaCap;
  sp.arena_top = gtop.as<uint64_t>();
  sp.counters = counters.as<unsigned long long>();
  sp.num_cus = device().num_cus;
  sp.ring_blocks = g_ring_blocks_per_cu * device().num_cus;  // at most three k_ring blocks fit a CU's LDS
  sp.copy_blocks = g_copy_blocks;
  sp.neighbor_blocks = g_neighbor_blocks;
  sp.avg_blocks = g_avg_blocks;
  sp.bucket_min_nodes = g_bucket_min_nodes;
  sp.accept_pipe_nodes = g_accept_pipe_nodes;
  sp.ring_max_cand = g_ring_max_cand;
  sp.pipe_form = g_pipe_form;
  // a group's buffers for levels of at most `maxNodes` nodes and `cand`
  // candidates in all
  auto group_params = [&](GroupMem& m, int64_t maxNodes, int64_t cand, SampleParams& p) -> int {
    // every candidate of a level is a point of the group's chunk records;
    // neighbour lists longer than a row (dense clusters): 4 entries per point
    const uint64_t scratchCap = sample_scratch_bound(cand, maxNodes, methodId);
    const uint64_t poolCap = (uint64_t)cand * 4 + 1024;
    if (m.work.alloc(sizeof(NodeWork) * maxNodes) || m.items.alloc(sizeof(uint32_t) * (maxNodes + 1)) ||
        m.outsegs.alloc(sizeof(OutSeg) * 9 * maxNodes) || m.segstart.alloc(sizeof(uint32_t) * (9 * maxNodes + 1)) ||
        m.pool.alloc(sizeof(uint32_t) * poolCap) || m.scratch.alloc(scratchCap) || m.tops.alloc(64))
      return 1;
    p = sp;
    p.scratch = m.scratch.as<uint8_t>();
    p.scratch_cap = scratchCap;
    p.scratch_top = m.tops.as<uint64_t>() + 1;
    p.item_start = m.items.as<uint32_t>();
    p.segs = m.outsegs.as<OutSeg>();
    p.seg_start = m.segstart.as<uint32_t>();
    p.nbr_pool = m.pool.as<uint32_t>();
    p.nbr_pool_cap = poolCap;
    p.nbr_pool_top = (unsigned long long*)(m.tops.as<uint64_t>() + 2);
    p.ring_ctl = (uint32_t*)(m.tops.as<uint64_t>() + 3);
    // the ring needs enough nodes per launch to fill the GPU; concurrent
    // streams share it
    p.ring_min_nodes = std::max(1, g_ring_min_nodes / S);
    return 0;
  };
  // each stream's buffers for the largest of its sets
  std::vector<SampleParams> gp(G);
  for (int k = 0; k < S; k++) {
    int64_t mn = 1, mc = 0;
    for (int g = k; g < G; g += S) {
      mn = std::max(mn, maxLevelNodes[g]);
      mc = std::max(mc, groupPoints[g]);
    }
    SampleParams base;
    if (group_params(gm[k], mn, mc, base)) return 1;
    for (int g = k; g < G; g += S) gp[g] = base;
  }
  if (!allLevels.empty())
    PCG_CHECK(hipMemcpyAsync(dlist.p, allLevels.data(), sizeof(int32_t) * allLevels.size(), hipMemcpyHostToDevice, s));
  PCG_CHECK(hipMemcpyAsync(dinfo.p, info.data(), sizeof(NodeInfo) * numNodes, hipMemcpyHostToDevice, s));
  PCG_CHECK(hipMemcpyAsync(dstate.p, state.data(), sizeof(NodeState) * numNodes, hipMemcpyHostToDevice, s));
  PCG_CHECK(hipMemsetAsync(gtop.p, 0, 64, s));
  PCG_CHECK(hipMemsetAsync(counters.p, 0, 128, s));
  int64_t sampled = 0;
  // the captured graph of the chunk-subtree sets, destroyed when the job
  // ends (after waiting for it); a capture left open by an error is closed
  struct SampleGraph {
    hipStream_t capturing = nullptr, launched = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int launch() {
      hipStream_t st = capturing;
      capturing = nullptr;
      PCG_CHECK(hipStreamEndCapture(st, &graph));
      PCG_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      PCG_CHECK(hipGraphLaunch(exec, st));
      launched = st;
      return 0;
    }
    ~SampleGraph() {
      if (capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(capturing, &g);
        if (g) (void)hipGraphDestroy(g);
      }
      if (launched) (void)hipStreamSynchronize(launched);
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
    }
  } sampleGraph;
  // one level of one list set on one stream
  auto run_level = [&](int lg, int l, GroupMem& m, const SampleParams& p, hipStream_t st) -> int {
    {
      const auto& lst = levels[lg][l];
      if (lst.empty()) return 0;
      sampled += (int64_t)lst.size();
      if (sample_zero(m.tops.as<uint64_t>() + 1, 2, st)) return 1;  // scratch and pool tops
      double tLevel = 0.0;
      if (levelStats) {
        PCG_CHECK(hipStreamSynchronize(st));
        tLevel = now_s();
      }
      if (g_plain_sampling) {
        // one launch sequence per node (the scratch of every node of the level
        // stays allocated until the level is done, as in the batched mode)
        for (size_t q = 0; q < lst.size(); q++)
          if (sample_level(dlist.as<int32_t>() + levelOffset[lg][l] + q, 1, dinfo.as<NodeInfo>(),
                           dstate.as<NodeState>(), m.work.as<NodeWork>() + q, p, st))
            return 1;
      } else if (sample_level(dlist.as<int32_t>() + levelOffset[lg][l], (int32_t)lst.size(), dinfo.as<NodeInfo>(),
                              dstate.as<NodeState>(), m.work.as<NodeWork>(), p, st)) {
        return 1;
      }
      if (levelStats) {
        // diagnostic summary of the level (PCG_LEVEL_STATS=1), costs a sync
        PCG_CHECK(hipStreamSynchronize(st));
        tLevel = now_s() - tLevel;
        std::vector<NodeWork> wk(lst.size());
        PCG_CHECK(stream_copy(wk.data(), m.work.p, sizeof(NodeWork) * wk.size(), hipMemcpyDeviceToHost));
        int64_t cand = 0, maxn = 0, flagged[6] = {0, 0, 0, 0, 0, 0};
        std::vector<int64_t> ns;
        for (const NodeWork& w : wk) {
          ns.push_back(w.n);
          cand += w.n;
          maxn = std::max<int64_t>(maxn, w.n);
          for (int b = 0; b < 6; b++) flagged[b] += (w.flags >> b) & 1;
        }
        unsigned long long hc7 = 0;
        PCG_CHECK(stream_copy(&hc7, counters.as<unsigned long long>() + 7, 8, hipMemcpyDeviceToHost));
        std::sort(ns.begin(), ns.end());
        {
          // the largest nodes outside the ring (cell path, exact sort)
          std::vector<std::pair<int64_t, uint32_t>> big;
          for (const NodeWork& w : wk)
            if (w.flags & (NW_CELL_PATH | NW_EXACT_SORT)) big.push_back({w.n, w.flags});
          std::sort(big.begin(), big.end(), [](auto& a, auto& b) { return a.first > b.first; });
          std::string txt;
          for (size_t i = 0; i < big.size() && i < 6; i++)
            txt += " " + std::to_string(big[i].first) + "/" + std::to_string(big[i].second);
          fprintf(stderr, "[%s %d] largest flagged nodes (candidates/flags):%s\n", g_sample_order ? "round" : "level",
                  g_sample_order ? maxLevel + 1 - l : l, txt.c_str());
        }
        fprintf(stderr, "[%s %d] %.3f ms, candidates per node p50 %lld p90 %lld p99 %lld\n",
                g_sample_order ? "round" : "level", g_sample_order ? maxLevel + 1 - l : l, 1e3 * tLevel,
                (long long)ns[ns.size() / 2], (long long)ns[ns.size() * 9 / 10], (long long)ns[ns.size() * 99 / 100]);
        fprintf(stderr,
                "[%s %d] nodes %zu candidates %lld max %lld exact-sort %lld seq-accept %lld seq-average %lld "
                "overflow %lld many-neighbours %lld cell-path %lld (long lists so far %llu)\n",
                g_sample_order ? "round" : "level", g_sample_order ? maxLevel + 1 - l : l, lst.size(),
                (long long)cand, (long long)maxn, (long long)flagged[0], (long long)flagged[1],
                (long long)flagged[2], (long long)flagged[3], (long long)flagged[4], (long long)flagged[5], hc7);
      }
    }
    return 0;
  };
  auto run_levels = [&](int lg, GroupMem& m, const SampleParams& p, hipStream_t st) -> int {
    for (int l = maxLevel; l >= 0; l--)
      if (run_level(lg, l, m, p, st)) return 1;
    return 0;
  };
  // ------------------------- completion order, phase 1 (host, meanwhile)
  // The walk of phase 1 depends on the tree only; which children it
  // completes depends on the sampling results, so the walk records every
  // child it would complete and the filter runs once the results are back.
  std::vector<uint8_t> sampledFlag(numNodes, 0);
  std::vector<int32_t> order;  // node ids in onNodeCompleted order
  std::vector<int32_t> phase1Cand;
  bool collectPhase1 = true;
  auto complete_children = [&](int32_t id) {
    if (!collectPhase1) {
      for (int c = 0; c < 8; c++) {
        int32_t ch = tree.nodes[id].children[c];
        if (ch >= 0 && (state[ch].flags & NS_COMPLETED)) order.push_back(ch);
      }
      return;
    }
    for (int c = 0; c < 8; c++) {
      int32_t ch = tree.nodes[id].children[c];
      if (ch >= 0) phase1Cand.push_back(ch);
    }
  };
  // sample(): post-order over not-yet-sampled children (sampler_poisson.h:27-36)
  // phase 1 samples each chunk subtree; afterwards the chunk roots are
  // leaves of the upper tree (their children were cleared, indexer.cpp:1675)
  std::function<void(int32_t, bool)> sample_rec = [&](int32_t id, bool phase1) {
    if (!phase1 && tree.nodes[id].chunkRoot) {
      sampledFlag[id] = 1;
      return;
    }
    for (int c = 0; c < 8; c++) {
      int32_t ch = tree.nodes[id].children[c];
      if (ch < 0 || sampledFlag[ch]) continue;
      sample_rec(ch, phase1);
    }
    sampledFlag[id] = 1;
    if (tree.nodes[id].hasChildren()) complete_children(id);
  };
  // While the GPU samples, the host also lays out hierarchy.bin for the
  // case that every node of the tree completes (SamplerPoisson and
  // SamplerRandom complete every child they sample; SamplerPoissonAverage
  // discards some): the layout then depends on the tree only, and after
  // sampling only the record order and the records' values are filled in.
  HierarchyPlan plan;
  bool planned = false;
  std::vector<int32_t> rank;
  std::vector<int64_t> recNp, recOff, recSize;
  std::vector<CopyDesc> copies;
  auto walk_phase1 = [&]() {
    for (int32_t id : chunkRoots) sample_rec(id, true);  // ascending chunk ID
    collectPhase1 = false;
    std::vector<uint64_t> keys(numNodes);
    for (int32_t i = 0; i < numNodes; i++) keys[i] = tree.nodes[i].key;
    planned = plan_hierarchy(keys, plan);
    // the per-node arrays of the assembly, touched now rather than after
    // sampling
    rank.assign(numNodes, -1);
    recNp.assign(numNodes, 0);
    recOff.assign(numNodes, 0);
    recSize.assign(numNodes, 0);
    if (!brotli) copies.assign(numNodes, CopyDesc{nullptr, 0, 0});
  };
  if (G == 1) {
    if (run_levels(0, gm[0], gp[0], s)) return 1;
    if (!sub) walk_phase1();
  } else {
    std::vector<hipStream_t> gs;
    if (group_streams(S, gs)) return 1;
    hipEvent_t ev_start = group_event(0);
    if (!ev_start) return 1;
    for (int k = 0; k < S; k++)
      if (!group_event(1 + k)) return 1;
    // With "sample_graph" the launch sequence of every set (streams forked
    // from and joined back to the library stream by events) is captured
    // into one graph, instantiated and launched on the library stream.
    const bool graph = g_sample_graph && !g_plain_sampling && !levelStats;
    if (graph) {
      PCG_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      sampleGraph.capturing = s;
    }
    PCG_CHECK(hipEventRecord(ev_start, s));
    // every group's work is enqueued before the library stream's waits, so
    // that a stream sharing a hardware queue with the library stream is
    // never queued behind a wait for another group
    // levels are enqueued round-robin over the groups so that every group
    // starts at once
    for (int k = 0; k < S; k++) PCG_CHECK(hipStreamWaitEvent(gs[k], ev_start, 0));
    for (int l = maxLevel; l >= 0; l--)
      for (int g = 0; g < G; g++)
        if (run_level(g, l, gm[g % S], gp[g], gs[g % S])) return 1;
    for (int k = 0; k < S; k++) PCG_CHECK(hipEventRecord(group_event(1 + k), gs[k]));
    for (int k = 0; k < S; k++) PCG_CHECK(hipStreamWaitEvent(s, group_event(1 + k), 0));
    if (graph && sampleGraph.launch()) return 1;
    walk_phase1();
    // the nodes above the chunk roots: their candidates are points the
    // chunk roots kept
    int64_t upperCand = 0;
    {
      std::vector<NodeState> st2(numNodes);
      PCG_CHECK(stream_copy(st2.data(), dstate.p, sizeof(NodeState) * numNodes, hipMemcpyDeviceToHost));
      for (int32_t cr : chunkRoots) upperCand += st2[cr].count;
    }
    SampleParams up;
    if (group_params(gm[0], maxLevelNodes[G], upperCand, up)) return 1;
    up.ring_min_nodes = g_ring_min_nodes;
    if (run_levels(G, gm[0], up, s)) return 1;
  }
  out.numSampled = sampled;
  PCG_CHECK(stream_copy(state.data(), dstate.p, sizeof(NodeState) * numNodes, hipMemcpyDeviceToHost));
  unsigned long long hc[16];
  PCG_CHECK(stream_copy(hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost));
  for (int i = 0; i < 8; i++) out.counters[i] = hc[i];
  if (hc[8] != 0) {
    set_error("equal-distance order check failed in " + std::to_string(hc[8]) + " runs");
    return 1;
  }
  if (hc[9] != 0) {
    set_error("pipelined acceptance gave up waiting in " + std::to_string(hc[9]) + " nodes");
    return 1;
  }
  if (hc[6] != 0) {
    set_error("sampling arena or scratch exhausted in " + std::to_string(hc[6]) + " nodes");
    return 1;
  }
  const double ts1 = now_s();
  out.t_sample = ts1 - tb1;

  if (sub) {
    // Sampler::sample's callbacks (sampler_poisson.h:247-294,
    // sampler_random.h:184-225, sampler_poisson_average.h:418-482): in
    // post-order over the visited nodes, each inner node's children in
    // index order.  Poisson / random: a child left without points that is a
    // leaf at that moment (its own children may have been unlinked) is
    // discarded and unlinked, then completed all the same (the first `if`
    // at sampler_poisson.h:264 has no `else`); Poisson-average: a child
    // without rejected points is unlinked without a callback.
    SubtreeResult& R = *subOut;
    R.ptr.assign(numNodes, nullptr);
    R.count.assign(numNodes, 0);
    R.children.assign((size_t)8 * numNodes, -1);
    R.events.clear();
    for (int32_t i = 0; i < numNodes; i++) {
      R.ptr[i] = state[i].ptr;
      R.count[i] = state[i].count;
      for (int c = 0; c < 8; c++) R.children[8 * i + c] = tree.nodes[i].children[c];
    }
    std::function<void(int32_t)> visit = [&](int32_t id) {
      const TNode& t = tree.nodes[id];
      if (!t.hasChildren()) return;
      for (int c = 0; c < 8; c++)
        if (t.children[c] >= 0) visit(t.children[c]);
      for (int c = 0; c < 8; c++) {
        const int32_t ch = t.children[c];
        if (ch < 0) continue;
        const NodeState& cs = state[ch];
        if (avg) {
          if (cs.flags & NS_DISCARDED) R.children[8 * id + c] = -1;
          else R.events.emplace_back(ch, 0);
          continue;
        }
        bool leafNow = true;
        for (int d = 0; d < 8; d++)
          if (R.children[8 * ch + d] >= 0) leafNow = false;
        if (cs.count == 0 && leafNow) {
          R.events.emplace_back(ch, 1);
          R.children[8 * id + c] = -1;
        }
        R.events.emplace_back(ch, 0);
      }
    };
    visit(0);
    out.numSampled = sampled;
    out.t_total = now_s() - t0;
    return 0;
  }

  // --------------------------------------------- canonical completion order
  // phase 1 was walked while the GPU sampled; keep the children that
  // completed
  for (int32_t id : phase1Cand)
    if (state[id].flags & NS_COMPLETED) order.push_back(id);
  // processChunkRoots (indexer.cpp:202-281): subtrees of the upper tree with
  // fewer than 5M chunk-root points become one task; tasks in pre-order
  {
    std::vector<int64_t> crPoints(numNodes, 0);
    std::vector<uint8_t> merged(numNodes, 0), hasFcr(numNodes, 0);
    for (int32_t id : chunkRoots) {
      crPoints[id] = (int64_t)(int32_t)state[id].accepted;
      hasFcr[id] = 1;
    }
    std::function<void(int32_t)> post = [&](int32_t id) {
      if (!tree.nodes[id].upper) return;
      bool leaf = true;
      for (int c = 0; c < 8; c++) {
        int32_t ch = tree.nodes[id].children[c];
        if (ch >= 0) { post(ch); leaf = false; }
      }
      if (leaf) return;
      int32_t np = 0;
      for (int c = 0; c < 8; c++) {
        int32_t ch = tree.nodes[id].children[c];
        if (ch >= 0) np += (int32_t)crPoints[ch];
      }
      crPoints[id] = np;
      if (np < 5000000) {
        merged[id] = 1;
        for (int c = 0; c < 8; c++) {
          int32_t ch = tree.nodes[id].children[c];
          if (ch >= 0 && hasFcr[ch]) hasFcr[id] = 1;
        }
      }
    };
    post(0);
    std::vector<int32_t> tasks;
    std::function<void(int32_t)> pre = [&](int32_t id) {
      if (hasFcr[id]) tasks.push_back(id);
      if (merged[id]) return;  // its children were cleared in the CR tree
      if (!tree.nodes[id].upper) return;
      for (int c = 0; c < 8; c++) {
        int32_t ch = tree.nodes[id].children[c];
        if (ch >= 0) pre(ch);
      }
    };
    pre(0);
    for (int32_t t : tasks) sample_rec(t, false);
  }
  if (chunkRoots.size() == 1) {
    // indexer.cpp:1736-1739: the single chunk root becomes the root
  } else if (!sampledFlag[0]) {
    sample_rec(0, false);
  }
  int32_t rootId = chunkRoots.size() == 1 ? chunkRoots[0] : 0;
  order.push_back(rootId);

  // byte offsets (Writer::writeAndUnload, indexer.cpp:1453-1516), per node
  for (int32_t id : order) recNp[id] = state[id].count;
  if (brotli) {
    std::vector<EncodeInput> encIn;
    std::vector<int32_t> encId;
    for (int32_t id : order) {
      const NodeState& st = state[id];
      if (st.count != 0) {
        encIn.push_back(EncodeInput{st.ptr, st.count});
        encId.push_back(id);
      }
    }
    std::vector<int64_t> sizes;
    if (encode_brotli(attributes, bpp, encIn, out.h_octree, sizes, s)) return 1;
    int64_t offset = 0;
    for (size_t i = 0; i < encId.size(); i++) {
      recOff[encId[i]] = offset;
      recSize[encId[i]] = sizes[i];
      offset += sizes[i];
    }
    out.octreeBytes = (uint64_t)offset;
    out.octreeOnHost = true;
  } else {
    size_t nc = 0;
    uint64_t offset = 0;
    if (copies.size() < order.size()) copies.resize(order.size());
    for (int32_t id : order) {
      const NodeState& st = state[id];
      if (st.count != 0) {
        const uint64_t bytes = (uint64_t)st.count * bpp;
        recOff[id] = (int64_t)offset;
        recSize[id] = (int64_t)bytes;
        copies[nc++] = CopyDesc{st.ptr, offset, bytes};
        offset += bytes;
      }
    }
    out.octreeBytes = offset;
    out.d_octree = (uint8_t*)workspace(WS_OCTREE, offset + 64);
    if (!out.d_octree) return 1;
    DevMem dcopies(WS_COPYDESC);
    if (dcopies.alloc(nc * sizeof(CopyDesc) + 16)) return 1;
    if (nc) {
      PCG_CHECK(hipMemcpyAsync(dcopies.p, copies.data(), nc * sizeof(CopyDesc), hipMemcpyHostToDevice, s));
      if (gather_ranges(dcopies.as<CopyDesc>(), (int64_t)nc, out.d_octree, s)) return 1;
    }
  }
  const double ta0 = now_s();
  out.t_order = ta0 - ts1;
  // hierarchy.bin and metadata.json on the host while the GPU gathers
  // octree.bin (root name is the chunk id when there is a single chunk)
  int32_t ranked = 0;
  for (size_t q = 0; q < order.size(); q++)
    if (rank[order[q]] < 0) {
      rank[order[q]] = (int32_t)q;
      ranked++;
    }
  out.completed.clear();
  if (planned && ranked == numNodes && order.size() == (size_t)numNodes) {
    out.hierarchy = fill_hierarchy(plan, rank.data(), recNp.data(), recOff.data(), recSize.data());
    out.firstChunkSize = plan.firstChunkSize;
  } else {
    // completion records in order (some nodes were discarded, or the tree
    // is deeper than node keys reach)
    out.completed.reserve(order.size());
    for (int32_t id : order) {
      CompletedNode cn;
      cn.key = tree.nodes[id].key;
      if (!cn.key) cn.name = tree.nodes[id].full_name();
      cn.numPoints = recNp[id];
      cn.byteOffset = recOff[id];
      cn.byteSize = recSize[id];
      out.completed.push_back(cn);
    }
    // every hierarchy chunk's root (names of 1 + 4k characters) must be
    // among the written nodes: HierarchyBuilder points to it from the batch
    // above (HierarchyBuilder.h:213-226 exits with 123 when it is missing;
    // :317-326 dereferences it).  Poisson-average unlinks a node without
    // writing it when its parent accepts all its points (e.g. a sparse
    // cluster in a large cube), which leaves its written descendants orphaned.
    {
      std::unordered_map<std::string, bool> written;
      for (const CompletedNode& c : out.completed) written[c.key ? node_name(c.key) : c.name] = true;
      for (const auto& kv : written)
        for (size_t len = 5; len < kv.first.size(); len += 4)
          if (!written.count(kv.first.substr(0, len))) {
            set_error("didn't find proxy node " + kv.first.substr(0, len) + " (node " + kv.first +
                      " was written, its hierarchy chunk root was not)");
            return 123;
          }
    }
    out.hierarchy = build_hierarchy_bin(out.completed, out.firstChunkSize);
  }
  out.metadata = create_metadata_json(attributes, dc.min, dc.max, spacing, name, pointsTotal, out.firstChunkSize,
                                      octreeDepth, encoding);
  const double ta1 = now_s();
  out.t_hierarchy = ta1 - ta0;
  PCG_CHECK(hipStreamSynchronize(s));
  out.t_assemble = now_s() - ts1;
  out.t_setup = tb0 - t0;
  out.t_total = now_s() - t0;
  if (const char* mr = getenv("PCG_MEM_REPORT"); mr && mr[0] == '1') {
    // diagnostic: device memory held by the library's buffers after the job
    fprintf(stderr, "[memory] %.3f GB in workspace buffers:", workspace_total() / 1e9);
    for (int k = 0; k < WS_COUNT; k++)
      if (workspace_bytes(k) >= (size_t)1 << 26) fprintf(stderr, " slot%d=%.2fGB", k, workspace_bytes(k) / 1e9);
    fprintf(stderr, "\n");
  }
  return 0;
}

}  // namespace pcg
