// hierarchy.bin and metadata.json writers.
//
// hierarchy.bin follows HierarchyFlusher (Converter/include/indexer.h:146-243)
// and HierarchyBuilder (Converter/include/HierarchyBuilder.h:26-354): nodes
// are grouped into batches by their first 5 name characters, each batch is
// split into chunks of 4 levels, chunks and nodes are ordered breadth-first,
// node types are LEAF / NORMAL / PROXY and proxies carry the absolute byte
// range of the chunk they continue in.  22-byte records: u8 type,
// u8 childMask, u32 numPoints, u64 byteOffset, u64 byteSize.  The batch file
// order of the reference is directory order; here it is ascending name.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <unordered_map>

#include "indexing.h"

namespace fs = std::filesystem;

namespace pcg {

namespace {

constexpr int kStep = 4;  // indexer.cpp:19 hierarchyStepSize
enum { T_NORMAL = 0, T_LEAF = 1, T_PROXY = 2 };

struct HNode {
  std::string name;
  int32_t numPoints = 0;
  uint8_t childMask = 0;
  uint64_t byteOffset = 0, byteSize = 0;
  int type = T_LEAF;
  uint64_t proxyByteOffset = 0, proxyByteSize = 0;
};
struct HChunk {
  std::string name;
  std::vector<std::shared_ptr<HNode>> nodes;
  int64_t byteOffset = 0;
};
struct HBatch {
  std::string name;
  int64_t byteSize = 0;
  std::vector<std::shared_ptr<HNode>> nodes;
  std::vector<std::shared_ptr<HChunk>> chunks;
  std::unordered_map<std::string, std::shared_ptr<HNode>> nodeMap;
  std::unordered_map<std::string, std::shared_ptr<HChunk>> chunkMap;
};

bool breadth_first_less(const std::string& a, const std::string& b) {
  if (a.size() != b.size()) return a.size() < b.size();
  return a < b;
}

std::shared_ptr<HBatch> load_batch(const std::string& name, const std::vector<const CompletedNode*>& recs) {
  auto batch = std::make_shared<HBatch>();
  batch->name = name;
  auto chunk_of = [&](const std::string& key) {
    auto it = batch->chunkMap.find(key);
    if (it != batch->chunkMap.end()) return it->second;
    auto c = std::make_shared<HChunk>();
    c->name = key;
    batch->chunkMap[key] = c;
    batch->chunks.push_back(c);
    return c;
  };
  for (const CompletedNode* r : recs) {
    auto node = std::make_shared<HNode>();
    // the 48-byte flush record keeps 31 name bytes, u32 numPoints, i32 byteSize
    node->name = r->name.substr(0, 31);
    node->numPoints = (int32_t)(uint32_t)r->numPoints;
    node->byteOffset = (uint64_t)r->byteOffset;
    node->byteSize = (uint64_t)(int64_t)(int32_t)r->byteSize;
    int chunkLevel = ((int)node->name.size() - 2) / 4;
    std::string key = node->name.substr(0, kStep * chunkLevel + 1);
    if (node->name == batch->name) key = node->name;
    chunk_of(key)->nodes.push_back(node);
    batch->nodes.push_back(node);
    batch->nodeMap[node->name] = node;
    bool isChunkKey = ((node->name.size() - 1) % kStep) == 0;
    bool isBatchSubChunk = node->name.size() > (size_t)kStep + 1;
    if (isChunkKey && isBatchSubChunk) chunk_of(node->name)->nodes.push_back(node);
  }
  std::sort(batch->chunks.begin(), batch->chunks.end(),
            [](const std::shared_ptr<HChunk>& a, const std::shared_ptr<HChunk>& b) {
              return breadth_first_less(a->name, b->name);
            });
  for (auto& node : batch->nodes) {
    node->type = T_LEAF;
    auto it = batch->nodeMap.find(node->name.substr(0, node->name.size() - 1));
    if (it != batch->nodeMap.end()) {
      it->second->type = T_NORMAL;
      it->second->childMask |= (uint8_t)(1 << (node->name.back() - '0'));
    }
  }
  for (auto& c : batch->chunks) {
    auto it = batch->nodeMap.find(c->name);
    if (it != batch->nodeMap.end()) it->second->type = T_PROXY;
  }
  for (auto& c : batch->chunks)
    std::sort(c->nodes.begin(), c->nodes.end(), [](const std::shared_ptr<HNode>& a, const std::shared_ptr<HNode>& b) {
      return breadth_first_less(a->name, b->name);
    });
  return batch;
}

void process_batch(HBatch& b) {
  int64_t off = 0;
  for (auto& c : b.chunks) {
    c->byteOffset = off;
    if (c->name != b.name) {
      auto proxy = b.nodeMap[c->name];
      proxy->type = T_PROXY;
      proxy->proxyByteOffset = (uint64_t)c->byteOffset;
      proxy->proxyByteSize = 22 * c->nodes.size();
    }
    off += 22 * (int64_t)c->nodes.size();
  }
  b.byteSize = off;
}

std::vector<uint8_t> serialize_batch(HBatch& b, int64_t bytesWritten) {
  size_t n = 0;
  for (auto& c : b.chunks) n += c->nodes.size();
  std::vector<uint8_t> out(22 * n);
  size_t k = 0;
  for (auto& c : b.chunks) {
    for (auto& node : c->nodes) {
      bool isProxy = node->type == T_PROXY && node->name != c->name;
      uint8_t type = (uint8_t)node->type;
      if (node->type == T_PROXY && !isProxy) type = T_NORMAL;
      uint64_t size = isProxy ? node->proxyByteSize : node->byteSize;
      uint64_t offset = isProxy ? (uint64_t)bytesWritten + node->proxyByteOffset : node->byteOffset;
      uint8_t* d = out.data() + 22 * k;
      uint32_t np = (uint32_t)node->numPoints;
      d[0] = type;
      d[1] = node->childMask;
      memcpy(d + 2, &np, 4);
      memcpy(d + 6, &offset, 8);
      memcpy(d + 14, &size, 8);
      k++;
    }
  }
  b.byteSize = (int64_t)out.size();
  return out;
}

}  // namespace

uint64_t node_key(const std::string& nm) {
  uint64_t k = 1;
  for (size_t c = 1; c < nm.size(); c++) k = (k << 3) | (uint64_t)(nm[c] - '0');
  return k;
}

std::string node_name(uint64_t key) {
  std::string s;
  for (; key > 1; key >>= 3) s.push_back((char)('0' + (key & 7)));
  s.push_back('r');
  return std::string(s.rbegin(), s.rend());
}

std::vector<uint8_t> build_hierarchy_bin_maps(const std::vector<CompletedNode>& recsIn, int64_t& firstChunkSize) {
  // records named by key get their names
  std::vector<CompletedNode> named;
  bool needNames = false;
  for (const auto& r : recsIn) needNames |= r.name.empty();
  if (needNames) {
    named = recsIn;
    for (auto& r : named)
      if (r.name.empty()) r.name = node_name(r.key);
  }
  const std::vector<CompletedNode>& recs = needNames ? named : recsIn;
  std::map<std::string, std::vector<const CompletedNode*>> groups;
  for (const auto& r : recs) {
    std::string key = r.name.size() <= (size_t)kStep + 1 ? std::string("r") : r.name.substr(0, kStep + 1);
    groups[key].push_back(&r);
    if (r.name.size() == (size_t)kStep + 1) groups[r.name].push_back(&r);
  }
  auto root = load_batch("r", groups["r"]);
  std::vector<uint8_t> file(22 * root->nodes.size(), 0);
  int64_t written = (int64_t)file.size();
  for (auto& kv : groups) {
    if (kv.first == "r") continue;
    auto b = load_batch(kv.first, kv.second);
    process_batch(*b);
    auto buf = serialize_batch(*b, written);
    if (b->nodes.size() > 1) {
      auto proxy = root->nodeMap[b->name];
      if (proxy) {
        proxy->type = T_PROXY;
        proxy->proxyByteOffset = (uint64_t)written;
        proxy->proxyByteSize = 22 * b->chunkMap[b->name]->nodes.size();
      }
    } else {
      auto leaf = root->nodeMap[b->name];
      if (leaf) leaf->type = T_LEAF;
    }
    file.insert(file.end(), buf.begin(), buf.end());
    written += (int64_t)buf.size();
  }
  auto rb = serialize_batch(*root, 0);
  memcpy(file.data(), rb.data(), rb.size());
  firstChunkSize = root->byteSize;
  return file;
}

// The same layout from integer node keys.  A name "r" + d digits becomes
// the marker key (1 << 3d) | digits (base 8), which orders breadth-first
// (shorter names first, then by name); one sort orders every record the way
// the batches, chunks and breadth-first node lists above do.  As in the
// record loop above (HierarchyBuilder.h:148-162) a node's type is reset to
// LEAF at its own record and turned NORMAL by children recorded after it, so
// the type depends on the record order while the child mask does not.
// Byte-identical to build_hierarchy_bin_maps; names deeper than 20 levels
// take that path.
namespace {
inline int key_depth(uint64_t k) { return (63 - __builtin_clzll(k)) / 3; }
inline uint64_t key_prefix(uint64_t k, int depth) { return k >> (3 * (key_depth(k) - depth)); }

struct HEntry {
  uint64_t batch, chunk, node;  // marker keys
  int32_t idx;                  // record index
};

void put_record(uint8_t* d, uint8_t type, uint8_t mask, uint32_t np, uint64_t offset, uint64_t size) {
  d[0] = type;
  d[1] = mask;
  memcpy(d + 2, &np, 4);
  memcpy(d + 6, &offset, 8);
  memcpy(d + 14, &size, 8);
}

// open-addressing map marker key -> record index
struct KeyIndex {
  std::vector<uint64_t> keys;
  std::vector<int32_t> vals;
  uint64_t mask = 0;
  explicit KeyIndex(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    keys.assign(cap, 0);
    vals.assign(cap, -1);
    mask = cap - 1;
  }
  static uint64_t h(uint64_t k) { return (k * 0x9E3779B97F4A7C15ull) >> 17; }
  void put(uint64_t k, int32_t v) {
    uint64_t s = h(k) & mask;
    while (keys[s] != 0 && keys[s] != k) s = (s + 1) & mask;
    keys[s] = k;
    vals[s] = v;
  }
  int32_t get(uint64_t k) const {
    uint64_t s = h(k) & mask;
    while (keys[s] != 0) {
      if (keys[s] == k) return vals[s];
      s = (s + 1) & mask;
    }
    return -1;
  }
};
}  // namespace

bool plan_hierarchy(const std::vector<uint64_t>& key, HierarchyPlan& plan) {
  const int32_t n = (int32_t)key.size();
  plan.items = n;
  plan.slots.clear();
  plan.parent.assign(n, -1);
  for (int32_t i = 0; i < n; i++)
    if (key[i] == 0 || key_depth(key[i]) > 20) return false;
  KeyIndex index((size_t)n);
  for (int32_t i = 0; i < n; i++) index.put(key[i], i);
  // child masks (present children), parents for the NORMAL / LEAF rule
  std::vector<uint8_t> mask(n, 0);
  for (int32_t i = 0; i < n; i++) {
    if (key[i] == 1) continue;
    const int32_t p = index.get(key[i] >> 3);
    if (p < 0) continue;
    mask[p] |= (uint8_t)(1u << (uint32_t)(key[i] & 7));
    plan.parent[i] = p;
  }
  // root batch: levels 0-4 breadth-first, one chunk "r"
  std::vector<int32_t> rootNodes;
  for (int32_t i = 0; i < n; i++)
    if (key_depth(key[i]) <= kStep) rootNodes.push_back(i);
  std::sort(rootNodes.begin(), rootNodes.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  // sub-batches: every node of level >= 4 under its level-4 prefix, in the
  // chunk of its 4-level group (the batch root in its own chunk); chunk roots
  // below the batch root are also in their own chunk.  Entries are bucketed
  // by batch (a level-4 key: 8^4 marker keys from 8^4 on) with a counting
  // sort, then each batch is sorted by (chunk, node).
  constexpr uint64_t kBatch0 = 1ull << (3 * kStep), kBatches = 1ull << (3 * kStep);
  std::vector<HEntry> tmp;
  tmp.reserve((size_t)n + n / 4);
  std::vector<uint32_t> bstart(kBatches + 1, 0);
  for (int32_t i = 0; i < n; i++) {
    const int d = key_depth(key[i]);
    if (d < kStep) continue;
    const uint64_t batch = key_prefix(key[i], kStep);
    const int cd = d == kStep ? kStep : kStep * ((d - 1) / kStep);
    tmp.push_back(HEntry{batch, key_prefix(key[i], cd), key[i], i});
    bstart[batch - kBatch0 + 1]++;
    if (d > kStep && d % kStep == 0) {
      tmp.push_back(HEntry{batch, key[i], key[i], i});
      bstart[batch - kBatch0 + 1]++;
    }
  }
  for (uint64_t b = 0; b < kBatches; b++) bstart[b + 1] += bstart[b];
  std::vector<HEntry> ent(tmp.size());
  {
    std::vector<uint32_t> fill(bstart.begin(), bstart.end() - 1);
    for (const HEntry& e : tmp) ent[fill[e.batch - kBatch0]++] = e;
  }
  for (uint64_t b = 0; b < kBatches; b++)
    if (bstart[b + 1] - bstart[b] > 1)
      std::sort(ent.begin() + bstart[b], ent.begin() + bstart[b + 1], [](const HEntry& x, const HEntry& y) {
        if (x.chunk != y.chunk) return x.chunk < y.chunk;
        return x.node < y.node;
      });
  plan.slots.resize(rootNodes.size() + ent.size());
  int64_t written = 22 * (int64_t)rootNodes.size();
  struct SubBatch { int64_t start = -1, firstChunk = 0, entries = 0; };
  std::unordered_map<int32_t, SubBatch> sub;
  std::vector<size_t> cstart;
  std::vector<int64_t> coff;
  size_t e0 = 0;
  while (e0 < ent.size()) {
    size_t e1 = e0;
    while (e1 < ent.size() && ent[e1].batch == ent[e0].batch) e1++;
    cstart.clear();
    for (size_t e = e0; e < e1; e++)
      if (e == e0 || ent[e].chunk != ent[e - 1].chunk) cstart.push_back(e);
    cstart.push_back(e1);
    const size_t nchunks = cstart.size() - 1;
    coff.assign(nchunks, 0);
    for (size_t c = 1; c < nchunks; c++) coff[c] = coff[c - 1] + 22 * (int64_t)(cstart[c] - cstart[c - 1]);
    auto chunk_of = [&](uint64_t k) -> size_t {
      size_t lo = 0, hi = nchunks;
      while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (ent[cstart[mid]].chunk < k) lo = mid + 1; else hi = mid;
      }
      return lo;
    };
    for (size_t e = e0; e < e1; e++) {
      const HEntry& x = ent[e];
      const bool chunkRoot = key_depth(x.node) % kStep == 0;  // every chunk key of a sub-batch
      const bool isProxy = chunkRoot && x.chunk != x.node;
      HierarchyPlan::Slot& sl = plan.slots[rootNodes.size() + e];
      sl.item = x.idx;
      sl.mask = mask[x.idx];
      sl.kind = chunkRoot ? (isProxy ? HierarchyPlan::S_PROXY : HierarchyPlan::S_NORMAL) : HierarchyPlan::S_PLAIN;
      sl.offset = sl.size = 0;
      if (isProxy) {
        const size_t own = chunk_of(x.node);
        sl.offset = (uint64_t)written + (uint64_t)coff[own];
        sl.size = 22ull * (uint64_t)(cstart[own + 1] - cstart[own]);
      }
    }
    SubBatch& sb = sub[ent[cstart[0]].idx];  // the batch root leads the first chunk
    sb.start = written;
    sb.firstChunk = 22 * (int64_t)(cstart[1] - cstart[0]);
    sb.entries = (int64_t)(e1 - e0);
    written += 22 * (int64_t)(e1 - e0);
    e0 = e1;
  }
  for (size_t q = 0; q < rootNodes.size(); q++) {
    const int32_t i = rootNodes[q];
    HierarchyPlan::Slot& sl = plan.slots[q];
    sl.item = i;
    sl.mask = mask[i];
    sl.kind = HierarchyPlan::S_PLAIN;
    sl.offset = sl.size = 0;
    if (key[i] == 1) sl.kind = HierarchyPlan::S_NORMAL;  // the chunk root "r" in its own chunk
    if (key_depth(key[i]) == kStep) {
      sl.mask = 0;  // its children are in its sub-batch
      sl.kind = HierarchyPlan::S_LEAF;
      auto it = sub.find(i);
      if (it != sub.end() && it->second.start >= 0 && it->second.entries > 1) {
        sl.kind = HierarchyPlan::S_PROXY;
        sl.offset = (uint64_t)it->second.start;
        sl.size = (uint64_t)it->second.firstChunk;
      }
    }
  }
  plan.firstChunkSize = 22 * (int64_t)rootNodes.size();
  return true;
}

std::vector<uint8_t> fill_hierarchy(const HierarchyPlan& plan, const int32_t* rank, const int64_t* numPoints,
                                    const int64_t* byteOffset, const int64_t* byteSize) {
  const int32_t n = plan.items;
  // NORMAL needs a child recorded after the node (HierarchyBuilder.h:148-162)
  std::vector<uint8_t> later(n, 0);
  for (int32_t i = 0; i < n; i++) {
    const int32_t p = plan.parent[i];
    if (p >= 0 && rank[i] > rank[p]) later[p] = 1;
  }
  std::vector<uint8_t> file(22 * plan.slots.size());
  for (size_t q = 0; q < plan.slots.size(); q++) {
    const HierarchyPlan::Slot& sl = plan.slots[q];
    const int32_t i = sl.item;
    uint8_t type;
    uint64_t offset = (uint64_t)byteOffset[i], size = (uint64_t)(int64_t)(int32_t)byteSize[i];
    switch (sl.kind) {
      case HierarchyPlan::S_PLAIN: type = (sl.mask && later[i]) ? T_NORMAL : T_LEAF; break;
      case HierarchyPlan::S_NORMAL: type = T_NORMAL; break;
      case HierarchyPlan::S_LEAF: type = T_LEAF; break;
      default: type = T_PROXY; offset = sl.offset; size = sl.size; break;
    }
    put_record(file.data() + 22 * q, type, sl.mask, (uint32_t)(int32_t)(uint32_t)numPoints[i], offset, size);
  }
  return file;
}

std::vector<uint8_t> build_hierarchy_bin(const std::vector<CompletedNode>& recs, int64_t& firstChunkSize) {
  const int32_t n = (int32_t)recs.size();
  std::vector<uint64_t> key(n);
  for (int32_t i = 0; i < n; i++) {
    const CompletedNode& r = recs[i];
    if (!r.key && (r.name.size() > 21 || r.name.empty())) return build_hierarchy_bin_maps(recs, firstChunkSize);
    key[i] = r.key ? r.key : node_key(r.name);
  }
  HierarchyPlan plan;
  if (!plan_hierarchy(key, plan)) return build_hierarchy_bin_maps(recs, firstChunkSize);
  std::vector<int32_t> rank(n);
  std::vector<int64_t> np(n), off(n), size(n);
  for (int32_t i = 0; i < n; i++) {
    rank[i] = i;
    np[i] = recs[i].numPoints;
    off[i] = recs[i].byteOffset;
    size[i] = recs[i].byteSize;
  }
  firstChunkSize = plan.firstChunkSize;
  return fill_hierarchy(plan, rank.data(), np.data(), off.data(), size.data());
}

std::string create_metadata_json(const Attributes& attributes, const Vec3& min, const Vec3& max, double spacing,
                                 const std::string& name, int64_t pointsTotal, int64_t firstChunkSize,
                                 int64_t octreeDepth, const std::string& encoding) {
  // indexer.cpp:354-524: tabs, key order and %.17g numbers
  auto d = [](double v) { return fmt17(v); };
  auto vec3 = [&](const Vec3& v) { return "[" + d(v.x) + ", " + d(v.y) + ", " + d(v.z) + "]"; };
  auto t = [](int n) { return std::string(n, '\t'); };
  auto s = [](const std::string& v) { return "\"" + v + "\""; };
  std::stringstream ss;
  ss << "{\n";
  ss << t(1) << s("version") << ": " << s("2.0") << ",\n";
  ss << t(1) << s("name") << ": " << s(name) << ",\n";
  ss << t(1) << s("description") << ": " << s("") << ",\n";
  ss << t(1) << s("points") << ": " << pointsTotal << ",\n";
  ss << t(1) << s("projection") << ": " << s("") << ",\n";
  ss << t(1) << s("hierarchy") << ": {\n";
  ss << t(2) << s("firstChunkSize") << ": " << firstChunkSize << ", \n";
  ss << t(2) << s("stepSize") << ": " << kStep << ", \n";
  ss << t(2) << s("depth") << ": " << octreeDepth << "\n";
  ss << t(1) << "},\n";
  ss << t(1) << s("offset") << ": " << vec3(attributes.posOffset) << ",\n";
  ss << t(1) << s("scale") << ": " << vec3(attributes.posScale) << ",\n";
  ss << t(1) << s("spacing") << ": " << d(spacing) << ",\n";
  ss << t(1) << s("boundingBox") << ": {\n";
  ss << t(2) << s("min") << ": " << vec3(min) << ", \n";
  ss << t(2) << s("max") << ": " << vec3(max) << "\n";
  ss << t(1) << "},\n";
  ss << t(1) << s("encoding") << ": " << s(encoding) << ",\n";
  ss << t(1) << s("attributes") << ": [\n";
  const auto& L = attributes.list;
  for (size_t i = 0; i < L.size(); i++) {
    const Attribute& a = L[i];
    if (i == 0) ss << t(2) << "{\n";
    ss << t(3) << s("name") << ": " << s(a.name) << ",\n";
    ss << t(3) << s("description") << ": " << s(a.description) << ",\n";
    ss << t(3) << s("size") << ": " << a.size << ",\n";
    ss << t(3) << s("numElements") << ": " << a.numElements << ",\n";
    ss << t(3) << s("elementSize") << ": " << a.elementSize << ",\n";
    ss << t(3) << s("type") << ": " << s(attribute_type_name(a.type)) << ",\n";
    bool empty = std::all_of(a.histogram.begin(), a.histogram.end(), [](int64_t h) { return h == 0; });
    if (a.size == 1 && !empty) {
      ss << t(3) << s("histogram") << ": [";
      for (size_t k = 0; k < a.histogram.size(); k++) ss << a.histogram[k] << (k + 1 < a.histogram.size() ? ", " : "");
      ss << "], \n";
    }
    auto vec = [&](const Vec3& v) {
      std::string r = "[" + d(v.x);
      if (a.numElements >= 2) r += ", " + d(v.y);
      if (a.numElements >= 3) r += ", " + d(v.z);
      return r + "]";
    };
    if (a.numElements >= 1 && a.numElements <= 3) {
      ss << t(3) << s("min") << ": " << vec(a.min) << ",\n";
      ss << t(3) << s("max") << ": " << vec(a.max) << ",\n";
      ss << t(3) << s("scale") << ": " << vec(a.scale) << ",\n";
      ss << t(3) << s("offset") << ": " << vec(a.offset) << "\n";
    }
    ss << t(2) << (i + 1 < L.size() ? "},{" : "}") << "\n";
  }
  ss << t(1) << "]\n";
  ss << "}\n";
  return ss.str();
}

int write_index_files(const IndexResult& r, const std::string& targetDir) {
  fs::create_directories(targetDir);
  {
    std::ofstream f(targetDir + "/octree.bin", std::ios::binary);
    if (r.octreeOnHost) {
      f.write((const char*)r.h_octree.data(), (std::streamsize)r.h_octree.size());
    } else {
      std::vector<uint8_t> host(r.octreeBytes);
      if (r.octreeBytes) PCG_CHECK(stream_copy(host.data(), r.d_octree, r.octreeBytes, hipMemcpyDeviceToHost));
      f.write((const char*)host.data(), (std::streamsize)host.size());
    }
  }
  {
    std::ofstream f(targetDir + "/hierarchy.bin", std::ios::binary);
    f.write((const char*)r.hierarchy.data(), (std::streamsize)r.hierarchy.size());
  }
  {
    std::ofstream f(targetDir + "/metadata.json", std::ios::binary);
    f << r.metadata;
  }
  return 0;
}

}  // namespace pcg
