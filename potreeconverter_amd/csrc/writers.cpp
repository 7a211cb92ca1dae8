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

std::vector<uint8_t> build_hierarchy_bin(const std::vector<CompletedNode>& recs, int64_t& firstChunkSize) {
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

std::string create_metadata_json(const Attributes& attributes, const Vec3& min, const Vec3& max, double spacing,
                                 const std::string& name, int64_t pointsTotal, int64_t firstChunkSize,
                                 int64_t octreeDepth) {
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
  ss << t(1) << s("encoding") << ": " << s("DEFAULT") << ",\n";
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
    std::vector<uint8_t> host(r.octreeBytes);
    if (r.octreeBytes) PCG_CHECK(stream_copy(host.data(), r.d_octree, r.octreeBytes, hipMemcpyDeviceToHost));
    std::ofstream f(targetDir + "/octree.bin", std::ios::binary);
    f.write((const char*)host.data(), (std::streamsize)host.size());
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
