// BROTLI node encoding (Converter/src/indexer.cpp:1110-1401, 1453-1516).
//
// Per node the reference builds one array per attribute (position as two
// 48-bit Morton codes of the coordinates relative to the node minimum, rgb
// as one Morton code, every other attribute as its raw bytes), orders the
// points by the position codes with std::sort, concatenates the arrays in
// attribute order and compresses them with BrotliEncoderCompress(quality 6,
// BROTLI_DEFAULT_WINDOW 22, BROTLI_DEFAULT_MODE 0).  The layout runs on the
// GPU (k_soa_layout, one block per node); brotli has no device form here and
// runs on host threads, one node at a time per thread.  The reference
// vendors brotli 1.0.9; the system's libbrotlienc.so.1 is that release and
// is bound with dlopen so that the library has no build-time dependency.
#include "encoding.h"

#include <dlfcn.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>

#include "converter.h"
#include "sampler.h"

namespace pcg {

namespace {

// BrotliEncoderCompress (brotli/encode.h of brotli 1.0.9)
typedef int (*BrotliCompressFn)(int quality, int lgwin, int mode, size_t input_size, const uint8_t* input,
                                size_t* encoded_size, uint8_t* encoded);

BrotliCompressFn brotli_fn() {
  static BrotliCompressFn fn = nullptr;
  static bool tried = false;
  if (!tried) {
    tried = true;
    void* h = dlopen("libbrotlienc.so.1", RTLD_NOW | RTLD_LOCAL);
    if (h) fn = (BrotliCompressFn)dlsym(h, "BrotliEncoderCompress");
  }
  return fn;
}

// indexer.cpp:1355-1390: output capacity input * 1.5 + 1000, up to five
// attempts, each failure growing it to (capacity + 1024) * 1.5
bool compress_one(BrotliCompressFn fn, const uint8_t* in, size_t size, std::vector<uint8_t>& out) {
  size_t cap = (size_t)(size * 1.5 + 1000);
  for (int attempt = 0; attempt < 5; attempt++) {
    out.resize(cap);
    size_t es = cap;
    if (fn(6, 22, 0, size, in, &es, out.data())) {
      out.resize(es);
      return true;
    }
    cap = (size_t)((cap + 1024) * 1.5);
  }
  return false;
}

}  // namespace

bool brotli_available() {
  if (brotli_fn()) return true;
  set_error("BROTLI encoding needs libbrotlienc.so.1, which could not be loaded");
  return false;
}

int encode_brotli(const Attributes& attributes, int32_t bpp, const std::vector<EncodeInput>& nodes,
                  std::vector<uint8_t>& image, std::vector<int64_t>& sizes, hipStream_t s) {
  image.clear();
  sizes.assign(nodes.size(), 0);
  if (!brotli_available()) return 1;
  BrotliCompressFn fn = brotli_fn();
  // columns in attribute order (indexer.cpp:1339-1354)
  EncParams E;
  memset(&E, 0, sizeof(E));
  E.bpp = bpp;
  E.pos_off = -1;
  int32_t off = 0, soaBytes = 0;
  for (const Attribute& a : attributes.list) {
    if (E.num_cols >= kMaxEncColumns) {
      set_error("BROTLI layout supports at most " + std::to_string(kMaxEncColumns) + " attributes");
      return 1;
    }
    EncColumn& c = E.cols[E.num_cols++];
    c.src = off;
    if (a.name == "position") {
      c.kind = EC_POSITION_MORTON;
      c.size = 16;
      E.pos_off = off;
    } else if (a.name == "rgb") {
      c.kind = EC_RGB_MORTON;
      c.size = 8;
    } else {
      c.kind = EC_RAW;
      c.size = a.size;
    }
    soaBytes += c.size;
    off += a.size;
  }
  if (E.pos_off < 0) {
    set_error("BROTLI encoding needs a position attribute");
    return 1;
  }
  std::vector<EncNode> enc(nodes.size());
  uint64_t outBytes = 0, scrBytes = 0;
  for (size_t i = 0; i < nodes.size(); i++) {
    enc[i] = EncNode{nodes[i].d_records, nodes[i].count, outBytes, scrBytes};
    outBytes += (uint64_t)nodes[i].count * soaBytes;
    scrBytes += enc_scratch_bytes(nodes[i].count);
  }
  if (nodes.empty()) return 0;
  void* dNodes = workspace(WS_ENC_NODES, enc.size() * sizeof(EncNode) + 16);
  uint8_t* dScratch = (uint8_t*)workspace(WS_ENC_SCRATCH, scrBytes + 64);
  uint8_t* dOut = (uint8_t*)workspace(WS_OCTREE, outBytes + 64);
  if (!dNodes || !dScratch || !dOut) return 1;
  PCG_CHECK(hipMemcpyAsync(dNodes, enc.data(), enc.size() * sizeof(EncNode), hipMemcpyHostToDevice, s));
  if (encode_layout((const EncNode*)dNodes, (int64_t)enc.size(), E, dScratch, dOut, s)) return 1;
  std::vector<uint8_t> layout(outBytes);
  PCG_CHECK(hipMemcpyAsync(layout.data(), dOut, outBytes, hipMemcpyDeviceToHost, s));
  PCG_CHECK(hipStreamSynchronize(s));

  // largest nodes first so that the pool ends together
  std::vector<size_t> todo(nodes.size());
  for (size_t i = 0; i < todo.size(); i++) todo[i] = i;
  std::stable_sort(todo.begin(), todo.end(), [&](size_t a, size_t b) { return nodes[a].count > nodes[b].count; });
  std::vector<std::vector<uint8_t>> packed(nodes.size());
  std::atomic<size_t> next{0};
  std::atomic<bool> failed{false};
  auto worker = [&]() {
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= todo.size() || failed.load()) return;
      const size_t i = todo[k];
      const uint8_t* in = layout.data() + enc[i].out_off;
      if (!compress_one(fn, in, (size_t)nodes[i].count * soaBytes, packed[i])) failed = true;
    }
  };
  unsigned hw = std::thread::hardware_concurrency();
  int threads = (int)std::min<size_t>(std::max(1u, std::min(hw, 16u)), nodes.size());
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  if (failed) {
    set_error("brotli compression of a node failed");
    return 1;
  }
  size_t total = 0;
  for (auto& p : packed) total += p.size();
  image.reserve(total);
  for (size_t i = 0; i < nodes.size(); i++) {
    sizes[i] = (int64_t)packed[i].size();
    image.insert(image.end(), packed[i].begin(), packed[i].end());
  }
  return 0;
}

}  // namespace pcg
