// Test-only driver: one Sampler::sample call of the reference samplers
// (Converter/include/sampler_poisson.h, sampler_poisson_average.h,
// sampler_random.h) on a subtree read from a file, with the callbacks
// recorded.  Compiled against the reference headers in place by
// oracle/build_ref.sh into oracle/_ref/ (never shipped).
//   ref_sampler <poisson|poisson_average|random> <in.bin> <out.bin>
//
// in.bin (little endian):
//   int32 num_nodes, int32 bpp, int32 rgb_offset, double scale[3],
//   double offset[3], double base_spacing, then per node: int32 children[8]
//   (-1 none), double min[3], double max[3], int32 level, int32 sampled,
//   int64 count; then the records of every node in node order.
//   Node 0 is named "r" + level zeros, a child its parent's name + digit.
// out.bin: per node int64 numPoints, int32 children[8] (-1 once unlinked),
//   int32 sampled, then its records; then int32 num_events and per event
//   int32 node, int32 kind (0 onNodeCompleted, 1 onNodeDiscarded).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "indexer.h"
#include "sampler_poisson.h"
#include "sampler_poisson_average.h"
#include "sampler_random.h"

// SamplerRandom seeds every leaf shuffle with system_clock::now()
// (sampler_random.h:102-104); fixed at REF_CLOCK_NS nanoseconds.
namespace std {
namespace chrono {
inline namespace _V2 {
system_clock::time_point system_clock::now() noexcept {
  const char* e = getenv("REF_CLOCK_NS");
  return time_point(duration(e ? strtoll(e, nullptr, 10) : 0));
}
}  // namespace _V2
}  // namespace chrono
}  // namespace std

template <class T>
static bool rd(FILE* f, T& v) {
  return fread(&v, sizeof(T), 1, f) == 1;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: ref_sampler <method> <in.bin> <out.bin>" << std::endl;
    return 2;
  }
  const std::string method = argv[1];
  FILE* f = fopen(argv[2], "rb");
  if (!f) return 2;
  int32_t n, bpp, rgbOffset;
  double scale[3], offset[3], baseSpacing;
  rd(f, n);
  rd(f, bpp);
  rd(f, rgbOffset);
  for (double& v : scale) rd(f, v);
  for (double& v : offset) rd(f, v);
  rd(f, baseSpacing);
  struct In {
    int32_t children[8];
    double mn[3], mx[3];
    int32_t level, sampled;
    int64_t count;
  };
  std::vector<In> in(n);
  for (auto& x : in) {
    for (int32_t& c : x.children) rd(f, c);
    for (double& v : x.mn) rd(f, v);
    for (double& v : x.mx) rd(f, v);
    rd(f, x.level);
    rd(f, x.sampled);
    rd(f, x.count);
  }
  std::vector<shared_ptr<Node>> nodes(n);
  std::vector<std::string> names(n);
  names[0] = "r" + std::string(in[0].level, '0');
  std::map<Node*, int32_t> idOf;
  for (int32_t i = 0; i < n; i++) {
    // parents precede their children in the file
    nodes[i] = make_shared<Node>(names[i], Vector3{in[i].mn[0], in[i].mn[1], in[i].mn[2]},
                                 Vector3{in[i].mx[0], in[i].mx[1], in[i].mx[2]});
    idOf[nodes[i].get()] = i;
    for (int c = 0; c < 8; c++)
      if (in[i].children[c] >= 0) names[in[i].children[c]] = names[i] + std::to_string(c);
  }
  for (int32_t i = 0; i < n; i++) {
    for (int c = 0; c < 8; c++)
      if (in[i].children[c] >= 0) nodes[i]->children[c] = nodes[in[i].children[c]];
    nodes[i]->sampled = in[i].sampled != 0;
    nodes[i]->numPoints = in[i].count;
    auto buf = make_shared<Buffer>(in[i].count * bpp);
    if (in[i].count > 0 && fread(buf->data, bpp, in[i].count, f) != (size_t)in[i].count) return 3;
    nodes[i]->points = buf;
  }
  fclose(f);

  vector<Attribute> list;
  list.emplace_back("position", 12, 3, 4, AttributeType::INT32);
  if (method == "poisson_average") {
    if (rgbOffset > 12) list.emplace_back("before rgb", rgbOffset - 12, 1, rgbOffset - 12, AttributeType::UNDEFINED);
    list.emplace_back("rgb", 6, 3, 2, AttributeType::UINT16);
    if (bpp > rgbOffset + 6) list.emplace_back("after rgb", bpp - rgbOffset - 6, 1, bpp - rgbOffset - 6, AttributeType::UNDEFINED);
  } else if (bpp > 12) {
    list.emplace_back("rest", bpp - 12, 1, bpp - 12, AttributeType::UNDEFINED);
  }
  Attributes attributes(list);
  attributes.posScale = Vector3{scale[0], scale[1], scale[2]};
  attributes.posOffset = Vector3{offset[0], offset[1], offset[2]};

  std::vector<std::pair<int32_t, int32_t>> events;
  auto completed = [&](Node* x) { events.emplace_back(idOf.at(x), 0); };
  auto discarded = [&](Node* x) { events.emplace_back(idOf.at(x), 1); };
  if (method == "poisson") {
    SamplerPoisson s;
    s.sample(nodes[0].get(), attributes, baseSpacing, completed, discarded);
  } else if (method == "poisson_average") {
    SamplerPoissonAverage s;
    s.sample(nodes[0].get(), attributes, baseSpacing, completed, discarded);
  } else if (method == "random") {
    SamplerRandom s;
    s.sample(nodes[0].get(), attributes, baseSpacing, completed, discarded);
  } else {
    return 2;
  }

  FILE* o = fopen(argv[3], "wb");
  if (!o) return 2;
  for (int32_t i = 0; i < n; i++) {
    Node* x = nodes[i].get();
    int64_t np = x->points ? x->numPoints : 0;
    fwrite(&np, 8, 1, o);
    for (int c = 0; c < 8; c++) {
      int32_t ch = x->children[c] ? idOf.at(x->children[c].get()) : -1;
      fwrite(&ch, 4, 1, o);
    }
    int32_t sm = x->sampled ? 1 : 0;
    fwrite(&sm, 4, 1, o);
    if (np > 0) fwrite(x->points->data, bpp, np, o);
  }
  int32_t ne = (int32_t)events.size();
  fwrite(&ne, 4, 1, o);
  for (auto& e : events) {
    fwrite(&e.first, 4, 1, o);
    fwrite(&e.second, 4, 1, o);
  }
  fclose(o);
  return 0;
}
