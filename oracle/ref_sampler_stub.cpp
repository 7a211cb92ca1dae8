// Test-only CPU stand-in of pcg_sample_subtree (include/potree_gpu.h): the
// reference samplers (Converter/include/sampler_*.h) on the flattened
// subtree, behind the same C ABI.  Linked, in place of libpotree_gpu.so,
// into oracle/_ref/ref_indexer_stub (oracle/build_ref.sh, with
// AddressSanitizer): the reference indexer running the integration glue
// of INTEGRATION.md on the CPU, so that the glue's handling of the nodes
// (flattening, write-back, unlinking, callbacks) is checked where no GPU
// is present.
#include "indexer.h"
#include "sampler_poisson.h"
#include "sampler_poisson_average.h"
#include "sampler_random.h"
#include "potree_gpu.h"
#include <map>
static std::string err;
extern "C" const char* pcg_last_error(void) { return err.c_str(); }
extern "C" int pcg_set_option(const char*, int64_t) { return 0; }
extern "C" int pcg_sample_subtree(const char* method, int32_t n, const int32_t* children, const double* bounds,
                       const int32_t* levels, const uint8_t* sampled, const int64_t* counts, const void* records,
                       int32_t bpp, int32_t rgb_offset, const double scale[3], const double offset[3],
                       double base_spacing, int64_t* out_counts, void* out_records, int64_t out_capacity,
                       int32_t* out_children, int32_t* events, int32_t* num_events) {
  std::vector<shared_ptr<Node>> nodes(n);
  std::vector<std::string> names(n);
  names[0] = "r" + std::string(levels[0], '0');
  std::map<Node*, int32_t> idOf;
  const uint8_t* rp = (const uint8_t*)records;
  for (int i = 0; i < n; i++) {
    nodes[i] = make_shared<Node>(names[i], Vector3{bounds[6*i], bounds[6*i+1], bounds[6*i+2]}, Vector3{bounds[6*i+3], bounds[6*i+4], bounds[6*i+5]});
    idOf[nodes[i].get()] = i;
    for (int c = 0; c < 8; c++) if (children[8*i+c] >= 0) { if (children[8*i+c] <= i) { err = "order"; return 1; } names[children[8*i+c]] = names[i] + std::to_string(c); }
  }
  for (int i = 0; i < n; i++) {
    for (int c = 0; c < 8; c++) if (children[8*i+c] >= 0) nodes[i]->children[c] = nodes[children[8*i+c]];
    nodes[i]->sampled = sampled && sampled[i];
    nodes[i]->numPoints = counts[i];
    auto buf = make_shared<Buffer>(counts[i] * bpp);
    memcpy(buf->data, rp, counts[i] * bpp); rp += counts[i] * bpp;
    nodes[i]->points = buf;
  }
  vector<Attribute> list;
  list.emplace_back("position", 12, 3, 4, AttributeType::INT32);
  if (rgb_offset > 12) list.emplace_back("a", rgb_offset - 12, 1, rgb_offset - 12, AttributeType::UNDEFINED);
  if (rgb_offset >= 12) list.emplace_back("rgb", 6, 3, 2, AttributeType::UINT16);
  int used = rgb_offset >= 12 ? rgb_offset + 6 : 12;
  if (bpp > used) list.emplace_back("b", bpp - used, 1, bpp - used, AttributeType::UNDEFINED);
  Attributes at(list);
  at.posScale = Vector3{scale[0], scale[1], scale[2]};
  at.posOffset = Vector3{offset[0], offset[1], offset[2]};
  std::vector<std::pair<int,int>> ev;
  auto comp = [&](Node* x) { ev.emplace_back(idOf.at(x), 0); };
  auto disc = [&](Node* x) { ev.emplace_back(idOf.at(x), 1); };
  std::string m = method;
  if (m == "poisson") { SamplerPoisson s; s.sample(nodes[0].get(), at, base_spacing, comp, disc); }
  else if (m == "poisson_average") { SamplerPoissonAverage s; s.sample(nodes[0].get(), at, base_spacing, comp, disc); }
  else { SamplerRandom s; s.sample(nodes[0].get(), at, base_spacing, comp, disc); }
  int64_t o = 0;
  for (int i = 0; i < n; i++) {
    int64_t np = nodes[i]->points ? nodes[i]->numPoints : 0;
    out_counts[i] = np;
    if (o + np > out_capacity) { err = "capacity"; return 1; }
    if (np) memcpy((uint8_t*)out_records + o * bpp, nodes[i]->points->data, np * bpp);
    o += np;
    for (int c = 0; c < 8; c++) out_children[8*i+c] = nodes[i]->children[c] ? idOf.at(nodes[i]->children[c].get()) : -1;
  }
  for (size_t e = 0; e < ev.size(); e++) { events[2*e] = ev[e].first; events[2*e+1] = ev[e].second; }
  *num_events = (int32_t)ev.size();
  return 0;
}
