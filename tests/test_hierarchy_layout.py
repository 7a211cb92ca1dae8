"""hierarchy.bin from node keys (the plan / fill split the indexer uses)
against the name-keyed batch and chunk maps that follow HierarchyBuilder
(Converter/include/HierarchyBuilder.h) one to one, on random trees, random
completion orders and record subsets, byte for byte.  Host code only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "potreeconverter_amd", "csrc")

HARNESS = r'''
#include <algorithm>
#include <cstdio>
#include <random>
#include <string>
#include <vector>
#include "indexing.h"
namespace pcg {
static std::string e;
void set_error(const std::string& m) { e = m; }
const std::string& last_error() { return e; }
Device& device() { static Device d; return d; }
}
using namespace pcg;
int main() {
  std::mt19937 rng(5);
  int bad = 0, cases = 0;
  for (int t = 0; t < 120; t++) {
    const size_t target = t < 20 ? 40 + 10 * t : 2000 + rng() % 20000;
    const size_t maxLen = t % 10 == 9 ? 26 : 4 + rng() % 14;  // some trees deeper than node keys
    std::vector<std::string> names{"r"};
    for (size_t i = 0; i < names.size() && names.size() < target; i++) {
      if (names[i].size() >= maxLen) continue;
      int m = rng() & 255;
      if (names[i].size() < 3) m |= 1 << (rng() & 7);
      for (int c = 0; c < 8; c++) if (m >> c & 1) names.push_back(names[i] + char('0' + c));
    }
    // completion order: a random permutation, or children before parents
    std::vector<size_t> ord(names.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = i;
    if (t % 3 == 0) std::shuffle(ord.begin(), ord.end(), rng);
    else std::reverse(ord.begin(), ord.end());
    // a random subset (some nodes not completed), except for the root
    const bool subset = t % 4 == 1;
    std::vector<CompletedNode> byKey, byName;
    int64_t off = 0;
    for (size_t i : ord) {
      if (subset && i != 0 && rng() % 5 == 0) continue;
      CompletedNode c;
      c.numPoints = rng() % 3 == 0 ? 0 : 1 + rng() % 20000;
      c.byteSize = c.numPoints * 26;
      c.byteOffset = c.numPoints ? off : 0;
      off += c.byteSize;
      c.name = names[i];
      byName.push_back(c);
      if (names[i].size() <= 21) { c.key = node_key(names[i]); c.name.clear(); }
      byKey.push_back(c);
    }
    int64_t f1 = 0, f2 = 0;
    auto a = build_hierarchy_bin(byKey, f1);
    auto b = build_hierarchy_bin_maps(byName, f2);
    cases++;
    if (a != b || f1 != f2) { bad++; printf("mismatch t=%d n=%zu\n", t, byName.size()); }
    // keys round trip
    for (size_t i = 0; i < names.size(); i += 97)
      if (names[i].size() <= 21 && node_name(node_key(names[i])) != names[i]) { bad++; break; }
  }
  printf("%d %d\n", cases, bad);
  return bad != 0;
}
'''


def test_key_layout_equals_name_maps(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "h"
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", CSRC, "-I", "/opt/rocm/include", str(src),
           os.path.join(CSRC, "writers.cpp"), os.path.join(CSRC, "attributes.cpp"), "-o", str(exe),
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    cases, bad = map(int, r.stdout.split()[-2:])
    assert cases == 120 and bad == 0
