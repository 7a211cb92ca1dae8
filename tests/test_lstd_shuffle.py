"""lstd_shuffle.h (the libstdc++ std::shuffle + std::default_random_engine
restatement used for SamplerRandom's leaf shuffle) against the system
std::shuffle, including the sizes around the switch between the paired and
the single draw per swap (n^2 against the generator's range 2^31 - 3)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r'''
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
#include "lstd_shuffle.h"
int main() {
  std::mt19937_64 rng(11);
  std::vector<long> sizes;
  for (long n = 0; n < 200; n++) sizes.push_back(n);
  for (long n : {1000L, 9999L, 10000L, 46338L, 46339L, 46340L, 46341L, 46342L, 65536L, 100001L}) sizes.push_back(n);
  for (int t = 0; t < 300; t++) sizes.push_back((long)(rng() % 30000));
  std::vector<unsigned> seeds = {0u, 1u, 7u, 2147483646u, 2147483647u, 2147483648u, 4294967295u, 123456789u};
  int bad = 0, cases = 0;
  for (size_t k = 0; k < sizes.size(); k++) {
    long n = sizes[k];
    unsigned seed = k < seeds.size() * 30 ? seeds[k % seeds.size()] : (unsigned)rng();
    std::vector<int> a(n), b(n);
    for (long i = 0; i < n; i++) a[i] = b[i] = (int)i;
    std::shuffle(a.begin(), a.end(), std::default_random_engine(seed));
    lstd::shuffle_swaps(n, seed, [&](int64_t i, int64_t j) { std::swap(b[i], b[j]); });
    cases++;
    if (a != b) { bad++; printf("mismatch n=%ld seed=%u\n", n, seed); }
  }
  // the generator alone
  std::default_random_engine e(42u);
  lstd::MinStd0 m(42u);
  for (int i = 0; i < 100000; i++) if (e() != m()) { bad++; break; }
  printf("%d %d\n", cases, bad);
  return bad != 0;
}
'''


def test_shuffle_matches_libstdcxx(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "potreeconverter_amd", "csrc"), str(src),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    cases, bad = map(int, r.stdout.split()[-2:])
    assert cases > 400 and bad == 0
