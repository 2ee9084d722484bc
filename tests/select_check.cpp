// Compares thg::nth_element_gcc (thea_amd/csrc/select.hpp) with the toolchain's std::nth_element on
// seeded arrays of 0..30 floats: continuous values, few distinct values (many ties) and small
// integer ranges, at the median (Thea's use) and at every other position. Prints
// "<trials> <mismatches> <heap-select hits>"; the heap-select branch (introselect's depth limit)
// is reached by a small fraction of the arrays.
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <random>
#include <vector>

static long heap_hits = 0;
#define THG_SEL_TRACE() (++heap_hits)
#include "select.hpp"

int main(int argc, char **argv) {
  const long trials = argc > 1 ? std::atol(argv[1]) : 2000000;
  std::mt19937_64 g(1);
  long bad = 0;
  for (long trial = 0; trial < trials; ++trial) {
    const int n = (int)(g() % 31);
    const unsigned long alpha = trial % 4 == 0 ? 3 : (trial % 4 == 1 ? 1000 : (1ul << 30));
    std::vector<float> a((size_t)n);
    for (auto &x : a) x = trial % 4 == 3 ? std::ldexp((float)(g() >> 40), -24) : (float)(g() % alpha);
    const int nth = (trial & 1) ? n / 2 : (int)(g() % (unsigned long)(n + 1));
    std::vector<float> b = a;
    std::nth_element(a.begin(), a.begin() + nth, a.end());
    thg::nth_element_gcc(thg::StridedF{b.data(), 1}, nth, n);
    if (a != b) ++bad;
  }
  std::printf("%ld %ld %ld\n", trials, bad, heap_hits);
  return 0;
}
