// Deterministic restatement of GNU libstdc++'s std::sort (introsort:
// median-of-three Hoare partition, 16-element threshold, heap sort past a
// depth limit of 2*floor(log2 n), final insertion sort), applied to an
// array of element handles with a strict-weak "key[a] < key[b]" order.
//
// The reference sorts Poisson candidates with std::sort, which is not
// stable: equal keys end up in an order fixed by the algorithm's moves.
// Reproducing those moves exactly is the only way to make equal-distance
// candidates resolve as in the reference, so this follows the published
// algorithm step for step (Musser 1997; the libstdc++ 11 parameters).
// Host and device (single lane) use the same code.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LSTD_HD __host__ __device__ inline
#else
#define LSTD_HD inline
#endif

namespace lstd {

struct KeyLess {
  const double* key;
  LSTD_HD bool operator()(uint32_t a, uint32_t b) const { return key[a] < key[b]; }
};

LSTD_HD int64_t floor_log2(int64_t n) {
  int64_t r = 0;
  while (n > 1) { n >>= 1; r++; }
  return r;
}

template <class C>
LSTD_HD void push_heap(uint32_t* f, int64_t hole, int64_t top, uint32_t value, const C& less) {
  int64_t parent = (hole - 1) / 2;
  while (hole > top && less(f[parent], value)) {
    f[hole] = f[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  f[hole] = value;
}

template <class C>
LSTD_HD void adjust_heap(uint32_t* f, int64_t hole, int64_t len, uint32_t value, const C& less) {
  const int64_t top = hole;
  int64_t second = hole;
  while (second < (len - 1) / 2) {
    second = 2 * (second + 1);
    if (less(f[second], f[second - 1])) second--;
    f[hole] = f[second];
    hole = second;
  }
  if ((len & 1) == 0 && second == (len - 2) / 2) {
    second = 2 * (second + 1);
    f[hole] = f[second - 1];
    hole = second - 1;
  }
  push_heap(f, hole, top, value, less);
}

template <class C>
LSTD_HD void heap_sort(uint32_t* f, int64_t len, const C& less) {
  if (len >= 2) {  // make_heap
    int64_t parent = (len - 2) / 2;
    while (true) {
      uint32_t v = f[parent];
      adjust_heap(f, parent, len, v, less);
      if (parent == 0) break;
      parent--;
    }
  }
  int64_t last = len;  // sort_heap
  while (last > 1) {
    --last;
    uint32_t v = f[last];
    f[last] = f[0];
    adjust_heap(f, 0, last, v, less);
  }
}

template <class C>
LSTD_HD void iter_swap(uint32_t* a, int64_t i, int64_t j) {
  uint32_t t = a[i];
  a[i] = a[j];
  a[j] = t;
}

template <class C>
LSTD_HD void move_median_to_first(uint32_t* a, int64_t result, int64_t x, int64_t y, int64_t z, const C& less) {
  if (less(a[x], a[y])) {
    if (less(a[y], a[z])) iter_swap<C>(a, result, y);
    else if (less(a[x], a[z])) iter_swap<C>(a, result, z);
    else iter_swap<C>(a, result, x);
  } else if (less(a[x], a[z])) {
    iter_swap<C>(a, result, x);
  } else if (less(a[y], a[z])) {
    iter_swap<C>(a, result, z);
  } else {
    iter_swap<C>(a, result, y);
  }
}

template <class C>
LSTD_HD int64_t unguarded_partition(uint32_t* a, int64_t first, int64_t last, int64_t pivot, const C& less) {
  while (true) {
    while (less(a[first], a[pivot])) ++first;
    --last;
    while (less(a[pivot], a[last])) --last;
    if (!(first < last)) return first;
    iter_swap<C>(a, first, last);
    ++first;
  }
}

template <class C>
LSTD_HD void unguarded_linear_insert(uint32_t* a, int64_t last, const C& less) {
  uint32_t val = a[last];
  int64_t next = last - 1;
  while (less(val, a[next])) {
    a[last] = a[next];
    last = next;
    --next;
  }
  a[last] = val;
}

template <class C>
LSTD_HD void insertion_sort(uint32_t* a, int64_t first, int64_t last, const C& less) {
  if (first == last) return;
  for (int64_t i = first + 1; i != last; ++i) {
    if (less(a[i], a[first])) {
      uint32_t val = a[i];
      for (int64_t k = i; k > first; --k) a[k] = a[k - 1];
      a[first] = val;
    } else {
      unguarded_linear_insert(a, i, less);
    }
  }
}

// The partition phase of introsort: after it, a[] is split into segments
// that are mutually ordered; the final insertion sort is stable, so the
// final order of equal keys is their order after this phase.
template <class C>
LSTD_HD void introsort_partition_phase(uint32_t* a, int64_t n, const C& less) {
  struct Frame { int64_t first, last, depth; };
  Frame stack[128];
  int sp = 0;
  stack[sp++] = Frame{0, n, 2 * floor_log2(n)};
  while (sp > 0) {
    Frame fr = stack[--sp];
    int64_t first = fr.first, last = fr.last, depth = fr.depth;
    while (last - first > 16) {
      if (depth == 0) {
        heap_sort(a + first, last - first, less);
        break;
      }
      --depth;
      int64_t mid = first + (last - first) / 2;
      move_median_to_first(a, first, first + 1, mid, last - 1, less);
      int64_t cut = unguarded_partition(a, first + 1, last, first, less);
      stack[sp++] = Frame{cut, last, depth};
      last = cut;
    }
  }
}

template <class C>
LSTD_HD void sort(uint32_t* a, int64_t n, const C& less) {
  if (n <= 0) return;
  introsort_partition_phase(a, n, less);
  if (n > 16) {
    insertion_sort(a, 0, 16, less);
    for (int64_t i = 16; i != n; ++i) unguarded_linear_insert(a, i, less);
  } else {
    insertion_sort(a, 0, n, less);
  }
}

}  // namespace lstd
