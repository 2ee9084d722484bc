// std::nth_element as GCC's libstdc++ performs it (introselect: median-of-three partitioning with a
// depth limit of 2 floor(log2 n), heap selection past the limit, insertion sort below four
// elements), restated for small float arrays on the host and the device.
//
// Thea's shape diameter reduction (Algorithms/SurfaceFeatures/Local/ShapeDiameter.hpp:230-252)
// pairs the values left in the order std::nth_element permutes them with weights kept in the
// original order, and sums them in that order, so its result depends on the exact permutation, not
// only on the selected element. The functions below reproduce that permutation; tests compare them
// with the toolchain's std::nth_element over random, duplicate-heavy and adversarial arrays.
//
// V is an indexable view (v[i] returns float&), so device code can keep the array in LDS with a
// stride. All comparisons are operator< on the values, as the default comparator.
#pragma once

#if defined(__HIPCC__)
#define THG_SEL_HD __host__ __device__ __forceinline__
#else
#define THG_SEL_HD inline
#endif

namespace thg {

// A strided view of floats: element i at p[i * stride].
struct StridedF {
  float *p;
  int stride;
  THG_SEL_HD float &operator[](int i) const { return p[i * stride]; }
};

template <class V> THG_SEL_HD void sel_swap(V v, int a, int b) {
  const float t = v[a];
  v[a] = v[b];
  v[b] = t;
}

// Sifts `value` into the max-heap v[first, first + len) from position `hole`: the hole first moves
// down along the larger children to a leaf, then `value` bubbles up from there.
template <class V> THG_SEL_HD void sel_adjust_heap(V v, int first, int hole, int len, float value) {
  const int top = hole;
  int child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (v[first + child] < v[first + child - 1]) --child;
    v[first + hole] = v[first + child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    v[first + hole] = v[first + child - 1];
    hole = child - 1;
  }
  int parent = (hole - 1) / 2;
  while (hole > top && v[first + parent] < value) {
    v[first + hole] = v[first + parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  v[first + hole] = value;
}

template <class V> THG_SEL_HD void sel_make_heap(V v, int first, int last) {
  const int len = last - first;
  if (len < 2) return;
  for (int parent = (len - 2) / 2;; --parent) {
    sel_adjust_heap(v, first, parent, len, v[first + parent]);
    if (parent == 0) return;
  }
}

// Keeps the (middle - first) smallest elements in a max-heap at the front.
template <class V> THG_SEL_HD void sel_heap_select(V v, int first, int middle, int last) {
  sel_make_heap(v, first, middle);
  for (int i = middle; i < last; ++i)
    if (v[i] < v[first]) {  // pop the heap's top into i, sift v[i]'s old value in
      const float value = v[i];
      v[i] = v[first];
      sel_adjust_heap(v, first, 0, middle - first, value);
    }
}

// Moves the median of v[a], v[b], v[c] to v[result].
template <class V> THG_SEL_HD void sel_median_to_first(V v, int result, int a, int b, int c) {
  if (v[a] < v[b]) {
    if (v[b] < v[c]) sel_swap(v, result, b);
    else if (v[a] < v[c]) sel_swap(v, result, c);
    else sel_swap(v, result, a);
  } else if (v[a] < v[c]) {
    sel_swap(v, result, a);
  } else if (v[b] < v[c]) {
    sel_swap(v, result, c);
  } else {
    sel_swap(v, result, b);
  }
}

// Hoare partition of v[first, last) around v[pivot] without bounds checks (the pivot's median-of-
// three placement guarantees both scans stop).
template <class V> THG_SEL_HD int sel_partition(V v, int first, int last, int pivot) {
  for (;;) {
    while (v[first] < v[pivot]) ++first;
    --last;
    while (v[pivot] < v[last]) --last;
    if (!(first < last)) return first;
    sel_swap(v, first, last);
    ++first;
  }
}

template <class V> THG_SEL_HD void sel_insertion_sort(V v, int first, int last) {
  if (first == last) return;
  for (int i = first + 1; i < last; ++i) {
    const float value = v[i];
    if (value < v[first]) {
      for (int k = i; k > first; --k) v[k] = v[k - 1];
      v[first] = value;
    } else {
      int j = i;
      while (value < v[j - 1]) {
        v[j] = v[j - 1];
        --j;
      }
      v[j] = value;
    }
  }
}

THG_SEL_HD int sel_floor_log2(int n) {
  int r = 0;
  while (n > 1) {
    n >>= 1;
    ++r;
  }
  return r;
}

// std::nth_element(v, v + nth, v + n).
template <class V> THG_SEL_HD void nth_element_gcc(V v, int nth, int n) {
  if (n == 0 || nth == n) return;
  int first = 0, last = n;
  int depth = 2 * sel_floor_log2(n);
  while (last - first > 3) {
    if (depth == 0) {
#ifdef THG_SEL_TRACE
      THG_SEL_TRACE();
#endif
      sel_heap_select(v, first, nth + 1, last);
      sel_swap(v, first, nth);
      return;
    }
    --depth;
    const int mid = first + (last - first) / 2;
    sel_median_to_first(v, first, first + 1, mid, last - 1);
    const int cut = sel_partition(v, first + 1, last, first);
    if (cut <= nth) first = cut;
    else last = cut;
  }
  sel_insertion_sort(v, first, last);
}

}  // namespace thg
