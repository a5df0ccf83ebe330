// The reference's ordering of multiprobe deltas by a bounded max-heap, used
// by the probe kernels when two deltas are exactly equal (the heap's
// mechanics then decide their order).
#pragma once
#include "common.hpp"

namespace culsh {

__device__ inline int ceil_log2_i(int x) {
  int l = 0;
  while ((1 << l) < x) ++l;
  return l;
}

// Sift-down of the bounded max-heap of delta indices (-1 = empty = +inf),
// cuLSH_Kernels.cu:144-175.
template <typename P>
__device__ inline void sift_deltas(P* perm, const double* d, int total, int levels) {
  int parent = 0;
  for (int level = 0; level < levels - 1; ++level) {
    int left = parent * 2 + 1, right = left + 1, mx = parent;
    if (left < total && perm[left] == -1)
      mx = left;
    else if (right < total && perm[right] == -1)
      mx = right;
    else {
      if (left < total && d[perm[left]] > d[perm[mx]]) mx = left;
      if (right < total && d[perm[right]] > d[perm[mx]]) mx = right;
    }
    if (mx == parent) break;
    const P tmp = perm[parent];
    perm[parent] = perm[mx];
    perm[mx] = tmp;
    parent = mx;
  }
}

// Heap select of the maxT smallest of d[0..twoM), then heapsort, into perm
// (ascending), cuLSH_Kernels.cu:202-214.  Run by one lane.
template <typename P>
__device__ inline void heap_order_deltas(const double* d, int twoM, int maxT, P* perm) {
  const int levels = ceil_log2_i(twoM + 1);
  for (int i = 0; i < maxT; ++i) perm[i] = -1;
  for (int i = 0; i < twoM; ++i)
    if (perm[0] == -1 || d[i] < d[perm[0]]) {
      perm[0] = (P)i;
      sift_deltas(perm, d, maxT, levels);
    }
  for (int i = maxT - 1; i; --i) {
    const P tmp = perm[0];
    perm[0] = perm[i];
    perm[i] = tmp;
    sift_deltas(perm, d, i, levels);
  }
}

}  // namespace culsh
