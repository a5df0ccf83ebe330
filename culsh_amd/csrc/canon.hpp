// Lexicographic bucket numbering on the device: the reference's bucket
// order (cuLSH_Indexing.cu:71-93 sorts the floored rows with stable passes
// from column M-1 down to column 0, so each table's buckets come out in
// lexicographic order of their code vectors).  The index keeps its buckets
// in key order; these kernels derive the lexicographic order of each table's
// buckets from the bucket codes, with the same stable LSD radix sort the
// index build uses, over keys that pack several code columns each:
//   k_code_range      smallest and largest code (the packing width)
//   k_lex_key         key of position i: columns [m0, m0 + ncols) of the
//                     bucket at perm[i], column m0 most significant
//   k_lex_table_key   key of position i: the bucket's table (last pass)
// and from the order, the per-table arrays the reference's getters return:
//   k_lex_renumber    global bucket -> its table-local lexicographic index
//   k_lex_sizes / k_lex_starts / k_lex_members / k_lex_codes
#pragma once
#include <climits>

#include "common.hpp"

namespace culsh {

__global__ void k_lex_init(int* range) {
  if (threadIdx.x == 0) {
    range[0] = INT_MAX;
    range[1] = INT_MIN;
  }
}

__global__ __launch_bounds__(256) void k_code_range(const int16_t* __restrict__ codes, long long n,
                                                    int* __restrict__ range) {
  int mn = INT_MAX, mx = INT_MIN;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const int c = codes[i];
    mn = min(mn, c);
    mx = max(mx, c);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    mn = min(mn, __shfl_xor(mn, off));
    mx = max(mx, __shfl_xor(mx, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&range[0], mn);
    atomicMax(&range[1], mx);
  }
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_lex_key(const uint32_t* __restrict__ perm, long long B,
                                                 const int16_t* __restrict__ codes, int M, int m0,
                                                 int ncols, int w, int cmin,
                                                 uint64_t* __restrict__ keys) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  const int16_t* c = codes + (long long)perm[i] * M + m0;
  uint64_t key = 0;
  for (int k = 0; k < ncols; ++k) key = (key << w) | (uint64_t)(uint32_t)((int)c[k] - cmin);
  keys[i] = key;
}

__global__ __launch_bounds__(256) void k_lex_table_key(const uint32_t* __restrict__ perm,
                                                       long long B,
                                                       const BucketRec* __restrict__ recs,
                                                       int tb, uint64_t* __restrict__ keys) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  keys[i] = recs[perm[i]].key >> (64 - tb);
}

__global__ __launch_bounds__(256) void k_lex_renumber(const uint32_t* __restrict__ perm,
                                                      long long B,
                                                      const BucketRec* __restrict__ recs, int tb,
                                                      const uint32_t* __restrict__ table_base,
                                                      int* __restrict__ renumber) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  const uint32_t g = perm[i];
  const int t = (int)(recs[g].key >> (64 - tb));
  renumber[g] = (int)(i - table_base[t]);
}

// Sizes in lexicographic order (all tables, table by table).
__global__ __launch_bounds__(256) void k_lex_sizes(const uint32_t* __restrict__ perm, long long B,
                                                   const BucketRec* __restrict__ recs,
                                                   uint32_t* __restrict__ sz) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < B) sz[i] = recs[perm[i]].size;
}

// Table-local starts from the global exclusive scan of the sizes: table t's
// members start at t * N (every table holds all N points).
__global__ __launch_bounds__(256) void k_lex_starts(const uint32_t* __restrict__ perm, long long B,
                                                    const BucketRec* __restrict__ recs, int tb,
                                                    long long N,
                                                    const uint32_t* __restrict__ gstart,
                                                    uint32_t* __restrict__ lstart) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  const long long t = (long long)(recs[perm[i]].key >> (64 - tb));
  lstart[i] = (uint32_t)((long long)gstart[i] - t * N);
}

// Members of the bucket at lexicographic position i, copied to its place in
// the concatenation (one thread per bucket; buckets are small on average).
__global__ __launch_bounds__(256) void k_lex_members(const uint32_t* __restrict__ perm, long long B,
                                                     const BucketRec* __restrict__ recs,
                                                     const uint32_t* __restrict__ sids,
                                                     const uint32_t* __restrict__ gstart,
                                                     uint32_t* __restrict__ total) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  const BucketRec r = recs[perm[i]];
  const uint32_t dst = gstart[i];
  for (uint32_t k = 0; k < r.size; ++k) total[dst + k] = sids[r.start + k];
}

// Codes as float, column-major per table ([B_t x M] at table_base[t] * M).
__global__ __launch_bounds__(256) void k_lex_codes(const uint32_t* __restrict__ perm, long long B,
                                                   const int16_t* __restrict__ codes, int M,
                                                   const BucketRec* __restrict__ recs, int tb,
                                                   const uint32_t* __restrict__ table_base,
                                                   float* __restrict__ out) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= B * M) return;
  const long long i = e / M;
  const int m = (int)(e - i * M);
  const uint32_t g = perm[i];
  const int t = (int)(recs[g].key >> (64 - tb));
  const long long b0 = table_base[t], Bt = (long long)table_base[t + 1] - b0;
  out[b0 * M + (long long)m * Bt + (i - b0)] = (float)codes[(long long)g * M + m];
}

}  // namespace culsh
