// Index-side kernels: projection -> floor -> code, FP64 boundary recheck,
// bucket keys, bucket heads, bucket arrays, key directory.
//
// Restates cuLSH_Indexing.cu:181-274 (indexTableData) for all L tables at
// once: the reference's per-table SGEMM + subtract + divide + floor
// (Indexing.cu:27-66, Kernels.cu:40-101), M-pass LSD row sort
// (Indexing.cu:71-93) and head-flag bucketing (Indexing.cu:98-144).
#pragma once
#include "common.hpp"

namespace culsh {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr float kU32 = 5.9604644775390625e-08f;  // 2^-24, FP32 unit roundoff
constexpr int kCodeMax = 32767;

// Upper bound of the Euclidean norm of each row (16 lanes per row).
__global__ __launch_bounds__(256) void k_row_norms(const float* __restrict__ X, long long N, int D,
                                                   float* __restrict__ nrm) {
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  const long long n = (long long)blockIdx.x * 16 + g;
  double acc = 0.0;
  if (n < N) {
    const float* x = X + n * D;
    for (int d = lig; d < D; d += 16) acc += (double)x[d] * (double)x[d];
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
  if (n < N && lig == 0) nrm[n] = (float)(sqrt(acc) * (1.0 + 0x1.0p-20));
}

// FP32 squared norm of each row (the norm expansion's |x|^2) and the largest
// row norm (as float bits; norms are non-negative so integer max orders
// them).
__global__ __launch_bounds__(256) void k_row_sqnorms(const float* __restrict__ X, long long N,
                                                     int D, float* __restrict__ xn2,
                                                     unsigned* __restrict__ max_norm_bits) {
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  const long long n = (long long)blockIdx.x * 16 + g;
  float acc = 0.f;
  if (n < N) {
    const float* x = X + n * D;
    for (int d = lig; d < D; d += 16) acc = fmaf(x[d], x[d], acc);
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
  if (n < N && lig == 0) {
    xn2[n] = acc;
    atomicMax(max_norm_bits, __float_as_uint(sqrtf(acc) * (1.0f + 1e-4f)));
  }
}

// Records projection (n, j) for FP64 re-evaluation.
__device__ inline void push_recheck(long long n, int j, unsigned long long* __restrict__ list,
                                    unsigned long long* __restrict__ count,
                                    unsigned long long cap) {
  unsigned long long i = atomicAdd(count, 1ull);
  if (i < cap) list[i] = ((unsigned long long)n << 24) | (unsigned long long)j;
}

// Epilogue of a projection tile element: v = (s - b) / W, code = floor(v),
// with the FP64 recheck decision.  `err_bound` bounds |s - x.a| (the dot
// product error of the fast path).  Any rounding of the epilogue itself is
// added below; the decision is taken in FP64.
__device__ inline void project_epilogue(float s, long long n, int j, long long N, int M, float bb,
                                        float W, double err_bound, int16_t* __restrict__ codes,
                                        unsigned long long* __restrict__ list,
                                        unsigned long long* __restrict__ count,
                                        unsigned long long cap, int* __restrict__ err) {
  const float v = (s - bb) / W;
  const double dv = (double)v;
  const double f = floor(dv);
  const double u = (double)kU32;
  const double eps = (err_bound + 2.0 * u * (fabs((double)s) + fabs((double)bb))) / (double)W +
                     4.0 * u * fabs(dv) + 1e-300;
  const double fr = dv - f;
  const int t = j / M, m = j - t * M;
  if (fr <= eps || 1.0 - fr <= eps) push_recheck(n, j, list, count, cap);
  if (f < -(double)(kCodeMax - 1) || f > (double)(kCodeMax - 1)) {
    atomicOr(err, 1);
    return;
  }
  codes[((long long)t * N + n) * M + m] = (int16_t)(int)f;
}

// Maps a linear workgroup id to (row tile, column tile) so that the column
// tiles of one row tile run on the same XCD (workgroups are dealt to the 8
// XCDs round-robin): the X tile is then fetched from HBM once per XCD L2.
__device__ inline bool xcd_tile(int nct, long long ntiles, long long& rt, int& ct) {
  const long long b = blockIdx.x;
  const long long per = (long long)gridDim.x / 8;
  const long long logical = (b & 7) * per + (b >> 3);
  if (logical >= ntiles) return false;
  rt = logical / nct;
  ct = (int)(logical - rt * nct);
  return true;
}

constexpr int kPjRows = 64, kPjCols = 64, kPjKc = 32;

// FP32 projection on MFMA (v_mfma_f32_16x16x4_f32): a 64-point x 64-column
// tile per workgroup, 16 rows per wave, D staged through LDS in chunks of 32.
// Ad is [Dp][Jp] (column j = t*M + m of table t), zero padded.
// Plain-form error bound: |s - x.a| <= 2 (D+2) u ||x|| ||a_j||.
__global__ __launch_bounds__(256) void k_project_f32(
    const float* __restrict__ X, long long N, int D, const float* __restrict__ Ad, int Jp, int J,
    int M, const float* __restrict__ bj, const float* __restrict__ naj,
    const float* __restrict__ nx, float W, int16_t* __restrict__ codes,
    unsigned long long* __restrict__ list, unsigned long long* __restrict__ count,
    unsigned long long cap, int* __restrict__ err) {
  __shared__ float Xs[kPjRows][kPjKc + 1];
  __shared__ float As[kPjKc][kPjCols + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nct = Jp / kPjCols;
  const long long nrt = (N + kPjRows - 1) / kPjRows;
  long long rt;
  int ct;
  if (!xcd_tile(nct, nrt * nct, rt, ct)) return;
  const long long n0 = rt * kPjRows;
  const int j0 = ct * kPjCols;
  floatx4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int d0 = 0; d0 < D; d0 += kPjKc) {
    for (int e = tid; e < kPjRows * kPjKc; e += 256) {
      const int r = e / kPjKc, c = e % kPjKc;
      const long long n = n0 + r;
      const int d = d0 + c;
      Xs[r][c] = (n < N && d < D) ? X[n * D + d] : 0.f;
    }
    for (int e = tid; e < kPjKc * kPjCols; e += 256) {
      const int r = e / kPjCols, c = e % kPjCols;
      const int d = d0 + r;
      As[r][c] = (d < D) ? Ad[(long long)d * Jp + j0 + c] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kPjKc; kk += 4) {
      const float a = Xs[w * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float b = As[kk + (lane >> 4)][c * 16 + (lane & 15)];
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const double ka = 2.0 * (double)(D + 2) * (double)kU32;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int j = j0 + c * 16 + (lane & 15);
    if (j >= J) continue;
    const float bb = bj[j];
    const double na = (double)naj[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long long n = n0 + w * 16 + (lane >> 4) * 4 + i;
      if (n >= N) continue;
      project_epilogue(acc[c][i], n, j, N, M, bb, W, ka * (double)nx[n] * na, codes, list, count,
                       cap, err);
    }
  }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// float -> bf16, round to nearest even (finite inputs).
__device__ inline uint16_t f2bf_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ inline float bf_bits2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Splits x into hi + lo bf16 parts: |x - hi - lo| <= 2^-16 |x|.
__device__ inline void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = f2bf_bits(x);
  lo = f2bf_bits(x - bf_bits2f(hi));
}

// Precision form: BF16x3 split-precision projection on MFMA
// (v_mfma_f32_16x16x32_bf16), x.a ~ xh.ah + xh.al + xl.ah with FP32
// accumulation.  Same tiling and epilogue as k_project_f32.  Error bound:
// each term is within 3.01 * 2^-16 |x_d a_d| of x_d a_d (dropped xl.al and
// the split residuals) and the 3D-term FP32 accumulation adds
// 3.1 D 2^-24 sum |x_d a_d|; with sum |x_d a_d| <= ||x|| ||a_j|| and a
// factor 2 margin.
__global__ __launch_bounds__(256) void k_project_bf16x3(
    const float* __restrict__ X, long long N, int D, const float* __restrict__ Ad, int Jp, int J,
    int M, const float* __restrict__ bj, const float* __restrict__ naj,
    const float* __restrict__ nx, float W, int16_t* __restrict__ codes,
    unsigned long long* __restrict__ list, unsigned long long* __restrict__ count,
    unsigned long long cap, int* __restrict__ err) {
  __shared__ float Xs[kPjRows][kPjKc + 1];
  __shared__ float As[kPjKc][kPjCols + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nct = Jp / kPjCols;
  const long long nrt = (N + kPjRows - 1) / kPjRows;
  long long rt;
  int ct;
  if (!xcd_tile(nct, nrt * nct, rt, ct)) return;
  const long long n0 = rt * kPjRows;
  const int j0 = ct * kPjCols;
  floatx4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int kb = (lane >> 4) * 8;  // this lane's 8 consecutive k of the 32
  for (int d0 = 0; d0 < D; d0 += kPjKc) {
    for (int e = tid; e < kPjRows * kPjKc; e += 256) {
      const int r = e / kPjKc, c = e % kPjKc;
      const long long n = n0 + r;
      const int d = d0 + c;
      Xs[r][c] = (n < N && d < D) ? X[n * D + d] : 0.f;
    }
    for (int e = tid; e < kPjKc * kPjCols; e += 256) {
      const int r = e / kPjCols, c = e % kPjCols;
      const int d = d0 + r;
      As[r][c] = (d < D) ? Ad[(long long)d * Jp + j0 + c] : 0.f;
    }
    __syncthreads();
    bf16x8 xh, xl;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      uint16_t h, l;
      split_bf16(Xs[w * 16 + (lane & 15)][kb + i], h, l);
      xh[i] = __builtin_bit_cast(__bf16, h);
      xl[i] = __builtin_bit_cast(__bf16, l);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      bf16x8 ah, al;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        uint16_t h, l;
        split_bf16(As[kb + i][c * 16 + (lane & 15)], h, l);
        ah[i] = __builtin_bit_cast(__bf16, h);
        al[i] = __builtin_bit_cast(__bf16, l);
      }
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, ah, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, al, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, ah, acc[c], 0, 0, 0);
    }
    __syncthreads();
  }
  const double kb3 = 2.0 * (3.01 * 0x1.0p-16 + 3.1 * (double)D * 0x1.0p-24);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int j = j0 + c * 16 + (lane & 15);
    if (j >= J) continue;
    const float bb = bj[j];
    const double na = (double)naj[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long long n = n0 + w * 16 + (lane >> 4) * 4 + i;
      if (n >= N) continue;
      project_epilogue(acc[c][i], n, j, N, M, bb, W, kb3 * (double)nx[n] * na, codes, list, count,
                       cap, err);
    }
  }
}

// FP64 projection of one (n, j), the oracle's operation order: sequential
// FMA over d ascending, then (acc - b) / W (cuLSH_Indexing.cu:43-62 in FP64).
__device__ inline double project64(const float* __restrict__ x, const float* __restrict__ a,
                                   int D, float b, float W) {
  double acc = 0.0;
  for (int d = 0; d < D; ++d) acc = fma((double)x[d], (double)a[d], acc);
  return (acc - (double)b) / (double)W;
}

// Re-evaluates the recorded projections in FP64 and patches their codes.
// Aorig is [L*M][D]: column j of the projection contiguous over d.
__global__ __launch_bounds__(256) void k_recheck(const float* __restrict__ X, long long N, int D,
                                                 const float* __restrict__ Aorig,
                                                 const float* __restrict__ bj, float W, int M,
                                                 const unsigned long long* __restrict__ list,
                                                 unsigned long long count,
                                                 int16_t* __restrict__ codes,
                                                 int* __restrict__ err) {
  const unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const unsigned long long e = list[i];
  const long long n = (long long)(e >> 24);
  const int j = (int)(e & 0xFFFFFFull);
  const double v = project64(X + n * D, Aorig + (long long)j * D, D, bj[j], W);
  const double f = floor(v);
  if (f < -(double)(kCodeMax - 1) || f > (double)(kCodeMax - 1)) {
    atomicOr(err, 1);
    return;
  }
  const int t = j / M, m = j - t * M;
  codes[((long long)t * N + n) * M + m] = (int16_t)(int)f;
}

// FP64 projection of every (n, j): the fallback when the recheck list
// overflows its capacity.
__global__ __launch_bounds__(256) void k_project_f64_all(const float* __restrict__ X, long long N,
                                                         int D, const float* __restrict__ Aorig,
                                                         const float* __restrict__ bj, float W,
                                                         int M, int J,
                                                         int16_t* __restrict__ codes,
                                                         int* __restrict__ err) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= N * J) return;
  const long long n = i / J;
  const int j = (int)(i - n * J);
  const double v = project64(X + n * D, Aorig + (long long)j * D, D, bj[j], W);
  const double f = floor(v);
  if (f < -(double)(kCodeMax - 1) || f > (double)(kCodeMax - 1)) {
    atomicOr(err, 1);
    return;
  }
  const int t = j / M, m = j - t * M;
  codes[((long long)t * N + n) * M + m] = (int16_t)(int)f;
}

// Bucket key of every (table, point): keys[t*N + n], vals = n.
__global__ __launch_bounds__(256) void k_keys(const int16_t* __restrict__ codes, long long N,
                                              int L, int M, const uint64_t* __restrict__ seeds,
                                              int tb, uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ vals) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)L * N) return;
  const int t = (int)(i / N);
  const long long n = i - (long long)t * N;
  const int16_t* c = codes + i * M;
  const uint64_t seed = seeds[t];
  uint64_t h = 0;
  for (int m = 0; m < M; ++m) h += code_term(seed, m, (int)c[m]);
  keys[i] = make_key(h, t, tb);
  vals[i] = (uint32_t)n;
}

// Head flags over the sorted (key, id) array.  Equal keys with different
// codes (a 64-bit hash collision) set *collision; the host then rebuilds
// with another hash seed.
__global__ __launch_bounds__(256) void k_heads(const uint64_t* __restrict__ keys,
                                               const uint32_t* __restrict__ ids, long long n_all,
                                               long long N, int M, int tb,
                                               const int16_t* __restrict__ codes,
                                               uint32_t* __restrict__ flag,
                                               int* __restrict__ collision) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_all) return;
  const bool head = (i == 0) || keys[i] != keys[i - 1];
  if (!head) {
    const int t = (int)(keys[i] >> (64 - tb));
    const int16_t* a = codes + ((long long)t * N + ids[i]) * M;
    const int16_t* b = codes + ((long long)t * N + ids[i - 1]) * M;
    for (int m = 0; m < M; ++m)
      if (a[m] != b[m]) {
        atomicOr(collision, 1);
        break;
      }
  }
  flag[i] = head ? 1u : 0u;
}

// Copies one M-code row, 8 bytes at a time when both rows allow it.
__device__ inline void copy_codes(int16_t* __restrict__ dst, const int16_t* __restrict__ src,
                                  int M) {
  if ((M & 3) == 0) {
    const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
    uint64_t* d = reinterpret_cast<uint64_t*>(dst);
    for (int i = 0; i < (M >> 2); ++i) d[i] = s[i];
  } else if ((M & 1) == 0) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    for (int i = 0; i < (M >> 1); ++i) d[i] = s[i];
  } else {
    for (int m = 0; m < M; ++m) dst[m] = src[m];
  }
}

// Bucket records from heads: key and start offset into the sorted ids, and
// codes (row-major [B][M]).
__global__ __launch_bounds__(256) void k_bucket_fill(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ ids,
    const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl, long long n_all,
    long long N, int M, int tb, const int16_t* __restrict__ codes, BucketRec* __restrict__ recs,
    int16_t* __restrict__ bcodes) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_all || !flag[i]) return;
  const uint32_t k = excl[i];
  recs[k].key = keys[i];
  recs[k].start = (uint32_t)i;
  const int t = (int)(keys[i] >> (64 - tb));
  copy_codes(bcodes + (long long)k * M, codes + ((long long)t * N + ids[i]) * M, M);
}

// Bucket sizes from consecutive starts.
__global__ __launch_bounds__(256) void k_rec_sizes(BucketRec* __restrict__ recs, long long B,
                                                   long long n_all) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k >= B) return;
  const uint32_t end = k + 1 < B ? recs[k + 1].start : (uint32_t)n_all;
  recs[k].size = end - recs[k].start;
}

// First bucket of each table (positions t*N in the sorted order) and the
// total bucket count at out[L].
__global__ void k_table_bases(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl,
                              long long N, int L, uint32_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L) out[t] = excl[(long long)t * N];
  if (t == L) {
    const long long last = (long long)L * N - 1;
    out[L] = excl[last] + flag[last];
  }
}

__device__ inline long long lower_bound_rec(const BucketRec* __restrict__ recs, long long B,
                                            uint64_t target) {
  long long lo = 0, hi = B;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (recs[mid].key < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Directory entry of every slot v in [0, nslots] (see dir_entry); slot
// nslots is the end sentinel (first = B, count 0).
__global__ __launch_bounds__(256) void k_directory(const BucketRec* __restrict__ recs, long long B,
                                                   int shift, long long nslots,
                                                   uint64_t* __restrict__ dir) {
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v > nslots) return;
  if (v == nslots) {
    dir[v] = dir_entry((uint32_t)B, 0, 0);
    return;
  }
  const long long first = lower_bound_rec(recs, B, (uint64_t)v << shift);
  const long long end =
      v + 1 == nslots ? B : lower_bound_rec(recs, B, (uint64_t)(v + 1) << shift);
  const uint32_t cnt = (uint32_t)(end - first);
  dir[v] = dir_entry((uint32_t)first, cnt ? recs[first].key : 0ull, cnt);
}

}  // namespace culsh
