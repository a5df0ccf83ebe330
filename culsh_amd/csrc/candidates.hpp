// Candidate stage of the search: union of the matched buckets with per-query
// dedup, direct FP32 squared L2, selection of the K-th distance, FP64
// re-evaluation near it, and the reference's top-K tie rule.  Restates
// cuLSH_SearchTables.cu:249-255,321-389 (candidates), cuLSH_Kernels.cu:301-348
// (distances) and cuLSH_Kernels.cu:356-456 (top-K).
//
// One workgroup (256 threads) per query.  query_candidates() works on buffers
// the caller places: k_candidates keeps everything in LDS (the common case);
// k_candidates_global serves the queries that overflow it from a per-
// workgroup slice of global scratch.
#pragma once
#include "common.hpp"
#include "sort.hpp"

namespace culsh {

constexpr int kHcap = 8192;   // LDS hash slots per query
constexpr int kUcap = 4096;   // unique candidates per query (LDS path)
constexpr int kScap = 1024;   // candidates re-evaluated in FP64 per query (LDS path)
constexpr int kScapG = 4096;  // same, global path
constexpr int kYcap = 1024;   // query dimensions staged in LDS
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// FP64 squared L2 in the oracle's order (sequential FMA over d ascending).
// Rows are fetched 32 floats at a time (8 independent 16-byte loads) ahead
// of the dependent FMA chain when D is a multiple of 32.
__device__ inline double dist64(const float* __restrict__ q, const float* __restrict__ x, int D) {
  double acc = 0.0;
  if ((D & 31) == 0) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    for (int d0 = 0; d0 < D; d0 += 32) {
      float4 xv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) xv[k] = x4[(d0 >> 2) + k];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float* xs = reinterpret_cast<const float*>(&xv[k]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double df = (double)q[d0 + 4 * k + e] - (double)xs[e];
          acc = fma(df, df, acc);
        }
      }
    }
    return acc;
  }
#pragma unroll 8
  for (int d = 0; d < D; ++d) {
    const double df = (double)q[d] - (double)x[d];
    acc = fma(df, df, acc);
  }
  return acc;
}

// Direct FP32 squared L2 of the U candidate rows ids[0..U) against q
// (cuLSH_Kernels.cu:319-346): 16 lanes per row, four rows per 16-lane group
// with their loads issued together.  Writes d32[u].
__device__ inline void l2_fp32_rows(const float* __restrict__ X, const float* __restrict__ q,
                                    int D, const uint32_t* ids, int U, float* d32) {
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  if ((D & 3) == 0) {
    const int D4 = D >> 2;
    const float4* q4 = reinterpret_cast<const float4*>(q);
    for (int u0 = g * 4; u0 < U; u0 += 64) {
      const float4* xr[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        xr[c] = reinterpret_cast<const float4*>(X + (long long)ids[min(u0 + c, U - 1)] * D);
      float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
      for (int f = lig; f < D4; f += 16) {
        float4 a[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = xr[c][f];
        const float4 b = q4[f];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float d0 = b.x - a[c].x, d1 = b.y - a[c].y, d2 = b.z - a[c].z,
                      d3 = b.w - a[c].w;
          acc[c] = fmaf(d0, d0, acc[c]);
          acc[c] = fmaf(d1, d1, acc[c]);
          acc[c] = fmaf(d2, d2, acc[c]);
          acc[c] = fmaf(d3, d3, acc[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 16);
      }
      if (lig == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (u0 + c < U) d32[u0 + c] = acc[c];
      }
    }
  } else {
    for (int u = g; u < U; u += 16) {
      const float* x = X + (long long)ids[u] * D;
      float acc = 0.f;
      for (int d = lig; d < D; d += 16) {
        const float df = q[d] - x[d];
        acc = fmaf(df, df, acc);
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
      if (lig == 0) d32[u] = acc;
    }
  }
}

// Where one query's working set lives.
struct CandBufs {
  uint32_t* hash;   // 2^hbits slots (open addressing)
  int hbits;
  uint32_t* ulist;  // unique candidate ids, capacity ucap
  unsigned ucap;
  float* d32;       // FP32 distances, capacity ucap (may alias hash)
  double* sd;       // FP64 distances of the near-K-th set, capacity scap
  uint32_t* sid;    // their ids
  unsigned scap;
  float* ys;        // query vector staging (kYcap floats, may alias hash), or null
};

// Processes query q.  Returns false (uniformly) when the query does not fit
// the buffers; its outputs are then not written.
__device__ inline bool query_candidates(
    int q, const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T,
    int K, const int* __restrict__ qb, const BucketRec* __restrict__ recs,
    const uint32_t* __restrict__ sids, const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, double rel_bound,
    unsigned long long* __restrict__ n_recheck, const CandBufs& B) {
  __shared__ uint32_t s_hist[256];
  __shared__ unsigned long long s_cnu;
  __shared__ unsigned s_count, s_ovf, s_scount, s_prefix, s_rem, s_misc[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const float* y = Qm + (long long)q * D;
  const uint32_t hslots = 1u << B.hbits, hmask = hslots - 1;

  __syncthreads();  // a previous query's readers of the shared counters are done
  if (tid == 0) {
    unsigned long long c = 0;
    for (int t = 0; t < L; ++t) c += cnu_qt[(long long)q * L + t];
    s_cnu = c;
    s_count = 0;
    s_ovf = 0;
    s_scount = 0;
  }
  for (uint32_t i = tid; i < hslots; i += 256) B.hash[i] = kEmpty;
  __syncthreads();
  const unsigned long long cnu = s_cnu;

  // Union of matched buckets (cuLSH_SearchTables.cu:321-364) with dedup: a
  // wave takes 64 (table, probe) entries, prefix-sums their bucket sizes and
  // spreads the member ids over its lanes.
  if (cnu > 0) {
    const int LT = L * T;
    for (int e0 = w * 64; e0 < LT; e0 += 256) {
      const int e = e0 + lane;
      uint32_t st = 0, sz = 0;
      if (e < LT) {
        const int t = e / T, p = e - t * T;
        const int k = qb[(long long)t * Q * T + (long long)q * T + p];
        if (k >= 0) {
          st = recs[k].start;
          sz = recs[k].size;
        }
      }
      uint32_t inc = sz;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t yv = __shfl_up(inc, off);
        if (lane >= off) inc += yv;
      }
      const uint32_t tot = __shfl(inc, 63);
      for (uint32_t b = 0; b < tot; b += 64) {
        const uint32_t i = b + lane;
        int lo = 0, hi = 63;
        while (lo < hi) {  // first lane whose inclusive prefix exceeds i
          const int mid = (lo + hi) >> 1;
          const uint32_t im = __shfl(inc, mid);
          if (im > i)
            hi = mid;
          else
            lo = mid + 1;
        }
        const uint32_t st_l = __shfl(st, lo);
        const uint32_t inc_l = __shfl(inc, lo);
        const uint32_t sz_l = __shfl(sz, lo);
        if (i < tot) {
          const uint32_t id = sids[st_l + (i - (inc_l - sz_l))];
          uint32_t h = (id * 0x9E3779B1u) >> (32 - B.hbits);
          for (uint32_t probe = 0;; ++probe) {
            if (probe == hslots) {
              s_ovf = 1;
              break;
            }
            const uint32_t old = atomicCAS(&B.hash[h], kEmpty, id);
            if (old == kEmpty) {
              const unsigned idx = atomicAdd(&s_count, 1u);
              if (idx < B.ucap)
                B.ulist[idx] = id;
              else
                s_ovf = 1;
              break;
            }
            if (old == id) break;
            h = (h + 1) & hmask;
          }
        }
      }
    }
  }
  __syncthreads();
  if (s_ovf) return false;
  const int U = (int)s_count;

  // The query vector moves to LDS-side staging (hash slots are free now).
  const bool y_staged = B.ys != nullptr && D <= kYcap;
  if (y_staged)
    for (int d = tid; d < D; d += 256) B.ys[d] = y[d];
  __syncthreads();
  if (y_staged)
    l2_fp32_rows(X, B.ys, D, B.ulist, U, B.d32);
  else
    l2_fp32_rows(X, y, D, B.ulist, U, B.d32);
  __syncthreads();

  // K-th smallest FP32 distance by radix select over the float bits
  // (non-negative floats order like their bit patterns).
  double thr = INFINITY;
  if (U > K) {
    if (tid == 0) {
      s_prefix = 0;
      s_rem = (unsigned)K;
    }
    uint32_t pmask = 0;
    for (int shiftb = 24; shiftb >= 0; shiftb -= 8) {
      s_hist[tid] = 0;
      __syncthreads();
      const uint32_t prefix = s_prefix;
      for (int u = tid; u < U; u += 256) {
        const uint32_t bits = __float_as_uint(B.d32[u]);
        if ((bits & pmask) == prefix) atomicAdd(&s_hist[(bits >> shiftb) & 255u], 1u);
      }
      __syncthreads();
      // The digit whose bin holds the rem-th smallest: bins are scanned in
      // parallel, exactly one thread finds rem inside its bin.
      const unsigned rem = s_rem;
      const uint32_t h = s_hist[tid];
      uint32_t tot;
      const uint32_t before = block_exclusive_scan_256(h, s_misc, tot);
      if (h > 0 && before < rem && rem <= before + h) {
        s_rem = rem - before;
        s_prefix = prefix | ((uint32_t)tid << shiftb);
      }
      pmask |= 255u << shiftb;
      __syncthreads();
    }
    const double dk32 = (double)__uint_as_float(s_prefix);
    thr = dk32 * ((1.0 + rel_bound) / (1.0 - rel_bound)) * (1.0 + 1e-12) + (double)D * 0x1.0p-120;
  }

  // Candidates within the error bound of the K-th distance, re-evaluated in
  // FP64 in the oracle's operation order.
  for (int u = tid; u < U; u += 256) {
    if ((double)B.d32[u] <= thr) {
      const unsigned i = atomicAdd(&s_scount, 1u);
      if (i < B.scap) {
        const uint32_t id = B.ulist[u];
        B.sd[i] = y_staged ? dist64(B.ys, X + (long long)id * D, D)
                           : dist64(y, X + (long long)id * D, D);
        B.sid[i] = id;
      } else {
        s_ovf = 1;
      }
    }
  }
  __syncthreads();
  if (s_ovf) return false;
  const int S = (int)s_scount;
  if (tid == 0) {
    cnu_q[q] = cnu;
    cu_q[q] = (unsigned long long)U;
    if (S > 0) atomicAdd(n_recheck, (unsigned long long)S);
  }
  // Candidates at or beyond FLT_MAX are never admitted by the reference's
  // heap (initialised to FLT_MAX, strict <, cuLSH_Kernels.cu:426).
  int P2 = 1;
  while (P2 < S) P2 <<= 1;
  for (int i = tid; i < P2; i += 256) {
    if (i >= S || B.sd[i] >= (double)FLT_MAX) {
      B.sd[i] = INFINITY;
      B.sid[i] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
  // Bitonic sort by (distance, id).
  for (int kk = 2; kk <= P2; kk <<= 1) {
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      for (int i = tid; i < P2; i += 256) {
        const int ixj = i ^ jj;
        if (ixj > i) {
          const bool upw = (i & kk) == 0;
          const double a = B.sd[i], b = B.sd[ixj];
          const uint32_t ia = B.sid[i], ib = B.sid[ixj];
          const bool gt = (a > b) || (a == b && ia > ib);
          if (gt == upw) {
            B.sd[i] = b;
            B.sd[ixj] = a;
            B.sid[i] = ib;
            B.sid[ixj] = ia;
          }
        }
      }
      __syncthreads();
    }
  }
  // Admitted entries form a prefix of length A.
  if (tid == 0) {
    s_misc[0] = 0;
    s_misc[1] = 0;
    s_misc[2] = 0;
    s_misc[3] = 0;
  }
  __syncthreads();
  for (int i = tid; i < S; i += 256)
    if (B.sd[i] != INFINITY && (i + 1 == S || B.sd[i + 1] == INFINITY))
      s_misc[0] = (unsigned)(i + 1);
  __syncthreads();
  const int A = (int)s_misc[0];
  int nout = A < K ? A : K;
  int Lt = nout, e0 = nout;
  if (U > K && A > K) {
    // Reference tie rule (closed form of the heap scan in descending id):
    // keep every candidate below d_K, then the r = K - |below| smallest ids
    // among the d_K-ties whose id is >= id*, where id* is the K-th largest
    // id among all candidates at or below d_K.
    const double dK = B.sd[K - 1];
    for (int i = tid; i < A; i += 256) {
      if (B.sd[i] == dK && (i == 0 || B.sd[i - 1] < dK)) s_misc[1] = (unsigned)i;
      if (B.sd[i] == dK && (i + 1 == A || B.sd[i + 1] > dK)) s_misc[2] = (unsigned)(i + 1);
    }
    __syncthreads();
    Lt = (int)s_misc[1];
    const int m0 = (int)s_misc[2];
    if (m0 > K) {
      for (int i = tid; i < m0; i += 256) {
        const uint32_t me = B.sid[i];
        int greater = 0;
        for (int j2 = 0; j2 < m0; ++j2) greater += B.sid[j2] > me;
        if (greater == K - 1) s_misc[3] = me;
      }
      __syncthreads();
      const uint32_t idstar = s_misc[3];
      for (int i = Lt + tid; i < m0; i += 256)
        if (B.sid[i] >= idstar && (i == Lt || B.sid[i - 1] < idstar)) s_misc[0] = (unsigned)i;
      __syncthreads();
      e0 = (int)s_misc[0];
    } else {
      e0 = Lt;
    }
    nout = K;
  }
  for (int k = tid; k < K; k += 256) {
    int id = -1;
    float dd = FLT_MAX;
    if (k < nout) {
      const int src = k < Lt ? k : e0 + (k - Lt);
      id = (int)B.sid[src];
      dd = (float)B.sd[src];
    }
    out_ids[(long long)q * K + k] = id;
    out_d[(long long)q * K + k] = dd;
  }
  __syncthreads();
  return true;
}

// One workgroup per query, all buffers in LDS (~50 KB).  After the union the
// hash slots hold d32 [kUcap], sd [kScap] (f64), sid [kScap] and the query.
// Queries that do not fit are appended to overflow[] for
// k_candidates_global.
__global__ __launch_bounds__(256) void k_candidates(
    const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const int* __restrict__ qb, const BucketRec* __restrict__ recs,
    const uint32_t* __restrict__ sids, const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, int* __restrict__ overflow,
    int* __restrict__ n_overflow, double rel_bound, unsigned long long* __restrict__ n_recheck) {
  __shared__ uint32_t s_hash[kHcap];
  __shared__ uint32_t s_u[kUcap];
  static_assert(kUcap + 3 * kScap + kYcap <= kHcap, "post-union LDS layout");
  CandBufs B;
  B.hash = s_hash;
  B.hbits = 13;
  B.ulist = s_u;
  B.ucap = kUcap;
  B.d32 = reinterpret_cast<float*>(s_hash);
  B.sd = reinterpret_cast<double*>(s_hash + kUcap);
  B.sid = s_hash + kUcap + 2 * kScap;
  B.scap = kScap;
  B.ys = reinterpret_cast<float*>(s_hash + kUcap + 3 * kScap);
  const int q = blockIdx.x;
  if (!query_candidates(q, X, Qm, Q, D, L, T, K, qb, recs, sids, cnu_qt, out_ids, out_d, cnu_q,
                        cu_q, rel_bound, n_recheck, B)) {
    if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1)] = q;
  }
}

// Largest summed bucket size (C_nu) among the listed queries.
__global__ __launch_bounds__(256) void k_max_cnu(const int* __restrict__ list, int n_list,
                                                 const unsigned long long* __restrict__ cnu_qt,
                                                 int L, unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_list) return;
  unsigned long long c = 0;
  for (int t = 0; t < L; ++t) c += cnu_qt[(long long)list[i] * L + t];
  atomicMax(out, c);
}

// The overflowed queries: workgroup b takes list[b], list[b + G], ...; its
// hash (2^hbits slots), id list and d32 (ucap each) live in global scratch
// at b * (2^hbits + 2 * ucap) words.  A query that still does not fit sets
// *err.
__global__ __launch_bounds__(256) void k_candidates_global(
    const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const int* __restrict__ qb, const BucketRec* __restrict__ recs,
    const uint32_t* __restrict__ sids, const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, const int* __restrict__ list, int n_list,
    uint32_t* __restrict__ scratch, int hbits, unsigned ucap, double rel_bound,
    unsigned long long* __restrict__ n_recheck, int* __restrict__ err) {
  __shared__ double s_sd[kScapG];
  __shared__ uint32_t s_sid[kScapG];
  __shared__ float s_y[kYcap];
  uint32_t* base = scratch + (size_t)blockIdx.x * ((1ull << hbits) + 2ull * ucap);
  CandBufs B;
  B.hash = base;
  B.hbits = hbits;
  B.ulist = base + (1u << hbits);
  B.ucap = ucap;
  B.d32 = reinterpret_cast<float*>(base + (1u << hbits) + ucap);
  B.sd = s_sd;
  B.sid = s_sid;
  B.scap = kScapG;
  B.ys = s_y;
  for (int i = blockIdx.x; i < n_list; i += gridDim.x) {
    if (!query_candidates(list[i], X, Qm, Q, D, L, T, K, qb, recs, sids, cnu_qt, out_ids, out_d,
                          cnu_q, cu_q, rel_bound, n_recheck, B)) {
      if (threadIdx.x == 0) atomicOr(err, 1);
    }
  }
}

}  // namespace culsh
