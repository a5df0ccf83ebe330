// Candidate stage of the search: union of the matched buckets with per-query
// dedup, approximate squared L2, selection of the K-th distance, FP64
// re-evaluation near it, and the reference's top-K tie rule.  Restates
// cuLSH_SearchTables.cu:249-255,321-389 (candidates), cuLSH_Kernels.cu:301-348
// (distances) and cuLSH_Kernels.cu:356-456 (top-K).
//
// Plain form: one workgroup (256 threads) per query, direct FP32 distances
// (k_candidates, all in LDS; k_candidates_global for queries that overflow
// it).  Precision form: one workgroup per group of 16 locality-ordered
// queries; the union of their candidates is read once and distances come
// from the norm expansion on FP32 MFMA (k_group_candidates).
#pragma once
#include "common.hpp"
#include "index_kernels.hpp"
#include "sort.hpp"

namespace culsh {

constexpr int kHcap = 8192;   // LDS hash slots per query
constexpr int kUcap = 4096;   // unique candidates per query (LDS path)
constexpr int kScap = 1024;   // candidates re-evaluated in FP64 per query (LDS paths)
constexpr int kScapG = 4096;  // same, global path
constexpr int kYcap = 1024;   // query dimensions staged in LDS
constexpr int kOrderMinQ = 16384;  // batches this large run in locality order
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr double kU32d = 0x1.0p-24;

// FP64 squared L2 in the oracle's order (sequential FMA over d ascending).
// When D is a multiple of 32 the row is fetched 32 floats at a time (8
// independent 16-byte loads), the next 32 in flight while the current ones
// go through the dependent FMA chain.
__device__ inline double dist64(const float* __restrict__ q, const float* __restrict__ x, int D) {
  double acc = 0.0;
  if ((D & 31) == 0) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    float4 cur[8], nxt[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) cur[k] = x4[k];
    for (int d0 = 0; d0 < D; d0 += 32) {
      const bool more = d0 + 32 < D;
      if (more) {
#pragma unroll
        for (int k = 0; k < 8; ++k) nxt[k] = x4[((d0 + 32) >> 2) + k];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float* xs = reinterpret_cast<const float*>(&cur[k]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double df = (double)q[d0 + 4 * k + e] - (double)xs[e];
          acc = fma(df, df, acc);
        }
      }
      if (more) {
#pragma unroll
        for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
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
// (cuLSH_Kernels.cu:319-346): a lane group per row (8 lanes when D is 32, 64
// or 96, else 16), R rows per group with their loads issued together (the
// candidate chunk).  Writes d32[u].
template <int R = 4>
__device__ inline void l2_fp32_rows(const float* __restrict__ X, const float* __restrict__ q,
                                    int D, const uint32_t* ids, int U, float* d32) {
  const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
  if ((D & 31) == 0 && D <= 96) {
    // D = 32, 64, 96: 8 lanes per row with D/32 16-byte loads each (no idle
    // lanes), R8 rows per group, so a workgroup has 32 R8 rows in flight.
    // Any summation order keeps the error bound the caller uses.
    constexpr int R8 = R >= 8 ? 6 : R;
    const int g8 = threadIdx.x >> 3, l8 = threadIdx.x & 7;
    const int nf = D >> 5;
    const float4* q4 = reinterpret_cast<const float4*>(q);
    float4 qv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) qv[k] = k < nf ? q4[l8 + 8 * k] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int u0 = g8 * R8; u0 < U; u0 += 32 * R8) {
      float4 xv[R8][3];
#pragma unroll
      for (int c = 0; c < R8; ++c) {
        const float4* xr =
            reinterpret_cast<const float4*>(X + (long long)ids[min(u0 + c, U - 1)] * D);
#pragma unroll
        for (int k = 0; k < 3; ++k)
          xv[c][k] = k < nf ? xr[l8 + 8 * k] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      float acc[R8];
#pragma unroll
      for (int c = 0; c < R8; ++c) {
        acc[c] = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k < nf) {
            const float d0 = qv[k].x - xv[c][k].x, d1 = qv[k].y - xv[c][k].y,
                        d2 = qv[k].z - xv[c][k].z, d3 = qv[k].w - xv[c][k].w;
            acc[c] = fmaf(d0, d0, acc[c]);
            acc[c] = fmaf(d1, d1, acc[c]);
            acc[c] = fmaf(d2, d2, acc[c]);
            acc[c] = fmaf(d3, d3, acc[c]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < R8; ++c) {
#pragma unroll
        for (int off = 4; off >= 1; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 8);
      }
      if (l8 == 0) {
#pragma unroll
        for (int c = 0; c < R8; ++c)
          if (u0 + c < U) d32[u0 + c] = acc[c];
      }
    }
  } else if ((D & 3) == 0) {
    const int D4 = D >> 2;
    const float4* q4 = reinterpret_cast<const float4*>(q);
    for (int u0 = g * R; u0 < U; u0 += 16 * R) {
      const float4* xr[R];
#pragma unroll
      for (int c = 0; c < R; ++c)
        xr[c] = reinterpret_cast<const float4*>(X + (long long)ids[min(u0 + c, U - 1)] * D);
      float acc[R];
#pragma unroll
      for (int c = 0; c < R; ++c) acc[c] = 0.f;
#pragma unroll 2
      for (int f = lig; f < D4; f += 16) {
        float4 a[R];
#pragma unroll
        for (int c = 0; c < R; ++c) a[c] = xr[c][f];
        const float4 b = q4[f];
#pragma unroll
        for (int c = 0; c < R; ++c) {
          const float d0 = b.x - a[c].x, d1 = b.y - a[c].y, d2 = b.z - a[c].z,
                      d3 = b.w - a[c].w;
          acc[c] = fmaf(d0, d0, acc[c]);
          acc[c] = fmaf(d1, d1, acc[c]);
          acc[c] = fmaf(d2, d2, acc[c]);
          acc[c] = fmaf(d3, d3, acc[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < R; ++c) {
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 16);
      }
      if (lig == 0) {
#pragma unroll
        for (int c = 0; c < R; ++c)
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

// FP32 squared L2 of the U candidate rows ids[0..U) of the FP16 copy X16
// ([N][D] half, D % 8 == 0) against q: each half is widened exactly and the
// arithmetic is the FP32 path's, so |d32 - d~| <= (D + 3) u d~ for the
// copy's distance d~.  4 lanes per row, each with the row's 16-byte chunks
// l, l + 4, ...; R rows per group (64 groups) with their loads issued
// together; q is read from LDS or L1 per chunk.  Writes d32[u].
__device__ inline float half_lo(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu));
}
__device__ inline float half_hi(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}
__device__ inline float sq8(const uint4 x, const float4 qa, const float4 qb, float acc) {
  float d;
  d = qa.x - half_lo(x.x); acc = fmaf(d, d, acc);
  d = qa.y - half_hi(x.x); acc = fmaf(d, d, acc);
  d = qa.z - half_lo(x.y); acc = fmaf(d, d, acc);
  d = qa.w - half_hi(x.y); acc = fmaf(d, d, acc);
  d = qb.x - half_lo(x.z); acc = fmaf(d, d, acc);
  d = qb.y - half_hi(x.z); acc = fmaf(d, d, acc);
  d = qb.z - half_lo(x.w); acc = fmaf(d, d, acc);
  d = qb.w - half_hi(x.w); acc = fmaf(d, d, acc);
  return acc;
}

template <int R, int CH>
__device__ inline void l2_fp16_rows_ch(const uint16_t* __restrict__ X16, const float* __restrict__ q,
                                       int D, const uint32_t* ids, int U, float* d32) {
  const int g4 = threadIdx.x >> 2, l4 = threadIdx.x & 3;
  const int nch = D >> 3;
  const float4* q4 = reinterpret_cast<const float4*>(q);
  for (int u0 = g4 * R; u0 < U; u0 += 64 * R) {
    uint4 xv[R][CH];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint4* xr = reinterpret_cast<const uint4*>(X16 + (long long)ids[min(u0 + r, U - 1)] * D);
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int c = l4 + 4 * k;
        xv[r][k] = c < nch ? xr[c] : make_uint4(0u, 0u, 0u, 0u);
      }
    }
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.f;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int c = l4 + 4 * k;
      if (c < nch) {
        const float4 qa = q4[2 * c], qb = q4[2 * c + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = sq8(xv[r][k], qa, qb, acc[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      acc[r] += __shfl_xor(acc[r], 2, 4);
      acc[r] += __shfl_xor(acc[r], 1, 4);
    }
    if (l4 == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (u0 + r < U) d32[u0 + r] = acc[r];
    }
  }
}

template <int R>
__device__ inline void l2_fp16_rows(const uint16_t* __restrict__ X16, const float* __restrict__ q,
                                    int D, const uint32_t* ids, int U, float* d32) {
  const int nch = D >> 3;
  if (nch <= 12) {
    l2_fp16_rows_ch<R, 3>(X16, q, D, ids, U, d32);
  } else if (nch <= 16) {
    l2_fp16_rows_ch<R, 4>(X16, q, D, ids, U, d32);
  } else {
    // Wide rows: 16 lanes per row, one chunk per lane at a time.
    const int g = threadIdx.x >> 4, lig = threadIdx.x & 15;
    const float4* q4 = reinterpret_cast<const float4*>(q);
    for (int u = g; u < U; u += 16) {
      const uint4* xr = reinterpret_cast<const uint4*>(X16 + (long long)ids[u] * D);
      float acc = 0.f;
      for (int c = lig; c < nch; c += 16) acc = sq8(xr[c], q4[2 * c], q4[2 * c + 1], acc);
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
      if (lig == 0) d32[u] = acc;
    }
  }
}

// Calls ins(id) for every member of every bucket a query matched
// (cuLSH_SearchTables.cu:321-340).  Wave w takes the (table, probe) entries
// w*64 + lane + 256 i; the member ranges of kE such rounds are loaded
// together, then per round the wave prefix-sums the bucket sizes and spreads
// the member ids over its lanes, kR rounds of 64 members with their id loads
// in flight together (and in flight while the previous batch is inserted).
// qrq[t * tstride + p] is the member range (start,
// size; size 0 when nothing matched) of the query's probe p of table t.
// All threads of the workgroup must call it.
template <class Ins>
__device__ inline void for_each_member(const uint2* __restrict__ qrq, long long tstride, int L,
                                       int T, const uint32_t* __restrict__ sids, Ins ins) {
  constexpr int kE = 8, kR = 8;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int LT = L * T;
  // The ids of one batch of kR rounds are inserted after the next batch's
  // loads are issued, so consecutive batches' loads overlap the inserts.
  uint32_t pend[kR];
  bool has_pend = false;  // uniform over the wave
  for (int p0 = w * 64; p0 < LT; p0 += 256 * kE) {
    uint32_t st[kE], sz[kE];
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = p0 + 256 * i + lane;
      st[i] = 0;
      sz[i] = 0;
      if (e < LT) {
        const int t = e / T, p = e - t * T;
        const uint2 r = qrq[(long long)t * tstride + p];
        st[i] = r.x;
        sz[i] = r.y;
      }
    }
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      if (p0 + 256 * i >= LT) break;  // uniform over the wave
      uint32_t inc = sz[i];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t yv = __shfl_up(inc, off);
        if (lane >= off) inc += yv;
      }
      const uint32_t tot = __shfl(inc, 63);
      for (uint32_t b = 0; b < tot; b += 64 * kR) {
        uint32_t id[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          id[r] = kEmpty;
          if (b + 64 * r < tot) {  // uniform over the wave
            const uint32_t m = b + 64 * r + lane;
            int lo = 0, hi = 63;
            while (lo < hi) {  // first lane whose inclusive prefix exceeds m
              const int mid = (lo + hi) >> 1;
              const uint32_t im = __shfl(inc, mid);
              if (im > m)
                hi = mid;
              else
                lo = mid + 1;
            }
            const uint32_t st_l = __shfl(st[i], lo);
            const uint32_t inc_l = __shfl(inc, lo);
            const uint32_t sz_l = __shfl(sz[i], lo);
            if (m < tot) id[r] = sids[st_l + (m - (inc_l - sz_l))];
          }
        }
        if (has_pend) {
#pragma unroll
          for (int r = 0; r < kR; ++r)
            if (pend[r] != kEmpty) ins(pend[r]);
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) pend[r] = id[r];
        has_pend = true;
      }
    }
  }
  if (has_pend) {
#pragma unroll
    for (int r = 0; r < kR; ++r)
      if (pend[r] != kEmpty) ins(pend[r]);
  }
}

__device__ inline unsigned long long query_cnu(int q, int L,
                                               const unsigned long long* __restrict__ cnu_qt) {
  unsigned long long c = 0;
  for (int t = 0; t < L; ++t) c += cnu_qt[(long long)q * L + t];
  return c;
}

// Exact top-K of query q over its U candidates ids[0..U) given approximate
// distances dapx[0..U) with |dapx - d| <= rel * d + abs_err -- or, when
// e2 > 0, distances of rows x~ with ||x - x~||^2 <= e2 and
// |dapx - d~| <= rel * d~ + abs_err for the rows' own distance d~ (the FP16
// copy; the bound is completed below).  Selects the
// K-th smallest dapx (radix select over the float bits; distances are >= 0),
// re-evaluates in FP64 (the oracle's order) every candidate whose dapx can
// still be at or below the true K-th distance, and applies the reference's
// tie rule to the FP64 values.  Writes the K outputs oid[0..K), od[0..K).
// Returns false (uniformly) when more than scap candidates need the FP64
// pass.  All threads of the workgroup must call it.
__device__ inline bool finish_topk(const float* __restrict__ X, const float* y, int D, int K,
                                   int U, const float* dapx, const uint32_t* ids, double rel,
                                   double abs_err, double e2, double* sd, uint32_t* sid,
                                   unsigned scap,
                                   int* __restrict__ oid, float* __restrict__ od,
                                   unsigned long long* __restrict__ n_recheck) {
  __shared__ uint32_t s_hist[256];
  __shared__ unsigned s_scount, s_ovf, s_prefix, s_rem, s_misc[4];
  const int tid = threadIdx.x;
  __syncthreads();
  if (tid == 0) {
    s_scount = 0;
    s_ovf = 0;
  }
  double thr = INFINITY;
  if (U > K && K == 1) {
    // The smallest approximate distance: a block minimum of the float bits
    // (non-negative, so their integer order is their numeric order).
    uint32_t m = 0xFFFFFFFFu;
    for (int u = tid; u < U; u += 256) m = min(m, __float_as_uint(fmaxf(dapx[u], 0.f)));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off));
    if ((tid & 63) == 0) s_misc[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) s_prefix = min(min(s_misc[0], s_misc[1]), min(s_misc[2], s_misc[3]));
    __syncthreads();
  } else if (U > K) {
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
        const uint32_t bits = __float_as_uint(fmaxf(dapx[u], 0.f));
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
  }
  if (U > K) {
    const double dk = (double)__uint_as_float(s_prefix);
    if (e2 > 0.0) {
      // |sqrt(d~) - sqrt(d)| <= ||x - x~|| = e, so |d~ - d| <= 2 e sqrt(d) + e^2
      // <= delta d + e^2 (1 + 1 / delta) for every delta > 0; delta =
      // sqrt(e2 / dk) (at most 1/4) keeps the band near the K-th distance
      // narrow.  With |dapx - d~| <= rel d~ + abs_err this gives
      // |dapx - d| <= rel' d + abs' for the rel', abs' below.
      const double delta = dk > e2 * 16.0 ? sqrt(e2 / dk) : 0.25;
      const double ext = e2 * (1.0 + 1.0 / delta);
      abs_err = (abs_err + ext * (1.0 + rel)) * (1.0 + 0x1.0p-40);
      rel = (rel + delta * (1.0 + rel)) * (1.0 + 0x1.0p-40);
    }
    thr = ((dk + abs_err) * ((1.0 + rel) / (1.0 - rel)) + abs_err) * (1.0 + 1e-12) +
          (double)D * 0x1.0p-120;
  } else {
    __syncthreads();
  }

  // Candidates that can be at or below the true K-th distance: their ids
  // first, then their FP64 distances, one row per thread (so a wave's rows
  // are re-read together rather than one chain per qualifying lane).
  for (int u = tid; u < U; u += 256) {
    if ((double)fmaxf(dapx[u], 0.f) <= thr) {
      const unsigned i = atomicAdd(&s_scount, 1u);
      if (i < scap)
        sid[i] = ids[u];
      else
        s_ovf = 1;
    }
  }
  __syncthreads();
  if (s_ovf) return false;
  const int S = (int)s_scount;
  for (int i = tid; i < S; i += 256) sd[i] = dist64(y, X + (long long)sid[i] * D, D);
  __syncthreads();
  if (tid == 0 && S > 0) atomicAdd(n_recheck, (unsigned long long)S);
  // Candidates at or beyond FLT_MAX are never admitted by the reference's
  // heap (initialised to FLT_MAX, strict <, cuLSH_Kernels.cu:426).
  int P2 = 1;
  while (P2 < S) P2 <<= 1;
  for (int i = tid; i < P2; i += 256) {
    if (i >= S || sd[i] >= (double)FLT_MAX) {
      sd[i] = INFINITY;
      sid[i] = 0xFFFFFFFFu;
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
          const double a = sd[i], b = sd[ixj];
          const uint32_t ia = sid[i], ib = sid[ixj];
          const bool gt = (a > b) || (a == b && ia > ib);
          if (gt == upw) {
            sd[i] = b;
            sd[ixj] = a;
            sid[i] = ib;
            sid[ixj] = ia;
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
    if (sd[i] != INFINITY && (i + 1 == S || sd[i + 1] == INFINITY)) s_misc[0] = (unsigned)(i + 1);
  __syncthreads();
  const int A = (int)s_misc[0];
  int nout = A < K ? A : K;
  int Lt = nout, e0 = nout;
  if (U > K && A > K) {
    // Reference tie rule (closed form of the heap scan in descending id):
    // keep every candidate below d_K, then the r = K - |below| smallest ids
    // among the d_K-ties whose id is >= id*, where id* is the K-th largest
    // id among all candidates at or below d_K.
    const double dK = sd[K - 1];
    for (int i = tid; i < A; i += 256) {
      if (sd[i] == dK && (i == 0 || sd[i - 1] < dK)) s_misc[1] = (unsigned)i;
      if (sd[i] == dK && (i + 1 == A || sd[i + 1] > dK)) s_misc[2] = (unsigned)(i + 1);
    }
    __syncthreads();
    Lt = (int)s_misc[1];
    const int m0 = (int)s_misc[2];
    if (m0 > K) {
      for (int i = tid; i < m0; i += 256) {
        const uint32_t me = sid[i];
        int greater = 0;
        for (int j2 = 0; j2 < m0; ++j2) greater += sid[j2] > me;
        if (greater == K - 1) s_misc[3] = me;
      }
      __syncthreads();
      const uint32_t idstar = s_misc[3];
      for (int i = Lt + tid; i < m0; i += 256)
        if (sid[i] >= idstar && (i == Lt || sid[i - 1] < idstar)) s_misc[0] = (unsigned)i;
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
      id = (int)sid[src];
      dd = (float)sd[src];
    }
    oid[k] = id;
    od[k] = dd;
  }
  __syncthreads();
  return true;
}

// Where one query's working set lives (plain form).
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
  // Optional FP16 copy of the dataset rows ([N][D] IEEE half, D % 8 == 0)
  // with e2 >= max over rows of ||x - x16||^2: the distance pass then reads
  // the copy (half the bytes) and the FP64 pass the FP32 rows.
  const uint16_t* x16 = nullptr;
  double e2 = 0.0;
  // Optional [3] shader-clock totals of the query's phases: union (member
  // expansion + dedup), FP32 distances, top-K (selection, FP64 recheck,
  // sort, tie rule); one atomic add per phase per query, or null.
  unsigned long long* phase = nullptr;
};

// Where one query's inputs and outputs live.  `slot` indexes its C_nu row
// (cnu_qt[slot * L + t]) and its cnu_q / cu_q entries.
struct QueryIO {
  const float* y;       // the query vector (D floats)
  const uint2* qrq;     // member range of its probe 0 of table 0; probe p of table t at t * tstride + p
  long long tstride;
  int T, K;
  int* oid;             // its K result ids and distances
  float* od;
  long long slot;
};

__device__ inline QueryIO search_io(int q, const float* Qm, int Q, int D, int T, int K,
                                    const uint2* qr, int* out_ids, float* out_d) {
  QueryIO io;
  io.y = Qm + (long long)q * D;
  io.qrq = qr + (long long)q * T;
  io.tstride = (long long)Q * T;
  io.T = T;
  io.K = K;
  io.oid = out_ids + (long long)q * K;
  io.od = out_d + (long long)q * K;
  io.slot = q;
  return io;
}

// Plain form, one query.  Returns false (uniformly) when the query does not
// fit the buffers; its outputs are then not written.
template <int R = 4>
__device__ inline bool query_candidates(const QueryIO& io, const float* __restrict__ X, int D,
                                        int L, const uint32_t* __restrict__ sids,
                                        const unsigned long long* __restrict__ cnu_qt,
                                        unsigned long long* __restrict__ cnu_q,
                                        unsigned long long* __restrict__ cu_q,
                                        unsigned long long* __restrict__ n_recheck,
                                        const CandBufs& B) {
  __shared__ unsigned long long s_cnu;
  __shared__ unsigned s_count, s_ovf;
  const int tid = threadIdx.x;
  const float* y = io.y;
  const int K = io.K;

  __syncthreads();  // a previous query's readers of the shared counters are done
  const bool stamp = B.phase != nullptr && tid == 0;
  unsigned long long t_mark = stamp ? (unsigned long long)clock64() : 0ull;
  auto phase_end = [&](int i) {
    if (stamp) {
      const unsigned long long now = (unsigned long long)clock64();
      atomicAdd(&B.phase[i], now - t_mark);
      t_mark = now;
    }
  };
  if (tid < 64) {
    unsigned long long c = 0;
    for (int t = tid; t < L; t += 64) c += cnu_qt[io.slot * L + t];
    c = wave_sum_u64(c);
    if (tid == 0) {
      s_cnu = c;
      s_count = 0;
      s_ovf = 0;
    }
  }
  __syncthreads();
  const unsigned long long cnu = s_cnu;
  // Hash slots: 2^hbits >= 2 C_nu (at least 256, at most the buffer's), so a
  // light query clears only what it uses.
  int hbits = 8;
  while (hbits < B.hbits && (1ull << hbits) < 2 * cnu) ++hbits;
  const uint32_t hslots = 1u << hbits, hmask = hslots - 1;
  for (uint32_t i = tid; i < hslots; i += 256) B.hash[i] = kEmpty;
  __syncthreads();
  if (cnu > 0) {
    // Hash insert; the lanes that inserted a new id append it to the list
    // with one counter update per wave.
    for_each_member(io.qrq, io.tstride, L, io.T, sids, [&](uint32_t id) {
      uint32_t h = (id * 0x9E3779B1u) >> (32 - hbits);
      bool fresh = false;
      for (uint32_t probe = 0;; ++probe) {
        if (probe == hslots) {
          s_ovf = 1;
          break;
        }
        const uint32_t old = atomicCAS(&B.hash[h], kEmpty, id);
        if (old == kEmpty) {
          fresh = true;
          break;
        }
        if (old == id) break;
        h = (h + 1) & hmask;
      }
      const unsigned long long bm = __ballot(fresh);
      if (bm != 0ull) {
        const int lane = threadIdx.x & 63, leader = __ffsll((long long)bm) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&s_count, (unsigned)__popcll(bm));
        base = __shfl(base, leader);
        if (fresh) {
          const unsigned idx = base + (unsigned)__popcll(bm & ((1ull << lane) - 1ull));
          if (idx < B.ucap)
            B.ulist[idx] = id;
          else
            s_ovf = 1;
        }
      }
    });
  }
  __syncthreads();
  phase_end(0);
  if (s_ovf) return false;
  const int U = (int)s_count;

  // The query vector moves to staging (the hash slots are free now).
  const bool y_staged = B.ys != nullptr && D <= kYcap;
  if (y_staged)
    for (int d = tid; d < D; d += 256) B.ys[d] = y[d];
  __syncthreads();
  const bool half = B.x16 != nullptr && y_staged;
  if (half)  // half the bytes per row: half the rows per lane group for as many in flight
    l2_fp16_rows<(R > 1 ? R / 2 : 1)>(B.x16, B.ys, D, B.ulist, U, B.d32);
  else if (y_staged)
    l2_fp32_rows<R>(X, B.ys, D, B.ulist, U, B.d32);
  else
    l2_fp32_rows<R>(X, y, D, B.ulist, U, B.d32);
  if (B.phase) {  // uniform: every thread takes the same branch
    __syncthreads();
    phase_end(1);
  }
  // Direct FP32: |d32 - d| <= (D + 3) u d; margin 2 (Higham ch. 3).
  const double rel = 2.0 * (double)(D + 8) * kU32d;
  const bool done = finish_topk(X, y_staged ? B.ys : y, D, K, U, B.d32, B.ulist, rel, 0.0,
                                half ? B.e2 : 0.0, B.sd, B.sid, B.scap, io.oid, io.od,
                                n_recheck);
  phase_end(2);
  if (!done) return false;
  if (tid == 0) {
    cnu_q[io.slot] = cnu;
    cu_q[io.slot] = (unsigned long long)U;
  }
  return true;
}

// One workgroup per query, all buffers in LDS (~50 KB, so three workgroups
// per CU; the launch bound holds the registers to that).  After the union the
// hash slots hold d32 [kUcap], sd [kScap] (f64), sid [kScap] and the query.
// Queries that do not fit are appended to overflow[] for
// k_candidates_global.  With `order` (a locality order of the queries),
// XCD x (which receives workgroups x, x+8, ...) takes the x-th eighth of the
// order in sequence, so queries sharing candidate rows run on one L2.
__global__ __launch_bounds__(256, 3) void k_candidates(
    const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const uint2* __restrict__ qr, const uint32_t* __restrict__ sids,
    const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, int* __restrict__ overflow,
    int* __restrict__ n_overflow, unsigned long long* __restrict__ n_recheck,
    const uint32_t* __restrict__ order, unsigned long long* __restrict__ phase,
    const uint16_t* __restrict__ x16, double e2) {
  __shared__ __align__(16) uint32_t s_hash[kHcap];
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
  B.phase = phase;
  B.x16 = x16;
  B.e2 = e2;
  int q = blockIdx.x;
  if (order) {
    const int per = (int)(gridDim.x >> 3);  // grid is a multiple of 8
    const int pos = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (pos >= Q) return;
    q = (int)order[pos];
  }
  if (!query_candidates<4>(search_io(q, Qm, Q, D, T, K, qr, out_ids, out_d), X, D, L, sids,
                           cnu_qt, cnu_q, cu_q, n_recheck, B)) {
    if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1)] = q;
  }
}

// Largest summed bucket size (C_nu) among the listed queries.
__global__ __launch_bounds__(256) void k_max_cnu(const int* __restrict__ list, int n_list,
                                                 const unsigned long long* __restrict__ cnu_qt,
                                                 int L, unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_list) return;
  atomicMax(out, query_cnu(list[i], L, cnu_qt));
}

// The overflowed queries: workgroup b takes list[b], list[b + G], ...; its
// hash (2^hbits slots), id list and d32 (ucap each) live in global scratch
// at b * (2^hbits + 2 * ucap) words.  A query that still does not fit sets
// *err.
__global__ __launch_bounds__(256) void k_candidates_global(
    const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const uint2* __restrict__ qr, const uint32_t* __restrict__ sids,
    const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, const int* __restrict__ list, int n_list,
    uint32_t* __restrict__ scratch, int hbits, unsigned ucap,
    unsigned long long* __restrict__ n_recheck, int* __restrict__ err,
    unsigned long long* __restrict__ phase) {
  __shared__ double s_sd[kScapG];
  __shared__ uint32_t s_sid[kScapG];
  __shared__ __align__(16) float s_y[kYcap];
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
  B.phase = phase;
  for (int i = blockIdx.x; i < n_list; i += gridDim.x) {
    if (!query_candidates(search_io(list[i], Qm, Q, D, T, K, qr, out_ids, out_d), X, D, L, sids,
                          cnu_qt, cnu_q, cu_q, n_recheck, B)) {
      if (threadIdx.x == 0) atomicOr(err, 1);
    }
  }
}

// Fixed working sets of the asynchronous search's overflow kernel: workgroups
// and the largest hash / id list of one query.
constexpr int kOvAsyncGrid = 16;
constexpr int kOvAsyncHbits = 21;
constexpr unsigned kOvAsyncUcap = 1u << 20;

// The overflowed queries of an asynchronous search: like k_candidates_global,
// but the list length is read on the device and every workgroup owns a
// fixed region of 2^hbits_max + 2 ucap_max words, so no host round trip
// sizes it.  A query with more than ucap_max non-unique candidates, or more
// FP64 rechecks than kScapG, sets *err (sticky: the host clears it).
__global__ __launch_bounds__(256) void k_candidates_overflow(
    const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const uint2* __restrict__ qr, const uint32_t* __restrict__ sids,
    const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, const int* __restrict__ list,
    const int* __restrict__ n_list, uint32_t* __restrict__ scratch, int hbits_max,
    unsigned ucap_max, unsigned long long* __restrict__ n_recheck, int* __restrict__ err) {
  __shared__ double s_sd[kScapG];
  __shared__ uint32_t s_sid[kScapG];
  __shared__ __align__(16) float s_y[kYcap];
  const int n = *n_list;
  uint32_t* base = scratch + (size_t)blockIdx.x * ((1ull << hbits_max) + 2ull * ucap_max);
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int q = list[e];
    const unsigned long long cnu = query_cnu(q, L, cnu_qt);
    if (cnu > ucap_max) {
      if (threadIdx.x == 0) atomicOr(err, 1);
      continue;
    }
    int hbits = 13;
    while ((1ull << hbits) < 2 * cnu) ++hbits;
    CandBufs B;
    B.hash = base;
    B.hbits = hbits;
    B.ulist = base + (1u << hbits_max);
    B.ucap = ucap_max;
    B.d32 = reinterpret_cast<float*>(base + (1u << hbits_max) + ucap_max);
    B.sd = s_sd;
    B.sid = s_sid;
    B.scap = kScapG;
    B.ys = s_y;
    if (!query_candidates(search_io(q, Qm, Q, D, T, K, qr, out_ids, out_d), X, D, L, sids,
                          cnu_qt, cnu_q, cu_q, n_recheck, B)) {
      if (threadIdx.x == 0) atomicOr(err, 1);
    }
  }
}

// ---------------------------------------------------------------------------
// Precision form: grouped queries, norm-expansion distances on MFMA.
// ---------------------------------------------------------------------------
constexpr int kGroup = 16;   // queries per group (the MFMA tile's M)
constexpr int kGRows = 64;   // union rows per MFMA step (4 waves x 16)
constexpr int kGDc = 64;     // dimensions staged per MFMA step

// Coarse locality key of each query: floor((a_m . q - b_m) / (4 W)) of the
// first min(6, M) hashes of table 0, ten bits each.  Only orders queries so
// that queries sharing candidates are grouped; results do not depend on it.
__global__ __launch_bounds__(256) void k_locality_key(const float* __restrict__ Qm, int Q, int D,
                                                      int M, const float* __restrict__ Aorig,
                                                      const float* __restrict__ bj, float W,
                                                      uint64_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= Q) return;
  const float* y = Qm + (long long)q * D;
  uint64_t key = 0;
  const int mm = M < 6 ? M : 6;
  for (int m = 0; m < 6; ++m) {
    int c = 0;
    if (m < mm) {
      const float* a = Aorig + (long long)m * D;
      float s = 0.f;
      for (int d = 0; d < D; ++d) s = fmaf(y[d], a[d], s);
      c = (int)floorf((s - bj[m]) / (4.f * W));
    }
    key = (key << 10) | (uint64_t)((c + 512) & 1023);
  }
  keys[q] = key;
  vals[q] = (uint32_t)q;
}

// Sum of C_nu over each group of kGroup consecutive queries of `order`; the
// largest goes to *out.
__global__ __launch_bounds__(256) void k_group_cnu(const uint32_t* __restrict__ order, int Q, int L,
                                                   const unsigned long long* __restrict__ cnu_qt,
                                                   unsigned long long* __restrict__ out) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const int ng = (Q + kGroup - 1) / kGroup;
  if (g >= ng) return;
  unsigned long long s = 0;
  for (int i = g * kGroup; i < min(Q, (g + 1) * kGroup); ++i) s += query_cnu(order[i], L, cnu_qt);
  atomicMax(out, s);
}

// One workgroup per group of up to 16 queries (taken in `order`), groups
// strided over the grid.  Per-workgroup global scratch of `region` words:
//   hash ids [2^hbits], complemented hash masks [2^hbits] (bit g clear =
//   query g has the id), union slots [qcap], per-query approximate distances
//   [qcap] and ids [qcap], union ids [qcap] and masks [qcap]
// where qcap >= the group's summed C_nu.  The scratch starts as all-ones
// bytes (empty ids, empty masks) and each group restores the slots it used,
// so no per-group clearing pass is needed.  Each union row is read once per
// d-chunk and multiplied against the group's queries on
// v_mfma_f32_16x16x4_f32:  d~ = |q|^2 + |x|^2 - 2 q.x  (FP32), with
// |d~ - d| <= 2 (D+4) u (|q| + xmax)^2 (Higham ch. 3, margin 2).  Then each
// query runs finish_topk on its own candidates.  Queries that need more
// FP64 rechecks than the LDS holds go to overflow[].
__global__ __launch_bounds__(256, 3) void k_group_candidates(
    const float* __restrict__ X, const float* __restrict__ xn2, float xmax,
    const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const uint2* __restrict__ qr, const uint32_t* __restrict__ sids,
    const unsigned long long* __restrict__ cnu_qt,
    const uint32_t* __restrict__ order, int* __restrict__ out_ids, float* __restrict__ out_d,
    unsigned long long* __restrict__ cnu_q, unsigned long long* __restrict__ cu_q,
    uint32_t* __restrict__ scratch, size_t region, int hbits, unsigned qcap,
    int* __restrict__ overflow, int* __restrict__ n_overflow,
    unsigned long long* __restrict__ n_recheck) {
  __shared__ float s_q[kGroup][kGDc + 4];
  __shared__ float s_x[kGRows][kGDc + 4];
  __shared__ uint32_t s_rid[kGRows];
  __shared__ double s_sd[kScap];
  __shared__ uint32_t s_sid[kScap];
  __shared__ float s_qn[kGroup], s_qnorm[kGroup];
  __shared__ int s_qid[kGroup];
  __shared__ unsigned s_qoff[kGroup], s_qcnt[kGroup], s_ucount, s_ovf;
  __shared__ unsigned long long s_cnu[kGroup];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t hslots = 1u << hbits, hmask = hslots - 1;
  uint32_t* hid = scratch + (size_t)blockIdx.x * region;
  uint32_t* hmk = hid + hslots;
  uint32_t* uslot = hmk + hslots;
  float* qd = reinterpret_cast<float*>(uslot + qcap);
  uint32_t* qi = uslot + 2ull * qcap;
  uint32_t* uid = uslot + 3ull * qcap;  // union ids, resolved once
  uint32_t* umk = uslot + 4ull * qcap;  // their membership masks
  const bool vec = (D & 3) == 0;
  const int ng = (Q + kGroup - 1) / kGroup;

  for (int grp = blockIdx.x; grp < ng; grp += gridDim.x) {
    const int g0 = grp * kGroup;
    const int gs = min(kGroup, Q - g0);
    __syncthreads();
    if (tid < kGroup) {
      s_qid[tid] = tid < gs ? (int)order[g0 + tid] : -1;
      s_cnu[tid] = tid < gs ? query_cnu((int)order[g0 + tid], L, cnu_qt) : 0ull;
      s_qcnt[tid] = 0;
    }
    if (tid == 0) {
      s_ucount = 0;
      s_ovf = 0;
    }
    __syncthreads();
    if (tid == 0) {
      unsigned off = 0;
      for (int g = 0; g < kGroup; ++g) {
        s_qoff[g] = off;
        off += (unsigned)s_cnu[g];
      }
    }
    // Query norms (FP32) for the expansion and its error bound.
    if (w == 0 && lane < kGroup) {
      float s = 0.f;
      if (lane < gs) {
        const float* y = Qm + (long long)s_qid[lane] * D;
        for (int d = 0; d < D; ++d) s = fmaf(y[d], y[d], s);
      }
      s_qn[lane] = s;
      s_qnorm[lane] = sqrtf(s);
    }
    // Union of the group's candidates; hash slot masks record membership.
    for (int g = 0; g < gs; ++g) {
      const int q = s_qid[g];
      for_each_member(qr + (long long)q * T, (long long)Q * T, L, T, sids, [&](uint32_t id) {
        uint32_t h = (id * 0x9E3779B1u) >> (32 - hbits);
        for (uint32_t probe = 0;; ++probe) {
          if (probe == hslots) {
            s_ovf = 1;
            break;
          }
          const uint32_t old = atomicCAS(&hid[h], kEmpty, id);
          if (old == kEmpty || old == id) {
            if (old == kEmpty) uslot[atomicAdd(&s_ucount, 1u)] = h;
            const uint32_t om = atomicAnd(&hmk[h], ~(1u << g));
            if (om & (1u << g)) atomicAdd(&s_qcnt[g], 1u);
            break;
          }
          h = (h + 1) & hmask;
        }
      });
    }
    __syncthreads();
    const int UG = (int)s_ucount;
    if (s_ovf) {  // cannot happen with qcap >= summed C_nu; fail safe
      for (int r = tid; r < UG; r += 256) {
        hid[uslot[r]] = kEmpty;
        hmk[uslot[r]] = 0xFFFFFFFFu;
      }
      if (tid < gs) overflow[atomicAdd(n_overflow, 1)] = s_qid[tid];
      continue;
    }
    unsigned cu[kGroup];
#pragma unroll
    for (int g = 0; g < kGroup; ++g) cu[g] = s_qcnt[g];
    for (int r = tid; r < UG; r += 256) {
      const uint32_t slot = uslot[r];
      uid[r] = hid[slot];
      umk[r] = ~hmk[slot];
    }
    __syncthreads();
    if (tid < kGroup) s_qcnt[tid] = 0;
    __syncthreads();

    // Union rows in chunks of 64: q.x on MFMA, then d~ for member pairs.
    for (int r0 = 0; r0 < UG; r0 += kGRows) {
      if (tid < kGRows) s_rid[tid] = r0 + tid < UG ? uid[r0 + tid] : uid[r0];
      __syncthreads();
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      for (int dc = 0; dc < D; dc += kGDc) {
        if (vec) {
          // 16 float4 per row chunk, 4 per thread, loads issued together.
          for (int e = tid; e < kGRows * (kGDc / 4); e += 256) {
            const int r = e >> 4, c4 = e & 15;
            const int d = dc + 4 * c4;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (d < D) v = *reinterpret_cast<const float4*>(X + (long long)s_rid[r] * D + d);
            *reinterpret_cast<float4*>(&s_x[r][4 * c4]) = v;
          }
        } else {
          for (int e = tid; e < kGRows * kGDc; e += 256) {
            const int r = e / kGDc, c = e % kGDc;
            const int d = dc + c;
            s_x[r][c] = d < D ? X[(long long)s_rid[r] * D + d] : 0.f;
          }
        }
        for (int e = tid; e < kGroup * kGDc; e += 256) {
          const int g = e / kGDc, c = e % kGDc;
          const int d = dc + c;
          s_q[g][c] = (g < gs && d < D) ? Qm[(long long)s_qid[g] * D + d] : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < kGDc; kk += 4) {
          const float a = s_q[lane & 15][kk + (lane >> 4)];
          const float b = s_x[w * 16 + (lane & 15)][kk + (lane >> 4)];
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
      }
      // lane: union row r0 + 16w + (lane & 15), queries 4 * (lane >> 4) + i
      const int r = r0 + w * 16 + (lane & 15);
      if (r < UG) {
        const uint32_t id = s_rid[w * 16 + (lane & 15)], mk = umk[r];
        const float xq = xn2[id];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int g = (lane >> 4) * 4 + i;
          if (g < gs && ((mk >> g) & 1u)) {
            const float dap = s_qn[g] + xq - 2.f * acc[i];
            const unsigned pos = s_qoff[g] + atomicAdd(&s_qcnt[g], 1u);
            qd[pos] = dap;
            qi[pos] = id;
          }
        }
      }
    }
    __syncthreads();

    for (int g = 0; g < gs; ++g) {
      const int q = s_qid[g];
      const double qb2 = (double)s_qnorm[g] + (double)xmax;
      const double abs_err = 2.0 * (double)(D + 4) * kU32d * qb2 * qb2 * (1.0 + 1e-6);
      const bool ok = finish_topk(X, Qm + (long long)q * D, D, K, (int)cu[g], qd + s_qoff[g],
                                  qi + s_qoff[g], 0.0, abs_err, 0.0, s_sd, s_sid, kScap,
                                  out_ids + (long long)q * K, out_d + (long long)q * K,
                                  n_recheck);
      if (tid == 0) {
        if (ok) {
          cnu_q[q] = s_cnu[g];
          cu_q[q] = cu[g];
        } else {
          overflow[atomicAdd(n_overflow, 1)] = q;
        }
      }
    }
    // Restore the used hash slots for the next group.
    for (int r = tid; r < UG; r += 256) {
      hid[uslot[r]] = kEmpty;
      hmk[uslot[r]] = 0xFFFFFFFFu;
    }
  }
}

}  // namespace culsh
