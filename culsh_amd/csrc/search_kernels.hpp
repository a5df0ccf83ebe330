// Query-side kernels.
//
// k_query_probe  one wave per (query, table): FP64 projection, deltas,
//                delta order, the multiprobe queue, probe keys, directory +
//                binary-search lookup with a full-code verify.  Restates
//                cuLSH_Querying.cu:171-225 (findTableQueryBins),
//                Querying.cu:63-119 + Kernels.cu:188-289 (multiprobe) and
//                Querying.cu:124-165 + Kernels.cu:124-138 (lookup).
// k_candidates   one workgroup per query: union of matched buckets with
//                dedup in an LDS hash set, direct FP32 squared L2, radix
//                select of the K-th distance, FP64 re-evaluation of every
//                candidate within the FP32 error bound of it, and the
//                reference's top-K tie rule.  Restates
//                cuLSH_SearchTables.cu:249-255,321-389 (candidates),
//                Kernels.cu:301-348 (distances), Kernels.cu:356-456 (top-K).
#pragma once
#include "common.hpp"
#include "delta_order.hpp"
#include "probe8.hpp"
#include "sort.hpp"

namespace culsh {

// One wave per (query, table).  J = ceil(T / 64) queue slots per lane.
// qb[t][q*T + p] receives the global bucket index of probe p (or -1);
// cnu_qt[q*L + t] the summed sizes of the matched buckets.
template <int J>
__global__ __launch_bounds__(256) void k_query_probe(
    const float* __restrict__ Qm, int Q, int D, int L, int M, int T,
    const double* __restrict__ At64, const double* __restrict__ b64, double W,
    const uint64_t* __restrict__ seeds, int tb, int shift, const uint64_t* __restrict__ dir,
    const BucketRec* __restrict__ recs,
    const int16_t* __restrict__ bcodes, int* __restrict__ qb,
    unsigned long long* __restrict__ cnu_qt) {
  __shared__ double s_ds[4][64];
  __shared__ double s_d[4][128];
  __shared__ int s_perm[4][64];
  __shared__ int s_base[4][64];
  __shared__ unsigned char s_rank[4][128];
  const int lane = threadIdx.x & 63;
  const int w = uniform_int(threadIdx.x >> 6);
  const long long pair = (long long)blockIdx.x * 4 + w;
  if (pair >= (long long)Q * L) return;
  const int q = (int)(pair / L);
  const int t = (int)(pair - (long long)q * L);
  double* ds = s_ds[w];
  int* perm = s_perm[w];
  int* sbase = s_base[w];
  unsigned char* rank = s_rank[w];

  // Projection in FP64 (the oracle's order), lane m < M.
  const float* y = Qm + (long long)q * D;
  double v = 0.0;
  if (lane < M) {
    const double* a = At64 + (long long)t * D * M + lane;
    double acc = 0.0;
#pragma unroll 4
    for (int d = 0; d < D; ++d) acc = fma((double)y[d], a[(long long)d * M], acc);
    v = (acc - b64[t * M + lane]) / W;
  }
  const double fl = floor(v);
  const int base = (int)fl;
  if (lane < M) sbase[lane] = base;
  rank[lane] = 255;
  rank[lane + 64] = 255;
  const uint64_t seed = seeds[t];
  const uint64_t basesum = wave_sum_u64(lane < M ? code_term(seed, lane, base) : 0ull);
  const int maxT = (2 * M < 64) ? 2 * M : 64;
  wave_lds_sync();

  if (T > 1) {
    // Deltas (cuLSH_Kernels.cu:280-287): index m = "-1 on m", M+m = "+1 on m".
    const double dn = v - fl;
    const double up = 1.0 - dn;
    int rdn = 0, rup = 0;
    bool tie = false;
    for (int j = 0; j < 2 * M; ++j) {
      const int src = j < M ? j : j - M;
      const double dj = shfl_f64(j < M ? dn : up, src);
      if (lane < M) {
        rdn += dj < dn;
        rup += dj < up;
        tie |= (dj == dn && j != lane) || (dj == up && j != M + lane);
      }
    }
    if (__ballot(tie) == 0ull) {
      // Distinct deltas: the ascending order is unique; rank = #smaller.
      if (lane < M) {
        if (rdn < maxT) {
          perm[rdn] = lane;
          ds[rdn] = dn;
          rank[lane] = (unsigned char)rdn;
        }
        if (rup < maxT) {
          perm[rup] = M + lane;
          ds[rup] = up;
          rank[M + lane] = (unsigned char)rup;
        }
      }
    } else {
      if (lane < M) {
        s_d[w][lane] = dn;
        s_d[w][M + lane] = up;
      }
      wave_lds_sync();
      if (lane == 0) heap_order_deltas(s_d[w], 2 * M, maxT, perm);
      wave_lds_sync();
      if (lane < maxT) {
        const int pi = perm[lane];
        ds[lane] = s_d[w][pi];
        rank[pi] = (unsigned char)lane;
      }
    }
    wave_lds_sync();
  }

  // Multiprobe queue (cuLSH_Kernels.cu:216-255): slot s lives in lane s%64,
  // register s/64.  pm holds the perturbation set of probe p likewise.
  uint64_t pm[J], qc[J];
  double qs[J];
#pragma unroll
  for (int jj = 0; jj < J; ++jj) {
    pm[jj] = 0ull;
    qc[jj] = 0ull;
    qs[jj] = DBL_MAX;
  }
  if (T > 1) {
    if (lane == 0) {
      qc[0] = 1ull;
      qs[0] = ds[0];
    }
    int size = 1;
    for (int p = 1; p < T; ++p) {
      double bs = DBL_MAX;
      int bi = 0x7fffffff;
#pragma unroll
      for (int jj = 0; jj < J; ++jj) {
        const int s = jj * 64 + lane;
        if (s < size && qs[jj] < bs) {
          bs = qs[jj];
          bi = s;
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double os = shfl_xor_f64(bs, off);
        const int oi = __shfl_xor(bi, off);
        if (os < bs || (os == bs && oi < bi)) {
          bs = os;
          bi = oi;
        }
      }
      const int bj = bi >> 6, bl = bi & 63;
      uint64_t mine = 0ull;
#pragma unroll
      for (int jj = 0; jj < J; ++jj)
        if (jj == bj) mine = qc[jj];
      const uint64_t probe = shfl_u64(mine, bl);
      if (lane == (p & 63)) {
#pragma unroll
        for (int jj = 0; jj < J; ++jj)
          if (jj == (p >> 6)) pm[jj] = probe;
      }
      const int maxPos = 63 - __clzll((long long)probe);
      if (maxPos < maxT - 1) {
        const double e = bs + ds[maxPos + 1];
        const uint64_t pe = probe | (1ull << (maxPos + 1));
        const double sh = e - ds[maxPos];
        const uint64_t ps = pe - (1ull << maxPos);
        if (lane == bl) {
#pragma unroll
          for (int jj = 0; jj < J; ++jj)
            if (jj == bj) {
              qc[jj] = pe;
              qs[jj] = e;
            }
        }
        if (lane == (size & 63)) {
#pragma unroll
          for (int jj = 0; jj < J; ++jj)
            if (jj == (size >> 6)) {
              qc[jj] = ps;
              qs[jj] = sh;
            }
        }
        ++size;
      }
    }
  }

  // Probe keys, lookup, verify.
  unsigned long long sizesum = 0;
  const long long qbBase = (long long)t * Q * T + (long long)q * T;
#pragma unroll
  for (int jj = 0; jj < J; ++jj) {
    const int p = jj * 64 + lane;
    if (p < T) {
      const uint64_t mask = pm[jj];
      uint64_t hs = basesum;
      uint64_t mm = mask;
      while (mm) {
        const int i = __ffsll((unsigned long long)mm) - 1;
        mm &= mm - 1;
        const int di = perm[i];
        const int dim = di < M ? di : di - M;
        const int partner = di < M ? di + M : di - M;
        const int pr = rank[partner];
        if (pr != 255 && ((mask >> pr) & 1ull)) continue;  // -1 and +1 on one dimension
        const int b0 = sbase[dim];
        hs += code_term(seed, dim, b0 + (di >= M ? 1 : -1)) - code_term(seed, dim, b0);
      }
      long long k = lookup_key(make_key(hs, t, tb), dir, shift, recs);
      if (k >= 0) {
        const int16_t* bc = bcodes + k * M;
        for (int m = 0; m < M; ++m) {
          const int rd = rank[m], ru = rank[M + m];
          int c = sbase[m];
          if (rd != 255 && ((mask >> rd) & 1ull)) c -= 1;
          if (ru != 255 && ((mask >> ru) & 1ull)) c += 1;
          if (c != (int)bc[m]) {
            k = -1;
            break;
          }
        }
      }
      qb[qbBase + p] = (int)k;
      if (k >= 0) sizesum += recs[k].size;
    }
  }
  sizesum = wave_sum_u64(sizesum);
  if (lane == 0) cnu_qt[(long long)q * L + t] = sizesum;
}

constexpr int kHcap = 8192;  // LDS hash slots per query
constexpr int kUcap = 4096;  // unique candidates per query (LDS path)
constexpr int kScap = 1024;  // candidates re-evaluated in FP64 per query
constexpr int kYcap = 1024;  // query dimensions staged in LDS
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

// One workgroup (256 threads) per query.  rel_bound bounds the relative
// error of the FP32 distance.  Queries whose unique candidates exceed the
// LDS capacity are appended to overflow[] (count in *n_overflow) and left
// for the caller.
__global__ __launch_bounds__(256) void k_candidates(
    const float* __restrict__ X, const float* __restrict__ Qm, int Q, int D, int L, int T, int K,
    const int* __restrict__ qb, const BucketRec* __restrict__ recs,
    const uint32_t* __restrict__ sids, const unsigned long long* __restrict__ cnu_qt,
    int* __restrict__ out_ids, float* __restrict__ out_d, unsigned long long* __restrict__ cnu_q,
    unsigned long long* __restrict__ cu_q, int* __restrict__ overflow,
    int* __restrict__ n_overflow, double rel_bound, unsigned long long* __restrict__ n_recheck) {
  __shared__ uint32_t s_hash[kHcap];  // hash set; then d32 [kUcap] + S arrays
  __shared__ uint32_t s_u[kUcap];
  __shared__ uint32_t s_hist[256];
  __shared__ unsigned long long s_cnu;
  __shared__ unsigned s_count, s_ovf, s_scount, s_prefix, s_rem, s_misc[4];
  float* s_d32 = reinterpret_cast<float*>(s_hash);
  double* s_sd = reinterpret_cast<double*>(s_hash + kUcap);            // kScap doubles
  uint32_t* s_sid = s_hash + kUcap + 2 * kScap;                         // kScap ids
  float* s_y = reinterpret_cast<float*>(s_hash + kUcap + 3 * kScap);     // kYcap floats
  static_assert(kUcap + 3 * kScap + kYcap <= kHcap, "post-insertion LDS layout");
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int q = blockIdx.x;
  const float* y = Qm + (long long)q * D;

  if (tid == 0) {
    unsigned long long c = 0;
    for (int t = 0; t < L; ++t) c += cnu_qt[(long long)q * L + t];
    s_cnu = c;
    s_count = 0;
    s_ovf = 0;
    s_scount = 0;
  }
  for (int i = tid; i < kHcap; i += 256) s_hash[i] = kEmpty;
  __syncthreads();
  const unsigned long long cnu = s_cnu;
  if (tid == 0) cnu_q[q] = cnu;

  // Union of matched buckets (cuLSH_SearchTables.cu:321-364) with dedup.
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
          uint32_t h = (id * 0x9E3779B1u) >> (32 - 13);
          for (int probe = 0; probe < kHcap; ++probe) {
            const uint32_t old = atomicCAS(&s_hash[h], kEmpty, id);
            if (old == kEmpty) {
              const unsigned idx = atomicAdd(&s_count, 1u);
              if (idx < (unsigned)kUcap)
                s_u[idx] = id;
              else
                s_ovf = 1;
              break;
            }
            if (old == id) break;
            h = (h + 1) & (kHcap - 1);
            if (probe == kHcap - 1) s_ovf = 1;
          }
        }
      }
    }
  }
  __syncthreads();
  if (s_ovf) {
    if (tid == 0) overflow[atomicAdd(n_overflow, 1)] = q;
    return;
  }
  const int U = (int)s_count;
  if (tid == 0) cu_q[q] = (unsigned long long)U;

  // The query vector moves to LDS (the hash slots are free now).
  const bool y_in_lds = D <= kYcap;
  if (y_in_lds)
    for (int d = tid; d < D; d += 256) s_y[d] = y[d];
  __syncthreads();
  if (y_in_lds)
    l2_fp32_rows(X, s_y, D, s_u, U, s_d32);
  else
    l2_fp32_rows(X, y, D, s_u, U, s_d32);
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
        const uint32_t bits = __float_as_uint(s_d32[u]);
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
    if ((double)s_d32[u] <= thr) {
      const unsigned i = atomicAdd(&s_scount, 1u);
      if (i < (unsigned)kScap) {
        const uint32_t id = s_u[u];
        s_sd[i] = y_in_lds ? dist64(s_y, X + (long long)id * D, D)
                           : dist64(y, X + (long long)id * D, D);
        s_sid[i] = id;
      } else {
        s_ovf = 1;
      }
    }
  }
  __syncthreads();
  if (s_ovf) {
    if (tid == 0) overflow[atomicAdd(n_overflow, 1)] = q;
    return;
  }
  const int S = (int)s_scount;
  if (tid == 0 && S > 0) atomicAdd(n_recheck, (unsigned long long)S);
  // Candidates at or beyond FLT_MAX are never admitted by the reference's
  // heap (initialised to FLT_MAX, strict <, cuLSH_Kernels.cu:426).
  int P2 = 1;
  while (P2 < S) P2 <<= 1;
  for (int i = tid; i < P2; i += 256) {
    if (i >= S || s_sd[i] >= (double)FLT_MAX) {
      s_sd[i] = INFINITY;
      s_sid[i] = 0xFFFFFFFFu;
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
          const double a = s_sd[i], b = s_sd[ixj];
          const uint32_t ia = s_sid[i], ib = s_sid[ixj];
          const bool gt = (a > b) || (a == b && ia > ib);
          if (gt == upw) {
            s_sd[i] = b;
            s_sd[ixj] = a;
            s_sid[i] = ib;
            s_sid[ixj] = ia;
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
    if (s_sd[i] != INFINITY && (i + 1 == S || s_sd[i + 1] == INFINITY)) s_misc[0] = (unsigned)(i + 1);
  __syncthreads();
  const int A = (int)s_misc[0];
  int nout = A < K ? A : K;
  int Lt = nout, e0 = nout;
  if (U > K && A > K) {
    // Reference tie rule (closed form of the heap scan in descending id):
    // keep every candidate below d_K, then the r = K - |below| smallest ids
    // among the d_K-ties whose id is >= id*, where id* is the K-th largest
    // id among all candidates at or below d_K.
    const double dK = s_sd[K - 1];
    for (int i = tid; i < A; i += 256) {
      if (s_sd[i] == dK && (i == 0 || s_sd[i - 1] < dK)) s_misc[1] = (unsigned)i;
      if (s_sd[i] == dK && (i + 1 == A || s_sd[i + 1] > dK)) s_misc[2] = (unsigned)(i + 1);
    }
    __syncthreads();
    Lt = (int)s_misc[1];
    const int m0 = (int)s_misc[2];
    if (m0 > K) {
      for (int i = tid; i < m0; i += 256) {
        const uint32_t me = s_sid[i];
        int greater = 0;
        for (int j2 = 0; j2 < m0; ++j2) greater += s_sid[j2] > me;
        if (greater == K - 1) s_misc[3] = me;
      }
      __syncthreads();
      const uint32_t idstar = s_misc[3];
      for (int i = Lt + tid; i < m0; i += 256)
        if (s_sid[i] >= idstar && (i == Lt || s_sid[i - 1] < idstar)) s_misc[0] = (unsigned)i;
      __syncthreads();
      e0 = (int)s_misc[0];
    } else {
      e0 = Lt;
    }
    nout = K;
  } else {
    Lt = nout;
    e0 = nout;
  }
  for (int k = tid; k < K; k += 256) {
    int id = -1;
    float dd = FLT_MAX;
    if (k < nout) {
      const int src = k < Lt ? k : e0 + (k - Lt);
      id = (int)s_sid[src];
      dd = (float)s_sd[src];
    }
    out_ids[(long long)q * K + k] = id;
    out_d[(long long)q * K + k] = dd;
  }
}

// Converts global bucket indices of qb to table-local ones through
// `renumber` (global bucket -> table-local lexicographic index).
__global__ __launch_bounds__(256) void k_local_buckets(const int* __restrict__ qb, long long QT,
                                                       int L, const int* __restrict__ renumber,
                                                       int* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= QT * L) return;
  const int k = qb[i];
  out[i] = k < 0 ? -1 : renumber[k];
}

}  // namespace culsh
