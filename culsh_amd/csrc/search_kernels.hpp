// Query-side kernels.
//
// k_query_probe  one wave per (query, table): FP64 projection, deltas,
//                delta order, the multiprobe queue, probe keys, filtered
//                directory lookup with a full-code verify.  Restates
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
#include "candidates.hpp"

namespace culsh {

// One wave per (query, table).  J = ceil(T / 64) queue slots per lane.
// qb[t][q*T + p] receives the global bucket index of probe p (or -1) when
// qb is given, qr[t][q*T + p] its member range (start, size; (0, 0) when
// nothing matched); cnu_qt[q*L + t] the summed sizes of the matched buckets.
template <int J>
__global__ __launch_bounds__(256) void k_query_probe(
    const float* __restrict__ Qm, int Q, int D, int L, int M, int T,
    const double* __restrict__ At64, const double* __restrict__ b64, double W,
    const uint64_t* __restrict__ seeds, ProbeIndex pix, int* __restrict__ qb,
    uint2* __restrict__ qr, unsigned long long* __restrict__ cnu_qt) {
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
  const uint2 tfw = pix.tfw[t];
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
      const uint64_t key = make_key(hs, t, pix.tb);
      const uint64_t fb = filter_bits(key);
      long long k = -1;
      uint2 range = make_uint2(0u, 0u);
      if ((pix.filter[filter_word(key, pix.tb, tfw)] & fb) == fb)
        k = lookup_unverified(key, pix.dir[key >> pix.shift], pix.dir, pix.shift, pix.recs);
      if (k >= 0) {
        const unsigned char* line = bucket_line(pix, (uint32_t)k);
        const int16_t* bc = line_codes(line);
        range = line_range(line);
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
      if (k < 0) range = make_uint2(0u, 0u);
      if (qb) qb[qbBase + p] = (int)k;
      qr[qbBase + p] = range;
      sizesum += range.y;
    }
  }
  sizesum = wave_sum_u64(sizesum);
  if (lane == 0) cnu_qt[(long long)q * L + t] = sizesum;
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
