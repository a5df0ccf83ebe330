// k_query_probe8: the query front of the search with one octet (8 lanes) per
// (query, table) and 32 queries of one table per workgroup.
//
// Same results as k_query_probe (one wave per pair) -- FP64 projection in the
// oracle's order, deltas, delta order, the reference's multiprobe queue,
// probe keys, filtered directory lookup and full-code verify
// (cuLSH_Querying.cu:63-225, cuLSH_Kernels.cu:124-289) -- but eight pairs
// share a wave and the queue's cross-lane steps are DPP moves inside the
// octet.
#pragma once
#include "common.hpp"
#include "delta_order.hpp"

#include <type_traits>

namespace culsh {

// Per-octet LDS scratch (bytes): ds[2M] f64 (deltas, then sorted deltas),
// perm[64] i8, rank[128] u8, base[M] i32.
__host__ __device__ inline int probe8_octet_bytes(int M) {
  return (16 * M + 64 + 128 + 4 * M + 15) & ~15;
}
// Dimensions per lane of the projection: ceil(M / 8) rounded up to one of
// the instantiated widths 1, 2, 3, 4, 8.
__host__ __device__ inline int probe8_mi(int M) {
  const int m = (M + 7) / 8;
  return m <= 4 ? m : 8;
}
// Dynamic LDS of one workgroup: 32 octet scratches.  Query vectors are read
// through L1 (the 8 lanes of an octet read the same address): staging them
// in LDS would cost more occupancy than it saves.
__host__ __device__ inline int probe8_lds_bytes(int D, int M) {
  (void)D;
  return 32 * probe8_octet_bytes(M);
}

// Does the probe code of `mask` (base code plus the mask's +-1 moves) equal
// the bucket code bc?  Compares the dimensions where the bucket differs from
// the base code with the dimensions the mask moves.
__device__ inline bool probe_matches(const int16_t* __restrict__ bc, const int* base,
                                     const signed char* perm, const unsigned char* rank, int M,
                                     uint64_t mask) {
  int nz = 0;
  for (int m = 0; m < M; ++m) nz += (int)bc[m] != base[m];
  int matched = 0;
  uint64_t mm = mask;
  while (mm) {
    const int i = __ffsll((unsigned long long)mm) - 1;
    mm &= mm - 1;
    const int di = perm[i];
    const int dim = di < M ? di : di - M;
    const int pr = rank[di < M ? di + M : di - M];
    if (pr != 255 && ((mask >> pr) & 1ull)) continue;  // -1 and +1 cancel
    if ((int)bc[dim] - base[dim] != (di >= M ? 1 : -1)) return false;
    ++matched;
  }
  return matched == nz;
}

// Key sum of probe `mask`: the base key sum with the moved dimensions'
// terms replaced; dt[di] is the change of the sum when delta di is applied
// (di = m: -1 on m, M + m: +1 on m).
__device__ inline uint64_t probe_hash(uint64_t basesum, const uint64_t* dt,
                                      const signed char* perm, const unsigned char* rank, int M,
                                      uint64_t mask) {
  uint64_t hs = basesum;
  uint64_t mm = mask;
  while (mm) {
    const int i = __ffsll((unsigned long long)mm) - 1;
    mm &= mm - 1;
    const int di = perm[i];
    const int pr = rank[di < M ? di + M : di - M];
    if (pr != 255 && ((mask >> pr) & 1ull)) continue;
    hs += dt[di];
  }
  return hs;
}

// One (query, table) pair per octet: FP64 projection of y (the oracle's
// order), deltas, delta order, the multiprobe queue, keys, lookup and verify.
// J = ceil(T / 8) queue slots per lane (slot s: lane s%8, register s/8);
// MI >= ceil(M / 8) dimensions per lane (lane l owns l + 8i; instantiated
// for MI = 1, 2, 3, 4, 8), a compile-time
// bound so that the projection loop is branch-free.  acol points at the
// table's column 0 of an f32 projection matrix with row stride astride
// (column m of row d at acol[d * astride + m]) and at least 8*MI readable
// columns per row (lanes past M read them and discard them): the [D][Jp]
// matrix in global memory, or the table's slice staged in LDS.  The FP64
// product of f32 operands is exact, so widening on read is the oracle's
// arithmetic.  orb is the octet's
// probe8_octet_bytes(M) of LDS.  Writes qrp[p] (member range of probe p's
// bucket: start, size; (0, 0) when nothing matched), qbp[p] (its bucket, or
// -1) when qbp is given, for p < T, and *cnu (summed sizes of the matched
// buckets) when active, and
// the query's 12-bit coarse-cell key to *cell_key when cell_key is given.
// Every lane of the octet calls it; T is uniform over the wave.
template <int J, int MI>
__device__ __forceinline__ void probe8_pair(const float* y, bool active, int t, int D, int M, int T,
                                   const float* __restrict__ acol0, int astride,
                                   const double* __restrict__ b64, double W,
                                   const uint64_t* __restrict__ seeds, const ProbeIndex& pix,
                                   unsigned char* orb,
                                   int* __restrict__ qbp, uint2* __restrict__ qrp,
                                   unsigned long long* __restrict__ cnu,
                                   unsigned* __restrict__ cell_key = nullptr) {
  const int lane8 = threadIdx.x & 7;
  double* ds = reinterpret_cast<double*>(orb);
  signed char* perm = reinterpret_cast<signed char*>(ds + 2 * M);
  unsigned char* rank = reinterpret_cast<unsigned char*>(perm + 64);
  int* base = reinterpret_cast<int*>(rank + 128);
  for (int r = lane8; r < 128; r += 8) rank[r] = 255;
  wave_lds_sync();

  // FP64 projection, the oracle's order; lane l owns dimensions l + 8i.
  bool valid[MI];
  double acc[MI];
#pragma unroll
  for (int i = 0; i < MI; ++i) {
    valid[i] = lane8 + 8 * i < M;
    acc[i] = 0.0;
  }
  const float* acol = acol0 + lane8;
#pragma unroll 8
  for (int d = 0; d < D; ++d) {
    const double yd = (double)y[d];
    const float* arow = acol + (long long)d * astride;
#pragma unroll
    for (int i = 0; i < MI; ++i) acc[i] = fma(yd, (double)arow[8 * i], acc[i]);
  }
  const uint64_t seed = seeds[t];
  double v[MI];
  uint64_t term = 0;
#pragma unroll
  for (int i = 0; i < MI; ++i) {
    const int m = lane8 + 8 * i;
    v[i] = valid[i] ? (acc[i] - b64[t * M + m]) / W : 0.0;
    const int c = (int)floor(v[i]);
    if (valid[i]) {
      base[m] = c;
      term += code_term(seed, m, c);
    }
  }
  const uint64_t basesum = octet_sum_u64(term);
  if (cell_key) {
    // Coarse cell of the query: floor(v_m / 4) of the first min(6, M)
    // hashes, folded into 12 bits (a locality key; results never depend on it).
    uint64_t ct = 0;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int m = lane8 + 8 * i;
      if (valid[i] && m < 6) ct += mix64(((uint64_t)m << 32) ^ (uint32_t)(int)floor(v[i] * 0.25));
    }
    ct = octet_sum_u64(ct);
    if (active && lane8 == 0) *cell_key = (unsigned)(mix64(ct) >> 52);
  }
  const int maxT = (2 * M < 64) ? 2 * M : 64;

  if (T > 1) {
    // Deltas (cuLSH_Kernels.cu:280-287): index m = "-1 on m", M+m = "+1 on m".
    double dn[MI], up[MI];
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int m = lane8 + 8 * i;
      dn[i] = v[i] - floor(v[i]);
      up[i] = 1.0 - dn[i];
      if (valid[i]) {
        ds[m] = dn[i];
        ds[M + m] = up[i];
      }
    }
    wave_lds_sync();
    // Rank of each delta = number of smaller deltas.  Equal deltas get equal
    // ranks, so the order is unique exactly when no two deltas below rank
    // maxT share a rank (equal deltas all at ranks >= maxT leave the first
    // maxT unchanged): every delta claims perm[rank], and a claim that does
    // not read back as its own marks a tie.
    int rdn[MI], rup[MI];
#pragma unroll
    for (int i = 0; i < MI; ++i) rdn[i] = rup[i] = 0;
#pragma unroll 4
    for (int j = 0; j < 2 * M; ++j) {
      const double dj = ds[j];
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        rdn[i] += dj < dn[i];
        rup[i] += dj < up[i];
      }
    }
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int m = lane8 + 8 * i;
      if (valid[i] && rdn[i] < maxT) perm[rdn[i]] = (signed char)m;
      if (valid[i] && rup[i] < maxT) perm[rup[i]] = (signed char)(M + m);
    }
    wave_lds_sync();
    bool tie = false;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      const int m = lane8 + 8 * i;
      tie |= valid[i] && rdn[i] < maxT && perm[rdn[i]] != (signed char)m;
      tie |= valid[i] && rup[i] < maxT && perm[rup[i]] != (signed char)(M + m);
    }
    const bool otie = octet_min_u32(tie ? 0u : 1u) == 0u;
    wave_lds_sync();  // every lane has read ds and perm before they are rewritten
    if (!otie) {
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        const int m = lane8 + 8 * i;
        if (valid[i]) {
          if (rdn[i] < maxT) {
            ds[rdn[i]] = dn[i];
            rank[m] = (unsigned char)rdn[i];
          }
          if (rup[i] < maxT) {
            ds[rup[i]] = up[i];
            rank[M + m] = (unsigned char)rup[i];
          }
        }
      }
    } else {
      // Exact ties: the reference heap's order (one lane), then the sorted
      // deltas from the lanes' registers.
      if (lane8 == 0) heap_order_deltas(ds, 2 * M, maxT, perm);
      wave_lds_sync();
      for (int r = lane8; r < maxT; r += 8) rank[perm[r]] = (unsigned char)r;
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        const int m = lane8 + 8 * i;
        if (valid[i]) {
          if (rank[m] != 255) ds[rank[m]] = dn[i];
          if (rank[M + m] != 255) ds[rank[M + m]] = up[i];
        }
      }
    }
  }
  wave_lds_sync();

  // Multiprobe queue (cuLSH_Kernels.cu:216-255).  Scores are non-negative
  // doubles, kept as their bit patterns (same order); empty slots hold ~0.
  // Sets as bit masks over the sorted deltas.  An entry made at step p has
  // its highest bit at most p, so T <= 32 (J <= 4) keeps them in 32 bits.
  using Code = typename std::conditional<(J <= 4), uint32_t, uint64_t>::type;
  constexpr int kTop = 8 * (int)sizeof(Code) - 1;
  uint64_t qs[J];
  Code qc[J], pm[J];
#pragma unroll
  for (int r = 0; r < J; ++r) {
    qs[r] = ~0ull;
    qc[r] = 0;
    pm[r] = 0;
  }
  if (T > 1) {
    if (lane8 == 0) {
      qc[0] = 1;
      qs[0] = (uint64_t)__double_as_longlong(ds[0]);
    }
    int size = 1;
    // Pop p = 8 r + j runs with r fixed at compile time, so the popped set
    // goes to register r of lane j and, while every earlier pop made a new
    // entry (size == p), so does the new shift entry (slot p).  Before pop p
    // at most p slots (0 .. p - 1) hold entries and the new one goes to slot
    // size <= p, so registers above r are empty: the scan and the writes
    // stop at register r.
#pragma unroll
    for (int r = 0; r < J; ++r) {
      for (int j = (r == 0 ? 1 : 0); j < 8; ++j) {
        const int p = r * 8 + j;
        if (p >= T) break;  // uniform over the wave
        // Pop: the lowest slot (s = 8 i + lane) among the minimal scores.
        uint64_t omin = ~0ull;
        Code probe = 0;
        int bi = 0;
#pragma unroll
        for (int i = 0; i <= r; ++i)
          if (qs[i] < omin) {
            omin = qs[i];
            bi = i;
            probe = qc[i];
          }
        uint32_t ws = (uint32_t)(bi * 8 + lane8);
        octet_argmin_u64(omin, ws, probe);
        const int wl = (int)(ws & 7), wi = (int)(ws >> 3);
        if (lane8 == j) pm[r] = probe;
        const int maxPos = kTop - (J <= 4 ? __clz((int)probe) : __clzll((long long)probe));
        if (maxPos < maxT - 1) {
          const double e = __longlong_as_double((long long)omin) + ds[maxPos + 1];
          const Code pe = probe | ((Code)1 << (maxPos + 1));
          const double sh = e - ds[maxPos];
          const Code ps = pe - ((Code)1 << maxPos);
          if (lane8 == wl) {
#pragma unroll
            for (int i = 0; i <= r; ++i)
              if (i == wi) {
                qc[i] = pe;
                qs[i] = (uint64_t)__double_as_longlong(e);
              }
          }
          if (size == p) {
            if (lane8 == j) {
              qc[r] = ps;
              qs[r] = (uint64_t)__double_as_longlong(sh);
            }
          } else if (lane8 == (size & 7)) {
#pragma unroll
            for (int i = 0; i <= r; ++i)
              if (i == (size >> 3)) {
                qc[i] = ps;
                qs[i] = (uint64_t)__double_as_longlong(sh);
              }
          }
          ++size;
        }
      }
    }
  }

  // Key sum changes of the 2M deltas, in the sorted deltas' slots (the
  // queue no longer reads them).
  uint64_t* dt = reinterpret_cast<uint64_t*>(ds);
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < MI; ++i) {
    const int m = lane8 + 8 * i;
    if (T > 1 && valid[i]) {
      const int b0 = base[m];
      const uint64_t t0 = code_term(seed, m, b0);
      dt[m] = code_term(seed, m, b0 - 1) - t0;
      dt[M + m] = code_term(seed, m, b0 + 1) - t0;
    }
  }
  wave_lds_sync();

  // Keys, lookup, verify: lane l handles probes l + 8r.  Each probe's key
  // filter word (in cache) is loaded first; only the keys the filter passes
  // load their directory entry, and a directory hit its bucket line.  All of
  // a lane's loads of one kind are issued before any is used.
  const uint2 tfw = pix.tfw[t];
  unsigned long long sizesum = 0;
  uint64_t key[J], ent[J];
#pragma unroll
  for (int r = 0; r < J; ++r) {
    const int p = r * 8 + lane8;
    key[r] = (active && p < T)
                 ? make_key(probe_hash(basesum, dt, perm, rank, M, pm[r]), t, pix.tb)
                 : 0ull;
  }
#pragma unroll
  for (int r = 0; r < J; ++r) {
    const int p = r * 8 + lane8;
    ent[r] = (active && p < T) ? pix.filter[filter_word(key[r], pix.tb, tfw)] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < J; ++r) {
    const int p = r * 8 + lane8;
    const uint64_t fb = filter_bits(key[r]);
    ent[r] = (active && p < T && (ent[r] & fb) == fb) ? pix.dir[key[r] >> pix.shift] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < J; ++r) {
    const int p = r * 8 + lane8;
    if (active && p < T) {
      long long k = lookup_unverified(key[r], ent[r], pix.dir, pix.shift, pix.recs);
      uint2 range = make_uint2(0u, 0u);
      if (k >= 0) {
        const unsigned char* line = bucket_line(pix, (uint32_t)k);
        const uint2 sr = line_range(line);  // in flight with the code loads
        if (probe_matches(line_codes(line), base, perm, rank, M, pm[r])) {
          sizesum += sr.y;
          range = sr;
        } else {
          k = -1;
        }
      }
      if (qbp) qbp[p] = (int)k;
      qrp[p] = range;
    }
  }
  sizesum = octet_sum_u64(sizesum);
  if (active && lane8 == 0) *cnu = sizesum;
}

// The search's octet kernel: 32 queries of one table per workgroup, tables
// in sequence (workgroup b: table b / ceil(Q/32)), so the workgroups in
// flight look up one or two tables' filters, directories and bucket lines at a
// time
// and those stay in cache.  qr[t][q*T + p], qb[t][q*T + p] (when qb is
// given), cnu_qt[q*L + t].  Needs
// probe8_lds_bytes(D, M) bytes of dynamic LDS.
template <int J, int MI>
__global__ __launch_bounds__(256) void k_query_probe8(
    const float* __restrict__ Qm, int Q, int D, int L, int M, int T,
    const float* __restrict__ Ad, int Jp, const double* __restrict__ b64, double W,
    const uint64_t* __restrict__ seeds, ProbeIndex pix, int* __restrict__ qb,
    uint2* __restrict__ qr, unsigned long long* __restrict__ cnu_qt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, oct = tid >> 3;
  const int nqb = (Q + 31) / 32;
  const int t = (int)blockIdx.x / nqb;
  const int q0 = ((int)blockIdx.x - t * nqb) * 32;
  const int q = q0 + oct;
  const bool active = q < Q;
  unsigned char* orb = smem + oct * probe8_octet_bytes(M);
  // Query vectors and projection columns are read through L1 (Ad has at
  // least 64 floats of padding after its last row).
  const int qa = active ? q : q0;
  probe8_pair<J, MI>(Qm + (long long)qa * D, active, t, D, M, T, Ad + t * M, Jp, b64, W,
                     seeds, pix, orb,
                     qb ? qb + (long long)t * Q * T + (long long)qa * T : nullptr,
                     qr + (long long)t * Q * T + (long long)qa * T, cnu_qt + (long long)qa * L + t);
}

}  // namespace culsh
