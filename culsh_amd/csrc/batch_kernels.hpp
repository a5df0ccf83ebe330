// Batching form: many independent requests (each with its own Q, K and T)
// searched by one persistent fused kernel.
//
// A batch is laid out in "batch order": requests sorted by T (descending),
// then K (descending), their queries consecutive.  Runs of equal T form
// classes; a class is cut into probe blocks of `qn` queries.  The work queue
// holds
//   probe items     (class, block, table group): qn x tn (query, table) pairs
//                   per pass, 32 pairs per workgroup (one octet each), over
//                   `tgroup` tables;
//   candidate items one per query, in batch order (heaviest T and K first).
// Workgroups claim items from one atomic ticket counter.  Every probe item
// has a lower ticket than every candidate item, so a candidate item whose
// probe block is unfinished waits only on items already claimed by running
// workgroups; the wait is bounded and reports a timeout instead of hanging.
// Per query, the work is exactly the search's: probe8_pair (cuLSH_Querying.cu
// :63-225) and query_candidates (cuLSH_SearchTables.cu:249-389,
// cuLSH_Kernels.cu:301-456), so results are identical to a search of each
// request alone.
#pragma once
#include "candidates.hpp"
#include "common.hpp"
#include "probe8.hpp"

namespace culsh {

// One request, as the host lays it out (batch order).
struct BatchReq {
  long long qrow0;  // first query row in the caller's query matrix
  long long out0;   // first result element in the caller's result arrays
  long long p0;     // first probe slot (within a table) of its queries
  int b0;           // first batch index of its queries
  int Q, K, T;
};

// A run of batch indices with equal T.
struct BatchClass {
  int b0, n, T;
  int blk0;   // first global probe block
  int item0;  // first probe item
  int pad;
};

// Per-batch header (copied from the host before each batch).
struct BatchHdr {
  const float* Qm;  // caller's queries, row-major [rows][D]
  int* out_ids;     // caller's results
  float* out_d;
  long long Ptot;   // probe slots per table
  int nreq, nclass, Qtot, nprobe;
  int qn, tn, tgroup, ntg;  // schedule: probe tile qn x tn, tables per item, items per block
  // Device-written counters (zero in the host copy).
  unsigned next_item;
  int n_overflow;
  int err;  // bit 0: a query exceeds the overflow path; bit 1: wait timeout
  int pad;
  unsigned long long n_recheck;
};

// Per-query descriptor, in batch order.
struct BatchQuery {
  long long qrow, out, poff;
  int K, T, blk, pad;
};

// Index and plan pointers of the persistent kernel.
struct BatchArgs {
  BatchHdr* hdr;
  const BatchClass* classes;
  const BatchQuery* bq;
  unsigned* blockdone;
  int* qb;
  unsigned long long* cnu_qt;
  unsigned long long* cnu_q;
  unsigned long long* cu_q;
  int* ovl;
  // Index.
  const float* X;
  const float* Ad;
  const double* b64;
  const uint64_t* seeds;
  const uint64_t* dir;
  const BucketRec* recs;
  const int16_t* bcodes;
  const uint32_t* sids;
  double W;
  int D, L, M, Jp, tb, shift;
};

constexpr unsigned kSpinLimit = 1u << 24;  // ~ seconds of s_sleep before a timeout

// Expands the request list into per-query descriptors; clears the probe
// block counters.
__global__ __launch_bounds__(256) void k_batch_expand(const BatchHdr* __restrict__ hdr,
                                                      const BatchReq* __restrict__ reqs,
                                                      const BatchClass* __restrict__ classes,
                                                      BatchQuery* __restrict__ bq,
                                                      unsigned* __restrict__ blockdone,
                                                      int nblk_cap) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int Qtot = hdr->Qtot, nreq = hdr->nreq, nclass = hdr->nclass, qn = hdr->qn;
  if (i < nblk_cap) blockdone[i] = 0u;
  if (i >= Qtot) return;
  int lo = 0, hi = nreq - 1;  // last request with b0 <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (reqs[mid].b0 <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  const BatchReq r = reqs[lo];
  int c = 0;
  while (c + 1 < nclass && classes[c + 1].b0 <= i) ++c;
  const int j = i - r.b0;
  BatchQuery d;
  d.qrow = r.qrow0 + j;
  d.out = r.out0 + (long long)j * r.K;
  d.poff = r.p0 + (long long)j * r.T;
  d.K = r.K;
  d.T = r.T;
  d.blk = classes[c].blk0 + (i - classes[c].b0) / qn;
  d.pad = 0;
  bq[i] = d;
}

// One probe item: queries [i0, i0 + nq) of one class (same T), tables
// [t_begin, t_end), qn x tn pairs per pass.
template <int J, int MI>
__device__ inline void batch_probe_item(const BatchArgs& a, const BatchHdr& h, int i0, int nq,
                                        int t_begin, int t_end, int T, unsigned char* pool,
                                        int pool_bytes) {
  __shared__ long long s_qrow[32];
  const int tid = threadIdx.x, oct = tid >> 3;
  const int D = a.D, M = a.M, qn = h.qn, tn = h.tn;
  if (tid < 32) s_qrow[tid] = tid < nq ? a.bq[i0 + tid].qrow : 0;
  __syncthreads();
  const int ys_bytes = (qn * D * 4 + 15) & ~15;
  const bool stage = ys_bytes + 32 * probe8_octet_bytes(M) <= pool_bytes;
  float* Ys = reinterpret_cast<float*>(pool);
  if (stage) {
    for (int e = tid; e < nq * D; e += 256) {
      const int r = e / D;
      Ys[e] = h.Qm[s_qrow[r] * D + (e - r * D)];
    }
  }
  __syncthreads();
  unsigned char* orb = pool + (stage ? ys_bytes : 0) + oct * probe8_octet_bytes(M);
  const int qi = oct % qn, ti = oct / qn;
  const int qa = qi < nq ? qi : 0;
  const int i = i0 + qa;
  const float* y = stage ? Ys + qa * D : h.Qm + s_qrow[qa] * D;
  const long long poff = a.bq[i].poff;
  for (int tt = t_begin; tt < t_end; tt += tn) {
    const int t = tt + ti;
    const bool active = qi < nq && t < t_end;
    const int te = active ? t : t_begin;
    probe8_pair<J, MI>(y, active, te, D, M, T, a.Ad, a.Jp, a.b64, a.W, a.seeds, a.tb, a.shift,
                       a.dir, a.recs, a.bcodes, orb, a.qb + (long long)te * h.Ptot + poff,
                       a.cnu_qt + (long long)i * a.L + te);
  }
}

// The persistent fused kernel.  MI: dimensions per lane of the probe
// projection (ceil(M / 8) rounded up to a power of two); R: candidate rows
// in flight per 16-lane group; OCC: workgroups per CU the register budget is
// bounded for (2: 256 registers per lane, 3: 168).  T <= 64.
template <int MI, int R, int OCC>
__global__ __launch_bounds__(256, OCC) void k_batch(BatchArgs a) {
  __shared__ __align__(16) uint32_t s_pool[kHcap + kUcap];
  __shared__ unsigned s_item;
  __shared__ int s_go;
  static_assert(kUcap + 3 * kScap + kYcap <= kHcap, "post-union LDS layout");
  const int tid = threadIdx.x;
  BatchHdr* hp = a.hdr;
  const BatchHdr h = *hp;
  const unsigned nitems = (unsigned)h.nprobe + (unsigned)h.Qtot;
  CandBufs B;
  B.hash = s_pool;
  B.hbits = 13;
  B.ulist = s_pool + kHcap;
  B.ucap = kUcap;
  B.d32 = reinterpret_cast<float*>(s_pool);
  B.sd = reinterpret_cast<double*>(s_pool + kUcap);
  B.sid = s_pool + kUcap + 2 * kScap;
  B.scap = kScap;
  B.ys = reinterpret_cast<float*>(s_pool + kUcap + 3 * kScap);

  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(&hp->next_item, 1u);
    __syncthreads();
    const unsigned it = s_item;
    if (it >= nitems) break;
    if (it < (unsigned)h.nprobe) {
      int c = 0;
      while (c + 1 < h.nclass && a.classes[c + 1].item0 <= (int)it) ++c;
      const BatchClass C = a.classes[c];
      const int local = (int)it - C.item0;
      const int bl = local / h.ntg, g = local - bl * h.ntg;
      const int i0 = C.b0 + bl * h.qn;
      const int nq = min(h.qn, C.b0 + C.n - i0);
      const int t_begin = g * h.tgroup, t_end = min(a.L, t_begin + h.tgroup);
      unsigned char* pool = reinterpret_cast<unsigned char*>(s_pool);
      const int pb = (int)sizeof(s_pool);
      if (C.T <= 8)
        batch_probe_item<1, MI>(a, h, i0, nq, t_begin, t_end, C.T, pool, pb);
      else if (C.T <= 16)
        batch_probe_item<2, MI>(a, h, i0, nq, t_begin, t_end, C.T, pool, pb);
      else if (C.T <= 32)
        batch_probe_item<4, MI>(a, h, i0, nq, t_begin, t_end, C.T, pool, pb);
      else
        batch_probe_item<8, MI>(a, h, i0, nq, t_begin, t_end, C.T, pool, pb);
      // Publish: every thread's qb / C_nu stores are visible device-wide
      // before the block counter moves.
      __threadfence();
      __syncthreads();
      if (tid == 0) atomicAdd(&a.blockdone[C.blk0 + bl], 1u);
    } else {
      const int i = (int)(it - (unsigned)h.nprobe);
      const BatchQuery d = a.bq[i];
      if (tid == 0) {
        int go = 1;
        unsigned spins = 0;
        while (atomicAdd(&a.blockdone[d.blk], 0u) < (unsigned)h.ntg) {
          __builtin_amdgcn_s_sleep(8);
          if (++spins == kSpinLimit) {
            atomicOr(&hp->err, 2);
            go = 0;
            break;
          }
        }
        s_go = go;
      }
      __syncthreads();
      __threadfence();  // acquire: the block's probe results
      if (!s_go) continue;
      QueryIO io;
      io.y = h.Qm + d.qrow * a.D;
      io.qbq = a.qb + d.poff;
      io.tstride = h.Ptot;
      io.T = d.T;
      io.K = d.K;
      io.oid = h.out_ids + d.out;
      io.od = h.out_d + d.out;
      io.slot = i;
      if (!query_candidates<R>(io, a.X, a.D, a.L, a.recs, a.sids, a.cnu_qt, a.cnu_q, a.cu_q,
                               &hp->n_recheck, B)) {
        if (tid == 0) a.ovl[atomicAdd(&hp->n_overflow, 1)] = i;
      }
    }
  }
}

// Queries that overflowed the LDS working set: workgroup b takes list
// entries b, b + G, ...; hash, id list and FP32 distances in global scratch
// (2^hbits_max + 2 ucap_max words per workgroup).  A query with more than
// ucap_max non-unique candidates, or more FP64 rechecks than kScapG, sets
// err bit 0.
__global__ __launch_bounds__(256) void k_batch_overflow(BatchArgs a, uint32_t* __restrict__ scratch,
                                                        int hbits_max, unsigned ucap_max) {
  __shared__ double s_sd[kScapG];
  __shared__ uint32_t s_sid[kScapG];
  __shared__ float s_y[kYcap];
  BatchHdr* hp = a.hdr;
  const int n = hp->n_overflow;
  const BatchHdr h = *hp;
  uint32_t* base = scratch + (size_t)blockIdx.x * ((1ull << hbits_max) + 2ull * ucap_max);
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int i = a.ovl[e];
    const unsigned long long cnu = query_cnu(i, a.L, a.cnu_qt);
    if (cnu > ucap_max) {
      if (threadIdx.x == 0) atomicOr(&hp->err, 1);
      continue;
    }
    int hbits = 13;
    while ((1ull << hbits) < 2 * cnu) ++hbits;
    const BatchQuery d = a.bq[i];
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
    QueryIO io;
    io.y = h.Qm + d.qrow * a.D;
    io.qbq = a.qb + d.poff;
    io.tstride = h.Ptot;
    io.T = d.T;
    io.K = d.K;
    io.oid = h.out_ids + d.out;
    io.od = h.out_d + d.out;
    io.slot = i;
    if (!query_candidates<4>(io, a.X, a.D, a.L, a.recs, a.sids, a.cnu_qt, a.cnu_q, a.cu_q,
                             &hp->n_recheck, B)) {
      if (threadIdx.x == 0) atomicOr(&hp->err, 1);
    }
  }
}

}  // namespace culsh
