// Batching form: many independent requests (each with its own Q, K and T)
// searched by one graph launch.  Two schedules, picked by the autotuner:
//   fused  one persistent kernel (k_batch), described below;
//   split  a grid-stride probe kernel over every request (k_batch_probe, at
//          the probe's own occupancy) followed by a persistent candidate
//          kernel (k_batch_cand) with a work queue over the queries.
//
// A batch is laid out in "batch order": requests sorted by T (descending),
// then K (descending), their queries consecutive.  Runs of equal T form
// classes; a class is cut into blocks of `bq` queries.  One work item is one
// block: the workgroup runs the probe stage for the block's queries over all
// L tables (qn x tn (query, table) pairs per pass, one octet each), then the
// candidate stage for each of its queries.  A block's probe results are
// produced and consumed by the same workgroup, so the only synchronisation
// is the workgroup barrier; workgroups claim blocks from one atomic ticket
// counter, heaviest class first.
// Per query, the work is exactly the search's: probe8_pair (cuLSH_Querying.cu
// :63-225) and query_candidates (cuLSH_SearchTables.cu:249-389,
// cuLSH_Kernels.cu:301-456), so results are identical to a search of each
// request alone.
#pragma once
#include "candidates.hpp"
#include "common.hpp"
#include "probe8.hpp"
#include "sort.hpp"

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
  int item0;   // first work item of the fused schedule (block of bq queries)
  int pitem0;  // first probe item of the split schedule
  int pad;
};

// Per-batch header (copied from the host before each batch).
struct BatchHdr {
  const float* Qm;  // caller's queries, row-major [rows][D]
  int* out_ids;     // caller's results
  float* out_d;
  long long Ptot;   // probe slots per table
  int nreq, nclass, Qtot, nitems;  // nitems: fused work items
  int nprobe;                      // split-schedule probe items
  int pj_lo[4], pj_hi[4];          // their ranges per queue width J = 1 << index
  int qn, tn, bq;  // schedule: probe tile qn x tn, queries per fused block
  // Device-written counters (zero in the host copy).
  unsigned next_item;      // fused schedule's work queue
  unsigned next_cand[8];   // split schedule's candidate queue, 8 interleaved shards
  int n_overflow;
  int err;  // bit 0: a query exceeds the overflow path
  int pad1;
  unsigned long long n_recheck;
};

// Per-query descriptor, in batch order.
struct BatchQuery {
  long long qrow, out, poff;
  int K, T;
};

// Index and plan pointers of the persistent kernel.
struct BatchArgs {
  BatchHdr* hdr;
  const BatchClass* classes;
  const BatchQuery* bq;
  uint2* qr;  // [L][Ptot] member ranges of the probes (start, size)
  unsigned long long* cnu_qt;
  unsigned long long* cnu_q;
  unsigned long long* cu_q;
  int* ovl;
  unsigned* lhist;  // [kLocBins] locality histogram, then bin offsets
  unsigned* lkey;   // [Qtot] coarse-cell keys (probe stage), then sort keys
  int* order;       // [Qtot] candidate-stage order (locality within class)
  // Index.
  const float* X;
  const uint16_t* X16;  // FP16 copy of the rows for the distance pass, or null
  double e2;            // its bound on ||x - x16||^2
  const float* Ad;
  const double* b64;
  const uint64_t* seeds;
  ProbeIndex pix;
  const uint32_t* sids;
  double W;
  int D, L, M, Jp;
};

// Expands the request list into per-query descriptors.
__global__ __launch_bounds__(256) void k_batch_expand(const BatchHdr* __restrict__ hdr,
                                                      const BatchReq* __restrict__ reqs,
                                                      BatchQuery* __restrict__ bq) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int Qtot = hdr->Qtot, nreq = hdr->nreq;
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
  const int j = i - r.b0;
  BatchQuery d;
  d.qrow = r.qrow0 + j;
  d.out = r.out0 + (long long)j * r.K;
  d.poff = r.p0 + (long long)j * r.T;
  d.K = r.K;
  d.T = r.T;
  bq[i] = d;
}

// Probe stage of queries [i0, i0 + nq) (same T), all tables, qn x tn pairs
// per pass.  Query vectors are staged in LDS when they fit beside the octet
// scratch, else read from global memory.
template <int J, int MI>
__device__ __forceinline__ void batch_probe(const BatchArgs& a, const BatchHdr& h, int i0, int nq,
                                            int T, unsigned char* pool, int pool_bytes) {
  __shared__ long long s_qrow[32];
  const int tid = threadIdx.x, oct = tid >> 3;
  const int D = a.D, M = a.M, L = a.L, qn = h.qn, tn = h.tn;
  const int ys_bytes = (qn * D * 4 + 15) & ~15;
  const bool stage = ys_bytes + 32 * probe8_octet_bytes(M) <= pool_bytes;
  float* Ys = reinterpret_cast<float*>(pool);
  unsigned char* orb = pool + (stage ? ys_bytes : 0) + oct * probe8_octet_bytes(M);
  const int qi = oct % qn, ti = oct / qn;
  for (int s0 = 0; s0 < nq; s0 += qn) {
    const int ns = min(qn, nq - s0);
    __syncthreads();  // the previous sub-block's readers of s_qrow / Ys are done
    if (tid < qn) s_qrow[tid] = tid < ns ? a.bq[i0 + s0 + tid].qrow : 0;
    __syncthreads();
    if (stage) {
      for (int e = tid; e < ns * D; e += 256) {
        const int r = e / D;
        Ys[e] = h.Qm[s_qrow[r] * D + (e - r * D)];
      }
      __syncthreads();
    }
    const int qa = qi < ns ? qi : 0;
    const int i = i0 + s0 + qa;
    const float* y = stage ? Ys + qa * D : h.Qm + s_qrow[qa] * D;
    const long long poff = a.bq[i].poff;
    for (int tt = 0; tt < L; tt += tn) {
      const int t = tt + ti;
      const bool active = qi < ns && t < L;
      const int te = active ? t : 0;
      probe8_pair<J, MI>(y, active, te, D, M, T, a.Ad + te * M, a.Jp, a.b64, a.W, a.seeds, a.pix,
                         orb, nullptr, a.qr + (long long)te * h.Ptot + poff,
                         a.cnu_qt + (long long)i * L + te);
    }
  }
}

// The persistent fused kernel.  MI: dimensions per lane of the probe
// projection (ceil(M / 8), rounded up to 1, 2, 3, 4 or 8); R: candidate rows
// in flight per 16-lane group; OCC: workgroups per CU the register budget is
// bounded for (2: 256 registers per lane, 3: 168).  T <= 64.
template <int MI, int R, int OCC>
__global__ __launch_bounds__(256, OCC) void k_batch(BatchArgs a) {
  __shared__ __align__(16) uint32_t s_pool[kHcap + kUcap];
  __shared__ unsigned s_item;
  static_assert(kUcap + 3 * kScap + kYcap <= kHcap, "post-union LDS layout");
  const int tid = threadIdx.x;
  BatchHdr* hp = a.hdr;
  const BatchHdr h = *hp;
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
  B.x16 = a.X16;
  B.e2 = a.e2;

  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(&hp->next_item, 1u);
    __syncthreads();
    const unsigned it = s_item;
    if (it >= (unsigned)h.nitems) break;
    int c = 0;
    while (c + 1 < h.nclass && a.classes[c + 1].item0 <= (int)it) ++c;
    const BatchClass C = a.classes[c];
    const int i0 = C.b0 + ((int)it - C.item0) * h.bq;
    const int nq = min(h.bq, C.b0 + C.n - i0);
    unsigned char* pool = reinterpret_cast<unsigned char*>(s_pool);
    const int pb = (int)sizeof(s_pool);
    if (C.T <= 8)
      batch_probe<1, MI>(a, h, i0, nq, C.T, pool, pb);
    else if (C.T <= 16)
      batch_probe<2, MI>(a, h, i0, nq, C.T, pool, pb);
    else if (C.T <= 32)
      batch_probe<4, MI>(a, h, i0, nq, C.T, pool, pb);
    else
      batch_probe<8, MI>(a, h, i0, nq, C.T, pool, pb);
    // The block's probe slots and C_nu rows were written by this workgroup;
    // query_candidates starts with a barrier.
    for (int i = i0; i < i0 + nq; ++i) {
      const BatchQuery d = a.bq[i];
      QueryIO io;
      io.y = h.Qm + d.qrow * a.D;
      io.qrq = a.qr + d.poff;
      io.tstride = h.Ptot;
      io.T = d.T;
      io.K = d.K;
      io.oid = h.out_ids + d.out;
      io.od = h.out_d + d.out;
      io.slot = i;
      if (!query_candidates<R>(io, a.X, a.D, a.L, a.sids, a.cnu_qt, a.cnu_q, a.cu_q,
                               &hp->n_recheck, B)) {
        if (tid == 0) a.ovl[atomicAdd(&hp->n_overflow, 1)] = i;
      }
    }
  }
}

// Split schedule, probe stage: probe items (class, block of qn queries, tn
// tables) -- 32 (query, table) pairs, one pass -- one per workgroup, items
// of heavier classes first and, within a class, table groups in sequence
// (the workgroups in flight look up few tables, as in k_query_probe8).  (A
// grid-stride loop over items keeps the projection's values live across
// iterations and doubles the registers.)  Needs batch_probe_lds(M) bytes of
// dynamic LDS.  Query vectors are read through L1, as in k_query_probe8.
__host__ __device__ inline int batch_probe_lds(int M) { return 32 * probe8_octet_bytes(M); }

// One launch per queue width J (index ji: J = 1 << ji; the host uses J = 4
// and 8) over that width's contiguous item range, so each launch has the
// registers of its own J; the
// grid covers the plan's largest possible range and the workgroups past
// this batch's range exit at once.
template <int MI, int J>
__global__ __launch_bounds__(256) void k_batch_probe(BatchArgs a) {
  constexpr int ji = J == 1 ? 0 : J == 2 ? 1 : J == 4 ? 2 : 3;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ long long s_qrow[32];
  const int tid = threadIdx.x, oct = tid >> 3;
  // Only the header fields this kernel needs (a full copy would crowd the
  // scalar registers).
  const BatchHdr* hp = a.hdr;
  const float* Qm = hp->Qm;
  const long long Ptot = hp->Ptot;
  const int nclass = hp->nclass, qn = hp->qn, tn = hp->tn;
  const int it_lo = hp->pj_lo[ji], it_hi = hp->pj_hi[ji];
  const int D = a.D, M = a.M, L = a.L;
  unsigned char* orb = smem + oct * probe8_octet_bytes(M);
  const int qi = oct % qn, ti = oct / qn;
  {
    const int it = it_lo + (int)blockIdx.x;
    if (it >= it_hi) return;
    int c = 0;
    while (c + 1 < nclass && a.classes[c + 1].pitem0 <= it) ++c;
    const BatchClass C = a.classes[c];
    const int local = it - C.pitem0;
    const int nbl = (C.n + qn - 1) / qn;  // the class's items: table groups in sequence
    const int g = local / nbl, bl = local - g * nbl;
    const int i0 = C.b0 + bl * qn;
    const int ns = min(qn, C.b0 + C.n - i0);
    if (tid < qn) s_qrow[tid] = tid < ns ? a.bq[i0 + tid].qrow : 0;
    __syncthreads();
    const int qa = qi < ns ? qi : 0;
    const int i = i0 + qa;
    const float* y = Qm + s_qrow[qa] * D;
    const long long poff = a.bq[i].poff;
    const int t = g * tn + ti;
    const bool active = qi < ns && t < L;
    const int te = active ? t : 0;
    probe8_pair<J, MI>(y, active, te, D, M, C.T, a.Ad + te * M, a.Jp, a.b64, a.W, a.seeds, a.pix,
                       orb, nullptr, a.qr + (long long)te * Ptot + poff,
                       a.cnu_qt + (long long)i * L + te, te == 0 ? a.lkey + i : nullptr);
  }
}

// Locality order of the split schedule's candidate stage: a counting sort
// of the key (class, 12-bit hash of the query's coarse cell: floor((a_m.y -
// b_m) / 4W) of the first min(6, M) hashes of table 0) -- queries in one
// coarse cell share most of their candidates, and the class (T) keeps heavy
// queries first.  Grids are fixed (graph capture); bins beyond the batch stay
// empty; batches under kLocMinQ queries keep batch order (the sort would
// cost more than it saves).  The order within a bin is whatever the atomics
// give: results do not depend on the order.
constexpr int kLocBins = 1 << 16;
constexpr int kLocChunk = 64;     // consecutive ordered queries per XCD claim
constexpr int kLocMinQ = 8192;    // smaller batches keep batch order

// Sort key of batch index i: its class, then the 12-bit coarse-cell key the
// probe stage left in lkey[i] (table 0).
__device__ inline unsigned batch_loc_key(const BatchArgs& a, const BatchHdr* hp, int i) {
  const int nclass = hp->nclass;
  int c = 0;
  while (c + 1 < nclass && a.classes[c + 1].b0 <= i) ++c;
  return ((unsigned)min(c, 15) << 12) | (a.lkey[i] & 0xFFFu);
}

__global__ __launch_bounds__(256) void k_batch_lzero(BatchArgs a) {
  if (a.hdr->Qtot < kLocMinQ) return;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < kLocBins; i += gridDim.x * 256)
    a.lhist[i] = 0u;
}

__global__ __launch_bounds__(256) void k_batch_lhist(BatchArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int Qtot = a.hdr->Qtot;
  if (i >= Qtot || Qtot < kLocMinQ) return;
  const unsigned k = batch_loc_key(a, a.hdr, i);
  a.lkey[i] = k;  // the full sort key, for the scatter
  atomicAdd(&a.lhist[k], 1u);
}

// Exclusive scan of the kLocBins counts in place (one workgroup, 256
// consecutive bins per thread).
__global__ __launch_bounds__(256) void k_batch_lscan(BatchArgs a) {
  __shared__ unsigned s_misc[4];
  if (a.hdr->Qtot < kLocMinQ) return;
  const int tid = threadIdx.x;
  constexpr int per = kLocBins / 256;
  unsigned* h = a.lhist + tid * per;
  unsigned sum = 0;
  for (int k = 0; k < per; ++k) sum += h[k];
  unsigned tot;
  unsigned run = block_exclusive_scan_256(sum, s_misc, tot);
  for (int k = 0; k < per; ++k) {
    const unsigned v = h[k];
    h[k] = run;
    run += v;
  }
}

__global__ __launch_bounds__(256) void k_batch_lscatter(BatchArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int Qtot = a.hdr->Qtot;
  if (i >= Qtot) return;
  if (Qtot < kLocMinQ) {
    a.order[i] = i;  // small batch: batch order
    return;
  }
  const unsigned pos = atomicAdd(&a.lhist[a.lkey[i]], 1u);
  a.order[pos] = i;
}

// Claims the next query of the split schedule's candidate queue.  The
// locality order is cut into chunks of kLocChunk; shard x (XCD x: workgroups
// x, x + 8, ...) takes chunks x, x + 8, ... in order, so neighbouring
// queries share an L2, and moves on to the next shard when its own runs dry
// (8 counters spread the queue's atomics).  Returns -1 when every shard is
// exhausted.  Called by thread 0.
__device__ inline int claim_cand(const BatchArgs& a, BatchHdr* hp, int Qtot, int& shard,
                                 int& tried) {
  while (tried < 8) {
    const unsigned t = atomicAdd(&hp->next_cand[shard], 1u);
    const long long p = ((long long)(t / kLocChunk) * 8 + shard) * kLocChunk + t % kLocChunk;
    if (p < Qtot) return a.order[p];
    // Chunks of this shard beyond Qtot: it is exhausted once its chunk
    // start passes the end.
    if (((long long)(t / kLocChunk) * 8 + shard) * kLocChunk >= Qtot) {
      shard = (shard + 1) & 7;
      ++tried;
    }
  }
  return -1;
}

// Split schedule, candidate stage: one query per work item, claimed from the
// sharded queue in locality order (heaviest classes first).
template <int R>
__global__ __launch_bounds__(256, 3) void k_batch_cand(BatchArgs a) {
  __shared__ __align__(16) uint32_t s_pool[kHcap + kUcap];
  __shared__ unsigned s_item;
  const int tid = threadIdx.x;
  BatchHdr* hp = a.hdr;
  const float* Qm = hp->Qm;
  int* out_ids = hp->out_ids;
  float* out_d = hp->out_d;
  const long long Ptot = hp->Ptot;
  const int Qtot = hp->Qtot;
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
  B.x16 = a.X16;
  B.e2 = a.e2;
  int shard = (int)(blockIdx.x & 7), tried = 0;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = (unsigned)claim_cand(a, hp, Qtot, shard, tried);
    __syncthreads();
    const int i = (int)s_item;
    if (i < 0) break;
    const BatchQuery d = a.bq[i];
    QueryIO io;
    io.y = Qm + d.qrow * a.D;
    io.qrq = a.qr + d.poff;
    io.tstride = Ptot;
    io.T = d.T;
    io.K = d.K;
    io.oid = out_ids + d.out;
    io.od = out_d + d.out;
    io.slot = i;
    if (!query_candidates<R>(io, a.X, a.D, a.L, a.sids, a.cnu_qt, a.cnu_q, a.cu_q,
                             &hp->n_recheck, B)) {
      if (tid == 0) a.ovl[atomicAdd(&hp->n_overflow, 1)] = i;
    }
  }
}

// Split schedule, candidate stage without persistence: one workgroup per
// query (the grid covers the plan's capacity; workgroups past the batch
// exit).  Workgroup b runs on XCD b % 8 and takes the (b / 8)-th position of
// that XCD's chunks of the locality order, so dispatch follows the order
// (heaviest classes first) and neighbouring queries share an L2.
template <int R>
__global__ __launch_bounds__(256, 3) void k_batch_cand_each(BatchArgs a) {
  __shared__ __align__(16) uint32_t s_pool[kHcap + kUcap];
  BatchHdr* hp = a.hdr;
  const int Qtot = hp->Qtot;
  const long long k = blockIdx.x >> 3, x = blockIdx.x & 7;
  const long long p = ((k / kLocChunk) * 8 + x) * kLocChunk + k % kLocChunk;
  if (p >= Qtot) return;
  const int i = a.order[p];
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
  B.x16 = a.X16;
  B.e2 = a.e2;
  const BatchQuery d = a.bq[i];
  QueryIO io;
  io.y = hp->Qm + d.qrow * a.D;
  io.qrq = a.qr + d.poff;
  io.tstride = hp->Ptot;
  io.T = d.T;
  io.K = d.K;
  io.oid = hp->out_ids + d.out;
  io.od = hp->out_d + d.out;
  io.slot = i;
  if (!query_candidates<R>(io, a.X, a.D, a.L, a.sids, a.cnu_qt, a.cnu_q, a.cu_q,
                           &hp->n_recheck, B)) {
    if (threadIdx.x == 0) a.ovl[atomicAdd(&hp->n_overflow, 1)] = i;
  }
}

// Queries that overflowed the LDS working set (or, with the FP16 rows, its
// FP64 re-evaluation set): workgroup b takes list
// entries b, b + G, ...; hash, id list and FP32 distances in global scratch
// (2^hbits_max + 2 ucap_max words per workgroup).  A query with more than
// ucap_max non-unique candidates, or more FP64 rechecks than kScapG, sets
// err bit 0.
__global__ __launch_bounds__(256) void k_batch_overflow(BatchArgs a, uint32_t* __restrict__ scratch,
                                                        int hbits_max, unsigned ucap_max) {
  __shared__ double s_sd[kScapG];
  __shared__ uint32_t s_sid[kScapG];
  __shared__ __align__(16) float s_y[kYcap];
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
    // FP32 rows here: a query whose FP16 band held more candidates than
    // the LDS path re-evaluates is finished as the search finishes it.
    B.x16 = nullptr;
    QueryIO io;
    io.y = h.Qm + d.qrow * a.D;
    io.qrq = a.qr + d.poff;
    io.tstride = h.Ptot;
    io.T = d.T;
    io.K = d.K;
    io.oid = h.out_ids + d.out;
    io.od = h.out_d + d.out;
    io.slot = i;
    if (!query_candidates<4>(io, a.X, a.D, a.L, a.sids, a.cnu_qt, a.cnu_q, a.cu_q,
                             &hp->n_recheck, B)) {
      if (threadIdx.x == 0) atomicOr(&hp->err, 1);
    }
  }
}

}  // namespace culsh
