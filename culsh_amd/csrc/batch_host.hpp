// Host side of the batching form (culsh_bt_*): request layout, capacities,
// the captured graph and the schedule autotuner.  Part of libculsh's single
// translation unit (included by culsh.hip after the index and search code,
// whose helpers and index state it uses).
#pragma once

struct culsh_bt {
  culsh_ht* ht = nullptr;
  culsh_schedule sched{32, 1, 32, 4, 3, 0};
  hipStream_t cap = nullptr;  // capture stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int ncu = 0;
  // Capacities of the device buffers the graph was captured with.
  int Qcap = 0, Rcap = 0;
  long long Pcap = 0;
  size_t off_cls = 0, off_req = 0, packed = 0;
  DevBuf<unsigned char> dpack;  // BatchHdr | BatchClass[Rcap] | BatchReq[Rcap]
  unsigned char* hpack = nullptr;  // pinned host image of dpack
  BatchHdr* hres = nullptr;        // pinned: header after the batch
  DevBuf<BatchQuery> dbq;
  DevBuf<uint2> dqr;  // probe member ranges
  DevBuf<int> dovl, dorder;
  DevBuf<unsigned> dlhist, dlkey;
  DevBuf<unsigned long long> dcnu_qt, dcnu_q, dcu_q;
  DevBuf<uint32_t> dscr;
  static constexpr int kOvGrid = 16;
  static constexpr int kHbitsMax = 21;
  static constexpr unsigned kUcapMax = 1u << 20;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  culsh_schedule graph_sched{0, 0, 0, 0, 0, 0};
  unsigned long long graph_gen = 0;  // the index's data_gen at capture
  bool pending = false;
  unsigned long long rechecked = 0;

  ~culsh_bt() {
    if (pending && ev1) (void)hipEventSynchronize(ev1);
    drop_graph();
    if (hpack) (void)hipHostFree(hpack);
    if (hres) (void)hipHostFree(hres);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (cap) (void)hipStreamDestroy(cap);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

bool same_schedule(const culsh_schedule& a, const culsh_schedule& b) {
  return a.probe_queries == b.probe_queries && a.probe_tables == b.probe_tables &&
         a.queries_per_item == b.queries_per_item && a.rows_in_flight == b.rows_in_flight &&
         a.wgs_per_cu == b.wgs_per_cu && a.fused == b.fused;
}

int check_schedule(const culsh_schedule& s) {
  const bool tile = (s.probe_queries == 32 && s.probe_tables == 1) ||
                    (s.probe_queries == 16 && s.probe_tables == 2) ||
                    (s.probe_queries == 8 && s.probe_tables == 4);
  if (!tile) return fail(CULSH_E_ARG, "probe tile must be 32x1, 16x2 or 8x4");
  if (s.queries_per_item < s.probe_queries || s.queries_per_item > 256 ||
      s.queries_per_item % s.probe_queries != 0)
    return fail(CULSH_E_ARG,
                "queries_per_item must be a multiple of probe_queries, at most 256");
  if (s.rows_in_flight != 2 && s.rows_in_flight != 4 && s.rows_in_flight != 8)
    return fail(CULSH_E_ARG, "rows_in_flight must be 2, 4 or 8");
  if (s.wgs_per_cu < 0 || s.wgs_per_cu > 3 || (s.fused && s.wgs_per_cu == 0))
    return fail(CULSH_E_ARG, "wgs_per_cu must be 1..3 (0: one workgroup per query, split only)");
  if (s.fused != 0 && s.fused != 1) return fail(CULSH_E_ARG, "fused must be 0 or 1");
  return CULSH_OK;
}

typedef void (*BatchKernel)(BatchArgs);

template <int MI, int OCC>
BatchKernel batch_kernel_r(int R) {
  if (R == 2) return k_batch<MI, 2, OCC>;
  if (R == 8) return k_batch<MI, 8, OCC>;
  return k_batch<MI, 4, OCC>;
}

template <int OCC>
BatchKernel batch_kernel_o(int M, int R) {
  if (M <= 8) return batch_kernel_r<1, OCC>(R);
  if (M <= 16) return batch_kernel_r<2, OCC>(R);
  if (M <= 24) return batch_kernel_r<3, OCC>(R);
  if (M <= 32) return batch_kernel_r<4, OCC>(R);
  return batch_kernel_r<8, OCC>(R);
}

// Three workgroups per CU need the tighter register bound.
BatchKernel batch_kernel(int M, int R, int wgs_per_cu) {
  return wgs_per_cu >= 3 ? batch_kernel_o<3>(M, R) : batch_kernel_o<2>(M, R);
}

template <int J>
BatchKernel batch_probe_kernel_j(int M) {
  if (M <= 8) return k_batch_probe<1, J>;
  if (M <= 16) return k_batch_probe<2, J>;
  if (M <= 24) return k_batch_probe<3, J>;
  if (M <= 32) return k_batch_probe<4, J>;
  return k_batch_probe<8, J>;
}

// Queue width index of a probe count (J = 1 << index slots per lane): the
// split schedule runs T <= 32 at J = 4 (the narrower widths take as many
// registers) and T <= 64 at J = 8.
int batch_ji(int T) { return T <= 32 ? 2 : 3; }

BatchKernel batch_cand_kernel(int R) {
  if (R == 2) return k_batch_cand<2>;
  if (R == 8) return k_batch_cand<8>;
  return k_batch_cand<4>;
}

BatchKernel batch_cand_each_kernel(int R) {
  if (R == 2) return k_batch_cand_each<2>;
  if (R == 8) return k_batch_cand_each<8>;
  return k_batch_cand_each<4>;
}

BatchArgs batch_args(culsh_bt* bt) {
  culsh_ht* ht = bt->ht;
  BatchArgs a;
  a.hdr = reinterpret_cast<BatchHdr*>(bt->dpack.p);
  a.classes = reinterpret_cast<const BatchClass*>(bt->dpack.p + bt->off_cls);
  a.bq = bt->dbq.p;
  a.qr = bt->dqr.p;
  a.cnu_qt = bt->dcnu_qt.p;
  a.cnu_q = bt->dcnu_q.p;
  a.cu_q = bt->dcu_q.p;
  a.ovl = bt->dovl.p;
  a.lhist = bt->dlhist.p;
  a.order = bt->dorder.p;
  a.lkey = bt->dlkey.p;
  a.X = ht->dX;
  a.X16 = ht->rows16_ok ? ht->dX16.p : nullptr;
  a.e2 = ht->rows16_e2;
  a.Ad = ht->dAd.p;
  a.b64 = ht->db64.p;
  a.seeds = ht->dseeds.p;
  a.pix = probe_index(ht);
  a.sids = ht->sids.p;
  a.W = (double)ht->W;
  a.D = ht->D;
  a.L = ht->L;
  a.M = ht->M;
  a.Jp = ht->Jp;
  return a;
}

// Grows the plan's buffers to hold a batch of Qtot queries, nreq requests
// and Ptot probe slots per table; drops the graph when anything moves.  On
// a failed allocation the capacities of the buffers it released are zeroed
// (the next reserve allocates them again) and no graph is left that
// references them.
int bt_reserve(culsh_bt* bt, int Qtot, int nreq, long long Ptot) {
  const int L = bt->ht->L;
  const bool grow_r = nreq > bt->Rcap || !bt->hpack, grow_q = Qtot > bt->Qcap,
             grow_p = Ptot > bt->Pcap, need_h = !bt->dlhist.p, need_s = !bt->dscr.p;
  if (!(grow_r || grow_q || grow_p || need_h || need_s)) return CULSH_OK;
  bt->drop_graph();
  const long long Pc = std::max(Ptot, std::min(2 * bt->Pcap, ((1ll << 31) - 1) / L));
  if (grow_p && (long long)L * Pc >= (1ll << 31))
    return fail(CULSH_E_LIMIT, "L * (sum of Q*T over the batch) exceeds 2^31 - 1");
  if (grow_r) {
    const int R = std::max(nreq, std::max(64, 2 * bt->Rcap));
    const size_t off_cls = (sizeof(BatchHdr) + 63) & ~size_t(63);
    const size_t off_req = (off_cls + R * sizeof(BatchClass) + 63) & ~size_t(63);
    const size_t packed = off_req + R * sizeof(BatchReq);
    bt->Rcap = 0;
    if (bt->hpack) HIP_TRY(hipHostFree(bt->hpack));
    bt->hpack = nullptr;
    TRY(bt->dpack.exact(packed));
    HIP_TRY(hipHostMalloc((void**)&bt->hpack, packed, hipHostMallocDefault));
    bt->Rcap = R;
    bt->off_cls = off_cls;
    bt->off_req = off_req;
    bt->packed = packed;
  }
  if (grow_q) {
    const int Qc = std::max(Qtot, std::min(2 * bt->Qcap, (1 << 30)));
    bt->Qcap = 0;
    TRY(bt->dbq.exact(Qc));
    TRY(bt->dcnu_qt.exact((size_t)Qc * L));
    TRY(bt->dcnu_q.exact(Qc));
    TRY(bt->dcu_q.exact(Qc));
    TRY(bt->dovl.exact(Qc));
    TRY(bt->dorder.exact(Qc));
    TRY(bt->dlkey.exact(Qc));
    bt->Qcap = Qc;
  }
  if (grow_p) {
    bt->Pcap = 0;
    TRY(bt->dqr.exact((size_t)L * Pc));
    bt->Pcap = Pc;
  }
  if (need_h) TRY(bt->dlhist.exact(kLocBins));
  if (need_s)
    TRY(bt->dscr.exact((size_t)culsh_bt::kOvGrid *
                       ((1ull << culsh_bt::kHbitsMax) + 2ull * culsh_bt::kUcapMax)));
  return CULSH_OK;
}

// The plan's work on stream s: descriptor upload, expansion, the probe and
// candidate stages, overflow kernel, header download.  Every pointer is a
// plan buffer, so a captured copy serves any batch within the capacities.
// With `ev` (eager runs only), ev[0..4] are recorded before the expansion,
// after the probe stage, after the locality order, after the candidate
// stage and at the end (the fused schedule's one kernel counts as the
// candidate stage).
void bt_enqueue(culsh_bt* bt, hipStream_t st, hipEvent_t* ev, hipError_t* err) {
  culsh_ht* ht = bt->ht;
  const culsh_schedule& s = bt->sched;
  const BatchArgs a = batch_args(bt);
  const int grid = bt->ncu * s.wgs_per_cu;
  const int egrid = (int)grid_for(bt->Qcap);
  // Split schedule: one workgroup per probe item, as many as a batch within
  // the plan's capacities can have (every class -- at most one per T <= 64 --
  // adds at most one partial block).
  const long long pcap = ((long long)(bt->Qcap + s.probe_queries - 1) / s.probe_queries +
                          std::min(bt->Rcap, 64)) *
                         ((ht->L + s.probe_tables - 1) / s.probe_tables);
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  const hipError_t m1 =
      hipMemcpyAsync(bt->dpack.p, bt->hpack, bt->packed, hipMemcpyHostToDevice, st);
  mark(0);
  hipLaunchKernelGGL(k_batch_expand, dim3(egrid), dim3(256), 0, st, a.hdr,
                     reinterpret_cast<const BatchReq*>(bt->dpack.p + bt->off_req), bt->dbq.p);
  if (s.fused) {
    mark(1);
    mark(2);
    hipLaunchKernelGGL(batch_kernel(ht->M, s.rows_in_flight, s.wgs_per_cu), dim3(grid),
                       dim3(256), 0, st, a);
  } else {
    // Wide queues first; a launch whose range is empty exits at once.
    const int plds = batch_probe_lds(ht->M);
    const dim3 pg((unsigned)pcap);
    hipLaunchKernelGGL(batch_probe_kernel_j<8>(ht->M), pg, dim3(256), plds, st, a);
    hipLaunchKernelGGL(batch_probe_kernel_j<4>(ht->M), pg, dim3(256), plds, st, a);
    mark(1);
    // Locality order of the candidate stage (counting sort, fixed grids).
    hipLaunchKernelGGL(k_batch_lzero, dim3(64), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_batch_lhist, dim3(egrid), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_batch_lscan, dim3(1), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_batch_lscatter, dim3(egrid), dim3(256), 0, st, a);
    mark(2);
    if (s.wgs_per_cu == 0) {
      // One workgroup per query position, chunks interleaved over the XCDs.
      const long long per = (bt->Qcap + 8ll * kLocChunk - 1) / (8ll * kLocChunk) * kLocChunk;
      hipLaunchKernelGGL(batch_cand_each_kernel(s.rows_in_flight), dim3((unsigned)(8 * per)),
                         dim3(256), 0, st, a);
    } else {
      hipLaunchKernelGGL(batch_cand_kernel(s.rows_in_flight), dim3(grid), dim3(256), 0, st, a);
    }
  }
  mark(3);
  hipLaunchKernelGGL(k_batch_overflow, dim3(culsh_bt::kOvGrid), dim3(256), 0, st, a, bt->dscr.p,
                     culsh_bt::kHbitsMax, culsh_bt::kUcapMax);
  const hipError_t m2 =
      hipMemcpyAsync(bt->hres, bt->dpack.p, sizeof(BatchHdr), hipMemcpyDeviceToHost, st);
  mark(4);
  hipError_t le = hipGetLastError();
  if (le == hipSuccess) le = m1 != hipSuccess ? m1 : m2;
  *err = le;
}

int bt_check_grid(culsh_bt* bt) {
  const culsh_schedule& s = bt->sched;
  const long long pcap = ((long long)(bt->Qcap + s.probe_queries - 1) / s.probe_queries +
                          std::min(bt->Rcap, 64)) *
                         ((bt->ht->L + s.probe_tables - 1) / s.probe_tables);
  if (pcap >= (1ll << 31)) return fail(CULSH_E_LIMIT, "batch probe grid too large");
  return CULSH_OK;
}

// Captures the plan's work into a graph on the capture stream.
int bt_capture(culsh_bt* bt) {
  TRY(bt_check_grid(bt));
  bt->drop_graph();
  HIP_TRY(hipStreamBeginCapture(bt->cap, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  bt_enqueue(bt, bt->cap, nullptr, &le);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(bt->cap, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(CULSH_E_DEVICE, "batch graph capture failed: %s",
                hipGetErrorString(le != hipSuccess ? le : ce));
  }
  bt->graph = g;
  HIP_TRY(hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
  bt->graph_sched = bt->sched;
  bt->graph_gen = bt->ht->data_gen;
  return CULSH_OK;
}

int bt_wait(culsh_bt* bt) {
  if (!bt->pending) return CULSH_OK;
  HIP_TRY(hipEventSynchronize(bt->ev1));
  bt->pending = false;
  const BatchHdr& r = *bt->hres;
  bt->rechecked = r.n_recheck;
  if (r.err & 1)
    return fail(CULSH_E_LIMIT,
                "batch: a query has more than %u non-unique candidates or more than %d "
                "candidates within the FP32 error bound of its K-th distance",
                culsh_bt::kUcapMax, kScapG);
  return CULSH_OK;
}

// Lays the batch out (batch order, classes, work items) into the pinned
// descriptor and makes sure the plan's buffers (and, with `capture`, its
// graph) fit it.  *empty is set when no request has queries.
int bt_prepare(culsh_bt* bt, const culsh_request* reqs, int nreq, const float* dQ, int* d_ids,
               float* d_dists, bool capture, bool* empty) {
  *empty = false;
  culsh_ht* ht = bt->ht;
  if (!ht->indexed || !ht->dX) return fail(CULSH_E_STATE, "the index holds no resident dataset");
  if (nreq <= 0 || !reqs) return fail(CULSH_E_ARG, "empty batch");
  if (!dQ || !d_ids || !d_dists) return fail(CULSH_E_ARG, "null device pointer");
  TRY(bt_wait(bt));
  TRY(ensure_rows16(ht));
  const culsh_schedule& s = bt->sched;
  // Batch order: T descending, then K descending, then request order.
  std::vector<int> ord;
  ord.reserve(nreq);
  long long Qsum = 0;
  for (int r = 0; r < nreq; ++r) {
    const culsh_request& q = reqs[r];
    if (q.Q < 0 || q.K <= 0 || q.T <= 0 || q.q_row < 0 || q.out_offset < 0)
      return fail(CULSH_E_ARG, "request %d: Q >= 0, K > 0, T > 0 and offsets >= 0 required", r);
    if (q.T > 64) return fail(CULSH_E_LIMIT, "request %d: T = %d exceeds 64 in a batch", r, q.T);
    if (q.Q > 0) ord.push_back(r);
    Qsum += q.Q;
  }
  if (Qsum >= (1ll << 30)) return fail(CULSH_E_LIMIT, "a batch holds 2^30 or more queries");
  if (ord.empty()) {
    *empty = true;
    return CULSH_OK;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
    if (reqs[x].T != reqs[y].T) return reqs[x].T > reqs[y].T;
    return reqs[x].K > reqs[y].K;
  });
  const int n = (int)ord.size();
  const int bq = s.queries_per_item;
  std::vector<BatchReq> vr(n);
  std::vector<BatchClass> vc;
  const int ntg = (ht->L + s.probe_tables - 1) / s.probe_tables;
  int b = 0, item = 0, pitem = 0;
  long long p = 0;
  for (int j = 0; j < n; ++j) {
    const culsh_request& q = reqs[ord[j]];
    if (vc.empty() || vc.back().T != q.T) {
      if (!vc.empty()) {
        item += (vc.back().n + bq - 1) / bq;
        pitem += (vc.back().n + s.probe_queries - 1) / s.probe_queries * ntg;
      }
      BatchClass c;
      c.b0 = b;
      c.n = 0;
      c.T = q.T;
      c.item0 = item;
      c.pitem0 = pitem;
      c.pad = 0;
      vc.push_back(c);
    }
    vc.back().n += q.Q;
    BatchReq& br = vr[j];
    br.qrow0 = q.q_row;
    br.out0 = q.out_offset;
    br.p0 = p;
    br.b0 = b;
    br.Q = q.Q;
    br.K = q.K;
    br.T = q.T;
    b += q.Q;
    p += (long long)q.Q * q.T;
  }
  item += (vc.back().n + bq - 1) / bq;
  pitem += (vc.back().n + s.probe_queries - 1) / s.probe_queries * ntg;
  TRY(bt_reserve(bt, b, n, p));
  // A graph holds the index's pointers: capture again after a re-index.
  if (capture &&
      (!bt->exec || !same_schedule(bt->graph_sched, s) || bt->graph_gen != ht->data_gen))
    TRY(bt_capture(bt));
  BatchHdr h;
  std::memset(&h, 0, sizeof(h));
  h.Qm = dQ;
  h.out_ids = d_ids;
  h.out_d = d_dists;
  h.Ptot = bt->Pcap;
  h.nreq = n;
  h.nclass = (int)vc.size();
  h.Qtot = b;
  h.nitems = item;
  h.nprobe = pitem;
  // Classes come in T-descending order, so each queue width's probe items
  // are one contiguous range.
  for (int k = 0; k < 4; ++k) {
    h.pj_lo[k] = 0;
    h.pj_hi[k] = 0;
  }
  for (size_t c = 0; c < vc.size(); ++c) {
    const int k = batch_ji(vc[c].T);
    const int lo = vc[c].pitem0;
    const int hi = c + 1 < vc.size() ? vc[c + 1].pitem0 : pitem;
    if (h.pj_hi[k] == h.pj_lo[k]) h.pj_lo[k] = lo;
    h.pj_hi[k] = hi;
  }
  h.qn = s.probe_queries;
  h.tn = s.probe_tables;
  h.bq = bq;
  std::memcpy(bt->hpack, &h, sizeof(h));
  std::memcpy(bt->hpack + bt->off_cls, vc.data(), vc.size() * sizeof(BatchClass));
  std::memcpy(bt->hpack + bt->off_req, vr.data(), vr.size() * sizeof(BatchReq));
  return CULSH_OK;
}

int bt_launch(culsh_bt* bt, const culsh_request* reqs, int nreq, const float* dQ, int* d_ids,
              float* d_dists, hipStream_t stream) {
  bool empty = false;
  TRY(bt_prepare(bt, reqs, nreq, dQ, d_ids, d_dists, true, &empty));
  if (empty) return CULSH_OK;
  HIP_TRY(hipEventRecord(bt->ev0, stream));
  HIP_TRY(hipGraphLaunch(bt->exec, stream));
  HIP_TRY(hipEventRecord(bt->ev1, stream));
  bt->pending = true;
  return CULSH_OK;
}

}  // namespace

extern "C" {

int culsh_bt_create(culsh_ht* ht, culsh_bt** out) {
  if (!out) return fail(CULSH_E_ARG, "null output pointer");
  if (!ht) return fail(CULSH_E_ARG, "null HashTables");
  auto* bt = new culsh_bt;
  bt->ht = ht;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&bt->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&bt->cap, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&bt->ev0);
  if (e == hipSuccess) e = hipEventCreate(&bt->ev1);
  if (e == hipSuccess) e = hipHostMalloc((void**)&bt->hres, sizeof(BatchHdr), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete bt;
    return fail(CULSH_E_DEVICE, "batch setup: %s", hipGetErrorString(e));
  }
  *out = bt;
  return CULSH_OK;
}

void culsh_bt_destroy(culsh_bt* bt) { delete bt; }

int culsh_bt_set_schedule(culsh_bt* bt, const culsh_schedule* s) {
  if (!bt || !s) return fail(CULSH_E_ARG, "null argument");
  TRY(check_schedule(*s));
  TRY(bt_wait(bt));
  bt->sched = *s;
  return CULSH_OK;
}

int culsh_bt_get_schedule(culsh_bt* bt, culsh_schedule* s) {
  if (!bt || !s) return fail(CULSH_E_ARG, "null argument");
  *s = bt->sched;
  return CULSH_OK;
}

int culsh_bt_search_device(culsh_bt* bt, const culsh_request* reqs, int nreq,
                           const float* d_queries, int* d_ids, float* d_dists, void* stream) {
  if (!bt) return fail(CULSH_E_ARG, "null batch");
  return bt_launch(bt, reqs, nreq, d_queries, d_ids, d_dists, (hipStream_t)stream);
}

int culsh_bt_wait(culsh_bt* bt) {
  if (!bt) return fail(CULSH_E_ARG, "null batch");
  return bt_wait(bt);
}

int culsh_bt_last_ms(culsh_bt* bt, float* ms) {
  if (!bt || !ms) return fail(CULSH_E_ARG, "null argument");
  TRY(bt_wait(bt));
  HIP_TRY(hipEventElapsedTime(ms, bt->ev0, bt->ev1));
  return CULSH_OK;
}

unsigned long long culsh_bt_rechecked(culsh_bt* bt) { return bt ? bt->rechecked : 0ull; }

int culsh_bt_profile(culsh_bt* bt, const culsh_request* reqs, int nreq, const float* d_queries,
                     int* d_ids, float* d_dists, void* stream, float* stage_ms) {
  if (!bt || !stage_ms) return fail(CULSH_E_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  bool empty = false;
  TRY(bt_prepare(bt, reqs, nreq, d_queries, d_ids, d_dists, false, &empty));
  for (int i = 0; i < 4; ++i) stage_ms[i] = 0.f;
  if (empty) return CULSH_OK;
  TRY(bt_check_grid(bt));
  hipEvent_t ev[5] = {};
  auto done = [&](int rc) {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    return rc;
  };
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return done(fail(CULSH_E_DEVICE, "hipEventCreate"));
  hipError_t le = hipSuccess;
  bt_enqueue(bt, st, ev, &le);
  if (le != hipSuccess) return done(fail(CULSH_E_DEVICE, "batch launch: %s", hipGetErrorString(le)));
  if (hipEventSynchronize(ev[4]) != hipSuccess) return done(fail(CULSH_E_DEVICE, "batch sync"));
  for (int i = 0; i < 4; ++i)
    if (hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]) != hipSuccess)
      return done(fail(CULSH_E_DEVICE, "hipEventElapsedTime"));
  const BatchHdr& r = *bt->hres;
  bt->rechecked = r.n_recheck;
  if (r.err & 1)
    return done(fail(CULSH_E_LIMIT, "batch: a query exceeded the candidate capacities"));
  return done(CULSH_OK);
}

int culsh_bt_autotune(culsh_bt* bt, const culsh_request* reqs, int nreq, const float* d_queries,
                      int* d_ids, float* d_dists, int reps, float* default_ms, float* tuned_ms) {
  if (!bt) return fail(CULSH_E_ARG, "null batch");
  reps = std::max(1, reps);
  auto timed = [&](const culsh_schedule& s, float* best) -> int {
    TRY(culsh_bt_set_schedule(bt, &s));
    *best = INFINITY;
    for (int r = 0; r < reps + 1; ++r) {  // the first run also captures the graph
      TRY(bt_launch(bt, reqs, nreq, d_queries, d_ids, d_dists, bt->cap));
      float ms = 0.f;
      TRY(culsh_bt_last_ms(bt, &ms));
      if (r > 0) *best = std::min(*best, ms);
    }
    return CULSH_OK;
  };
  // The starting point is a fixed naive schedule (one fused kernel, 32x1
  // probe tile, 32 queries per item, 4 rows in flight, one workgroup per
  // CU), not the library's default, so tuned vs default measures the search.
  const culsh_schedule def{32, 1, 32, 4, 1, 1};
  culsh_schedule best = def;
  float tbest = 0.f;
  TRY(timed(def, &tbest));
  if (default_ms) *default_ms = tbest;
  const int tiles[3][2] = {{32, 1}, {16, 2}, {8, 4}};
  const int qpis[4] = {8, 16, 32, 64};
  const int rows[3] = {2, 4, 8};
  const int wgs[4] = {0, 1, 2, 3};
  bool improved = false;
  auto consider = [&](const culsh_schedule& s) -> int {
    if (same_schedule(s, best)) return CULSH_OK;
    float t = 0.f;
    TRY(timed(s, &t));
    if (t < 0.99f * tbest) {
      tbest = t;
      best = s;
      improved = true;
    }
    return CULSH_OK;
  };
  // Coordinate sweeps until one changes nothing (at most three).
  for (int pass = 0; pass < 3; ++pass) {
    improved = false;
    for (int f = 0; f < 2; ++f) {
      culsh_schedule s = best;
      s.fused = f;
      if (f && s.wgs_per_cu == 0) s.wgs_per_cu = 3;
      TRY(consider(s));
    }
    for (int v : wgs) {
      culsh_schedule s = best;
      s.wgs_per_cu = v;
      if (s.fused && v == 0) continue;
      TRY(consider(s));
    }
    for (auto& tl : tiles) {
      culsh_schedule s = best;
      s.probe_queries = tl[0];
      s.probe_tables = tl[1];
      s.queries_per_item = std::max(tl[0], s.queries_per_item / tl[0] * tl[0]);
      TRY(consider(s));
    }
    for (int v : qpis) {
      culsh_schedule s = best;
      if (v < s.probe_queries) continue;
      s.queries_per_item = v;
      TRY(consider(s));
    }
    for (int v : rows) {
      culsh_schedule s = best;
      s.rows_in_flight = v;
      TRY(consider(s));
    }
    if (!improved) break;
  }
  TRY(culsh_bt_set_schedule(bt, &best));
  if (tuned_ms) *tuned_ms = tbest;
  return CULSH_OK;
}

}  // extern "C"
