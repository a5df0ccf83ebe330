// Serving form (culsh_srv_*): a real-time request loop over device-resident
// indexes with continuous batching.  Part of libculsh's single translation
// unit (included by culsh.hip after the batching form, whose plans and
// launch path it drives).
//
// Each index owns `slots_per_index` slots; a slot is a culsh_bt with its own
// stream, plan buffers and a captured graph, sized at creation from the free
// HBM.  The loop moves arrived requests into per-index FIFO queues; a free
// slot takes every queued request of its index that still fits its
// capacity (queries, probe slots, requests) -- head-of-line order, so a
// request that does not fit behind others waits for the next free slot --
// and launches the batch at once.  A request larger than an empty slot
// grows that slot when the HBM free at that moment allows it, or else runs
// in pieces over successive batches.  Batches of different slots overlap on
// the device.  Completion times come from the slots' HIP events, measured
// against an event recorded when the run starts, so latencies are on the
// GPU timeline; with host IO they include the query upload and the result
// download of each batch.
#pragma once
#include <chrono>
#include <deque>

struct culsh_srv {
  struct Piece {
    int r;       // request
    int q0, nq;  // its queries [q0, q0 + nq)
  };
  struct Slot {
    culsh_bt* bt = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t io_done = nullptr;  // after the result download (host IO runs)
    bool busy = false;
    int Qcap = 0, Rcap = 0;  // this slot's capacities (it may grow)
    long long Pcap = 0;
    std::vector<Piece> pieces;  // the batch in flight
  };
  std::vector<culsh_ht*> ht;
  culsh_srv_config cfg{};
  std::vector<std::vector<Slot>> slots;  // [index][slot]
  int maxL = 0;
  long long slot_bytes = 0;  // HBM of one slot at its initial capacity
  // Plain mode: one SearchTables per index.
  std::vector<culsh_st*> st;
  hipStream_t plain_stream = nullptr;
  hipEvent_t start = nullptr, plain_done = nullptr;

  ~culsh_srv() {
    for (auto& v : slots)
      for (auto& s : v) {
        if (s.bt) culsh_bt_destroy(s.bt);
        if (s.io_done) (void)hipEventDestroy(s.io_done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
      }
    for (auto* p : st) culsh_st_destroy(p);
    if (plain_stream) (void)hipStreamDestroy(plain_stream);
    if (start) (void)hipEventDestroy(start);
    if (plain_done) (void)hipEventDestroy(plain_done);
  }
};

namespace {

// HBM one slot reserves for Qcap queries and Pcap probe slots per table
// (the plan buffers of bt_reserve plus the overflow scratch).
long long slot_bytes_for(int L, long long Qcap, long long Pcap, int Rcap) {
  const long long ovf = (long long)culsh_bt::kOvGrid *
                        ((1ll << culsh_bt::kHbitsMax) + 2ll * culsh_bt::kUcapMax) * 4;
  return Qcap * ((long long)sizeof(BatchQuery) + 8ll * L + 8 + 8 + 4) + 8ll * L * Pcap +
         (long long)Rcap * (sizeof(BatchReq) + sizeof(BatchClass)) + ovf;
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p * (double)(v.size() - 1);
  const size_t lo = (size_t)pos;
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

struct SrvCounters {
  long long batches = 0, batched_pieces = 0, deferred = 0, split = 0, grown = 0;
};

void fill_stats(const std::vector<double>& lat, const culsh_srv_request* reqs, int n,
                double last_done_ms, const SrvCounters& c, long long slot_bytes, bool host_io,
                culsh_srv_stats* s) {
  std::memset(s, 0, sizeof(*s));
  s->p50_ms = pct(lat, 0.50);
  s->p99_ms = pct(lat, 0.99);
  double sum = 0.0, mx = 0.0;
  long long q = 0;
  for (int r = 0; r < n; ++r) {
    sum += lat[r];
    mx = std::max(mx, lat[r]);
    q += reqs[r].Q;
  }
  s->mean_ms = n ? sum / n : 0.0;
  s->max_ms = mx;
  const double span_s = (last_done_ms - (n ? reqs[0].arrival_s * 1e3 : 0.0)) * 1e-3;
  s->queries_per_s = span_s > 0 ? (double)q / span_s : 0.0;
  s->requests_per_s = span_s > 0 ? (double)n / span_s : 0.0;
  s->batches = c.batches;
  s->queries = q;
  s->requests = n;
  s->mean_batch_requests = c.batches ? (double)c.batched_pieces / (double)c.batches : 0.0;
  s->slot_bytes = slot_bytes;
  s->deferred = c.deferred;
  s->split_requests = c.split;
  s->grown_slots = c.grown;
  s->host_io = host_io ? 1 : 0;
}

int srv_check(culsh_srv* srv, const culsh_srv_request* reqs, int n) {
  double prev = -1.0;
  for (int r = 0; r < n; ++r) {
    const culsh_srv_request& q = reqs[r];
    if (q.index < 0 || q.index >= (int)srv->ht.size())
      return fail(CULSH_E_ARG, "request %d: index %d out of range", r, q.index);
    if (q.Q <= 0 || q.K <= 0 || q.T <= 0 || q.T > 64 || q.q_row < 0 || q.out_offset < 0)
      return fail(CULSH_E_ARG, "request %d: Q, K > 0, 0 < T <= 64, offsets >= 0 required", r);
    if (q.arrival_s < prev) return fail(CULSH_E_ARG, "requests must be sorted by arrival");
    prev = q.arrival_s;
  }
  return CULSH_OK;
}

// Host arrays of a run with query upload and result download (all null for
// a run on device-resident queries and results).
struct HostIO {
  const float* const* q = nullptr;
  int* const* ids = nullptr;
  float* const* dists = nullptr;
  bool on() const { return q != nullptr; }
};

// Query rows [q0, q0 + nq) of request r: host -> device.
int io_upload(const HostIO& io, const culsh_srv_request& q, int D, int q0, int nq,
              const float* const* dQ, hipStream_t s) {
  const size_t off = (size_t)(q.q_row + q0) * D;
  HIP_TRY(hipMemcpyAsync(const_cast<float*>(dQ[q.index]) + off, io.q[q.index] + off,
                         (size_t)nq * D * 4, hipMemcpyHostToDevice, s));
  return CULSH_OK;
}

// Results of queries [q0, q0 + nq) of request r: device -> host.
int io_download(const HostIO& io, const culsh_srv_request& q, int q0, int nq,
                int* const* d_ids, float* const* d_dists, hipStream_t s) {
  const size_t off = (size_t)q.out_offset + (size_t)q0 * q.K, cnt = (size_t)nq * q.K;
  HIP_TRY(hipMemcpyAsync(io.ids[q.index] + off, d_ids[q.index] + off, cnt * 4,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(io.dists[q.index] + off, d_dists[q.index] + off, cnt * 4,
                         hipMemcpyDeviceToHost, s));
  return CULSH_OK;
}

// The uploads and downloads of one batch's pieces (all of index i), merged
// into contiguous ranges: requests that arrived in sequence usually hold
// adjacent query rows and adjacent result slots.
int io_batch(const HostIO& io, const culsh_srv_request* reqs,
             const std::vector<culsh_srv::Piece>& pieces, int i, int D, const float* const* dQ,
             int* const* d_ids, float* const* d_dists, hipStream_t s, bool upload) {
  long long a0 = -1, n = 0;  // current range (rows, or result elements)
  auto flush = [&]() -> int {
    if (n <= 0) return CULSH_OK;
    if (upload) {
      HIP_TRY(hipMemcpyAsync(const_cast<float*>(dQ[i]) + a0 * D, io.q[i] + a0 * D,
                             (size_t)n * D * 4, hipMemcpyHostToDevice, s));
    } else {
      HIP_TRY(hipMemcpyAsync(io.ids[i] + a0, d_ids[i] + a0, (size_t)n * 4,
                             hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(io.dists[i] + a0, d_dists[i] + a0, (size_t)n * 4,
                             hipMemcpyDeviceToHost, s));
    }
    return CULSH_OK;
  };
  for (const auto& pc : pieces) {
    const culsh_srv_request& q = reqs[pc.r];
    const long long b = upload ? q.q_row + pc.q0 : q.out_offset + (long long)pc.q0 * q.K;
    const long long m = upload ? pc.nq : (long long)pc.nq * q.K;
    if (n > 0 && b == a0 + n) {
      n += m;
    } else {
      TRY(flush());
      a0 = b;
      n = m;
    }
  }
  return flush();
}

// One request at a time through the search path (the plain baseline).
int srv_run_plain(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                  const float* const* dQ, int* const* d_ids, float* const* d_dists,
                  const HostIO& io, std::vector<double>& lat, double* last_done_ms) {
  using clk = std::chrono::steady_clock;
  hipStream_t s = srv->plain_stream;
  HIP_TRY(hipEventRecord(srv->start, s));
  HIP_TRY(hipEventSynchronize(srv->start));
  const auto t0 = clk::now();
  for (int r = 0; r < n; ++r) {
    const culsh_srv_request& q = reqs[r];
    while (std::chrono::duration<double>(clk::now() - t0).count() < q.arrival_s) {
    }
    culsh_st* st = srv->st[q.index];
    culsh_ht* ht = srv->ht[q.index];
    TRY(culsh_st_reset(st, ht, q.K, q.T));
    if (io.on()) TRY(io_upload(io, q, ht->D, 0, q.Q, dQ, s));
    TRY(run_search(st, dQ[q.index] + q.q_row * ht->D, ht->dX, q.Q, d_ids[q.index] + q.out_offset,
                   d_dists[q.index] + q.out_offset, s, false));
    if (io.on()) TRY(io_download(io, q, 0, q.Q, d_ids, d_dists, s));
    HIP_TRY(hipEventRecord(srv->plain_done, s));
    HIP_TRY(hipEventSynchronize(srv->plain_done));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, srv->start, srv->plain_done));
    lat[r] = (double)ms - q.arrival_s * 1e3;
    *last_done_ms = ms;
  }
  return CULSH_OK;
}

// Waits for every slot's batch in flight and frees the slots, keeping the
// current error message: a run starts, and a failed run ends, with no batch
// of it left on the device.
void srv_drain(culsh_srv* srv) {
  const std::string keep = g_err;
  for (auto& v : srv->slots)
    for (auto& sl : v) {
      if (sl.bt && sl.bt->pending) (void)bt_wait(sl.bt);
      if (sl.busy && sl.stream) (void)hipStreamSynchronize(sl.stream);
      sl.busy = false;
      sl.pieces.clear();
    }
  g_err = keep;
}

// Grows a free slot so that one batch can hold Qn queries and Pn probe
// slots per table, when the HBM free at this moment (times the configured
// fraction) covers the extra buffers.  Returns false (nothing changed) when
// it does not, or the probe slots would exceed the 31-bit addressing.
bool srv_try_grow(culsh_srv* srv, culsh_srv::Slot& sl, long long Qn, long long Pn, int* rc) {
  *rc = CULSH_OK;
  if (srv->cfg.no_growth) return false;
  const int L = sl.bt->ht->L;
  Qn = std::max<long long>(Qn, 2ll * sl.Qcap);
  Pn = std::max<long long>(Pn, 2ll * sl.Pcap);
  if (Qn >= (1ll << 30) || (long long)L * Pn >= (1ll << 31)) {
    Qn = std::min<long long>(Qn, (1ll << 30) - 1);
    Pn = std::min<long long>(Pn, ((1ll << 31) - 1) / L);
  }
  if (Qn <= sl.Qcap && Pn <= sl.Pcap) return false;
  const int Rn = std::max(sl.Rcap, (int)std::min<long long>(Qn, 4096));
  const long long extra = slot_bytes_for(L, Qn, Pn, Rn) - slot_bytes_for(L, sl.Qcap, sl.Pcap, sl.Rcap);
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return false;
  if ((double)extra > (double)freeb * srv->cfg.hbm_fraction) return false;
  *rc = bt_reserve(sl.bt, (int)Qn, Rn, Pn);
  if (*rc == CULSH_OK) *rc = bt_capture(sl.bt);
  if (*rc != CULSH_OK) return false;
  sl.Qcap = (int)Qn;
  sl.Pcap = Pn;
  sl.Rcap = Rn;
  return true;
}

int srv_run_batched_loop(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                         const float* const* dQ, int* const* d_ids, float* const* d_dists,
                         const HostIO& io, std::vector<double>& lat, double* last_done_ms,
                         SrvCounters& c) {
  using clk = std::chrono::steady_clock;
  const int nidx = (int)srv->ht.size();
  std::vector<std::deque<culsh_srv::Piece>> queue(nidx);
  std::vector<int> left(n);  // queries of a request not yet completed
  for (int r = 0; r < n; ++r) left[r] = reqs[r].Q;
  std::vector<culsh_request> breq;
  HIP_TRY(hipEventRecord(srv->start, srv->slots[0][0].stream));
  HIP_TRY(hipEventSynchronize(srv->start));
  const auto t0 = clk::now();
  int next = 0, done = 0;
  while (done < n) {
    const double now = std::chrono::duration<double>(clk::now() - t0).count();
    while (next < n && reqs[next].arrival_s <= now) {
      queue[reqs[next].index].push_back({next, 0, reqs[next].Q});
      ++next;
    }
    // Completions: a request is done when its last piece is.
    for (int i = 0; i < nidx; ++i)
      for (auto& sl : srv->slots[i]) {
        if (!sl.busy) continue;
        hipEvent_t ev = io.on() ? sl.io_done : sl.bt->ev1;
        const hipError_t q = hipEventQuery(ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) HIP_TRY(q);
        TRY(bt_wait(sl.bt));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, srv->start, ev));
        for (const auto& pc : sl.pieces) {
          left[pc.r] -= pc.nq;
          if (left[pc.r] == 0) {
            lat[pc.r] = (double)ms - reqs[pc.r].arrival_s * 1e3;
            ++done;
          }
        }
        *last_done_ms = std::max(*last_done_ms, (double)ms);
        sl.pieces.clear();
        sl.busy = false;
      }
    // Dispatch: every free slot takes what fits of its index's queue, in
    // arrival order.  A request that does not fit behind others waits for
    // the next free slot; one that does not fit an empty slot grows the
    // slot (when the free HBM allows) or runs in pieces over several
    // batches (results per query are independent of the batch).
    for (int i = 0; i < nidx; ++i)
      for (auto& sl : srv->slots[i]) {
        if (sl.busy || queue[i].empty()) continue;
        long long qsum = 0, psum = 0;
        breq.clear();
        while (!queue[i].empty()) {
          culsh_srv::Piece& pc = queue[i].front();
          const culsh_srv_request& q = reqs[pc.r];
          long long fit = std::min<long long>(sl.Qcap - qsum, (sl.Pcap - psum) / q.T);
          if ((int)breq.size() + 1 > sl.Rcap) fit = 0;
          int take = 0;
          if (fit >= pc.nq) {
            take = pc.nq;
          } else if (breq.empty()) {
            int rc = CULSH_OK;
            if (srv_try_grow(srv, sl, pc.nq, (long long)pc.nq * q.T, &rc)) {
              ++c.grown;
              continue;  // re-evaluate against the new capacity
            }
            TRY(rc);
            take = (int)std::max<long long>(1, fit);
            ++c.split;
          } else {
            ++c.deferred;
            break;
          }
          culsh_request b;
          b.Q = take;
          b.K = q.K;
          b.T = q.T;
          b.reserved = 0;
          b.q_row = q.q_row + pc.q0;
          b.out_offset = q.out_offset + (long long)pc.q0 * q.K;
          breq.push_back(b);
          sl.pieces.push_back({pc.r, pc.q0, take});
          qsum += take;
          psum += (long long)take * q.T;
          if (take == pc.nq) {
            queue[i].pop_front();
          } else {
            pc.q0 += take;
            pc.nq -= take;
            break;
          }
        }
        if (breq.empty()) continue;
        sl.busy = true;  // from here on the slot holds work of this run
        if (io.on())
          TRY(io_batch(io, reqs, sl.pieces, i, srv->ht[i]->D, dQ, d_ids, d_dists, sl.stream,
                       true));
        TRY(bt_launch(sl.bt, breq.data(), (int)breq.size(), dQ[i], d_ids[i], d_dists[i],
                      sl.stream));
        if (io.on()) {
          TRY(io_batch(io, reqs, sl.pieces, i, srv->ht[i]->D, dQ, d_ids, d_dists, sl.stream,
                       false));
          HIP_TRY(hipEventRecord(sl.io_done, sl.stream));
        }
        ++c.batches;
        c.batched_pieces += (long long)breq.size();
      }
    std::this_thread::yield();
  }
  return CULSH_OK;
}

int srv_run_batched(culsh_srv* srv, const culsh_srv_request* reqs, int n, const float* const* dQ,
                    int* const* d_ids, float* const* d_dists, const HostIO& io,
                    std::vector<double>& lat, double* last_done_ms, SrvCounters& c) {
  srv_drain(srv);
  const int rc =
      srv_run_batched_loop(srv, reqs, n, dQ, d_ids, d_dists, io, lat, last_done_ms, c);
  if (rc != CULSH_OK) srv_drain(srv);
  return rc;
}

int srv_run(culsh_srv* srv, const culsh_srv_request* reqs, int n, const float* const* d_queries,
            int* const* d_ids, float* const* d_dists, const HostIO& io, double* latency_ms,
            culsh_srv_stats* stats) {
  if (!srv || (!reqs && n > 0) || !d_queries || !d_ids || !d_dists)
    return fail(CULSH_E_ARG, "invalid arguments");
  if (n == 0) return CULSH_OK;
  TRY(srv_check(srv, reqs, n));
  std::vector<double> lat(n, 0.0);
  double last = 0.0;
  SrvCounters c;
  if (srv->cfg.batching) {
    TRY(srv_run_batched(srv, reqs, n, d_queries, d_ids, d_dists, io, lat, &last, c));
  } else {
    TRY(srv_run_plain(srv, reqs, n, d_queries, d_ids, d_dists, io, lat, &last));
    c.batches = n;
    c.batched_pieces = n;
  }
  if (latency_ms) std::copy(lat.begin(), lat.end(), latency_ms);
  if (stats) fill_stats(lat, reqs, n, last, c, srv->slot_bytes, io.on(), stats);
  return CULSH_OK;
}

}  // namespace

extern "C" {

int culsh_srv_create(culsh_ht* const* indexes, int nindex, const culsh_srv_config* cfg,
                     culsh_srv** out) {
  if (!out || !indexes || nindex <= 0 || !cfg) return fail(CULSH_E_ARG, "invalid arguments");
  if (cfg->slots_per_index < 1 || cfg->hbm_fraction <= 0.0 || cfg->hbm_fraction > 1.0 ||
      cfg->max_batch_queries < 0)
    return fail(CULSH_E_ARG, "slots_per_index >= 1, 0 < hbm_fraction <= 1 required");
  for (int i = 0; i < nindex; ++i)
    if (!indexes[i] || !indexes[i]->indexed || !indexes[i]->dX)
      return fail(CULSH_E_STATE, "index %d holds no resident dataset", i);
  auto* srv = new culsh_srv;
  srv->ht.assign(indexes, indexes + nindex);
  srv->cfg = *cfg;
  auto bail = [&](int rc) {
    std::string keep = g_err;
    delete srv;
    g_err = keep;
    return rc;
  };
  hipError_t e = hipEventCreate(&srv->start);
  if (e == hipSuccess) e = hipEventCreate(&srv->plain_done);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&srv->plain_stream, hipStreamNonBlocking);
  if (e != hipSuccess)
    return bail(fail(CULSH_E_DEVICE, "server setup: %s", hipGetErrorString(e)));
  int maxL = 0;
  for (auto* h : srv->ht) maxL = std::max(maxL, h->L);
  if (cfg->batching) {
    // The indexes' FP16 row copies (the batches' distance pass) come first,
    // so the slots are sized from what is left.
    for (auto* h : srv->ht) {
      const int rc = ensure_rows16(h);
      if (rc != CULSH_OK) return bail(rc);
    }
    // Initial slot capacity from the free HBM now: Q queries and 32 Q probe
    // slots per table.  Slots grow later by the free HBM at that moment.
    size_t freeb = 0, totalb = 0;
    e = hipMemGetInfo(&freeb, &totalb);
    if (e != hipSuccess) return bail(fail(CULSH_E_DEVICE, "hipMemGetInfo: %s", hipGetErrorString(e)));
    const int nslots = nindex * cfg->slots_per_index;
    const long long share = (long long)((double)freeb * cfg->hbm_fraction) / nslots;
    // Probe slots of all tables are addressed with 31-bit offsets.
    long long Q = std::min<long long>(cfg->max_batch_queries > 0 ? cfg->max_batch_queries : 4096,
                                      ((1ll << 31) - 1) / (32ll * maxL));
    while (Q > 1 && slot_bytes_for(maxL, Q, 32 * Q, (int)std::min<long long>(Q, 4096)) > share)
      Q /= 2;
    const long long need = slot_bytes_for(maxL, Q, 32 * Q, (int)std::min<long long>(Q, 4096));
    if (need > share)
      return bail(fail(CULSH_E_MEMORY, "free HBM share %lld B is below one slot's minimum",
                       share));
    srv->maxL = maxL;
    srv->slot_bytes = need;
    srv->slots.resize(nindex);
    for (int i = 0; i < nindex; ++i) {
      srv->slots[i].resize(cfg->slots_per_index);
      for (auto& sl : srv->slots[i]) {
        int rc = culsh_bt_create(srv->ht[i], &sl.bt);
        if (rc != CULSH_OK) return bail(rc);
        e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&sl.io_done);
        if (e != hipSuccess) return bail(fail(CULSH_E_DEVICE, "%s", hipGetErrorString(e)));
        sl.Qcap = (int)Q;
        sl.Pcap = 32 * Q;
        sl.Rcap = (int)std::min<long long>(Q, 4096);
        rc = bt_reserve(sl.bt, sl.Qcap, sl.Rcap, sl.Pcap);
        if (rc == CULSH_OK) rc = bt_capture(sl.bt);
        if (rc != CULSH_OK) return bail(rc);
      }
    }
  } else {
    srv->st.resize(nindex, nullptr);
    for (int i = 0; i < nindex; ++i) {
      int rc = culsh_st_create(&srv->st[i]);
      if (rc == CULSH_OK) rc = culsh_st_reset(srv->st[i], srv->ht[i], 1, 1);
      if (rc == CULSH_OK) rc = st_setup(srv->st[i]);
      if (rc != CULSH_OK) return bail(rc);
    }
  }
  *out = srv;
  return CULSH_OK;
}

void culsh_srv_destroy(culsh_srv* srv) { delete srv; }

int culsh_srv_run(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                  const float* const* d_queries, int* const* d_ids, float* const* d_dists,
                  double* latency_ms, culsh_srv_stats* stats) {
  return srv_run(srv, reqs, n, d_queries, d_ids, d_dists, HostIO{}, latency_ms, stats);
}

int culsh_srv_run_host(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                       const float* const* h_queries, int* const* h_ids, float* const* h_dists,
                       const float* const* d_queries, int* const* d_ids, float* const* d_dists,
                       double* latency_ms, culsh_srv_stats* stats) {
  if (!h_queries || !h_ids || !h_dists) return fail(CULSH_E_ARG, "null host arrays");
  HostIO io;
  io.q = h_queries;
  io.ids = h_ids;
  io.dists = h_dists;
  return srv_run(srv, reqs, n, d_queries, d_ids, d_dists, io, latency_ms, stats);
}

}  // extern "C"
