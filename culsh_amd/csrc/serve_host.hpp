// Serving form (culsh_srv_*): a real-time request loop over device-resident
// indexes with continuous batching.  Part of libculsh's single translation
// unit (included by culsh.hip after the batching form, whose plans and
// launch path it drives).
//
// Each index owns `slots_per_index` slots; a slot is a culsh_bt with its own
// stream, buffers sized once at creation from the free HBM, and a captured
// graph.  The loop moves arrived requests into per-index FIFO queues; a free
// slot takes every queued request of its index that still fits its
// capacity (queries, probe slots, requests) -- head-of-line order, so a
// request that does not fit waits for the next free slot -- and launches the
// batch at once.  Batches of different slots overlap on the device.
// Completion times come from the slots' HIP events, measured against an
// event recorded when the run starts, so latencies are on the GPU timeline.
#pragma once
#include <chrono>
#include <deque>

struct culsh_srv {
  struct Slot {
    culsh_bt* bt = nullptr;
    hipStream_t stream = nullptr;
    bool busy = false;
    std::vector<int> reqs;  // request ids of the batch in flight
  };
  std::vector<culsh_ht*> ht;
  culsh_srv_config cfg{};
  std::vector<std::vector<Slot>> slots;  // [index][slot]
  int Qcap = 0, Rcap = 0;
  long long Pcap = 0, slot_bytes = 0;
  // Plain mode: one SearchTables per index.
  std::vector<culsh_st*> st;
  hipStream_t plain_stream = nullptr;
  hipEvent_t start = nullptr, plain_done = nullptr;

  ~culsh_srv() {
    for (auto& v : slots)
      for (auto& s : v) {
        if (s.bt) culsh_bt_destroy(s.bt);
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
  return Qcap * ((long long)sizeof(BatchQuery) + 8ll * L + 8 + 8 + 4) + 4ll * L * Pcap +
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

void fill_stats(const std::vector<double>& lat, const culsh_srv_request* reqs, int n,
                double last_done_ms, long long batches, long long batched_reqs, long long deferred,
                long long slot_bytes, culsh_srv_stats* s) {
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
  s->batches = batches;
  s->queries = q;
  s->requests = n;
  s->mean_batch_requests = batches ? (double)batched_reqs / (double)batches : 0.0;
  s->slot_bytes = slot_bytes;
  s->deferred = deferred;
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
    if (srv->cfg.batching &&
        (q.Q > srv->Qcap || (long long)q.Q * q.T > srv->Pcap))
      return fail(CULSH_E_LIMIT, "request %d (Q=%d, T=%d) exceeds a slot's capacity", r, q.Q,
                  q.T);
  }
  return CULSH_OK;
}

// One request at a time through the search path (the plain baseline).
int srv_run_plain(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                  const float* const* dQ, int* const* d_ids, float* const* d_dists,
                  std::vector<double>& lat, double* last_done_ms) {
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
    TRY(run_search(st, dQ[q.index] + q.q_row * ht->D, ht->dX, q.Q, d_ids[q.index] + q.out_offset,
                   d_dists[q.index] + q.out_offset, s, false));
    HIP_TRY(hipEventRecord(srv->plain_done, s));
    HIP_TRY(hipEventSynchronize(srv->plain_done));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, srv->start, srv->plain_done));
    lat[r] = (double)ms - q.arrival_s * 1e3;
    *last_done_ms = ms;
  }
  return CULSH_OK;
}

int srv_run_batched(culsh_srv* srv, const culsh_srv_request* reqs, int n,
                    const float* const* dQ, int* const* d_ids, float* const* d_dists,
                    std::vector<double>& lat, double* last_done_ms, long long* batches,
                    long long* batched_reqs, long long* deferred) {
  using clk = std::chrono::steady_clock;
  const int nidx = (int)srv->ht.size();
  std::vector<std::deque<int>> queue(nidx);
  std::vector<culsh_request> breq;
  HIP_TRY(hipEventRecord(srv->start, srv->slots[0][0].stream));
  HIP_TRY(hipEventSynchronize(srv->start));
  const auto t0 = clk::now();
  int next = 0, done = 0;
  while (done < n) {
    const double now = std::chrono::duration<double>(clk::now() - t0).count();
    while (next < n && reqs[next].arrival_s <= now) {
      queue[reqs[next].index].push_back(next);
      ++next;
    }
    // Completions.
    for (int i = 0; i < nidx; ++i)
      for (auto& sl : srv->slots[i]) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.bt->ev1);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) HIP_TRY(q);
        TRY(bt_wait(sl.bt));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, srv->start, sl.bt->ev1));
        for (int r : sl.reqs) lat[r] = (double)ms - reqs[r].arrival_s * 1e3;
        *last_done_ms = std::max(*last_done_ms, (double)ms);
        done += (int)sl.reqs.size();
        sl.reqs.clear();
        sl.busy = false;
      }
    // Dispatch: every free slot takes what fits of its index's queue.
    for (int i = 0; i < nidx; ++i)
      for (auto& sl : srv->slots[i]) {
        if (sl.busy || queue[i].empty()) continue;
        long long qsum = 0, psum = 0;
        breq.clear();
        while (!queue[i].empty()) {
          const culsh_srv_request& q = reqs[queue[i].front()];
          if (qsum + q.Q > srv->Qcap || psum + (long long)q.Q * q.T > srv->Pcap ||
              (int)breq.size() + 1 > srv->Rcap) {
            ++*deferred;
            break;
          }
          culsh_request b;
          b.Q = q.Q;
          b.K = q.K;
          b.T = q.T;
          b.reserved = 0;
          b.q_row = q.q_row;
          b.out_offset = q.out_offset;
          breq.push_back(b);
          sl.reqs.push_back(queue[i].front());
          queue[i].pop_front();
          qsum += q.Q;
          psum += (long long)q.Q * q.T;
        }
        TRY(bt_launch(sl.bt, breq.data(), (int)breq.size(), dQ[i], d_ids[i], d_dists[i],
                      sl.stream));
        sl.busy = true;
        ++*batches;
        *batched_reqs += (long long)breq.size();
      }
    std::this_thread::yield();
  }
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
    // Slot capacity from free HBM: Q queries and 32 Q probe slots per table.
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
    srv->Qcap = (int)Q;
    srv->Pcap = 32 * Q;
    srv->Rcap = (int)std::min<long long>(Q, 4096);
    srv->slot_bytes = need;
    srv->slots.resize(nindex);
    for (int i = 0; i < nindex; ++i) {
      srv->slots[i].resize(cfg->slots_per_index);
      for (auto& sl : srv->slots[i]) {
        int rc = culsh_bt_create(srv->ht[i], &sl.bt);
        if (rc != CULSH_OK) return bail(rc);
        e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e != hipSuccess) return bail(fail(CULSH_E_DEVICE, "%s", hipGetErrorString(e)));
        rc = bt_reserve(sl.bt, srv->Qcap, srv->Rcap, srv->Pcap);
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
  if (!srv || (!reqs && n > 0) || !d_queries || !d_ids || !d_dists)
    return fail(CULSH_E_ARG, "invalid arguments");
  if (n == 0) return CULSH_OK;
  TRY(srv_check(srv, reqs, n));
  std::vector<double> lat(n, 0.0);
  double last = 0.0;
  long long batches = 0, breqs = 0, deferred = 0;
  if (srv->cfg.batching)
    TRY(srv_run_batched(srv, reqs, n, d_queries, d_ids, d_dists, lat, &last, &batches, &breqs,
                        &deferred));
  else
    TRY(srv_run_plain(srv, reqs, n, d_queries, d_ids, d_dists, lat, &last));
  if (latency_ms) std::copy(lat.begin(), lat.end(), latency_ms);
  if (stats) {
    if (!srv->cfg.batching) {
      batches = n;
      breqs = n;
    }
    fill_stats(lat, reqs, n, last, batches, breqs, deferred, srv->slot_bytes, stats);
  }
  return CULSH_OK;
}

}  // extern "C"
