#!/usr/bin/env python3
"""Benchmark of the LSH index/query path on one MI355X.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--config all|c1|c2|c3|c4|c5]
                  [--form plain|precision]

Default (--config all), the metric on its own workloads, all on one GPU:

  headline (value)  c4 batching form: the c3 index (N=4M, D=96, L=32, M=20,
                    W=2) queried by 256 requests with their own Q
                    (log-uniform in [1, 4096]), K in {1, 10, 100} and T in
                    {1, 8, 64}; a step is the whole batch as one graph launch
                    with queries and results in HBM; `chosen_vs_plain` is
                    against the plain path (one SearchTables search per
                    request, back to back on one stream); recall@10 over
                    queries of the K >= 10 requests; `roofline` is the
                    candidate kernel (k_batch_cand) timed with HIP events on
                    its stream in eager runs of the same batch
  c3_plain          the c3 search (Q=50,000, K=100, T=64), plain form
  c3_precision      the same in the precision form (BF16x3 index projection,
                    FP16-row distance pass, FP64 rechecks): rechecked
                    fractions, ids and distances of 300 queries against the
                    oracle, distance error against FP64
  c5_serve          serving form: two resident indexes (2M x 128 and
                    1M x 960), Poisson streams at 2k-20k requests/s, Q
                    log-uniform in [1, 512], 70/30 over the indexes,
                    K in {10, 100}, T in {16, 64}; continuous batching vs one
                    request at a time; p50/p99 with queries and results in
                    HBM and with the query upload + result download
  cpu_baseline      the oracle (FP64 OpenMP restatement of the reference
                    algorithm) on the host cores: search of the c4 requests
                    (all 256 by default, about 7.5 s a run on 16 cores)
                    against its own c3 index (median of 3), index build
                    timed separately

--config c1|c2|c3 give one search line, c4 the batching line alone, c5 the
serving line alone (used for profiling runs).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The oracle's OpenMP threads stay on the cores they start on (set before any
# OpenMP runtime starts).
os.environ.setdefault("OMP_PROC_BIND", "close")

HBM_PEAK_GBS = 8000.0
METRIC = "queries/s at fixed recall@10 (chosen vs plain, same GPU); p50/p99 serve latency"
PROFILES = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _measured_traffic(workload_key, kernel):
    """HBM bytes per dispatch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a 1 GiB
    copy), or None."""
    try:
        with open(PROFILES) as f:
            t = json.load(f)
        return int(t[workload_key][kernel]["hbm_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def _roofline(kernels, dom, workload_key=None):
    """kernels: name -> (ms per launch, algorithmic bytes per launch)."""
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = _measured_traffic(workload_key, dom) if workload_key else None
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
           "kernel": dom, "kernel_ms": round(dom_ms, 4), "algorithmic_bytes": dom_bytes,
           "other_kernels": {k: {"ms": round(v[0], 4), "algorithmic_bytes": v[1],
                                 "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 2) if v[0] > 0
                                 else None,
                                 "traffic": (_measured_traffic(workload_key, k)
                                             if workload_key else None)}
                             for k, v in kernels.items() if k != dom}}
    if traffic is not None:
        out["traffic_source"] = ("profiles/traffic.json [%s]: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                 "per dispatch, scaled by a 1 GiB copy" % workload_key)
        out["traffic_GBps"] = round(traffic / (dom_ms * 1e-3) / 1e9, 2)
    return out


def _search_bytes(D, L, T, K, cnu, cu, rechecked=0, row_bytes=None):
    """Bytes of a search: probe (query read + one key and offset pair per
    probe) and candidates (bucket id lists, one read of every unique row,
    results).  With the defaults this is the algorithmic work per query of
    the f32 layout, B_q = 4 D + 16 L T + 4 C_nu + 4 D C_u + 8 K, the same for
    every form; row_bytes (2 D for an FP16 copy of the rows) and rechecked
    (FP32 rows re-read by the FP64 pass, 4 D each) give the bytes a form
    actually reads."""
    import numpy as np
    cnu = np.asarray(cnu, np.float64)
    cu = np.asarray(cu, np.float64)
    Q = cnu.size
    rb = 4.0 * D if row_bytes is None else float(row_bytes)
    return (Q * (4.0 * D + 16.0 * L * T),
            float((4.0 * cnu + rb * cu + 8.0 * K).sum()) + 4.0 * D * float(rechecked))


def _base(args, value, unit, elapsed, steps, workload, extra_cfg, dtype="f32"):
    return {
        "metric": METRIC, "value": round(value, 1), "unit": unit, "n_gpus": 1,
        "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / max(steps, 1) * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded Philox generator, the workload's distribution)",
        "config": {"workload": workload, **extra_cfg},
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median(v):
    v = sorted(v)
    return v[len(v) // 2]


def _recall(X, Y, ids, n):
    import oracle
    n = min(n, Y.shape[0])
    return {"value": round(oracle.recall_at_k(X, Y[:n], ids[:n], 10), 4), "queries": n,
            "ground_truth": "exact FP64 brute force (oracle), tie-aware"}


def _index_twice(ht, dX):
    """Builds the index twice: the first build of a process also pays the
    one-time load of the library's GPU code at its first kernel launch; the
    second is the steady-state build time.  Returns (first_ms, second_ms)."""
    import torch
    out = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert ht.index_device(dX), ht.last_error
        torch.cuda.synchronize()
        out.append((time.perf_counter() - t0) * 1e3)
    return out[0], out[1]


class C3Index:
    """The c3 workload's data and its device-resident index (built once)."""

    def __init__(self, form="plain"):
        import torch
        from culsh_amd import FORM_PLAIN, FORM_PRECISION, HashTables
        from culsh_amd import workload as wl
        self.cfg = cfg = wl.CONFIGS["c3"]
        t0 = time.perf_counter()
        self.X, self.Y, self.A, self.b = wl.make_data(cfg)
        log(f"c3 data {time.perf_counter() - t0:.1f} s")
        self.dX = torch.from_numpy(self.X).cuda()
        self.ht = HashTables()
        assert self.ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, self.A, self.b), self.ht.last_error
        if form != "plain":
            assert self.ht.set_form(FORM_PRECISION), self.ht.last_error
        self.form = FORM_PLAIN if form == "plain" else FORM_PRECISION
        self.index_first_ms, self.index_ms = _index_twice(self.ht, self.dX)
        log(f"c3 index {self.index_ms:.1f} ms (first build in the process "
            f"{self.index_first_ms:.1f} ms)")
        self._oracle = None

    def oracle(self):
        """The oracle's own c3 index (built once; its build time is kept)."""
        if self._oracle is None:
            import oracle
            t0 = time.perf_counter()
            self._oracle = oracle.Index(self.X, self.cfg.L, self.cfg.M, self.cfg.W, self.A, self.b)
            self.oracle_index_s = time.perf_counter() - t0
            log(f"oracle c3 index {self.oracle_index_s:.1f} s")
        return self._oracle


def bench_search_cfg(args, cfg, X, Y, A, b, ht, index_ms, form, steps, traffic_key=None):
    """Device search of the workload's Q queries, `steps` times back to back;
    `traffic_key` names the measured-traffic entry of profiles/traffic.json
    (default: the workload and --form)."""
    import numpy as np
    import torch

    from culsh_amd import FORM_PLAIN, SearchTables
    N, D = X.shape
    Q, K, T, L = Y.shape[0], cfg.K, cfg.T, cfg.L
    st = SearchTables()
    assert st.reset(ht, K, T), st.last_error
    if form != FORM_PLAIN:
        assert st.set_form(form), st.last_error
    dY = torch.from_numpy(Y).cuda()
    ids = torch.empty((Q, K), dtype=torch.int32, device="cuda")
    dists = torch.empty((Q, K), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(args.warmup):
        assert st.search_device(dY, ids, dists, stream), st.last_error
    assert st.wait(), st.last_error
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        assert st.search_device(dY, ids, dists, stream), st.last_error
    assert st.wait(), st.last_error
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gpu_ids = ids.cpu().numpy()
    # Per-kernel device times (HIP events on the search stream), same steps.
    probe_ms = cand_ms = 0.0
    for _ in range(steps):
        assert st.search_device(dY, ids, dists, stream), st.last_error
        p, c = st.getKernelMs()
        probe_ms += p
        cand_ms += c
    # Candidate counts (identical on every step) for the algorithmic bytes.
    assert st.search(Y, Q, X), st.last_error
    assert np.array_equal(st.getKnnIds(), gpu_ids)
    cnu, cu = st.getQueryCandidates()
    probe_bytes, cand_bytes = _search_bytes(D, L, T, K, cnu, cu)
    from culsh_amd import FORM_PRECISION_GROUPED
    cand_kernel = "k_group_candidates" if form == FORM_PRECISION_GROUPED else "k_candidates"
    out = {
        "value": round(Q * steps / elapsed, 1), "unit": "queries/s",
        "ms_per_step": round(elapsed / steps * 1e3, 4), "steps": steps,
        "probe_ms": round(probe_ms / steps, 4), "candidates_ms": round(cand_ms / steps, 4),
        "index_ms": round(index_ms, 2), "index_rechecked": ht.getRechecked(),
        "search_rechecked": st.getRechecked(), "selectivity": st.getSelectivity(),
        "total_candidates": {"nonunique": int(cnu.sum()), "unique": int(cu.sum())},
        "roofline": _roofline({"k_query_probe8": (probe_ms / steps, probe_bytes),
                               cand_kernel: (cand_ms / steps, cand_bytes)},
                              cand_kernel if cand_ms >= probe_ms else "k_query_probe8",
                              traffic_key or f"{cfg.name.split('-')[0]}/{args.form}"),
    }
    return out, gpu_ids


def c3_precision(args, c3, n_check=300):
    """c3 in the precision form (its workload's form): BF16x3 index
    projection, FP16-row distance pass, FP64 rechecks; rechecked fractions,
    ids and distances of a spread sample against the oracle's c3 index, and
    the distances' error against FP64 (numpy)."""
    import numpy as np
    import torch

    from culsh_amd import FORM_PRECISION, HashTables, SearchTables
    cfg = c3.cfg
    ht = HashTables()
    assert ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, c3.A, c3.b), ht.last_error
    assert ht.set_form(FORM_PRECISION), ht.last_error
    first_ms, index_ms = _index_twice(ht, c3.dX)
    out, _ = bench_search_cfg(args, cfg, c3.X, c3.Y, c3.A, c3.b, ht, index_ms, FORM_PRECISION,
                              max(3, args.steps // 2), "c3/precision")
    out["index_first_ms"] = round(first_ms, 2)
    # The counts above come from a host-API search of the host array, which
    # reads the FP32 rows; rebuilt from that array, the index's resident rows
    # are the host array and its search takes the FP16 distance pass, as the
    # timed device searches did: its FP64 re-evaluations and algorithmic
    # bytes (2 D per unique row, 4 D per re-evaluation) replace them.
    assert ht.index(c3.X), ht.last_error
    st = SearchTables()
    assert st.reset(ht, cfg.K, cfg.T), st.last_error
    assert st.set_form(FORM_PRECISION), st.last_error
    assert st.search(c3.Y, c3.Y.shape[0], c3.X), st.last_error
    cnu, cu = st.getQueryCandidates()
    out["search_rechecked"] = st.getRechecked()
    pb, cb = _search_bytes(cfg.D, cfg.L, cfg.T, cfg.K, cnu, cu)
    _, cb_read = _search_bytes(cfg.D, cfg.L, cfg.T, cfg.K, cnu, cu, st.getRechecked(),
                               2 * cfg.D)
    out["roofline"] = _roofline({"k_query_probe8": (out["probe_ms"], pb),
                                 "k_candidates": (out["candidates_ms"], cb)},
                                "k_candidates", "c3/precision")
    out["roofline"]["read_model_bytes"] = cb_read
    out["roofline"]["read_model_GBps"] = round(cb_read / (out["candidates_ms"] * 1e6), 2)
    out["index_rechecked_fraction"] = round(
        out["index_rechecked"] / (cfg.N * cfg.L * cfg.M), 6)
    out["search_rechecked_fraction"] = round(
        out["search_rechecked"] / max(1, out["total_candidates"]["unique"]), 6)
    sel = np.linspace(0, c3.Y.shape[0] - 1, n_check).astype(np.int64)
    Ys = np.ascontiguousarray(c3.Y[sel])
    dq = torch.from_numpy(Ys).cuda()
    ids = torch.empty((n_check, cfg.K), dtype=torch.int32, device="cuda")
    ds = torch.empty((n_check, cfg.K), dtype=torch.float32, device="cuda")
    assert st.search_device(dq, ids, ds), st.last_error
    assert st.wait(), st.last_error
    gi, gd = ids.cpu().numpy(), ds.cpu().numpy()
    o = c3.oracle().search(Ys, c3.X, cfg.K, cfg.T)
    valid = gi >= 0
    d64 = np.zeros(gi.shape, np.float64)
    for q in range(n_check):
        v = gi[q][valid[q]]
        diff = c3.X[v].astype(np.float64) - Ys[q].astype(np.float64)
        d64[q, :v.size] = (diff * diff).sum(axis=1)
    err = np.abs(gd.astype(np.float64) - d64)[valid]
    ulp = np.spacing(gd.astype(np.float32))[valid].astype(np.float64)
    out["oracle_check"] = {
        "queries": n_check, "ids_equal": bool(np.array_equal(gi, o["ids"])),
        "dists_equal": bool(np.array_equal(gd, o["dists"])),
        "max_abs_dist_err_vs_fp64": float(err.max()) if err.size else 0.0,
        "max_dist_err_in_fp32_ulps": round(float((err / ulp).max()), 4) if err.size else 0.0}
    out["workload"] = cfg.describe() + " (precision form)"
    del st, ht
    return out


def bench_search(args):
    """c1-c3: one search line."""
    import numpy as np
    import torch

    from culsh_amd import FORM_PLAIN, FORM_PRECISION, FORM_PRECISION_GROUPED, HashTables
    from culsh_amd import workload as wl
    cfg = wl.CONFIGS[args.config]
    form = {"plain": FORM_PLAIN, "precision": FORM_PRECISION,
            "grouped": FORM_PRECISION_GROUPED}[args.form]
    X, Y, A, b = wl.make_data(cfg)
    N, D = X.shape
    ht = HashTables()
    assert ht.reset(N, D, cfg.L, cfg.M, cfg.W, A, b), ht.last_error
    if form != FORM_PLAIN:
        assert ht.set_form(FORM_PRECISION), ht.last_error  # BF16x3 index projection
    dX = torch.from_numpy(X).cuda()
    index_first_ms, index_ms = _index_twice(ht, dX)
    sub, gpu_ids = bench_search_cfg(args, cfg, X, Y, A, b, ht, index_ms, form, args.steps)
    sub["index_first_ms"] = round(index_first_ms, 2)
    out = _base(args, sub["value"], "queries/s", sub["ms_per_step"] * 1e-3 * args.steps,
                args.steps, cfg.describe(),
                {"form": args.form, "N": N, "D": D, "L": cfg.L, "M": cfg.M, "W": cfg.W,
                 "Q": cfg.Q, "K": cfg.K, "T": cfg.T})
    out.update({k: v for k, v in sub.items() if k not in ("value", "unit", "ms_per_step",
                                                          "steps")})
    if args.cpu_queries > 0 or args.recall_queries > 0:
        import oracle
        if args.recall_queries > 0 and cfg.K >= 10:
            out["recall@10"] = _recall(X, Y, gpu_ids, args.recall_queries)
        if args.cpu_queries > 0:
            ix = oracle.Index(X, cfg.L, cfg.M, cfg.W, A, b)
            qs = min(args.cpu_queries, cfg.Q)
            runs = []
            for _ in range(3):
                t0 = time.perf_counter()
                ref = ix.search(Y[:qs], X, cfg.K, cfg.T)
                runs.append(time.perf_counter() - t0)
            out["cpu_baseline"] = {
                "value": round(qs / _median(runs), 1), "unit": "queries/s",
                "cores": oracle.num_threads(), "cpu_model": _cpu_model(), "kind": "port",
                "sample": f"oracle search of the first {qs} of the {cfg.Q} queries against its "
                          f"own index of the full dataset, median of 3 (index build untimed)",
                "runs_s": [round(r, 3) for r in runs],
                "ids_match_gpu": bool(np.array_equal(ref["ids"], gpu_ids[:qs])),
            }
    return out


def batch_workload(X, cfg):
    """The c4 requests laid out back to back and their queries."""
    from culsh_amd import layout_requests
    from culsh_amd import workload as wl
    shapes = wl.batch_requests()
    reqs, rows, total = layout_requests(shapes)
    Y = wl.queries_near(X, rows, cfg.query_noise, 4002)
    return reqs, rows, total, Y


def bench_batch_on(args, c3, steps):
    """c4 on the c3 index: the batch as one graph launch vs one search per
    request; returns the line's fields."""
    import numpy as np
    import torch

    from culsh_amd import Batch, SearchTables, layout_requests
    from culsh_amd import workload as wl
    cfg = c3.cfg
    reqs, rows, total, Y = batch_workload(c3.X, cfg)
    dY = torch.from_numpy(Y).cuda()
    ht = c3.ht
    s = torch.cuda.current_stream().cuda_stream

    # Plain baseline: one SearchTables search per request, back to back.
    st = SearchTables()
    alone = []

    def plain_pass(keep):
        for Q, K, T, row, off in reqs:
            assert st.reset(ht, K, T)
            i = torch.empty((Q, K), dtype=torch.int32, device="cuda")
            d = torch.empty((Q, K), dtype=torch.float32, device="cuda")
            assert st.search_device(dY[row:row + Q], i, d, s), st.last_error
            if keep:
                alone.append(i)
    plain_pass(True)
    torch.cuda.synchronize()
    plain_runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        plain_pass(False)
        assert st.wait(), st.last_error
        torch.cuda.synchronize()
        plain_runs.append(time.perf_counter() - t0)
    plain_s = _median(plain_runs)
    log(f"c4 plain per-request {plain_s * 1e3:.2f} ms")

    # A second request mix for the autotuner: the serving workload's shapes
    # (Q log-uniform in [1, 512], K {10, 100}, T {16, 64}) on the c3 index
    # (tuned first, so the c4 batch's own runs are the last of its kernels).
    rng = np.random.default_rng(4005)
    shapes2 = [(int(q), int(k), int(t)) for q, k, t in
               zip(wl.log_uniform_q(rng, 256, 512), rng.choice([10, 100], 256),
                   rng.choice([16, 64], 256))]
    reqs2, rows2, total2 = layout_requests(shapes2)
    Y2 = wl.queries_near(c3.X, rows2, cfg.query_noise, 4006)
    dY2 = torch.from_numpy(Y2).cuda()
    ids2 = torch.empty(total2, dtype=torch.int32, device="cuda")
    d2 = torch.empty(total2, dtype=torch.float32, device="cuda")
    bt2 = Batch(ht)
    tune2 = bt2.autotune(reqs2, dY2, ids2, d2, reps=2)
    assert tune2 is not None, bt2.last_error
    second_mix = {
        "workload": "256 requests, Q log-uniform [1,512], K {10,100}, T {16,64} on the c3 index",
        "queries": rows2, "naive_default_ms": round(tune2[0], 3), "tuned_ms": round(tune2[1], 3),
        "schedule": bt2.get_schedule()}
    del bt2
    log(f"second mix autotune {tune2[0]:.2f} -> {tune2[1]:.2f} ms")
    bt = Batch(ht)
    ids = torch.empty(total, dtype=torch.int32, device="cuda")
    dists = torch.empty(total, dtype=torch.float32, device="cuda")
    tune = bt.autotune(reqs, dY, ids, dists, reps=2)
    assert tune is not None, bt.last_error
    log(f"c4 autotune naive {tune[0]:.2f} ms -> tuned {tune[1]:.2f} ms {bt.get_schedule()}")
    for _ in range(args.warmup):
        assert bt.search_device(reqs, dY, ids, dists, s, wait=False), bt.last_error
    assert bt.wait(), bt.last_error
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        assert bt.search_device(reqs, dY, ids, dists, s, wait=False), bt.last_error
    assert bt.wait(), bt.last_error
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = rows * steps / elapsed
    h = ids.cpu().numpy()
    same = all(np.array_equal(h[off:off + Q * K].reshape(Q, K), a.cpu().numpy())
               for (Q, K, T, row, off), a in zip(reqs, alone))
    # Stage times of the same batch: eager runs of the graph's kernels with
    # HIP events on the launch stream.
    prof = [bt.profile(reqs, dY, ids, dists, s) for _ in range(max(3, min(steps, 10)))]
    assert all(p is not None for p in prof), bt.last_error
    stage = {k: sum(p[k] for p in prof) / len(prof) for k in prof[0]}

    # Algorithmic bytes of the batch (the f32-layout work per query) from
    # each request's candidate counts, and the bytes the batch actually
    # reads: the FP16 copy of the rows (D % 8 == 0 and every value within
    # the half range, as for this data) plus the FP32 rows of the candidates
    # its FP64 pass re-evaluates.
    batch_rechecked = bt.getRechecked()
    pb = cb = cb_read = 0.0
    for Q, K, T, row, off in reqs:
        assert st.reset(ht, K, T)
        assert st.search(Y[row:row + Q], Q, c3.X), st.last_error
        cnu, cu = st.getQueryCandidates()
        p_, c_ = _search_bytes(cfg.D, cfg.L, T, K, cnu, cu)
        _, r_ = _search_bytes(cfg.D, cfg.L, T, K, cnu, cu, 0,
                              2 * cfg.D if cfg.D % 8 == 0 else 4 * cfg.D)
        pb += p_
        cb += c_
        cb_read += r_
    cb_read += 4.0 * cfg.D * batch_rechecked
    sched = bt.get_schedule()
    cand_name = "k_batch_cand" if not sched["fused"] else "k_batch"
    out = {
        "value": value, "elapsed": elapsed, "rows": rows,
        "workload": ("c4-batch-256-requests-on-c3-index: N=4000000 D=96 L=32 M=20 W=2.0, "
                     "Q log-uniform [1,4096], K {1,10,100}, T {1,8,64}"),
        "config": {"form": "batching", "requests": len(reqs), "queries": rows, "N": cfg.N,
                   "D": cfg.D, "L": cfg.L, "M": cfg.M, "W": cfg.W},
        "plain_per_request": {"value": round(rows / plain_s, 1), "unit": "queries/s",
                              "ms_per_batch": round(plain_s * 1e3, 3),
                              "runs_ms": [round(r * 1e3, 3) for r in plain_runs]},
        "chosen_vs_plain": round(value / (rows / plain_s), 3),
        "results_equal_plain": bool(same),
        "fp64_rechecked": int(batch_rechecked),
        "distance_rows": ("FP16 copy of the rows for the distance pass (widened exactly to "
                          "f32), every candidate that can reach the top K re-evaluated in "
                          "FP64 from the f32 rows: results bit-identical to the plain path "
                          "(results_equal_plain) and the FP64 oracle"),
        "autotune": {"naive_default_ms": round(tune[0], 3), "tuned_ms": round(tune[1], 3),
                     "naive_default": {"probe_queries": 32, "probe_tables": 1,
                                       "queries_per_item": 32, "rows_in_flight": 4,
                                       "wgs_per_cu": 1, "fused": 1},
                     "schedule": sched},
        "stage_ms": {k: round(v, 4) for k, v in stage.items()},
        "roofline": _roofline({cand_name: (stage["candidates"], cb),
                               "k_batch_probe": (stage["probe"], pb)}, cand_name,
                              "c4/batching"),
        "graph_ms": round(elapsed / steps * 1e3, 4),
        "_ids": h, "_reqs": reqs, "_Y": Y,
    }
    out["roofline"]["read_model_bytes"] = cb_read
    out["roofline"]["read_model_GBps"] = round(cb_read / (stage["candidates"] * 1e6), 2)
    out["roofline"]["note"] = ("kernel time: eager runs of the batch (same kernels, HIP events "
                               "on the launch stream); algorithmic bytes: the candidate stage's "
                               "f32-layout work of all requests (4 C_nu + 4 D C_u + 8 K per "
                               "query, as for the plain path); read_model_bytes: what the batch "
                               "reads (4 C_nu + 2 D C_u + 8 K with the FP16 copy of the rows, "
                               "+ 4 D per FP64 re-evaluation); traffic: measured HBM bytes")

    out["autotune"]["second_mix"] = second_mix
    return out


def c4_recall(c3, b, n):
    """Tie-aware recall@10 over the first query of the K >= 10 requests."""
    import numpy as np
    import oracle
    qs, got = [], []
    for Q, K, T, row, off in b["_reqs"]:
        if K >= 10 and len(qs) < n:
            qs.append(b["_Y"][row])
            got.append(b["_ids"][off:off + 10])
    r = oracle.recall_at_k(c3.X, np.stack(qs), np.stack(got), 10)
    return {"value": round(r, 4), "queries": len(qs),
            "ground_truth": "exact FP64 brute force (oracle), tie-aware; first query of each "
                            "K >= 10 request"}


def c4_cpu_baseline(c3, b, min_queries):
    """The oracle on the first c4 requests holding at least `min_queries`
    queries (all of them by default; median of 3), ids compared with the
    batch's; index build timed on the full c3 dataset and, median of 3,
    on a 500,000-point slice."""
    import numpy as np
    import oracle
    ix = c3.oracle()
    sel, nq = [], 0
    for r in b["_reqs"]:
        if nq >= min_queries:
            break
        sel.append(r)
        nq += r[0]
    runs, match = [], True
    for rep in range(3):
        t0 = time.perf_counter()
        outs = [ix.search(b["_Y"][row:row + Q], c3.X, K, T) for Q, K, T, row, off in sel]
        runs.append(time.perf_counter() - t0)
        if rep == 0:
            match = all(np.array_equal(o["ids"], b["_ids"][off:off + Q * K].reshape(Q, K))
                        for o, (Q, K, T, row, off) in zip(outs, sel))
    n_slice = 500_000
    idx_runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.Index(c3.X[:n_slice], c3.cfg.L, c3.cfg.M, c3.cfg.W, c3.A, c3.b)
        idx_runs.append(time.perf_counter() - t0)
    return {
        "value": round(nq / _median(runs), 1), "unit": "queries/s",
        "cores": oracle.num_threads(), "cpu_model": _cpu_model(), "kind": "port",
        "sample": (f"oracle (FP64 OpenMP restatement) search of "
                   + ("all" if len(sel) == len(b["_reqs"]) else "the first")
                   + f" {len(sel)} c4 requests ({nq} queries) against its own c3 index, "
                   "median of 3 runs"),
        "runs_s": [round(r, 3) for r in runs],
        "ids_match_gpu": bool(match),
        "index": {"full_c3_build_s": round(c3.oracle_index_s, 2),
                  "full_c3_points_per_s": round(c3.cfg.N / c3.oracle_index_s, 1),
                  "slice_points": n_slice,
                  "slice_build_s_median_of_3": round(_median(idx_runs), 3),
                  "slice_runs_s": [round(r, 3) for r in idx_runs]},
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
    }


def bench_batch(args):
    """c4 alone (profiling runs)."""
    c3 = C3Index()
    b = bench_batch_on(args, c3, args.steps)
    out = _base(args, b["value"], "queries/s", b["elapsed"], args.steps, b["workload"],
                b["config"])
    out.update({k: v for k, v in b.items()
                if k not in ("value", "elapsed", "rows", "workload", "config")
                and not k.startswith("_")})
    return out


def serve_indexes():
    import torch
    from culsh_amd import HashTables
    from culsh_amd import workload as wl
    hts, data = [], []
    for name in ("c5a", "c5b"):
        cfg = wl.CONFIGS[name]
        t0 = time.perf_counter()
        X, _, A, b = wl.make_data(cfg, Q=1)
        log(f"{name} data {time.perf_counter() - t0:.1f} s")
        dX = torch.from_numpy(X).cuda()
        ht = HashTables()
        assert ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, A, b), ht.last_error
        _, index_ms = _index_twice(ht, dX)
        hts.append(ht)
        data.append((cfg, X, A, b, dX, index_ms))
    return hts, data


def bench_serve_on(args, hts, data, rates):
    """Serving runs per rate: continuous batching vs one request at a time,
    device-resident IO, then (batching and plain) with host IO."""
    import numpy as np
    import torch

    from culsh_amd import Server
    from culsh_amd import workload as wl
    n = args.serve_requests
    per_rate = []
    keep_top = None
    for rate in rates:
        stream = wl.serve_stream(rate, n, seed=5004)
        rows, outs, reqs = [0, 0], [0, 0], []
        for a, ix, Q, K, T in stream:
            reqs.append((a, ix, Q, K, T, rows[ix], outs[ix]))
            rows[ix] += Q
            outs[ix] += Q * K
        Ys = [wl.queries_near(data[i][1], rows[i], data[i][0].query_noise, 5002 + 100 * i)
              for i in range(2)]
        dY = [torch.from_numpy(Y).cuda() for Y in Ys]
        ids = [torch.empty(o, dtype=torch.int32, device="cuda") for o in outs]
        ds = [torch.empty(o, dtype=torch.float32, device="cuda") for o in outs]
        hY = [torch.from_numpy(Y).pin_memory() for Y in Ys]
        hI = [torch.empty(o, dtype=torch.int32).pin_memory() for o in outs]
        hD = [torch.empty(o, dtype=torch.float32).pin_memory() for o in outs]
        row = {"rate_req_s": rate, "requests": n}
        keep = {}
        for batching in (True, False):
            for host in (False, True):
                srv = Server(hts, batching=batching, slots_per_index=2, hbm_fraction=0.25)
                # Warm-up: one full pass of the same stream, so the measured
                # pass sees a server whose streams, graphs and pinned-buffer
                # mappings are all in use (nothing is cached between passes).
                srv.run(reqs, dY, ids, ds, host=(hY, hI, hD) if host else None)
                res = srv.run(reqs, dY, ids, ds, host=(hY, hI, hD) if host else None)
                assert res is not None, srv.last_error
                lat, st = res
                if host:
                    got = [t.numpy().copy() for t in hI]
                else:
                    got = [t.cpu().numpy() for t in ids]
                keep[(batching, host)] = got
                key = ("batching" if batching else "plain") + ("_host_io" if host else "")
                row[key] = {
                    "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
                    "mean_ms": round(st["mean_ms"], 4),
                    "queries_per_s": round(st["queries_per_s"], 1),
                    "batches": st["batches"],
                    "mean_batch_requests": round(st["mean_batch_requests"], 3),
                    "split_requests": st["split_requests"], "grown_slots": st["grown_slots"]}
                del srv
        ref = keep[(True, False)]
        row["results_equal"] = all(all(np.array_equal(a, b) for a, b in zip(ref, v))
                                   for v in keep.values())
        per_rate.append(row)
        log(f"serve {rate} req/s: batching p99 {row['batching']['p99_ms']:.3f} ms "
            f"(host io {row['batching_host_io']['p99_ms']:.3f}), equal {row['results_equal']}")
        keep_top = (row, Ys, ref, reqs)
    return per_rate, keep_top


def serve_summary(args, hts, data, rates):
    import numpy as np
    per_rate, (row, Ys, res, reqs) = bench_serve_on(args, hts, data, rates)
    out = {
        "workload": ("c5-serve: Poisson 2k-20k req/s, Q log-uniform [1,512], 70/30 over "
                     "(a) 2M x 128 L=24 M=20 W=3.0 and (b) 1M x 960 L=16 M=28 W=3.0, K {10,100}, "
                     "T {16,64}"),
        "rate_req_s": row["rate_req_s"],
        "p50_ms": row["batching"]["p50_ms"], "p99_ms": row["batching"]["p99_ms"],
        "p50_ms_host_io": row["batching_host_io"]["p50_ms"],
        "p99_ms_host_io": row["batching_host_io"]["p99_ms"],
        "queries_per_s": row["batching"]["queries_per_s"],
        "plain_per_request": row["plain"],
        "plain_per_request_host_io": row["plain_host_io"],
        "chosen_vs_plain": round(row["batching"]["queries_per_s"] /
                                 row["plain"]["queries_per_s"], 3),
        "per_rate": per_rate,
        "index_ms": {"c5a": round(data[0][5], 1), "c5b": round(data[1][5], 1)},
        "latency_note": ("request latency = completion on the GPU timeline minus arrival; "
                         "*_host_io includes the batch's query upload and result download "
                         "(pinned host memory); each measured pass follows one full warm-up "
                         "pass of the same stream on the same server"),
    }
    if args.recall_queries > 0:
        import oracle
        rec = {}
        for i, name in enumerate(("c5a", "c5b")):
            nq = min(args.recall_queries, 50)
            sel = [r for r in reqs if r[1] == i][:nq]
            qs, got = [], []
            for a, ix, Q, K, T, rw, off in sel:
                qs.append(Ys[i][rw])
                got.append(res[i][off:off + K])
            got = np.stack([g[:10] for g in got])
            rec[name] = round(oracle.recall_at_k(data[i][1], np.stack(qs), got, 10), 4)
        out["recall@10"] = {
            "value": rec, "queries_per_index": min(args.recall_queries, 50),
            "ground_truth": "exact FP64 brute force (oracle), tie-aware; first query of the "
                            "first requests",
            "note": ("c5b's fixed parameters (L=16, M=28, W=3 on a D=960 mixture, sigma 0.05) "
                     "put almost every query alone in its buckets, so its recall@10 is far "
                     "below 0.9 for every form; c5a's W=3.0 was chosen for recall@10 >= 0.9")}
    return out


def bench_serve(args):
    """c5 alone (profiling runs)."""
    hts, data = serve_indexes()
    rates = [2000, 5000, 10000, 20000]
    t0 = time.perf_counter()
    sv = serve_summary(args, hts, data, rates)
    elapsed = time.perf_counter() - t0
    out = _base(args, sv["queries_per_s"], "queries/s", elapsed, len(rates), sv["workload"],
                {"form": "serving", "slots_per_index": 2, "rates": rates})
    out.update({k: v for k, v in sv.items() if k not in ("workload", "queries_per_s")})
    return out


def bench_all(args):
    """The default line: c4 headline, c3 plain and c5 serving sub-objects,
    CPU baseline."""
    import torch
    t_all = time.perf_counter()
    c3 = C3Index()
    # c3 plain search (Q = 50,000, K = 100, T = 64).
    from culsh_amd import FORM_PLAIN
    c3s, c3_ids = bench_search_cfg(args, c3.cfg, c3.X, c3.Y, c3.A, c3.b, c3.ht, c3.index_ms,
                                   FORM_PLAIN, max(3, args.steps // 2))
    c3s["index_first_ms"] = round(c3.index_first_ms, 2)
    log(f"c3 plain {c3s['value']:.0f} q/s")
    # c4 batching (headline).
    b = bench_batch_on(args, c3, args.steps)
    log(f"c4 batch {b['value']:.0f} q/s, {b['chosen_vs_plain']}x plain")
    out = _base(args, b["value"], "queries/s", b["elapsed"], args.steps, b["workload"],
                b["config"])
    out.update({k: v for k, v in b.items()
                if k not in ("value", "elapsed", "rows", "workload", "config")
                and not k.startswith("_")})
    if args.recall_queries > 0:
        out["recall@10"] = c4_recall(c3, b, args.recall_queries)
        c3s["recall@10"] = _recall(c3.X, c3.Y, c3_ids, args.recall_queries)
    c3s["workload"] = c3.cfg.describe()
    out["c3_plain"] = c3s
    if args.cpu_queries > 0:
        out["cpu_baseline"] = c4_cpu_baseline(c3, b, args.cpu_queries)
        log(f"cpu baseline {out['cpu_baseline']['value']} q/s")
        # c3 in its own (precision) form, checked against the oracle's index.
        out["c3_precision"] = c3_precision(args, c3)
        oc = out["c3_precision"]["oracle_check"]
        log(f"c3 precision {out['c3_precision']['value']:.0f} q/s, index "
            f"{out['c3_precision']['index_ms']} ms, ids equal oracle {oc['ids_equal']}")
    # Free the c3 index before the serving indexes.
    del c3, b
    torch.cuda.empty_cache()
    hts, data = serve_indexes()
    out["c5_serve"] = serve_summary(args, hts, data, [2000, 5000, 10000, 20000])
    out["bench_wall_s"] = round(time.perf_counter() - t_all, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="all", choices=["all", "c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--form", default="plain", choices=["plain", "precision", "grouped"],
                    help="c1-c3: plain (FP32 projection), precision (BF16x3 index projection), "
                         "grouped (BF16x3 + grouped norm-expansion distances)")
    ap.add_argument("--cpu-queries", type=int, default=150000,
                    help="queries in the oracle's timed CPU sample, whole requests "
                         "(the default covers all of c4; 0 = skip)")
    ap.add_argument("--recall-queries", type=int, default=100)
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (5 times): the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the LSH path here is single-GPU; see the sharding design).", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.config == "all":
        out = bench_all(args)
    elif args.config == "c4":
        out = bench_batch(args)
    elif args.config == "c5":
        out = bench_serve(args)
    else:
        out = bench_search(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
