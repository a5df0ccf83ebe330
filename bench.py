#!/usr/bin/env python3
"""Benchmark of the LSH search path on one MI355X.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--config c2|c1|c3|c4|c5]
                  [--form plain|precision]

c1-c3 (default c2): a step is one SearchTables search of the workload's Q
queries against the device-resident index (queries already in HBM, results
written to HBM) through the C ABI's device entry point; the index is built
once, untimed.

c4 (batching form): the c3 index queried by 256 requests with their own
Q (log-uniform in [1, 4096]), K in {1, 10, 100} and T in {1, 8, 64}; a step
is the whole batch as one graph launch, after the schedule autotuner.  The
plain baseline runs the same requests one SearchTables search at a time.

c5 (serving form): two resident indexes (2M x 128 and 1M x 960), Poisson
request streams at 2k-20k requests/s, Q log-uniform in [1, 512], 70/30 over
the indexes, K in {10, 100}, T in {16, 64}; continuous batching vs one
request at a time through the search path.  `value` is the served
queries/s at the highest rate.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "queries/s at fixed recall@10 (chosen vs plain, same GPU); p50/p99 serve latency"


PROFILES = os.path.join(ROOT, "profiles", "traffic.json")


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


def _roofline(kernels, workload_key=None):
    """kernels: name -> (ms, algorithmic bytes); the dominant one is reported."""
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = _measured_traffic(workload_key, dom) if workload_key else None
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
           "kernel": dom, "kernel_ms": round(dom_ms, 4), "algorithmic_bytes": dom_bytes,
           "other_kernels": {k: {"ms": round(v[0], 4), "algorithmic_bytes": v[1],
                                 "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 2),
                                 "traffic": (_measured_traffic(workload_key, k)
                                             if workload_key else None)}
                             for k, v in kernels.items() if k != dom}}
    if traffic is not None:
        out["traffic_source"] = ("profiles/traffic.json [%s]: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                 "per dispatch, scaled by a 1 GiB copy" % workload_key)
        out["traffic_GBps"] = round(traffic / (dom_ms * 1e-3) / 1e9, 2)
    return out


def _search_bytes(D, L, T, K, cnu, cu):
    """Algorithmic bytes of a search: probe (query read + one key and offset
    pair per probe) and candidates (bucket id lists, unique rows, results)."""
    import numpy as np
    cnu = np.asarray(cnu, np.float64)
    cu = np.asarray(cu, np.float64)
    Q = cnu.size
    return Q * (4.0 * D + 16.0 * L * T), float((4.0 * cnu + 4.0 * D * cu + 8.0 * K).sum())


def _base(args, value, unit, elapsed, workload, extra_cfg, dtype="f32"):
    return {
        "metric": METRIC, "value": round(value, 1), "unit": unit, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / max(args.steps, 1) * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded Philox generator, the workload's distribution)",
        "config": {"workload": workload, **extra_cfg},
    }


def bench_search(args):
    import numpy as np
    import torch

    from culsh_amd import FORM_PLAIN, FORM_PRECISION, HashTables, SearchTables
    from culsh_amd import workload as wl

    cfg = wl.CONFIGS[args.config]
    form = FORM_PLAIN if args.form == "plain" else FORM_PRECISION
    X, Y, A, b = wl.make_data(cfg)
    N, D = X.shape
    Q, K, T, L = cfg.Q, cfg.K, cfg.T, cfg.L

    ht = HashTables()
    assert ht.reset(N, D, L, cfg.M, cfg.W, A, b), ht.last_error
    if form != FORM_PLAIN:
        assert ht.set_form(form), ht.last_error
    dX = torch.from_numpy(X).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    assert ht.index_device(dX), ht.last_error
    torch.cuda.synchronize()
    index_ms = (time.perf_counter() - t0) * 1e3

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
    # Timed steps: back to back on the stream (the plain form's device search
    # never waits for the device), one synchronisation at the end.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assert st.search_device(dY, ids, dists, stream), st.last_error
    assert st.wait(), st.last_error
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = Q * args.steps / elapsed
    gpu_ids = ids.cpu().numpy()
    # Per-kernel device times (HIP events on the search stream), same steps.
    probe_ms = cand_ms = 0.0
    for _ in range(args.steps):
        assert st.search_device(dY, ids, dists, stream), st.last_error
        p, c = st.getKernelMs()
        probe_ms += p
        cand_ms += c

    # Candidate counts (identical on every step) for the algorithmic bytes.
    assert st.search(Y, Q, X), st.last_error
    assert np.array_equal(st.getKnnIds(), gpu_ids)
    cnu, cu = st.getQueryCandidates()
    probe_bytes, cand_bytes = _search_bytes(D, L, T, K, cnu, cu)
    out = _base(args, value, "queries/s", elapsed, cfg.describe(),
                {"form": args.form, "N": N, "D": D, "L": L, "M": cfg.M, "W": cfg.W, "Q": Q,
                 "K": K, "T": T})
    out["index_ms"] = round(index_ms, 2)
    out["index_rechecked"] = ht.getRechecked()
    out["search_rechecked"] = st.getRechecked()
    out["selectivity"] = st.getSelectivity()
    out["total_candidates"] = {"nonunique": int(cnu.sum()), "unique": int(cu.sum())}
    out["roofline"] = _roofline({
        "k_query_probe8": (probe_ms / args.steps, probe_bytes),
        "k_candidates" if form == FORM_PLAIN else "k_group_candidates":
            (cand_ms / args.steps, cand_bytes),
    }, f"{args.config}/{args.form}")

    if args.cpu_queries > 0 or args.recall_queries > 0:
        import oracle
        ix = oracle.Index(X, L, cfg.M, cfg.W, A, b)
        if args.cpu_queries > 0:
            qs = min(args.cpu_queries, Q)
            t0 = time.perf_counter()
            ref = ix.search(Y[:qs], X, K, T)
            cpu_s = time.perf_counter() - t0
            out["cpu_baseline"] = {
                "value": round(qs / cpu_s, 1), "unit": "queries/s",
                "cores": oracle.num_threads(), "kind": "port",
                "sample": (f"oracle (FP64 OpenMP restatement) search of the first {qs} of the "
                           f"{Q} queries against its own index of the full dataset (index "
                           f"build untimed)"),
                "ids_match_gpu": bool(np.array_equal(ref["ids"], gpu_ids[:qs])),
            }
        if args.recall_queries > 0 and K >= 10:
            nq = min(args.recall_queries, Q)
            out["recall@10"] = {"value": oracle.recall_at_k(X, Y[:nq], gpu_ids[:nq], 10),
                                "queries": nq,
                                "ground_truth": "exact FP64 brute force (oracle), tie-aware"}
    return out


def bench_batch(args):
    """c4: 256 requests on the c3 index as one graph launch vs one search each."""
    import numpy as np
    import torch

    from culsh_amd import Batch, HashTables, SearchTables, layout_requests
    from culsh_amd import workload as wl

    cfg = wl.CONFIGS["c3"]
    X, _, A, b = wl.make_data(cfg, Q=1)
    shapes = wl.batch_requests()
    reqs, rows, total = layout_requests(shapes)
    Y = wl.queries_near(X, rows, cfg.query_noise, 4002)
    dX = torch.from_numpy(X).cuda()
    dY = torch.from_numpy(Y).cuda()
    ht = HashTables()
    assert ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, A, b), ht.last_error
    assert ht.index_device(dX), ht.last_error
    s = torch.cuda.current_stream().cuda_stream

    # Plain baseline: one SearchTables search per request.
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
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 4)):
        plain_pass(False)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / max(1, args.steps // 4)

    bt = Batch(ht)
    ids = torch.empty(total, dtype=torch.int32, device="cuda")
    dists = torch.empty(total, dtype=torch.float32, device="cuda")
    tune = bt.autotune(reqs, dY, ids, dists, reps=2)
    assert tune is not None, bt.last_error
    for _ in range(args.warmup):
        assert bt.search_device(reqs, dY, ids, dists, s, wait=False), bt.last_error
    torch.cuda.synchronize()
    dev_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assert bt.search_device(reqs, dY, ids, dists, s, wait=False), bt.last_error
        dev_ms += bt.last_ms()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = rows * args.steps / elapsed
    h = ids.cpu().numpy()
    same = all(np.array_equal(h[off:off + Q * K].reshape(Q, K), a.cpu().numpy())
               for (Q, K, T, row, off), a in zip(reqs, alone))

    # Algorithmic bytes of the batch from each request's candidate counts.
    pb = cb = 0.0
    for Q, K, T, row, off in reqs:
        assert st.reset(ht, K, T)
        assert st.search(Y[row:row + Q], Q, X), st.last_error
        cnu, cu = st.getQueryCandidates()
        p_, c_ = _search_bytes(cfg.D, cfg.L, T, K, cnu, cu)
        pb += p_
        cb += c_
    out = _base(args, value, "queries/s", elapsed,
                "c4-batch-256-requests-on-c3-index: Q log-uniform [1,4096], K {1,10,100}, "
                "T {1,8,64}",
                {"form": "batching", "requests": len(reqs), "queries": rows, "N": cfg.N,
                 "D": cfg.D, "L": cfg.L, "M": cfg.M, "W": cfg.W})
    out["plain_per_request"] = {"value": round(rows / plain_s, 1), "unit": "queries/s",
                                "ms_per_batch": round(plain_s * 1e3, 3)}
    out["chosen_vs_plain"] = round(value / (rows / plain_s), 3)
    out["results_equal_plain"] = bool(same)
    out["autotune"] = {"default_ms": round(tune[0], 3), "tuned_ms": round(tune[1], 3),
                       "schedule": bt.get_schedule()}
    # The whole graph (probe + candidate stages) against the search bytes.
    out["roofline"] = _roofline({"batch graph (k_batch_probe + k_batch_cand)":
                                 (dev_ms / args.steps, pb + cb)})
    out["roofline"]["note"] = "bytes: probe + candidate algorithmic bytes of all requests"
    if args.cpu_queries > 0:
        import oracle
        ix = oracle.Index(X, cfg.L, cfg.M, cfg.W, A, b)
        done, t_cpu, r = 0, 0.0, 0
        while done < args.cpu_queries and r < len(reqs):
            Q, K, T, row, off = reqs[r]
            t0 = time.perf_counter()
            ref = ix.search(Y[row:row + Q], X, K, T)
            t_cpu += time.perf_counter() - t0
            assert np.array_equal(ref["ids"], h[off:off + Q * K].reshape(Q, K))
            done += Q
            r += 1
        out["cpu_baseline"] = {"value": round(done / t_cpu, 1), "unit": "queries/s",
                               "cores": oracle.num_threads(), "kind": "port",
                               "sample": f"oracle search of the first {r} requests ({done} "
                                         f"queries), ids equal to the batch's"}
    return out


def bench_serve(args):
    """c5: Poisson request streams over two resident indexes."""
    import numpy as np
    import torch

    from culsh_amd import HashTables, Server
    from culsh_amd import workload as wl

    hts, data = [], []
    for name in ("c5a", "c5b"):
        cfg = wl.CONFIGS[name]
        X, _, A, b = wl.make_data(cfg, Q=1)
        dX = torch.from_numpy(X).cuda()
        ht = HashTables()
        assert ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, A, b), ht.last_error
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert ht.index_device(dX), ht.last_error
        torch.cuda.synchronize()
        hts.append(ht)
        data.append((cfg, X, A, b, dX, (time.perf_counter() - t0) * 1e3))
    n = args.serve_requests
    rates = [2000, 5000, 10000, 20000]
    per_rate = []
    top = None
    t_all = time.perf_counter()
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
        row = {"rate_req_s": rate, "requests": n}
        keep = {}
        for batching in (True, False):
            srv = Server(hts, batching=batching, slots_per_index=2, hbm_fraction=0.25)
            srv.run(reqs[:100], dY, ids, ds)  # warm-up
            lat, st = srv.run(reqs, dY, ids, ds)
            keep[batching] = [t.cpu().numpy() for t in ids]
            row["batching" if batching else "plain"] = {
                "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
                "mean_ms": round(st["mean_ms"], 4), "queries_per_s": round(st["queries_per_s"], 1),
                "batches": st["batches"], "mean_batch_requests": round(st["mean_batch_requests"], 3)}
            del srv
        row["results_equal"] = all(np.array_equal(a, b) for a, b in zip(keep[True], keep[False]))
        per_rate.append(row)
        top = (row, Ys, keep[True], reqs)
    elapsed = time.perf_counter() - t_all
    row, Ys, res, reqs = top
    value = row["batching"]["queries_per_s"]
    out = _base(args, value, "queries/s", elapsed,
                "c5-serve: Poisson 2k-20k req/s, Q log-uniform [1,512], 70/30 over "
                "(a) 2M x 128 L=24 M=20 W=3.0 and (b) 1M x 960 L=16 M=28 W=3.0, K {10,100}, "
                "T {16,64}",
                {"form": "serving", "slots_per_index": 2, "rates": rates})
    out["steps"] = len(rates)
    out["p50_ms"] = row["batching"]["p50_ms"]
    out["p99_ms"] = row["batching"]["p99_ms"]
    out["plain_per_request"] = row["plain"]
    out["chosen_vs_plain"] = round(value / row["plain"]["queries_per_s"], 3)
    out["per_rate"] = per_rate
    out["index_ms"] = {"c5a": round(data[0][5], 1), "c5b": round(data[1][5], 1)}
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
        out["recall@10"] = {"value": rec, "queries_per_index": min(args.recall_queries, 50),
                            "ground_truth": "exact FP64 brute force (oracle), tie-aware; "
                                            "first query of the first requests"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--form", default="plain", choices=["plain", "precision"])
    ap.add_argument("--cpu-queries", type=int, default=2000,
                    help="queries in the oracle's timed CPU sample (0 = skip)")
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
    if args.config == "c4":
        out = bench_batch(args)
    elif args.config == "c5":
        out = bench_serve(args)
    else:
        out = bench_search(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
