#!/usr/bin/env python3
"""Benchmark of the LSH search path on one MI355X.

A step is one SearchTables search of the workload's Q queries against the
device-resident index (queries already in HBM, ids/distances written to
HBM), through the C ABI's device entry point.  The index is built once,
untimed.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--config c2] [--form plain]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--form", default="plain", choices=["plain", "precision"])
    ap.add_argument("--cpu-queries", type=int, default=2000,
                    help="queries in the oracle's timed CPU sample (0 = skip)")
    ap.add_argument("--recall-queries", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the LSH path here is single-GPU; see the sharding design).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from culsh_amd import FORM_PLAIN, FORM_PRECISION, HashTables, SearchTables
    from culsh_amd import workload as wl

    cfg = wl.CONFIGS[args.config]
    form = FORM_PLAIN if args.form == "plain" else FORM_PRECISION
    torch.cuda.set_device(0)
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
    torch.cuda.synchronize()
    probe_ms = cand_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assert st.search_device(dY, ids, dists, stream), st.last_error
        p, c = st.getKernelMs()
        probe_ms += p
        cand_ms += c
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = Q * args.steps / elapsed
    gpu_ids = ids.cpu().numpy()

    # Candidate counts (identical on every step) for the algorithmic bytes.
    assert st.search(Y, Q, X), st.last_error
    assert np.array_equal(st.getKnnIds(), gpu_ids)
    cnu, cu = st.getQueryCandidates()
    cnu = cnu.astype(np.float64)
    cu = cu.astype(np.float64)
    probe_bytes = Q * (4.0 * D + 16.0 * L * T)
    cand_bytes = float((4.0 * cnu + 4.0 * D * cu + 8.0 * K).sum())
    kernels = {
        "k_query_probe": (probe_ms / args.steps, probe_bytes),
        "k_candidates": (cand_ms / args.steps, cand_bytes),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": dom, "kernel_ms": round(dom_ms, 4), "algorithmic_bytes": dom_bytes,
                "other_kernels": {k: {"ms": round(v[0], 4), "algorithmic_bytes": v[1],
                                      "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 2)}
                                  for k, v in kernels.items() if k != dom}}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox generator, the workload's distribution)",
        "config": {"workload": cfg.describe(), "form": args.form, "N": N, "D": D, "L": L,
                   "M": cfg.M, "W": cfg.W, "Q": Q, "K": K, "T": T},
        "index_ms": round(index_ms, 2),
        "selectivity": st.getSelectivity(),
        "total_candidates": {"nonunique": int(cnu.sum()), "unique": int(cu.sum())},
        "roofline": roofline,
    }

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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
