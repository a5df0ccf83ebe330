"""Writes the committed golden fixtures (inputs and oracle outputs).

The reference ships no golden vectors, so these small fixtures are produced
by the build's FP64 oracle (itself pinned by tests/test_oracle.py) and
guard both the oracle and the HIP path against regressions.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from culsh_amd import workload as wl  # noqa: E402

CASES = {
    # name: (N, D, L, M, W, Q, K, T, seed, kind)
    "mixture_T1": (2000, 16, 4, 6, 1.0, 50, 5, 1, 101, "mixture"),
    "mixture_T8": (2000, 16, 4, 6, 1.0, 50, 5, 8, 102, "mixture"),
    "ints_ties_T12": (1500, 8, 5, 4, 3.0, 60, 16, 12, 103, "ints"),
}


def make(name):
    N, D, L, M, W, Q, K, T, seed, kind = CASES[name]
    if kind == "mixture":
        X = wl.mixture(N, D, 20, 0.2, seed)
        Y = wl.queries_near(X, Q, 0.05, seed + 1)
    else:
        rng = np.random.default_rng(seed)
        X = rng.integers(0, 4, size=(N, D)).astype(np.float32)
        Y = rng.integers(0, 4, size=(Q, D)).astype(np.float32)
    A, b = wl.projection(L, D, M, W, seed + 2)
    ix = oracle.Index(X, L, M, W, A, b)
    r = ix.search(Y, X, K, T, want_buckets=True)
    tot, start, size, codes = ix.table(0)
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"), X=X, Y=Y, A=A, b=b,
        params=np.array([N, D, L, M, Q, K, T], np.int64), W=np.float32(W),
        ids=r["ids"], dists=r["dists"], cnu=r["cnu"], cu=r["cu"],
        query_buckets=r["query_buckets"], t0_total=tot, t0_start=start, t0_size=size,
        t0_codes=codes)


if __name__ == "__main__":
    for name in CASES:
        make(name)
        print("wrote", name)
