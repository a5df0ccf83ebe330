"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must reproduce them on the CPU, the HIP path on the GPU."""
import glob
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _load(path):
    z = np.load(path)  # allow_pickle defaults to False
    N, D, L, M, Q, K, T = (int(v) for v in z["params"])
    return z, (N, D, L, M, float(z["W"]), Q, K, T)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    z, (N, D, L, M, W, Q, K, T) = _load(path)
    ix = oracle.Index(z["X"], L, M, W, z["A"], z["b"])
    r = ix.search(z["Y"], z["X"], K, T, want_buckets=True)
    for key in ("ids", "dists", "cnu", "cu", "query_buckets"):
        np.testing.assert_array_equal(r[key], z[key], err_msg=key)
    tot, start, size, codes = ix.table(0)
    np.testing.assert_array_equal(tot, z["t0_total"])
    np.testing.assert_array_equal(codes, z["t0_codes"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_hip_path_reproduces_golden(path):
    from lshtest import run_gpu
    z, (N, D, L, M, W, Q, K, T) = _load(path)
    print("golden", os.path.basename(path), flush=True)
    g = run_gpu(z["X"], z["Y"], z["A"], z["b"], L, M, W, K, T)
    for key in ("ids", "dists", "cnu", "cu", "query_buckets"):
        np.testing.assert_array_equal(g[key], z[key], err_msg=key)
    ht = g["ht"]
    np.testing.assert_array_equal(ht.getBuckets_totalIndices(0), z["t0_total"])
    np.testing.assert_array_equal(ht.getBuckets_startingIndices(0), z["t0_start"])
    np.testing.assert_array_equal(ht.getBuckets_sizes(0), z["t0_size"])
    np.testing.assert_array_equal(ht.getBuckets_codes(0), z["t0_codes"])
