"""Shared helpers for the parity tests: run the HIP path (through the C ABI)
and the oracle on the same inputs and compare them."""
import struct

import numpy as np

import oracle
from culsh_amd import HashTables, SearchTables


def run_gpu(X, Y, A, b, L, M, W, K, T, form=0, want_buckets=True, st_form=None):
    """form: the index projection form (0 plain, 1 precision); st_form: the
    search form (defaults to form; 2 = grouped norm-expansion distances)."""
    N, D = X.shape
    st_form = form if st_form is None else st_form
    ht = HashTables()
    assert ht.reset(N, D, L, M, W, A, b), ht.last_error
    if form:
        assert ht.set_form(form), ht.last_error
    assert ht.index(X), ht.last_error
    st = SearchTables()
    assert st.reset(ht, K, T), st.last_error
    if st_form:
        assert st.set_form(st_form), st.last_error
    assert st.search(Y, Y.shape[0], X), st.last_error
    cnu, cu = st.getQueryCandidates()
    out = dict(ids=st.getKnnIds(), dists=st.getKnnDistances(), cnu=cnu, cu=cu, ht=ht, st=st)
    if want_buckets:
        out["query_buckets"] = np.stack([st.getQueryBuckets(t) for t in range(L)])
    return out


def run_oracle(X, Y, A, b, L, M, W, K, T, want_buckets=True):
    ix = oracle.Index(X, L, M, W, A, b)
    r = ix.search(Y, X, K, T, want_buckets=want_buckets)
    r["index"] = ix
    return r


def assert_buckets_equal(ht, ix, L):
    for t in range(L):
        tot, start, size, codes = ix.table(t)
        assert ht.getBuckets(t) == ix.buckets(t), f"table {t}: bucket count"
        np.testing.assert_array_equal(ht.getBuckets_sizes(t), size, err_msg=f"table {t} sizes")
        np.testing.assert_array_equal(ht.getBuckets_startingIndices(t), start,
                                      err_msg=f"table {t} starts")
        np.testing.assert_array_equal(ht.getBuckets_totalIndices(t), tot,
                                      err_msg=f"table {t} members")
        np.testing.assert_array_equal(ht.getBuckets_codes(t), codes, err_msg=f"table {t} codes")


def assert_search_equal(g, o, check_buckets=True):
    bad = np.nonzero((g["ids"] != o["ids"]).any(axis=1))[0]
    assert bad.size == 0, (f"{bad.size} queries differ, first q={bad[0]}: gpu {g['ids'][bad[0]]} "
                           f"oracle {o['ids'][bad[0]]}")
    np.testing.assert_array_equal(g["dists"], o["dists"])
    np.testing.assert_array_equal(g["cnu"], o["cnu"])
    np.testing.assert_array_equal(g["cu"], o["cu"])
    if check_buckets:
        np.testing.assert_array_equal(g["query_buckets"], o["query_buckets"])


def recall_at(ids, true_ids, true_d, X, Y, k=10):
    """Tie-aware recall@k: a returned id counts if its exact distance is <=
    the exact k-th distance."""
    Q = ids.shape[0]
    hit = 0
    for q in range(Q):
        dk = true_d[q, k - 1]
        for i in ids[q, :k]:
            if i >= 0 and oracle.dist(Y[q], X[i]) <= dk:
                hit += 1
    return hit / (Q * k)


def write_reference_format(path, ix, A, b, N, D, L, M, W):
    """cuLSH_HashTables.cu:415-433: int N, D, L, M; float W; per table A_t, b_t;
    unsigned B[L]; per table totalIndices[N], startingIndices[B], sizes[B],
    codes[B*M] (column-major)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<iiiif", N, D, L, M, W))
        for t in range(L):
            f.write(np.ascontiguousarray(A[t], np.float32).tobytes())   # (M, D) == [D x M] col-major
            f.write(np.ascontiguousarray(b[t], np.float32).tobytes())
        f.write(np.array([ix.buckets(t) for t in range(L)], np.uint32).tobytes())
        for t in range(L):
            tot, start, size, codes = ix.table(t)
            f.write(tot.astype(np.uint32).tobytes())
            f.write(start.astype(np.uint32).tobytes())
            f.write(size.astype(np.uint32).tobytes())
            f.write(np.ascontiguousarray(codes.T, np.float32).tobytes())
