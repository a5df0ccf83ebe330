"""ORACLE — test infrastructure only.

ctypes bindings to ``oracle/liboracle.so``, the FP64 OpenMP C++ restatement
of the cuLSH indexing and query algorithm (``oracle/oracle.cpp``).  It is the
correctness checker for the HIP path and the CPU baseline that ``bench.py``
times.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module; the shipped package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_ll, c_f, c_vp = ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_void_p
        L.orc_project.argtypes = [_f32p, c_ll, c_int, _f32p, _f32p, c_int, c_f, _f64p]
        L.orc_project.restype = None
        L.orc_index_build.argtypes = [c_int, c_int, c_int, c_int, c_f, _f32p, _f32p, _f32p]
        L.orc_index_build.restype = c_vp
        L.orc_index_free.argtypes = [c_vp]
        L.orc_index_free.restype = None
        L.orc_index_buckets.argtypes = [c_vp, c_int]
        L.orc_index_buckets.restype = ctypes.c_uint
        L.orc_index_table.argtypes = [c_vp, c_int, _u32p, _u32p, _u32p, _f32p]
        L.orc_index_table.restype = None
        L.orc_probe_sequence.argtypes = [_f64p, c_int, c_int, _i32p, _u64p, _i32p]
        L.orc_probe_sequence.restype = c_int
        L.orc_topk_heap.argtypes = [_f64p, _u32p, c_int, c_int, _i32p, _f64p]
        L.orc_topk_heap.restype = None
        L.orc_dist.argtypes = [_f32p, _f32p, c_int]
        L.orc_dist.restype = ctypes.c_double
        L.orc_search.argtypes = [c_vp, _f32p, c_int, _f32p, c_int, c_int,
                                 c_vp, c_vp, c_vp, c_vp, c_vp]
        L.orc_search.restype = c_int
        L.orc_bruteforce_knn.argtypes = [_f32p, c_ll, c_int, _f32p, c_int, c_int, _i32p, _f64p]
        L.orc_bruteforce_knn.restype = None
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = c_int
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def project(X, A_t, b_t, W):
    """FP64 projections v[n, m] of X [N, D] onto one table (A_t [D x M] col-major)."""
    X = _c(X, np.float32)
    N, D = X.shape
    A_t = _c(A_t, np.float32).reshape(-1)
    b_t = _c(b_t, np.float32).reshape(-1)
    M = b_t.size
    v = np.empty((N, M), np.float64)
    lib().orc_project(X, N, D, A_t, b_t, M, W, v)
    return v


def probe_sequence(v, T):
    """Multiprobe sequence of one (query, table): (codes [T, M], masks [T], perm)."""
    v = _c(v, np.float64).reshape(-1)
    M = v.size
    codes = np.empty((T, M), np.int32)
    masks = np.zeros(T, np.uint64)
    perm = np.full(64, -1, np.int32)
    maxT = lib().orc_probe_sequence(v, M, T, codes, masks, perm)
    return codes, masks, perm[:maxT]


def topk_heap(dist, ids, K):
    """Literal reference heap top-K over ascending ids -> (ids [K], dists [K])."""
    dist = _c(dist, np.float64)
    ids = _c(ids, np.uint32)
    oi = np.empty(K, np.int32)
    od = np.empty(K, np.float64)
    lib().orc_topk_heap(dist, ids, int(dist.size), K, oi, od)
    return oi, od


def dist(q, x):
    q = _c(q, np.float32)
    x = _c(x, np.float32)
    return lib().orc_dist(q, x, q.size)


class Index:
    """FP64 oracle index: HashTables::reset(N,D,L,M,W,A,b) + index(X)."""

    def __init__(self, X, L, M, W, A, b):
        X = _c(X, np.float32)
        self.N, self.D = X.shape
        self.L, self.M, self.W = L, M, float(W)
        A = _c(A, np.float32).reshape(-1)
        b = _c(b, np.float32).reshape(-1)
        self._h = lib().orc_index_build(self.N, self.D, L, M, self.W, A, b, X)
        if not self._h:
            raise ValueError("oracle index: invalid parameters")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_index_free(self._h)
            self._h = None

    def buckets(self, t):
        return int(lib().orc_index_buckets(self._h, t))

    def table(self, t):
        """(totalIndices [N], startingIndices [B], sizes [B], codes float [B, M])."""
        B = self.buckets(t)
        total = np.empty(self.N, np.uint32)
        start = np.empty(B, np.uint32)
        size = np.empty(B, np.uint32)
        codes = np.empty(B * self.M, np.float32)
        lib().orc_index_table(self._h, t, total, start, size, codes)
        return total, start, size, codes.reshape(self.M, B).T.copy()

    def search(self, queries, X, K, T, want_buckets=False):
        """Returns dict(ids [Q,K], dists [Q,K], cnu [Q], cu [Q], query_buckets [L,Q,T]?)."""
        queries = _c(queries, np.float32)
        X = _c(X, np.float32)
        Q = queries.shape[0]
        ids = np.empty((Q, K), np.int32)
        dists = np.empty((Q, K), np.float32)
        cnu = np.empty(Q, np.uint64)
        cu = np.empty(Q, np.uint64)
        qb = np.empty((self.L, Q, T), np.int32) if want_buckets else None
        rc = lib().orc_search(self._h, queries, Q, X, K, T,
                              ids.ctypes.data, dists.ctypes.data,
                              qb.ctypes.data if qb is not None else None,
                              cnu.ctypes.data, cu.ctypes.data)
        if rc != 0:
            raise ValueError("oracle search: invalid arguments")
        out = dict(ids=ids, dists=dists, cnu=cnu, cu=cu)
        if qb is not None:
            out["query_buckets"] = qb
        return out


def bruteforce_knn(X, queries, K):
    X = _c(X, np.float32)
    queries = _c(queries, np.float32)
    Q = queries.shape[0]
    ids = np.empty((Q, K), np.int32)
    d = np.empty((Q, K), np.float64)
    lib().orc_bruteforce_knn(X, X.shape[0], X.shape[1], queries, Q, K, ids, d)
    return ids, d


def recall_at_k(X, queries, ids, k=10):
    """Tie-aware recall@k of `ids` [Q, >=k] against exact FP64 brute force: a
    returned id counts when its exact distance is <= the exact k-th distance."""
    true_ids, true_d = bruteforce_knn(X, queries, k)
    hit = 0
    for q in range(queries.shape[0]):
        dk = true_d[q, k - 1]
        for i in ids[q, :k]:
            if i >= 0 and dist(queries[q], X[i]) <= dk:
                hit += 1
    return hit / (queries.shape[0] * k)


def num_threads():
    return int(lib().orc_num_threads())
