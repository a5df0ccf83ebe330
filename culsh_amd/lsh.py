"""Python mirror of the cuLSH API (cuLSH_HashTables.cu:51-108,
cuLSH_SearchTables.cu:42-107) over libculsh.so.

Same method names and argument meaning as the reference classes; like the
reference, methods report failure by returning False and printing
``[Error] <message>`` to stdout (cuLSH.h:79), and the last message is kept in
``last_error``.  Arrays: a dataset or query set is a C-contiguous float32
array of shape (N, D) -- the reference's column-major [D x N]; ``A`` has
shape (L, M, D) (each table column-major [D x M]) and ``b`` (L, M).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (FORM_PLAIN, FORM_PRECISION, FORM_PRECISION_GROUPED,  # noqa: F401
                   Request, Schedule, SrvConfig, SrvRequest, SrvStats, Times)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


_NP = {ctypes.c_float: np.float32, ctypes.c_uint: np.uint32, ctypes.c_int: np.int32,
       ctypes.c_ulonglong: np.uint64}


def _view(ptr, ctype, count):
    """Copy of `count` elements at a library-owned host pointer."""
    if not ptr or count == 0:
        return np.empty(0, dtype=_NP[ctype])
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(count,)).copy()


class _Base:
    last_error = ""

    def _ok(self, rc):
        if rc != 0:
            self.last_error = _lib.last_error()
            print(f"[Error] {self.last_error}", flush=True)
            return False
        self.last_error = ""
        return True


class HashTables(_Base):
    """cuLSH::HashTables: projection parameters and the device-resident index."""

    def __init__(self):
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        if self._L.culsh_ht_create(ctypes.byref(h)) != 0:
            raise RuntimeError(_lib.last_error())
        self._h = h
        self._dataset = None

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.culsh_ht_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    # reset(N, D, L, M, W[, A, b]) / reset(filename)  (cuLSH_HashTables.cu:190-249)
    def reset(self, N, D=None, L=None, M=None, W=None, A=None, b=None, seed=0, debug=None):
        if isinstance(N, (str, bytes)):
            return self.load(N)
        if A is None or b is None:
            return self._ok(self._L.culsh_ht_reset_random(self._h, int(N), int(D), int(L),
                                                          int(M), float(W), int(seed), debug))
        A = _f32(A).reshape(int(L), int(M) * int(D))
        b = _f32(b).reshape(int(L), int(M))
        self._A, self._b = A, b
        pa = (ctypes.c_void_p * int(L))(*[A[t].ctypes.data for t in range(int(L))])
        pb = (ctypes.c_void_p * int(L))(*[b[t].ctypes.data for t in range(int(L))])
        return self._ok(self._L.culsh_ht_reset(self._h, int(N), int(D), int(L), int(M), float(W),
                                               pa, pb, debug))

    def index(self, matrix, debug=None):
        """index(matrix): host [D x N] dataset (numpy (N, D) float32)."""
        X = _f32(matrix)
        self._dataset = X
        return self._ok(self._L.culsh_ht_index(self._h, X.ctypes.data, debug))

    def index_device(self, tensor, debug=None):
        """Index a dataset already in HBM (a contiguous float32 torch CUDA tensor (N, D))."""
        self._dataset_dev = tensor
        return self._ok(self._L.culsh_ht_index_device(self._h, tensor.data_ptr(), debug))

    def save(self, filename, debug=None):
        return self._ok(self._L.culsh_ht_save(self._h, str(filename).encode(), debug))

    def load(self, filename, debug=None):
        return self._ok(self._L.culsh_ht_load(self._h, str(filename).encode(), debug))

    def set_form(self, form):
        return self._ok(self._L.culsh_ht_set_form(self._h, int(form)))

    def getN(self):
        return self._L.culsh_ht_N(self._h)

    def getD(self):
        return self._L.culsh_ht_D(self._h)

    def getL(self):
        return self._L.culsh_ht_L(self._h)

    def getM(self):
        return self._L.culsh_ht_M(self._h)

    def getW(self):
        return self._L.culsh_ht_W(self._h)

    def getA(self, table):
        p = self._L.culsh_ht_A(self._h, table)
        return None if not p else _view(p, ctypes.c_float, self.getD() * self.getM())

    def getb(self, table):
        p = self._L.culsh_ht_b(self._h, table)
        return None if not p else _view(p, ctypes.c_float, self.getM())

    def getBuckets(self, table=None):
        if table is None:
            p = self._L.culsh_ht_buckets_all(self._h)
            return None if not p else _view(p, ctypes.c_uint, self.getL())
        return int(self._L.culsh_ht_buckets(self._h, table))

    def getBuckets_totalIndices(self, table):
        p = self._L.culsh_ht_buckets_total_indices(self._h, table)
        return None if not p else _view(p, ctypes.c_uint, self.getN())

    def getBuckets_startingIndices(self, table):
        p = self._L.culsh_ht_buckets_starting_indices(self._h, table)
        return None if not p else _view(p, ctypes.c_uint, self.getBuckets(table))

    def getBuckets_sizes(self, table):
        p = self._L.culsh_ht_buckets_sizes(self._h, table)
        return None if not p else _view(p, ctypes.c_uint, self.getBuckets(table))

    def getBuckets_codes(self, table):
        """Bucket codes as (B, M) float32 (the reference's column-major [B x M])."""
        B, M = self.getBuckets(table), self.getM()
        p = self._L.culsh_ht_buckets_codes(self._h, table)
        return None if not p else _view(p, ctypes.c_float, B * M).reshape(M, B).T.copy()

    def getRechecked(self):
        return int(self._L.culsh_ht_rechecked(self._h))


class SearchTables(_Base):
    """cuLSH::SearchTables: multiprobe kNN search against a HashTables index."""

    def __init__(self):
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        if self._L.culsh_st_create(ctypes.byref(h)) != 0:
            raise RuntimeError(_lib.last_error())
        self._h = h
        self._ht = None
        self.K = self.T = self.Q = 0

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.culsh_st_destroy(self._h)
            self._h = ctypes.c_void_p()

    def reset(self, hashtables, K, T=1):
        self._ht = hashtables
        self.K, self.T = int(K), int(T)
        return self._ok(self._L.culsh_st_reset(self._h, hashtables.handle, self.K, self.T))

    def set_form(self, form):
        return self._ok(self._L.culsh_st_set_form(self._h, int(form)))

    def search(self, queries, Q=None, dataset=None, debug=None):
        """search(queries, Q, dataset): host [D x Q] queries (numpy (Q, D) float32)."""
        Y = _f32(queries)
        Q = Y.shape[0] if Q is None else int(Q)
        if dataset is None:
            dataset = self._ht._dataset if self._ht is not None else None
        if dataset is None:
            self.last_error = "search needs the indexed dataset"
            print(f"[Error] {self.last_error}", flush=True)
            return False
        X = dataset if (isinstance(dataset, np.ndarray) and dataset.dtype == np.float32
                        and dataset.flags.c_contiguous) else _f32(dataset)
        self.Q = Q
        return self._ok(self._L.culsh_st_search(self._h, Y.ctypes.data, Q, X.ctypes.data, debug))

    def search_device(self, queries, ids, dists, stream=0):
        """Device-resident search: torch CUDA tensors queries (Q, D) f32, ids (Q, K) i32,
        dists (Q, K) f32; runs on `stream` (an int handle, 0 = null stream)."""
        Q = queries.shape[0]
        return self._ok(self._L.culsh_st_search_device(self._h, queries.data_ptr(), Q,
                                                       ids.data_ptr(), dists.data_ptr(),
                                                       ctypes.c_void_p(stream)))

    def set_pass_limit(self, max_queries):
        """Caps the queries per pass (0: sized from the free device memory)."""
        return self._ok(self._L.culsh_st_set_pass_limit(self._h, int(max_queries)))

    def search_device_noalloc(self, queries, ids, dists, status, stream=0):
        """Allocation-free device search (plain form): like search_device, with
        the working buffers in the library's fixed device array; `status` is a
        device int32 tensor set non-zero when a query exceeded the fixed
        candidate caps (its results are then not written)."""
        return self._ok(self._L.culsh_st_search_device_noalloc(
            self._h, queries.data_ptr(), int(queries.shape[0]), ids.data_ptr(), dists.data_ptr(),
            status.data_ptr(), ctypes.c_void_p(stream)))

    def wait(self):
        """Waits for the last device search; False (and the message) if it or an
        earlier asynchronous search failed."""
        return self._ok(self._L.culsh_st_wait(self._h))

    def getKnnIds(self):
        p = self._L.culsh_st_knn_ids(self._h)
        return None if not p else _view(p, ctypes.c_int, self.Q * self.K).reshape(self.Q, self.K)

    def getKnnDistances(self):
        p = self._L.culsh_st_knn_dists(self._h)
        return None if not p else _view(p, ctypes.c_float, self.Q * self.K).reshape(self.Q,
                                                                                    self.K)

    getKnnDists = getKnnDistances

    def getQueryBuckets(self, table):
        p = self._L.culsh_st_query_buckets(self._h, table)
        return None if not p else _view(p, ctypes.c_int, self.Q * self.T).reshape(self.Q,
                                                                                  self.T)

    def getSelectivity(self):
        return float(self._L.culsh_st_selectivity(self._h))

    def getNonUniquePercentage(self):
        return float(self._L.culsh_st_nonunique_pct(self._h))

    def getTotalCandidates_unique(self):
        return int(self._L.culsh_st_total_unique(self._h))

    def getTotalCandidates_nonUnique(self):
        return int(self._L.culsh_st_total_nonunique(self._h))

    def getQueryCandidates(self):
        """(C_nu [Q], C_u [Q]) of the last search."""
        a = self._L.culsh_st_query_cnu(self._h)
        b = self._L.culsh_st_query_cu(self._h)
        if not a or not b:
            return None, None
        return _view(a, ctypes.c_ulonglong, self.Q), _view(b, ctypes.c_ulonglong, self.Q)

    def getRechecked(self):
        return int(self._L.culsh_st_rechecked(self._h))

    def getKernelMs(self):
        """(probe kernel ms, candidate kernel ms) of the last search."""
        a, b = ctypes.c_float(), ctypes.c_float()
        if not self._ok(self._L.culsh_st_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b))):
            return None
        return a.value, b.value

    def getTimes(self):
        t = Times()
        self._L.culsh_st_times(self._h, ctypes.byref(t))
        return {f: getattr(t, f) for f, _ in Times._fields_}


def _requests(reqs):
    """[(Q, K, T, q_row, out_offset), ...] -> ctypes array of Request."""
    arr = (Request * len(reqs))()
    for i, (Q, K, T, row, off) in enumerate(reqs):
        arr[i] = Request(int(Q), int(K), int(T), 0, int(row), int(off))
    return arr


def layout_requests(shapes):
    """Packs requests [(Q, K, T), ...] back to back: returns the request
    tuples (Q, K, T, q_row, out_offset), the total query rows and the total
    result elements."""
    out, row, off = [], 0, 0
    for Q, K, T in shapes:
        out.append((Q, K, T, row, off))
        row += Q
        off += Q * K
    return out, row, off


class Batch(_Base):
    """Batching form: independent requests, each with its own Q, K and T,
    searched against one index by one graph launch (persistent fused kernel).
    Each request's results equal a SearchTables search of it alone."""

    def __init__(self, hashtables):
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        if self._L.culsh_bt_create(hashtables.handle, ctypes.byref(h)) != 0:
            raise RuntimeError(_lib.last_error())
        self._h = h
        self._ht = hashtables

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.culsh_bt_destroy(self._h)
            self._h = ctypes.c_void_p()

    def set_schedule(self, probe_queries=32, probe_tables=1, queries_per_item=32,
                     rows_in_flight=4, wgs_per_cu=3, fused=0):
        s = Schedule(probe_queries, probe_tables, queries_per_item, rows_in_flight, wgs_per_cu,
                     fused)
        return self._ok(self._L.culsh_bt_set_schedule(self._h, ctypes.byref(s)))

    def get_schedule(self):
        s = Schedule()
        self._ok(self._L.culsh_bt_get_schedule(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Schedule._fields_}

    def search_device(self, reqs, queries, ids, dists, stream=0, wait=True):
        """reqs: [(Q, K, T, q_row, out_offset)]; queries: device (rows, D) f32 tensor;
        ids / dists: flat device i32 / f32 tensors holding every request's [K x Q]."""
        arr = _requests(reqs)
        ok = self._ok(self._L.culsh_bt_search_device(self._h, arr, len(reqs), queries.data_ptr(),
                                                     ids.data_ptr(), dists.data_ptr(),
                                                     ctypes.c_void_p(stream)))
        if ok and wait:
            ok = self.wait()
        return ok

    def wait(self):
        return self._ok(self._L.culsh_bt_wait(self._h))

    def last_ms(self):
        v = ctypes.c_float()
        return v.value if self._ok(self._L.culsh_bt_last_ms(self._h, ctypes.byref(v))) else None

    def getRechecked(self):
        return int(self._L.culsh_bt_rechecked(self._h))

    def profile(self, reqs, queries, ids, dists, stream=0):
        """One eager run of the batch's kernels (no graph) with HIP events
        between stages on `stream`; returns the stage times in ms
        {probe, locality, candidates, rest} or None on failure."""
        arr = _requests(reqs)
        ms = (ctypes.c_float * 4)()
        if not self._ok(self._L.culsh_bt_profile(self._h, arr, len(reqs), queries.data_ptr(),
                                                 ids.data_ptr(), dists.data_ptr(),
                                                 ctypes.c_void_p(stream), ms)):
            return None
        return dict(zip(("probe", "locality", "candidates", "rest"), list(ms)))

    def autotune(self, reqs, queries, ids, dists, reps=2):
        """Returns (naive default ms, tuned ms); the tuned schedule stays set."""
        arr = _requests(reqs)
        a, b = ctypes.c_float(), ctypes.c_float()
        if not self._ok(self._L.culsh_bt_autotune(self._h, arr, len(reqs), queries.data_ptr(),
                                                  ids.data_ptr(), dists.data_ptr(), int(reps),
                                                  ctypes.byref(a), ctypes.byref(b))):
            return None
        return a.value, b.value


class Server(_Base):
    """Serving form: a real-time request loop over device-resident indexes.
    batching=True: continuous batching (slots of the batching form, sized
    from the free HBM; a request larger than a slot grows it when the HBM
    free at that moment allows, unless grow=False, or runs in pieces);
    batching=False: one request at a time through the search path (the plain
    baseline)."""

    def __init__(self, indexes, batching=True, slots_per_index=2, hbm_fraction=0.5,
                 max_batch_queries=0, grow=True):
        self._L = _lib.lib()
        self._hts = list(indexes)
        arr = (ctypes.c_void_p * len(self._hts))(*[h.handle.value for h in self._hts])
        cfg = SrvConfig(int(slots_per_index), int(bool(batching)), float(hbm_fraction),
                        int(max_batch_queries), int(not grow))
        h = ctypes.c_void_p()
        if self._L.culsh_srv_create(arr, len(self._hts), ctypes.byref(cfg), ctypes.byref(h)) != 0:
            raise RuntimeError(_lib.last_error())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.culsh_srv_destroy(self._h)
            self._h = ctypes.c_void_p()

    def run(self, reqs, queries, ids, dists, host=None):
        """reqs: [(arrival_s, index, Q, K, T, q_row, out_offset)] sorted by arrival;
        queries / ids / dists: per index, device tensors.  host: optional
        (queries, ids, dists) per index as pinned CPU tensors: each batch then
        uploads its query rows and downloads its results, and latencies
        include both.  Returns (latencies_ms numpy array, stats dict) or None
        on failure."""
        n = len(reqs)
        arr = (SrvRequest * n)()
        for i, (a, ix, Q, K, T, row, off) in enumerate(reqs):
            arr[i] = SrvRequest(float(a), int(ix), int(Q), int(K), int(T), int(row), int(off))
        k = len(self._hts)

        def ptrs(ts):
            return (ctypes.c_void_p * k)(*[t.data_ptr() for t in ts])

        lat = np.zeros(n, np.float64)
        st = SrvStats()
        latp = lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if host is None:
            rc = self._L.culsh_srv_run(self._h, arr, n, ptrs(queries), ptrs(ids), ptrs(dists),
                                       latp, ctypes.byref(st))
        else:
            hq, hi, hd = host
            rc = self._L.culsh_srv_run_host(self._h, arr, n, ptrs(hq), ptrs(hi), ptrs(hd),
                                            ptrs(queries), ptrs(ids), ptrs(dists), latp,
                                            ctypes.byref(st))
        if not self._ok(rc):
            return None
        return lat, {f: getattr(st, f) for f, _ in SrvStats._fields_}
