"""ctypes binding of libculsh.so (C ABI declared in include/culsh_c.h).

Loading fails loudly when the library is not built: there is no fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libculsh.so")

_lib = None

c_int, c_uint, c_float, c_double = ctypes.c_int, ctypes.c_uint, ctypes.c_float, ctypes.c_double
c_ull, c_vp, c_char_p = ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_char_p
P = ctypes.POINTER


class Times(ctypes.Structure):
    _fields_ = [("buckets", c_float), ("collect", c_float), ("sort", c_float),
                ("unique", c_float), ("distances", c_float), ("knn", c_float)]


class Request(ctypes.Structure):
    _fields_ = [("Q", c_int), ("K", c_int), ("T", c_int), ("reserved", c_int),
                ("q_row", ctypes.c_longlong), ("out_offset", ctypes.c_longlong)]


class Schedule(ctypes.Structure):
    _fields_ = [("probe_queries", c_int), ("probe_tables", c_int), ("queries_per_item", c_int),
                ("rows_in_flight", c_int), ("wgs_per_cu", c_int), ("fused", c_int)]


class SrvRequest(ctypes.Structure):
    _fields_ = [("arrival_s", c_double), ("index", c_int), ("Q", c_int), ("K", c_int),
                ("T", c_int), ("q_row", ctypes.c_longlong), ("out_offset", ctypes.c_longlong)]


class SrvConfig(ctypes.Structure):
    _fields_ = [("slots_per_index", c_int), ("batching", c_int), ("hbm_fraction", c_double),
                ("max_batch_queries", ctypes.c_longlong), ("no_growth", c_int)]


class SrvStats(ctypes.Structure):
    _fields_ = [("p50_ms", c_double), ("p99_ms", c_double), ("mean_ms", c_double),
                ("max_ms", c_double), ("queries_per_s", c_double), ("requests_per_s", c_double),
                ("batches", ctypes.c_longlong), ("queries", ctypes.c_longlong),
                ("requests", ctypes.c_longlong), ("mean_batch_requests", c_double),
                ("slot_bytes", ctypes.c_longlong), ("deferred", ctypes.c_longlong),
                ("split_requests", ctypes.c_longlong), ("grown_slots", ctypes.c_longlong),
                ("host_io", c_int)]


# name -> (restype, argtypes); every symbol include/culsh_c.h declares.
SIGNATURES = {
    "culsh_last_error": (c_char_p, []),
    "culsh_version": (c_char_p, []),
    "culsh_ht_create": (c_int, [P(c_vp)]),
    "culsh_ht_destroy": (None, [c_vp]),
    "culsh_ht_reset_random": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_float, c_ull, c_vp]),
    "culsh_ht_reset": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_float, P(c_vp), P(c_vp),
                               c_vp]),
    "culsh_ht_load": (c_int, [c_vp, c_char_p, c_vp]),
    "culsh_ht_save": (c_int, [c_vp, c_char_p, c_vp]),
    "culsh_ht_index": (c_int, [c_vp, c_vp, c_vp]),
    "culsh_ht_index_device": (c_int, [c_vp, c_vp, c_vp]),
    "culsh_ht_N": (c_int, [c_vp]),
    "culsh_ht_D": (c_int, [c_vp]),
    "culsh_ht_L": (c_int, [c_vp]),
    "culsh_ht_M": (c_int, [c_vp]),
    "culsh_ht_W": (c_float, [c_vp]),
    "culsh_ht_A": (c_vp, [c_vp, c_int]),
    "culsh_ht_b": (c_vp, [c_vp, c_int]),
    "culsh_ht_buckets": (c_uint, [c_vp, c_int]),
    "culsh_ht_buckets_all": (c_vp, [c_vp]),
    "culsh_ht_buckets_total_indices": (c_vp, [c_vp, c_int]),
    "culsh_ht_buckets_starting_indices": (c_vp, [c_vp, c_int]),
    "culsh_ht_buckets_sizes": (c_vp, [c_vp, c_int]),
    "culsh_ht_buckets_codes": (c_vp, [c_vp, c_int]),
    "culsh_ht_rechecked": (c_ull, [c_vp]),
    "culsh_ht_set_form": (c_int, [c_vp, c_int]),
    "culsh_st_create": (c_int, [P(c_vp)]),
    "culsh_st_destroy": (None, [c_vp]),
    "culsh_st_reset": (c_int, [c_vp, c_vp, c_int, c_int]),
    "culsh_st_search": (c_int, [c_vp, c_vp, c_int, c_vp, c_vp]),
    "culsh_st_search_device": (c_int, [c_vp, c_vp, c_int, c_vp, c_vp, c_vp]),
    "culsh_st_set_pass_limit": (c_int, [c_vp, c_int]),
    "culsh_st_search_device_noalloc": (c_int, [c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp]),
    "culsh_st_wait": (c_int, [c_vp]),
    "culsh_st_set_form": (c_int, [c_vp, c_int]),
    "culsh_st_knn_ids": (c_vp, [c_vp]),
    "culsh_st_knn_dists": (c_vp, [c_vp]),
    "culsh_st_query_buckets": (c_vp, [c_vp, c_int]),
    "culsh_st_selectivity": (c_float, [c_vp]),
    "culsh_st_nonunique_pct": (c_float, [c_vp]),
    "culsh_st_total_unique": (c_ull, [c_vp]),
    "culsh_st_total_nonunique": (c_ull, [c_vp]),
    "culsh_st_times": (None, [c_vp, P(Times)]),
    "culsh_st_query_cnu": (c_vp, [c_vp]),
    "culsh_st_query_cu": (c_vp, [c_vp]),
    "culsh_st_rechecked": (c_ull, [c_vp]),
    "culsh_st_kernel_ms": (c_int, [c_vp, P(c_float), P(c_float)]),
    "culsh_bt_create": (c_int, [c_vp, P(c_vp)]),
    "culsh_bt_destroy": (None, [c_vp]),
    "culsh_bt_set_schedule": (c_int, [c_vp, P(Schedule)]),
    "culsh_bt_get_schedule": (c_int, [c_vp, P(Schedule)]),
    "culsh_bt_search_device": (c_int, [c_vp, P(Request), c_int, c_vp, c_vp, c_vp, c_vp]),
    "culsh_bt_wait": (c_int, [c_vp]),
    "culsh_bt_last_ms": (c_int, [c_vp, P(c_float)]),
    "culsh_bt_rechecked": (c_ull, [c_vp]),
    "culsh_bt_profile": (c_int, [c_vp, P(Request), c_int, c_vp, c_vp, c_vp, c_vp, P(c_float)]),
    "culsh_bt_autotune": (c_int, [c_vp, P(Request), c_int, c_vp, c_vp, c_vp, c_int, P(c_float),
                                  P(c_float)]),
    "culsh_srv_create": (c_int, [P(c_vp), c_int, P(SrvConfig), P(c_vp)]),
    "culsh_srv_destroy": (None, [c_vp]),
    "culsh_srv_run": (c_int, [c_vp, P(SrvRequest), c_int, P(c_vp), P(c_vp), P(c_vp), P(c_double),
                              P(SrvStats)]),
    "culsh_srv_run_host": (c_int, [c_vp, P(SrvRequest), c_int, P(c_vp), P(c_vp), P(c_vp),
                                   P(c_vp), P(c_vp), P(c_vp), P(c_double), P(SrvStats)]),
}

FORM_PLAIN = 0
FORM_PRECISION = 1
FORM_PRECISION_GROUPED = 2


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libculsh.so is not built ({LIB_PATH}); run `python -m culsh_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().culsh_last_error().decode()
