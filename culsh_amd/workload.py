"""Seeded synthetic workloads for the LSH index/query path.

Thin ctypes wrapper over ``libculsh_workload.so`` (``csrc/workload.cpp``:
Philox4x32-10 counter-based draws, Box-Muller normals, Gamma(2, theta) as a
sum of two exponentials).  Arrays use the cuLSH layout: a dataset is
column-major float ``[D x N]``, i.e. a C-contiguous numpy array of shape
``(N, D)``; ``A`` is ``(L, M, D)`` (per table, column-major ``[D x M]``) and
``b`` is ``(L, M)``.

Seeds: data ``1000c+1``, queries ``1000c+2``, A/b ``1000c+3`` for config c.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libculsh_workload.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(_LIB_PATH)
        c_ll, c_int, c_d, c_u64 = (ctypes.c_longlong, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_ulonglong)
        L.culsh_gen_mixture.argtypes = [_f32p, c_ll, c_int, c_int, c_d, c_d, c_d, c_u64]
        L.culsh_gen_gamma_ints.argtypes = [_f32p, c_ll, c_int, c_d, c_d, c_d, c_u64]
        L.culsh_gen_queries_near.argtypes = [_f32p, c_int, _f32p, c_ll, c_int, c_d, c_u64]
        L.culsh_gen_projection.argtypes = [_f32p, _f32p, c_int, c_int, c_int, ctypes.c_float,
                                           c_u64]
        L.culsh_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint * 4),
                                          ctypes.POINTER(ctypes.c_uint * 2),
                                          ctypes.POINTER(ctypes.c_uint * 4)]
        for f in ("culsh_gen_mixture", "culsh_gen_gamma_ints", "culsh_gen_queries_near",
                  "culsh_gen_projection", "culsh_philox4x32_10"):
            getattr(L, f).restype = None
        _lib = L
    return _lib


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint * 4)(*ctr)
    k = (ctypes.c_uint * 2)(*key)
    o = (ctypes.c_uint * 4)()
    lib().culsh_philox4x32_10(ctypes.byref(c), ctypes.byref(k), ctypes.byref(o))
    return list(o)


def mixture(N, D, n_centers, sigma, seed, lo=-1.0, hi=1.0):
    X = np.empty((N, D), np.float32)
    lib().culsh_gen_mixture(X, N, D, n_centers, sigma, lo, hi, seed)
    return X


def gamma_ints(N, D, seed, p_zero=0.6, theta=25.0, maxv=255.0):
    X = np.empty((N, D), np.float32)
    lib().culsh_gen_gamma_ints(X, N, D, p_zero, theta, maxv, seed)
    return X


def queries_near(X, Q, noise, seed):
    X = np.ascontiguousarray(X, np.float32)
    N, D = X.shape
    Y = np.empty((Q, D), np.float32)
    lib().culsh_gen_queries_near(Y, Q, X, N, D, noise, seed)
    return Y


def projection(L, D, M, W, seed):
    """A (L, M, D) float32 with N(0,1) entries, b (L, M) with entries in [0, W)."""
    A = np.empty((L, M, D), np.float32)
    b = np.empty((L, M), np.float32)
    lib().culsh_gen_projection(A, b, L, D, M, W, seed)
    return A, b


@dataclass
class Config:
    name: str
    N: int
    D: int
    L: int
    M: int
    W: float
    Q: int
    K: int
    T: int
    data: str                       # "mixture" | "gamma_ints"
    n_centers: int = 0
    sigma: float = 0.0
    query: str = "near"             # "near" (data point + noise) | "same" (drawn like data)
    query_noise: float = 0.01
    seed_base: int = 0
    extra: dict = field(default_factory=dict)

    def describe(self):
        return (f"{self.name}: N={self.N} D={self.D} L={self.L} M={self.M} W={self.W} "
                f"Q={self.Q} K={self.K} T={self.T} data={self.data}")


# The five workloads of the LSH path (sizes from the benchmark definition).
CONFIGS = {
    "c1": Config("c1-mixture-100k", N=100_000, D=128, L=8, M=16, W=4.0, Q=1_000, K=10, T=1,
                 data="mixture", n_centers=100, sigma=0.1, seed_base=1000),
    "c2": Config("c2-gamma-1M-multiprobe", N=1_000_000, D=128, L=20, M=24, W=400.0, Q=10_000,
                 K=10, T=32, data="gamma_ints", query="same", seed_base=2000),
    "c3": Config("c3-mixture-4M-precision", N=4_000_000, D=96, L=32, M=20, W=2.0, Q=50_000,
                 K=100, T=64, data="mixture", n_centers=1000, sigma=0.05, seed_base=3000),
    # Serving indexes: (a) 2M x 128 (W chosen for recall@10 >= 0.9), (b) 1M x 960.
    "c5a": Config("c5a-mixture-2M-serve", N=2_000_000, D=128, L=24, M=20, W=3.0, Q=0, K=10,
                  T=16, data="mixture", n_centers=1000, sigma=0.05, seed_base=5000),
    "c5b": Config("c5b-mixture-1M-d960-serve", N=1_000_000, D=960, L=16, M=28, W=3.0, Q=0,
                  K=10, T=16, data="mixture", n_centers=500, sigma=0.05, seed_base=5100),
}


def log_uniform_q(rng, n, qmax):
    """Q = floor(exp(U ln(qmax + 1))), U ~ U[0, 1): log-uniform on [1, qmax]."""
    return np.floor(np.exp(rng.uniform(0.0, np.log(qmax + 1.0), n))).astype(np.int64).clip(1, qmax)


def batch_requests(n=256, qmax=4096, Ks=(1, 10, 100), Ts=(1, 8, 64), seed=4004):
    """The batching workload's requests against the c3 index: (Q, K, T) with Q
    log-uniform in [1, qmax], K and T uniform over their sets."""
    rng = np.random.default_rng(seed)
    Q = log_uniform_q(rng, n, qmax)
    K = rng.choice(np.asarray(Ks), n)
    T = rng.choice(np.asarray(Ts), n)
    return [(int(q), int(k), int(t)) for q, k, t in zip(Q, K, T)]


def serve_stream(rate, n, qmax=512, p_a=0.7, Ks=(10, 100), Ts=(16, 64), seed=5004):
    """The serving workload's request stream: Poisson arrivals at `rate`
    requests/s; per request (arrival s, index 0=a / 1=b, Q, K, T) with Q
    log-uniform in [1, qmax]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    idx = (rng.uniform(size=n) >= p_a).astype(np.int64)
    Q = log_uniform_q(rng, n, qmax)
    K = rng.choice(np.asarray(Ks), n)
    T = rng.choice(np.asarray(Ts), n)
    return [(float(a), int(i), int(q), int(k), int(tt)) for a, i, q, k, tt in zip(t, idx, Q, K, T)]


def make_data(cfg, N=None, Q=None):
    """(X [N, D], queries [Q, D], A (L, M, D), b (L, M)) for a config, optionally shrunk."""
    N = cfg.N if N is None else N
    Q = cfg.Q if Q is None else Q
    s = cfg.seed_base
    if cfg.data == "mixture":
        X = mixture(N, cfg.D, cfg.n_centers, cfg.sigma, s + 1)
    elif cfg.data == "gamma_ints":
        X = gamma_ints(N, cfg.D, s + 1)
    else:
        raise ValueError(cfg.data)
    if cfg.query == "near":
        Y = queries_near(X, Q, cfg.query_noise, s + 2)
    else:
        Y = gamma_ints(Q, cfg.D, s + 2)
    A, b = projection(cfg.L, cfg.D, cfg.M, cfg.W, s + 3)
    return X, Y, A, b
