"""HIP path vs the FP64 oracle on the same seeded inputs (through the C ABI).

Bar: bit-exact ids, distances ((float) of the FP64 distance), per-query
candidate counts, query buckets and bucket arrays (lexicographic numbering,
as the reference produces them)."""
import numpy as np
import pytest

import oracle
from culsh_amd import workload as wl
from lshtest import assert_buckets_equal, assert_search_equal, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _small(seed, N=2000, D=16, L=4, M=6, W=1.0, Q=50, centers=20, sigma=0.2):
    X = wl.mixture(N, D, centers, sigma, seed)
    Y = wl.queries_near(X, Q, 0.05, seed + 1)
    A, b = wl.projection(L, D, M, W, seed + 2)
    return X, Y, A, b


@pytest.mark.parametrize("T", [1, 8, 40])
@pytest.mark.parametrize("K", [1, 5, 33])
def test_small_mixture(T, K):
    print(f"small mixture T={T} K={K}", flush=True)
    X, Y, A, b = _small(11)
    g = run_gpu(X, Y, A, b, 4, 6, 1.0, K, T)
    o = run_oracle(X, Y, A, b, 4, 6, 1.0, K, T)
    assert_buckets_equal(g["ht"], o["index"], 4)
    assert_search_equal(g, o)


@pytest.mark.parametrize("D", [32, 64, 96])
def test_row_widths_of_the_8_lane_distance_path(D):
    """D = 32, 64, 96 take the candidate stage's 8-lanes-per-row FP32
    distances (1, 2, 3 loads per lane); a few hundred candidates per query
    (wide W) so every row round and the near-K-th FP64 recheck run."""
    print("8-lane rows D=%d" % D, flush=True)
    X, Y, A, b = _small(900 + D, N=6000, D=D, L=4, M=6, W=3.0, Q=60, centers=12, sigma=0.2)
    g = run_gpu(X, Y, A, b, 4, 6, 3.0, 20, 8)
    o = run_oracle(X, Y, A, b, 4, 6, 3.0, 20, 8)
    assert_search_equal(g, o)
    cu = np.asarray(g["cu"])
    assert (cu > 64).sum() >= 10 and (cu <= 4096).sum() >= 10  # rows in several rounds, LDS path


def test_tie_heavy_integers():
    """Small integer data: many exact distance ties at the K-th boundary, so the
    reference's tie rule decides which ids are returned."""
    print("tie-heavy integers", flush=True)
    rng = np.random.default_rng(5)
    N, D, L, M, W, Q = 3000, 8, 6, 4, 3.0, 200
    X = rng.integers(0, 4, size=(N, D)).astype(np.float32)
    Y = rng.integers(0, 4, size=(Q, D)).astype(np.float32)
    A, b = wl.projection(L, D, M, W, 77)
    for K, T in [(10, 1), (16, 12), (3, 64), (50, 5)]:
        g = run_gpu(X, Y, A, b, L, M, W, K, T)
        o = run_oracle(X, Y, A, b, L, M, W, K, T)
        assert_search_equal(g, o)


def test_wide_probe_queue_and_many_dims():
    """T > 64 (two queue slots per lane), M > 32 (64-delta truncation), D not a
    multiple of 4."""
    print("wide queue", flush=True)
    X = wl.mixture(4000, 37, 10, 0.3, 21)
    Y = wl.queries_near(X, 40, 0.05, 22)
    L, M, W = 3, 40, 2.0
    A, b = wl.projection(L, 37, M, W, 23)
    for K, T in [(7, 100), (4, 65)]:
        g = run_gpu(X, Y, A, b, L, M, W, K, T)
        o = run_oracle(X, Y, A, b, L, M, W, K, T)
        assert_buckets_equal(g["ht"], o["index"], L)
        assert_search_equal(g, o)


def test_wave_kernel_paths():
    """T > 128 and D*M beyond the octet kernel's LDS staging use the
    wave-per-(query, table) probe kernel."""
    print("wave kernel paths", flush=True)
    X = wl.mixture(3000, 24, 10, 0.3, 71)
    Y = wl.queries_near(X, 30, 0.05, 72)
    A, b = wl.projection(2, 24, 6, 1.0, 73)
    g = run_gpu(X, Y, A, b, 2, 6, 1.0, 5, 150)
    o = run_oracle(X, Y, A, b, 2, 6, 1.0, 5, 150)
    assert_search_equal(g, o)
    X = wl.mixture(2000, 520, 10, 0.05, 74)
    Y = wl.queries_near(X, 30, 0.01, 75)
    A, b = wl.projection(2, 520, 40, 4.0, 76)
    g = run_gpu(X, Y, A, b, 2, 40, 4.0, 5, 20)
    o = run_oracle(X, Y, A, b, 2, 40, 4.0, 5, 20)
    assert_search_equal(g, o)


def test_large_candidate_sets_use_global_path():
    """Wide buckets (large W) give tens of thousands of unique candidates per
    query, beyond the LDS working set; K=1500 exceeds the LDS recheck set.
    Both go through the global-memory candidate kernel."""
    print("large candidate sets", flush=True)
    X = wl.mixture(30000, 16, 3, 0.3, 81)
    Y = wl.queries_near(X, 40, 0.05, 82)
    A, b = wl.projection(3, 16, 3, 8.0, 83)
    for K, T in [(10, 4), (1500, 2)]:
        g = run_gpu(X, Y, A, b, 3, 3, 8.0, K, T)
        o = run_oracle(X, Y, A, b, 3, 3, 8.0, K, T)
        assert int(o["cu"].max()) > 4096
        assert_search_equal(g, o)


def test_locality_ordered_large_batch():
    """Batches of 16384+ queries run the candidate kernel in locality order
    (each XCD takes a contiguous run of it); Q = 16387 leaves a partial
    last run."""
    print("locality-ordered batch Q=16387", flush=True)
    X = wl.mixture(20000, 16, 40, 0.2, 101)
    Y = wl.queries_near(X, 16387, 0.05, 102)
    A, b = wl.projection(4, 16, 6, 1.0, 103)
    g = run_gpu(X, Y, A, b, 4, 6, 1.0, 5, 8)
    o = run_oracle(X, Y, A, b, 4, 6, 1.0, 5, 8)
    assert_search_equal(g, o)


@pytest.mark.parametrize("form", [0, 1])
def test_wide_rows_take_the_fp64_projection(form):
    """D = 960 and a small W: more than 2^20 projections fall within the
    FP32 / BF16x3 error bound of a bucket edge, so the index is projected by
    the tiled FP64 kernel (the oracle's order, no recheck)."""
    print("wide rows FP64 projection form", form, flush=True)
    X = wl.mixture(20000, 960, 50, 0.05, 111)
    Y = wl.queries_near(X, 60, 0.01, 112)
    L, M, W = 16, 28, 0.25
    A, b = wl.projection(L, 960, M, W, 113)
    g = run_gpu(X, Y, A, b, L, M, W, 10, 16, form=form)
    assert g["ht"].getRechecked() > (1 << 20)
    o = run_oracle(X, Y, A, b, L, M, W, 10, 16)
    assert_buckets_equal(g["ht"], o["index"], L)
    assert_search_equal(g, o)


def test_exact_delta_ties():
    """Duplicated projection columns give exactly tied deltas: the delta order
    then follows the reference heap's mechanics."""
    print("delta ties", flush=True)
    X = wl.mixture(3000, 12, 8, 0.3, 31)
    Y = wl.queries_near(X, 60, 0.05, 32)
    L, M, W = 2, 6, 1.5
    A, b = wl.projection(L, 12, M, W, 33)
    A[:, 3] = A[:, 1]
    b[:, 3] = b[:, 1]
    A[:, 5] = A[:, 0]
    b[:, 5] = b[:, 0]
    g = run_gpu(X, Y, A, b, L, M, W, 6, 30)
    o = run_oracle(X, Y, A, b, L, M, W, 6, 30)
    assert_search_equal(g, o)


def test_delta_ties_around_the_probe_window():
    """M=40: 80 deltas, of which the first 64 feed the queue.  Duplicated
    columns tie a dimension's "-1" and "+1" deltas with another's; one of the
    two tied pairs usually ranks inside the first 64 and the other often
    beyond, so both the heap-order path and the plain sort are exercised."""
    print("delta ties M=40", flush=True)
    X = wl.mixture(4000, 16, 8, 0.3, 41)
    Y = wl.queries_near(X, 80, 0.05, 42)
    L, M, W = 2, 40, 1.5
    A, b = wl.projection(L, 16, M, W, 43)
    for src, dst in ((1, 3), (0, 5), (7, 30), (12, 39), (20, 21)):
        A[:, dst] = A[:, src]
        b[:, dst] = b[:, src]
    g = run_gpu(X, Y, A, b, L, M, W, 5, 64)
    o = run_oracle(X, Y, A, b, L, M, W, 5, 64)
    assert_search_equal(g, o)


def test_config1_full():
    """c1: N=100k, D=128, L=8, M=16, W=4, Q=1k, K=10, T=1 -- ids exact."""
    cfg = wl.CONFIGS["c1"]
    print(cfg.describe(), flush=True)
    X, Y, A, b = wl.make_data(cfg)
    g = run_gpu(X, Y, A, b, cfg.L, cfg.M, cfg.W, cfg.K, cfg.T)
    o = run_oracle(X, Y, A, b, cfg.L, cfg.M, cfg.W, cfg.K, cfg.T)
    assert_buckets_equal(g["ht"], o["index"], cfg.L)
    assert_search_equal(g, o)


def test_config2_full():
    """c2: N=1M, D=128 integer data, L=20, M=24, W=400, Q=10k, K=10, T=32."""
    cfg = wl.CONFIGS["c2"]
    print(cfg.describe(), flush=True)
    X, Y, A, b = wl.make_data(cfg)
    g = run_gpu(X, Y, A, b, cfg.L, cfg.M, cfg.W, cfg.K, cfg.T)
    o = run_oracle(X, Y, A, b, cfg.L, cfg.M, cfg.W, cfg.K, cfg.T)
    assert_search_equal(g, o)


def test_config3_slice():
    """c3 parameters (D=96, L=32, M=20, W=2, K=100, T=64) on a 400k-point,
    1000-query slice of its mixture."""
    cfg = wl.CONFIGS["c3"]
    print(cfg.describe(), "slice N=400000 Q=1000", flush=True)
    X, Y, A, b = wl.make_data(cfg, N=400_000, Q=1_000)
    g = run_gpu(X, Y, A, b, cfg.L, cfg.M, cfg.W, cfg.K, cfg.T)
    o = run_oracle(X, Y, A, b, cfg.L, cfg.M, cfg.W, cfg.K, cfg.T)
    assert_search_equal(g, o)


@pytest.mark.parametrize("st_form", [1, 2])
@pytest.mark.parametrize("case", ["small_T8", "small_T40_K33", "ties", "c1", "c3_slice",
                                  "large", "coarse_half", "d_not_multiple_of_8"])
def test_precision_form_matches_oracle(case, st_form):
    """Precision form: BF16x3 MFMA projection for the index; the search with
    the distance pass on the FP16 copy of the resident rows (1) or grouped
    norm-expansion MFMA distances (2); FP64 rechecks keep every result
    bit-exact.  coarse_half: values near 1000, whose half spacing (0.5) is
    large next to the distances (wide FP64 bands); d_not_multiple_of_8: no
    FP16 copy, FP32 rows."""
    print("precision", case, "search form", st_form, flush=True)
    if case in ("small_T8", "small_T40_K33"):
        X, Y, A, b = _small(91)
        L, M, W = 4, 6, 1.0
        K, T = (5, 8) if case == "small_T8" else (33, 40)
    elif case == "ties":
        rng = np.random.default_rng(92)
        X = rng.integers(0, 4, size=(3000, 8)).astype(np.float32)
        Y = rng.integers(0, 4, size=(200, 8)).astype(np.float32)
        L, M, W, K, T = 6, 4, 3.0, 16, 12
        A, b = wl.projection(L, 8, M, W, 93)
    elif case == "c1":
        cfg = wl.CONFIGS["c1"]
        X, Y, A, b = wl.make_data(cfg)
        L, M, W, K, T = cfg.L, cfg.M, cfg.W, cfg.K, cfg.T
    elif case == "c3_slice":
        cfg = wl.CONFIGS["c3"]
        X, Y, A, b = wl.make_data(cfg, N=400_000, Q=1_000)
        L, M, W, K, T = cfg.L, cfg.M, cfg.W, cfg.K, cfg.T
    elif case == "coarse_half":
        X = (wl.mixture(8000, 24, 20, 0.2, 84) + 1000.0).astype(np.float32)
        Y = wl.queries_near(X, 150, 0.05, 85)
        A, b = wl.projection(5, 24, 8, 1.5, 86)
        L, M, W, K, T = 5, 8, 1.5, 20, 16
    elif case == "d_not_multiple_of_8":
        X = wl.mixture(8000, 20, 20, 0.2, 87)
        Y = wl.queries_near(X, 150, 0.05, 88)
        A, b = wl.projection(5, 20, 8, 1.5, 89)
        L, M, W, K, T = 5, 8, 1.5, 20, 16
    else:
        X = wl.mixture(30000, 16, 3, 0.3, 81)
        Y = wl.queries_near(X, 40, 0.05, 82)
        A, b = wl.projection(3, 16, 3, 8.0, 83)
        L, M, W, K, T = 3, 3, 8.0, 1500, 2
    g = run_gpu(X, Y, A, b, L, M, W, K, T, form=1, st_form=st_form)
    o = run_oracle(X, Y, A, b, L, M, W, K, T)
    assert_buckets_equal(g["ht"], o["index"], L)
    assert_search_equal(g, o)
    print("  index rechecks", g["ht"].getRechecked(), "search rechecks",
          g["st"].getRechecked(), flush=True)


def test_device_api_matches_host_api():
    import torch
    print("device api", flush=True)
    X, Y, A, b = _small(41, N=5000, Q=300)
    g = run_gpu(X, Y, A, b, 4, 6, 1.0, 9, 16, want_buckets=False)
    ht, st = g["ht"], g["st"]
    dY = torch.from_numpy(Y).cuda()
    ids = torch.empty((Y.shape[0], 9), dtype=torch.int32, device="cuda")
    d = torch.empty((Y.shape[0], 9), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert st.search_device(dY, ids, d), st.last_error
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy(), g["ids"])
    np.testing.assert_array_equal(d.cpu().numpy(), g["dists"])


def test_device_search_overflow_runs_on_the_device():
    """The asynchronous device search handles queries beyond the LDS working
    set with its fixed-size device kernel (no host round trip) and matches
    the host API; a query beyond that kernel's 2^20 candidates is reported by
    wait(), and the object keeps working."""
    import torch
    from culsh_amd import HashTables, SearchTables
    print("device search overflow", flush=True)
    X = wl.mixture(30000, 16, 3, 0.3, 81)
    Y = wl.queries_near(X, 40, 0.05, 82)
    A, b = wl.projection(3, 16, 3, 8.0, 83)
    g = run_gpu(X, Y, A, b, 3, 3, 8.0, 10, 4, want_buckets=False)
    assert int(g["cu"].max()) > 4096
    st = g["st"]
    dY = torch.from_numpy(Y).cuda()
    ids = torch.empty((Y.shape[0], 10), dtype=torch.int32, device="cuda")
    d = torch.empty((Y.shape[0], 10), dtype=torch.float32, device="cuda")
    assert st.search_device(dY, ids, d), st.last_error
    assert st.wait(), st.last_error
    np.testing.assert_array_equal(ids.cpu().numpy(), g["ids"])
    np.testing.assert_array_equal(d.cpu().numpy(), g["dists"])
    # Every point in one bucket per table: 4 x 300,000 > 2^20 candidates.
    rng = np.random.default_rng(7)
    Xb = rng.standard_normal((300000, 4)).astype(np.float32)
    Ab = np.zeros((4, 1, 4), np.float32)  # every code is exactly 0
    bb = np.zeros((4, 1), np.float32)
    ht = HashTables()
    assert ht.reset(300000, 4, 4, 1, 1e6, Ab, bb)
    assert ht.index(Xb)
    st2 = SearchTables()
    assert st2.reset(ht, 5, 1)
    qy = torch.from_numpy(Xb[:2].copy()).cuda()
    i2 = torch.empty((2, 5), dtype=torch.int32, device="cuda")
    d2 = torch.empty((2, 5), dtype=torch.float32, device="cuda")
    assert st2.search_device(qy, i2, d2)
    assert not st2.wait()
    assert "asynchronous" in st2.last_error
    assert st2.wait()
    assert st2.search(Xb[:2], 2, Xb), st2.last_error
    assert (st2.getKnnIds()[:, 0] == np.arange(2)).all()


def test_index_device_resident_dataset():
    import torch
    print("index from device", flush=True)
    from culsh_amd import HashTables, SearchTables
    X, Y, A, b = _small(51, N=3000, Q=80)
    o = run_oracle(X, Y, A, b, 4, 6, 1.0, 5, 8)
    dX = torch.from_numpy(X).cuda()
    ht = HashTables()
    assert ht.reset(X.shape[0], X.shape[1], 4, 6, 1.0, A, b)
    assert ht.index_device(dX), ht.last_error
    st = SearchTables()
    assert st.reset(ht, 5, 8)
    assert st.search(Y, Y.shape[0], X), st.last_error
    np.testing.assert_array_equal(st.getKnnIds(), o["ids"])


def test_random_reset_statistics():
    """reset(N,D,L,M,W) draws A ~ N(0,1) and b = (U - min U) W
    (cuLSH_HashTables.cu:122-187): one b per table is exactly 0, all in [0, W)."""
    from culsh_amd import HashTables, SearchTables
    print("random reset", flush=True)
    X = wl.mixture(2000, 16, 10, 0.2, 61)
    ht = HashTables()
    assert ht.reset(2000, 16, 5, 6, 2.5, seed=1234)
    A = np.stack([ht.getA(t) for t in range(5)])
    bb = np.stack([ht.getb(t) for t in range(5)])
    assert abs(A.mean()) < 0.1 and abs(A.std() - 1) < 0.1
    assert (bb.min(axis=1) == 0).all() and (bb < 2.5).all() and (bb >= 0).all()
    assert ht.index(X)
    st = SearchTables()
    assert st.reset(ht, 4, 4)
    assert st.search(X[:50], 50, X)
    # A data point used as a query finds itself at distance 0.
    assert (st.getKnnIds()[:, 0] == np.arange(50)).all()


def test_precision_search_after_reindex_uses_the_new_dataset():
    """One SearchTables in the grouped precision form across re-indexing: index(X1),
    then index(X2) of the same shape (its device copy usually lands at the
    same address), then index_device on one tensor whose contents change in
    place.  Every search must use the current dataset's row norms."""
    import torch
    from culsh_amd import HashTables, SearchTables
    print("precision reindex", flush=True)
    N, D, L, M, W, K, T = 6000, 24, 4, 6, 2.0, 10, 8
    A, b = wl.projection(L, D, M, W, 601)
    X1 = wl.mixture(N, D, 20, 0.2, 602)
    X2 = wl.mixture(N, D, 20, 0.2, 603) * 1.5
    Y = wl.queries_near(X2, 80, 0.05, 604)
    ht = HashTables()
    assert ht.reset(N, D, L, M, W, A, b)
    assert ht.set_form(1)
    st = SearchTables()
    assert st.reset(ht, K, T)
    assert st.set_form(2)
    for X in (X1, X2, X1):
        assert ht.index(X), ht.last_error
        assert st.search(Y, Y.shape[0], X), st.last_error
        o = oracle.Index(X, L, M, W, A, b).search(Y, X, K, T)
        np.testing.assert_array_equal(st.getKnnIds(), o["ids"])
        np.testing.assert_array_equal(st.getKnnDistances(), o["dists"])
    dX = torch.from_numpy(X1.copy()).cuda()
    ids = torch.empty((Y.shape[0], K), dtype=torch.int32, device="cuda")
    d = torch.empty((Y.shape[0], K), dtype=torch.float32, device="cuda")
    dY = torch.from_numpy(Y).cuda()
    for X in (X1, X2):
        dX.copy_(torch.from_numpy(X))
        assert ht.index_device(dX), ht.last_error
        assert st.search_device(dY, ids, d), st.last_error
        assert st.wait(), st.last_error
        o = oracle.Index(X, L, M, W, A, b).search(Y, X, K, T)
        np.testing.assert_array_equal(ids.cpu().numpy(), o["ids"])
        np.testing.assert_array_equal(d.cpu().numpy(), o["dists"])


def test_search_dataset_upload_follows_the_host_pointer():
    """After index_device, search(queries, Q, dataset) uploads the host
    dataset once and reuses it while the same array is passed; another array
    is uploaded again (distances follow the dataset passed, as the
    reference's search does)."""
    import torch
    from culsh_amd import HashTables, SearchTables
    print("search dataset upload", flush=True)
    X, Y, A, b = _small(701, N=4000, Q=60)
    ix = oracle.Index(X, 4, 6, 1.0, A, b)
    dX = torch.from_numpy(X).cuda()
    ht = HashTables()
    assert ht.reset(X.shape[0], X.shape[1], 4, 6, 1.0, A, b)
    assert ht.index_device(dX), ht.last_error
    st = SearchTables()
    assert st.reset(ht, 5, 8)
    X2 = (X * 0.5).astype(np.float32)
    for data in (X, X, X2, X):
        assert st.search(Y, Y.shape[0], data), st.last_error
        o = ix.search(Y, data, 5, 8)
        np.testing.assert_array_equal(st.getKnnIds(), o["ids"])
        np.testing.assert_array_equal(st.getKnnDistances(), o["dists"])


def test_search_in_passes_equals_one_pass():
    """Queries split into passes (here forced to 7 queries per pass) give the
    one-pass results: ids, distances, candidate counts and query buckets,
    through the host and the device APIs; timesStruct reports every phase."""
    import torch
    from culsh_amd import SearchTables
    print("search in passes", flush=True)
    X, Y, A, b = _small(801, N=6000, Q=101)
    o = run_oracle(X, Y, A, b, 4, 6, 1.0, 7, 12)
    g = run_gpu(X, Y, A, b, 4, 6, 1.0, 7, 12)
    st = SearchTables()
    assert st.reset(g["ht"], 7, 12)
    assert st.set_pass_limit(7)
    assert st.search(Y, Y.shape[0], X), st.last_error
    cnu, cu = st.getQueryCandidates()
    gp = dict(ids=st.getKnnIds(), dists=st.getKnnDistances(), cnu=cnu, cu=cu,
              query_buckets=np.stack([st.getQueryBuckets(t) for t in range(4)]))
    assert_search_equal(gp, o)
    t = st.getTimes()
    print("  times", t, flush=True)
    assert t["buckets"] > 0 and t["collect"] > 0 and t["distances"] > 0 and t["knn"] > 0
    assert t["unique"] == 0 and t["sort"] >= 0
    dY = torch.from_numpy(Y).cuda()
    ids = torch.empty((Y.shape[0], 7), dtype=torch.int32, device="cuda")
    d = torch.empty((Y.shape[0], 7), dtype=torch.float32, device="cuda")
    assert st.search_device(dY, ids, d), st.last_error
    assert st.wait(), st.last_error
    np.testing.assert_array_equal(ids.cpu().numpy(), o["ids"])
    np.testing.assert_array_equal(d.cpu().numpy(), o["dists"])
