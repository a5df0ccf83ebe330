"""Batching form: one graph launch of the persistent fused kernel over many
requests with their own Q, K and T.  Bar: every request's ids and distances
are bit-identical to a SearchTables search of that request alone (itself
bit-exact with the oracle), under every schedule the autotuner can pick."""
import numpy as np
import pytest

from culsh_amd import Batch, HashTables, SearchTables, layout_requests
from culsh_amd import workload as wl
from lshtest import run_oracle

pytestmark = pytest.mark.gpu


def _index(X, A, b, L, M, W):
    ht = HashTables()
    assert ht.reset(X.shape[0], X.shape[1], L, M, W, A, b), ht.last_error
    assert ht.index(X), ht.last_error
    return ht


def _mixed_shapes(rng, n, qmax, Ks=(1, 10, 100), Ts=(1, 8, 64)):
    Q = np.floor(np.exp(rng.uniform(0, np.log(qmax + 1), n))).astype(int)
    K = rng.choice(Ks, n)
    T = rng.choice(Ts, n)
    return [(int(q), int(k), int(t)) for q, k, t in zip(Q, K, T)]


def _run_batch(bt, reqs, dY, total, sched=None):
    import torch
    if sched is not None:
        assert bt.set_schedule(**sched), bt.last_error
    ids = torch.full((max(total, 1),), -7, dtype=torch.int32, device="cuda")
    d = torch.full((max(total, 1),), -7.0, dtype=torch.float32, device="cuda")
    assert bt.search_device(reqs, dY, ids, d), bt.last_error
    return ids.cpu().numpy(), d.cpu().numpy()


def _alone(ht, Y, reqs):
    """Each request searched alone through SearchTables (device API)."""
    import torch
    out = []
    st = SearchTables()
    for Q, K, T, row, _ in reqs:
        if Q == 0:
            out.append((np.empty((0, K), np.int32), np.empty((0, K), np.float32)))
            continue
        assert st.reset(ht, K, T)
        dq = torch.from_numpy(np.ascontiguousarray(Y[row:row + Q])).cuda()
        i = torch.empty((Q, K), dtype=torch.int32, device="cuda")
        d = torch.empty((Q, K), dtype=torch.float32, device="cuda")
        assert st.search_device(dq, i, d), st.last_error
        torch.cuda.synchronize()
        out.append((i.cpu().numpy(), d.cpu().numpy()))
    return out


def _check(reqs, ids, dists, ref):
    for r, ((Q, K, T, row, off), (ri, rd)) in enumerate(zip(reqs, ref)):
        gi = ids[off:off + Q * K].reshape(Q, K)
        gd = dists[off:off + Q * K].reshape(Q, K)
        bad = np.nonzero((gi != ri).any(axis=1))[0]
        assert bad.size == 0, (f"request {r} (Q={Q} K={K} T={T}): {bad.size} queries differ, "
                               f"first {bad[0]}: batch {gi[bad[0]]} alone {ri[bad[0]]}")
        np.testing.assert_array_equal(gd, rd, err_msg=f"request {r} distances")


def _setup(seed=201, N=20000, D=24, L=6, M=8, W=1.5, n=40, qmax=300):
    import torch
    rng = np.random.default_rng(seed)
    X = wl.mixture(N, D, 30, 0.2, seed)
    shapes = _mixed_shapes(rng, n, qmax)
    shapes[3] = (0, 10, 8)  # an empty request
    reqs, rows, total = layout_requests(shapes)
    Y = wl.queries_near(X, rows, 0.05, seed + 1)
    A, b = wl.projection(L, D, M, W, seed + 2)
    ht = _index(X, A, b, L, M, W)
    return X, Y, A, b, ht, reqs, total, torch.from_numpy(Y).cuda()


def test_batch_matches_alone_and_oracle():
    print("batch vs alone vs oracle", flush=True)
    X, Y, A, b, ht, reqs, total, dY = _setup()
    bt = Batch(ht)
    ids, dists = _run_batch(bt, reqs, dY, total)
    _check(reqs, ids, dists, _alone(ht, Y, reqs))
    # Oracle on a few requests directly.
    for r in (0, 1, 2, 5):
        Q, K, T, row, off = reqs[r]
        if Q == 0:
            continue
        o = run_oracle(X, Y[row:row + Q], A, b, 6, 8, 1.5, K, T, want_buckets=False)
        np.testing.assert_array_equal(ids[off:off + Q * K].reshape(Q, K), o["ids"])
        np.testing.assert_array_equal(dists[off:off + Q * K].reshape(Q, K), o["dists"])


SCHEDULES = [
    dict(probe_queries=32, probe_tables=1, queries_per_item=32, rows_in_flight=4, wgs_per_cu=3,
         fused=0),
    dict(probe_queries=16, probe_tables=2, queries_per_item=32, rows_in_flight=8, wgs_per_cu=0,
         fused=0),
    dict(probe_queries=16, probe_tables=2, queries_per_item=16, rows_in_flight=2, wgs_per_cu=1,
         fused=0),
    dict(probe_queries=8, probe_tables=4, queries_per_item=64, rows_in_flight=8, wgs_per_cu=2,
         fused=0),
    dict(probe_queries=32, probe_tables=1, queries_per_item=32, rows_in_flight=4, wgs_per_cu=2,
         fused=1),
    dict(probe_queries=16, probe_tables=2, queries_per_item=16, rows_in_flight=2, wgs_per_cu=1,
         fused=1),
    dict(probe_queries=8, probe_tables=4, queries_per_item=64, rows_in_flight=8, wgs_per_cu=3,
         fused=1),
    dict(probe_queries=32, probe_tables=1, queries_per_item=64, rows_in_flight=8, wgs_per_cu=3,
         fused=1),
    dict(probe_queries=16, probe_tables=2, queries_per_item=48, rows_in_flight=4, wgs_per_cu=2,
         fused=1),
]


def test_batch_schedules_give_identical_results():
    print("batch schedules", flush=True)
    X, Y, A, b, ht, reqs, total, dY = _setup(seed=211, n=30)
    bt = Batch(ht)
    ref = None
    for s in SCHEDULES:
        print("  schedule", s, flush=True)
        ids, dists = _run_batch(bt, reqs, dY, total, s)
        if ref is None:
            ref = (ids, dists)
            _check(reqs, ids, dists, _alone(ht, Y, reqs))
        else:
            np.testing.assert_array_equal(ids, ref[0])
            np.testing.assert_array_equal(dists, ref[1])


def test_batch_graph_reuse_and_growth():
    """A small batch, a larger one (buffers grow, the graph is re-captured),
    then the small one again through the re-captured graph."""
    import torch
    print("batch graph reuse", flush=True)
    X, Y, A, b, ht, reqs, total, dY = _setup(seed=221, n=12, qmax=50)
    bt = Batch(ht)
    ids1, d1 = _run_batch(bt, reqs, dY, total)
    rng = np.random.default_rng(5)
    shapes2 = _mixed_shapes(rng, 60, 600)
    reqs2, rows2, total2 = layout_requests(shapes2)
    Y2 = wl.queries_near(X, rows2, 0.05, 222)
    dY2 = torch.from_numpy(Y2).cuda()
    ids2, d2 = _run_batch(bt, reqs2, dY2, total2)
    _check(reqs2, ids2, d2, _alone(ht, Y2, reqs2))
    ids3, d3 = _run_batch(bt, reqs, dY, total)
    np.testing.assert_array_equal(ids1, ids3)
    np.testing.assert_array_equal(d1, d3)
    _check(reqs, ids3, d3, _alone(ht, Y, reqs))


def test_batch_overflow_path():
    """Wide buckets: queries with tens of thousands of candidates go through
    the batch's global-memory kernel, and K=1500 needs more FP64 rechecks
    than LDS holds."""
    import torch
    print("batch overflow", flush=True)
    X = wl.mixture(30000, 16, 3, 0.3, 81)
    A, b = wl.projection(3, 16, 3, 8.0, 83)
    ht = _index(X, A, b, 3, 3, 8.0)
    shapes = [(20, 10, 4), (5, 1500, 2), (7, 1, 1), (30, 100, 8)]
    reqs, rows, total = layout_requests(shapes)
    Y = wl.queries_near(X, rows, 0.05, 82)
    bt = Batch(ht)
    ids, dists = _run_batch(bt, reqs, torch.from_numpy(Y).cuda(), total)
    _check(reqs, ids, dists, _alone(ht, Y, reqs))


def test_batch_offsets_and_rows_are_callers():
    """Requests may point anywhere in the caller's query and result arrays."""
    import torch
    print("batch offsets", flush=True)
    X, Y, A, b, ht, _, _, dY = _setup(seed=231, n=4, qmax=20)
    rows = Y.shape[0]
    reqs = [(5, 3, 8, rows - 5, 100), (4, 2, 1, 0, 0), (5, 3, 8, 2, 40)]
    total = 200
    bt = Batch(ht)
    ids, dists = _run_batch(bt, reqs, dY, total)
    _check(reqs, ids, dists, _alone(ht, Y, reqs))
    untouched = np.ones(total, bool)
    for Q, K, _, _, off in reqs:
        untouched[off:off + Q * K] = False
    assert (ids[untouched] == -7).all()


def test_batch_rejects_large_T():
    import torch
    print("batch limits", flush=True)
    X, Y, A, b, ht, _, _, dY = _setup(seed=241, n=4, qmax=20)
    bt = Batch(ht)
    ids = torch.empty(100, dtype=torch.int32, device="cuda")
    d = torch.empty(100, dtype=torch.float32, device="cuda")
    assert not bt.search_device([(3, 2, 65, 0, 0)], dY, ids, d)
    assert "exceeds 64" in bt.last_error
    assert not bt.search_device([(3, 0, 8, 0, 0)], dY, ids, d)
    assert not bt.set_schedule(probe_queries=4, probe_tables=8)
    assert not bt.set_schedule(probe_queries=16, probe_tables=2, queries_per_item=24)
    assert not bt.set_schedule(wgs_per_cu=0, fused=1)


def test_batch_autotune_keeps_results():
    print("batch autotune", flush=True)
    X, Y, A, b, ht, reqs, total, dY = _setup(seed=251, n=20, qmax=200)
    import torch
    bt = Batch(ht)
    ids = torch.empty(total, dtype=torch.int32, device="cuda")
    d = torch.empty(total, dtype=torch.float32, device="cuda")
    r = bt.autotune(reqs, dY, ids, d, reps=1)
    assert r is not None, bt.last_error
    print("  default %.3f ms tuned %.3f ms" % r, bt.get_schedule(), flush=True)
    assert r[1] <= r[0]
    ids2, d2 = _run_batch(bt, reqs, dY, total)
    _check(reqs, ids2, d2, _alone(ht, Y, reqs))


def test_batch_after_reindex_searches_the_new_index():
    """A plan whose graph was captured on one index build searches the next
    build of the same HashTables (another dataset of the same shape): the
    graph is captured again, with the new index's buffers and FP16 rows."""
    import torch
    print("batch after re-index", flush=True)
    X, Y, A, b, ht, reqs, total, dY = _setup(seed=261, n=10, qmax=120)
    bt = Batch(ht)
    ids, dists = _run_batch(bt, reqs, dY, total)
    _check(reqs, ids, dists, _alone(ht, Y, reqs))
    X2 = wl.mixture(X.shape[0], X.shape[1], 30, 0.2, 262)
    Y2 = wl.queries_near(X2, Y.shape[0], 0.05, 263)
    assert ht.index(X2), ht.last_error
    ids2, d2 = _run_batch(bt, reqs, torch.from_numpy(Y2).cuda(), total)
    _check(reqs, ids2, d2, _alone(ht, Y2, reqs))


@pytest.mark.parametrize("case", ["d_not_multiple_of_8", "beyond_half_range", "coarse_half"])
def test_batch_distance_rows(case):
    """The batch's distance pass reads an FP16 copy of the rows when D is a
    multiple of 8 and every value fits a half; results stay identical to the
    FP32 search: with D = 20 (FP32 rows), with values beyond the half range
    (FP32 rows), and with values near 1000, whose half spacing (0.5) is
    large next to the distances (wide bands of FP64 re-evaluation)."""
    import torch
    print("batch distance rows", case, flush=True)
    rng = np.random.default_rng(271)
    D = 20 if case == "d_not_multiple_of_8" else 24
    X = wl.mixture(8000, D, 20, 0.2, 272)
    if case == "beyond_half_range":
        X = (X * 1.0e5).astype(np.float32)
    elif case == "coarse_half":
        X = (X + 1000.0).astype(np.float32)
    W = {"d_not_multiple_of_8": 1.5, "beyond_half_range": 1.5e5, "coarse_half": 1.5}[case]
    L, M = 5, 8
    shapes = _mixed_shapes(rng, 12, 150)
    reqs, rows, total = layout_requests(shapes)
    Y = wl.queries_near(X, rows, 0.05 * (1.0e5 if case == "beyond_half_range" else 1.0), 273)
    A, b = wl.projection(L, D, M, W, 274)
    ht = _index(X, A, b, L, M, W)
    bt = Batch(ht)
    ids, dists = _run_batch(bt, reqs, torch.from_numpy(Y).cuda(), total)
    _check(reqs, ids, dists, _alone(ht, Y, reqs))
    Q, K, T, row, off = next(r for r in reqs if r[0] > 0)
    o = run_oracle(X, Y[row:row + Q], A, b, L, M, W, K, T, want_buckets=False)
    np.testing.assert_array_equal(ids[off:off + Q * K].reshape(Q, K), o["ids"])
    np.testing.assert_array_equal(dists[off:off + Q * K].reshape(Q, K), o["dists"])
