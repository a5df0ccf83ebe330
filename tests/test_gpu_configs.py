"""The workloads at their full sizes on the GPU, against the FP64 oracle.

  c3  N=4M, D=96, L=32, M=20, W=2: all 50,000 queries searched on the
      device; 300 of them (spread over the set) compared with the oracle's
      search of its own full index -- ids, distances, candidate counts and
      query buckets (lexicographic numbering)
  c4  256 requests (Q log-uniform in [1, 4096], K {1, 10, 100}, T {1, 8, 64})
      on the c3 index as one batch: every request equals its plain search
      alone, and the first queries of 48 requests equal the oracle's
  c5  both serving indexes (2M x 128 L=24 M=20 W=3; 1M x 960 L=16 M=28 W=3),
      a Poisson stream served with continuous batching: every request equals
      the one-at-a-time run, and requests of both indexes equal the oracle's

The oracle indexes are built once per module (tens of seconds on the host).
"""
import numpy as np
import pytest

import oracle
from culsh_amd import workload as wl
from lshtest import assert_search_equal

pytestmark = pytest.mark.gpu


class _Ns:
    pass


@pytest.fixture(scope="module")
def c3():
    import torch
    from culsh_amd import HashTables
    cfg = wl.CONFIGS["c3"]
    print(cfg.describe(), flush=True)
    n = _Ns()
    n.cfg = cfg
    n.X, n.Y, n.A, n.b = wl.make_data(cfg)
    n.dX = torch.from_numpy(n.X).cuda()
    n.ht = HashTables()
    assert n.ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, n.A, n.b), n.ht.last_error
    assert n.ht.index_device(n.dX), n.ht.last_error
    print("  oracle index of the full c3 dataset", flush=True)
    n.ix = oracle.Index(n.X, cfg.L, cfg.M, cfg.W, n.A, n.b)
    return n


def test_config3_full_index(c3):
    import torch
    from culsh_amd import SearchTables
    cfg = c3.cfg
    Q, K, T = cfg.Q, cfg.K, cfg.T
    st = SearchTables()
    assert st.reset(c3.ht, K, T)
    dY = torch.from_numpy(c3.Y).cuda()
    ids = torch.empty((Q, K), dtype=torch.int32, device="cuda")
    d = torch.empty((Q, K), dtype=torch.float32, device="cuda")
    assert st.search_device(dY, ids, d), st.last_error
    assert st.wait(), st.last_error
    sel = np.arange(0, Q, Q // 300)[:300]
    o = c3.ix.search(c3.Y[sel], c3.X, K, T, want_buckets=True)
    gi, gd = ids.cpu().numpy()[sel], d.cpu().numpy()[sel]
    bad = np.nonzero((gi != o["ids"]).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} of 300 queries differ from the oracle"
    np.testing.assert_array_equal(gd, o["dists"])
    # The drop-in host search of the same 300 queries: candidate counts and
    # query buckets in the reference's lexicographic numbering too.
    Ys = np.ascontiguousarray(c3.Y[sel])
    assert st.search(Ys, Ys.shape[0], c3.X), st.last_error
    cnu, cu = st.getQueryCandidates()
    g = dict(ids=st.getKnnIds(), dists=st.getKnnDistances(), cnu=cnu, cu=cu,
             query_buckets=np.stack([st.getQueryBuckets(t) for t in range(cfg.L)]))
    assert_search_equal(g, o)
    print(f"  c3 full: 300 queries equal the oracle; mean C_u {cu.mean():.0f}", flush=True)


def test_config4_batch_on_the_c3_index(c3):
    import torch
    from culsh_amd import Batch, SearchTables, layout_requests
    shapes = wl.batch_requests()
    reqs, rows, total = layout_requests(shapes)
    Y = wl.queries_near(c3.X, rows, c3.cfg.query_noise, 4002)
    dY = torch.from_numpy(Y).cuda()
    ids = torch.empty(total, dtype=torch.int32, device="cuda")
    d = torch.empty(total, dtype=torch.float32, device="cuda")
    bt = Batch(c3.ht)
    assert bt.search_device(reqs, dY, ids, d), bt.last_error
    hi, hd = ids.cpu().numpy(), d.cpu().numpy()
    st = SearchTables()
    for Q, K, T, row, off in reqs:
        assert st.reset(c3.ht, K, T)
        i = torch.empty((Q, K), dtype=torch.int32, device="cuda")
        dd = torch.empty((Q, K), dtype=torch.float32, device="cuda")
        assert st.search_device(dY[row:row + Q], i, dd), st.last_error
        assert st.wait(), st.last_error
        np.testing.assert_array_equal(hi[off:off + Q * K].reshape(Q, K), i.cpu().numpy())
        np.testing.assert_array_equal(hd[off:off + Q * K].reshape(Q, K), dd.cpu().numpy())
    # Oracle on the first queries of 48 requests spread over the batch.
    nsel = 0
    for r in range(0, len(reqs), len(reqs) // 48):
        Q, K, T, row, off = reqs[r]
        q = min(Q, 4)
        o = c3.ix.search(Y[row:row + q], c3.X, K, T)
        np.testing.assert_array_equal(hi[off:off + q * K].reshape(q, K), o["ids"])
        np.testing.assert_array_equal(hd[off:off + q * K].reshape(q, K), o["dists"])
        nsel += 1
    print(f"  c4: {len(reqs)} requests ({rows} queries) equal alone; {nsel} requests "
          f"checked against the oracle", flush=True)


@pytest.fixture(scope="module")
def c5():
    import torch
    from culsh_amd import HashTables
    n = _Ns()
    n.hts, n.X, n.dX, n.ix, n.cfg = [], [], [], [], []
    for name in ("c5a", "c5b"):
        cfg = wl.CONFIGS[name]
        print(cfg.describe(), flush=True)
        X, _, A, b = wl.make_data(cfg, Q=1)
        dX = torch.from_numpy(X).cuda()
        ht = HashTables()
        assert ht.reset(cfg.N, cfg.D, cfg.L, cfg.M, cfg.W, A, b), ht.last_error
        assert ht.index_device(dX), ht.last_error
        print(f"  oracle index of {name}", flush=True)
        n.ix.append(oracle.Index(X, cfg.L, cfg.M, cfg.W, A, b))
        n.hts.append(ht)
        n.X.append(X)
        n.dX.append(dX)
        n.cfg.append(cfg)
    return n


def _serve_layout(c5, rate, n, seed):
    import torch
    stream = wl.serve_stream(rate, n, seed=seed)
    rows, outs, reqs = [0, 0], [0, 0], []
    for a, ix, Q, K, T in stream:
        reqs.append((a, ix, Q, K, T, rows[ix], outs[ix]))
        rows[ix] += Q
        outs[ix] += Q * K
    Ys = [wl.queries_near(c5.X[i], rows[i], c5.cfg[i].query_noise, seed + 1 + i)
          for i in range(2)]
    dY = [torch.from_numpy(Y).cuda() for Y in Ys]
    ids = [torch.empty(o, dtype=torch.int32, device="cuda") for o in outs]
    ds = [torch.empty(o, dtype=torch.float32, device="cuda") for o in outs]
    return reqs, Ys, dY, ids, ds, outs


def test_config5_served_results_equal_alone_and_the_oracle(c5):
    from culsh_amd import Server
    reqs, Ys, dY, ids, ds, outs = _serve_layout(c5, 20000, 600, 5104)
    res = {}
    for batching in (True, False):
        srv = Server(c5.hts, batching=batching, slots_per_index=2, hbm_fraction=0.25)
        r = srv.run(reqs, dY, ids, ds)
        assert r is not None, srv.last_error
        lat, st = r
        res[batching] = ([t.cpu().numpy() for t in ids], [t.cpu().numpy() for t in ds])
        if batching:
            assert st["batches"] < len(reqs)  # requests shared batches
            print(f"  served p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms, "
                  f"{st['mean_batch_requests']:.2f} requests per batch", flush=True)
        del srv
    for i in range(2):
        np.testing.assert_array_equal(res[True][0][i], res[False][0][i])
        np.testing.assert_array_equal(res[True][1][i], res[False][1][i])
    # The oracle on the first queries of 40 requests of each index.
    for i in range(2):
        sel = [r for r in reqs if r[1] == i][:40]
        for a, ix, Q, K, T, row, off in sel:
            q = min(Q, 3)
            o = c5.ix[i].search(Ys[i][row:row + q], c5.X[i], K, T)
            np.testing.assert_array_equal(res[True][0][i][off:off + q * K].reshape(q, K),
                                          o["ids"])
            np.testing.assert_array_equal(res[True][1][i][off:off + q * K].reshape(q, K),
                                          o["dists"])


def test_serving_with_host_io_equals_device_resident(c5):
    """Host IO runs (query upload and result download per batch, pinned host
    memory) return the device-resident run's results in host memory."""
    import torch
    from culsh_amd import Server
    reqs, Ys, dY, ids, ds, outs = _serve_layout(c5, 10000, 300, 5204)
    hY = [torch.from_numpy(Y).pin_memory() for Y in Ys]
    hI = [torch.full((o,), -7, dtype=torch.int32).pin_memory() for o in outs]
    hD = [torch.empty(o, dtype=torch.float32).pin_memory() for o in outs]
    srv = Server(c5.hts, batching=True, slots_per_index=2, hbm_fraction=0.25)
    r = srv.run(reqs, dY, ids, ds)
    assert r is not None, srv.last_error
    dev = [t.cpu().numpy() for t in ids], [t.cpu().numpy() for t in ds]
    for t in ids:
        t.fill_(-3)
    for t in dY:
        t.zero_()  # the host run must upload the queries itself
    r = srv.run(reqs, dY, ids, ds, host=(hY, hI, hD))
    assert r is not None, srv.last_error
    assert r[1]["host_io"] == 1
    for i in range(2):
        np.testing.assert_array_equal(hI[i].numpy(), dev[0][i])
        np.testing.assert_array_equal(hD[i].numpy(), dev[1][i])
    # Plain (one request at a time) with host IO as well.
    srv2 = Server(c5.hts, batching=False)
    for t in hI:
        t.fill_(-7)
    r = srv2.run(reqs, dY, ids, ds, host=(hY, hI, hD))
    assert r is not None, srv2.last_error
    for i in range(2):
        np.testing.assert_array_equal(hI[i].numpy(), dev[0][i])


def test_config3_search_beyond_31_bit_probe_slots(c3):
    """1,100,000 queries at L = 32, T = 64 address more than 2^31 probe slots:
    the device search runs them in two passes (cuLSH_SearchTables.cu:283-311
    batches by memory the same way); the first and last 500 queries equal
    their search alone."""
    import torch
    from culsh_amd import SearchTables
    cfg = c3.cfg
    K, T = 10, cfg.T
    Q = 1_100_000
    reps = (Q + cfg.Q - 1) // cfg.Q
    dY = torch.from_numpy(c3.Y).cuda().repeat(reps, 1)[:Q].contiguous()
    st = SearchTables()
    assert st.reset(c3.ht, K, T)
    ids = torch.empty((Q, K), dtype=torch.int32, device="cuda")
    d = torch.empty((Q, K), dtype=torch.float32, device="cuda")
    assert st.search_device(dY, ids, d), st.last_error
    assert st.wait(), st.last_error
    for sl in (slice(0, 500), slice(Q - 500, Q)):
        i2 = torch.empty((500, K), dtype=torch.int32, device="cuda")
        d2 = torch.empty((500, K), dtype=torch.float32, device="cuda")
        assert st.search_device(dY[sl].contiguous(), i2, d2), st.last_error
        assert st.wait(), st.last_error
        np.testing.assert_array_equal(ids[sl].cpu().numpy(), i2.cpu().numpy())
        np.testing.assert_array_equal(d[sl].cpu().numpy(), d2.cpu().numpy())
