"""Serving form: real-time request loop with continuous batching over two
device-resident indexes.  Bar: every request's ids and distances in the mix
are bit-identical to a SearchTables search of that request alone."""
import numpy as np
import pytest

from culsh_amd import HashTables, SearchTables, Server
from culsh_amd import workload as wl

pytestmark = pytest.mark.gpu


def _index(X, L, M, W, seed):
    A, b = wl.projection(L, X.shape[1], M, W, seed)
    ht = HashTables()
    assert ht.reset(X.shape[0], X.shape[1], L, M, W, A, b), ht.last_error
    assert ht.index(X), ht.last_error
    return ht


def _setup(n=300, rate=5000.0, qmax=64, seed=301):
    import torch
    Xa = wl.mixture(20000, 32, 40, 0.2, seed)
    Xb = wl.mixture(5000, 400, 20, 0.05, seed + 1)
    hts = [_index(Xa, 6, 8, 1.5, seed + 2), _index(Xb, 4, 10, 2.0, seed + 3)]
    stream = wl.serve_stream(rate, n, qmax=qmax, seed=seed + 4)
    rows, outs, reqs = [0, 0], [0, 0], []
    for a, ix, Q, K, T in stream:
        reqs.append((a, ix, Q, K, T, rows[ix], outs[ix]))
        rows[ix] += Q
        outs[ix] += Q * K
    Ys = [wl.queries_near(X, max(r, 1), 0.05, seed + 5 + i)
          for i, (X, r) in enumerate(zip((Xa, Xb), rows))]
    dY = [torch.from_numpy(Y).cuda() for Y in Ys]
    return hts, reqs, Ys, dY, outs


def _buffers(outs):
    import torch
    ids = [torch.full((max(o, 1),), -7, dtype=torch.int32, device="cuda") for o in outs]
    ds = [torch.full((max(o, 1),), -7.0, dtype=torch.float32, device="cuda") for o in outs]
    return ids, ds


def _check_alone(hts, reqs, Ys, ids, ds):
    import torch
    st = SearchTables()
    hi = [t.cpu().numpy() for t in ids]
    hd = [t.cpu().numpy() for t in ds]
    for r, (_, ix, Q, K, T, row, off) in enumerate(reqs):
        assert st.reset(hts[ix], K, T)
        dq = torch.from_numpy(np.ascontiguousarray(Ys[ix][row:row + Q])).cuda()
        i = torch.empty((Q, K), dtype=torch.int32, device="cuda")
        d = torch.empty((Q, K), dtype=torch.float32, device="cuda")
        assert st.search_device(dq, i, d), st.last_error
        torch.cuda.synchronize()
        np.testing.assert_array_equal(hi[ix][off:off + Q * K].reshape(Q, K), i.cpu().numpy(),
                                      err_msg=f"request {r} ids")
        np.testing.assert_array_equal(hd[ix][off:off + Q * K].reshape(Q, K), d.cpu().numpy(),
                                      err_msg=f"request {r} distances")


@pytest.mark.parametrize("batching", [True, False])
def test_serve_mix_equals_alone(batching):
    print("serve batching" if batching else "serve plain", flush=True)
    hts, reqs, Ys, dY, outs = _setup()
    srv = Server(hts, batching=batching, slots_per_index=2, hbm_fraction=0.1)
    ids, ds = _buffers(outs)
    r = srv.run(reqs, dY, ids, ds)
    assert r is not None, srv.last_error
    lat, stats = r
    print("  ", {k: (round(v, 3) if isinstance(v, float) else v) for k, v in stats.items()},
          flush=True)
    assert stats["requests"] == len(reqs) and (lat > 0).all()
    assert stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"] + 1e-9
    if batching:
        assert 1 <= stats["batches"] <= len(reqs)
    _check_alone(hts, reqs, Ys, ids, ds)


def test_serve_admission_by_slot_memory():
    """Slots sized for 256 queries (8192 probe slots per table): requests
    that do not fit the forming batch wait for the next free slot; results
    are unchanged."""
    print("serve admission", flush=True)
    # Arrivals ~1 ns apart: the queue holds ~4000 queries when the first slot
    # frees, so batches are capacity-bound regardless of timing.
    hts, reqs, Ys, dY, outs = _setup(n=200, rate=1e9, qmax=100, seed=311)
    srv = Server(hts, batching=True, slots_per_index=1, hbm_fraction=0.1, max_batch_queries=256)
    ids, ds = _buffers(outs)
    r = srv.run(reqs, dY, ids, ds)
    assert r is not None, srv.last_error
    lat, stats = r
    print("  ", stats, flush=True)
    assert stats["deferred"] > 0 and stats["mean_batch_requests"] > 1
    _check_alone(hts, reqs, Ys, ids, ds)


def test_serve_runs_requests_larger_than_a_slot():
    """Slots of 4 queries: a request of up to 100 queries either grows its
    slot (the free HBM allows it) or, with growth off, runs in pieces over
    several batches; results equal the request alone either way."""
    print("serve oversized requests", flush=True)
    hts, reqs, Ys, dY, outs = _setup(n=40, rate=20000.0, qmax=100, seed=321)
    assert max(r[2] for r in reqs) > 4
    for grow in (False, True):
        srv = Server(hts, batching=True, slots_per_index=1, hbm_fraction=0.1,
                     max_batch_queries=4, grow=grow)
        ids, ds = _buffers(outs)
        r = srv.run(reqs, dY, ids, ds)
        assert r is not None, srv.last_error
        lat, stats = r
        print("  grow" if grow else "  pieces", stats, flush=True)
        if grow:
            assert stats["grown_slots"] > 0
        else:
            assert stats["split_requests"] > 0 and stats["grown_slots"] == 0
        assert (lat > 0).all()
        _check_alone(hts, reqs, Ys, ids, ds)
    srv = Server(hts, batching=True, slots_per_index=1, hbm_fraction=0.1, max_batch_queries=4)
    ids, ds = _buffers(outs)
    assert srv.run([(0.0, 2, 1, 1, 1, 0, 0)], dY, ids, ds) is None  # no index 2
    assert "out of range" in srv.last_error


def test_serve_run_after_a_failed_run():
    """A run that fails while other slots hold batches (a request with more
    than 2^20 candidates: every point in one bucket) leaves no slot busy; the
    next, smaller run on the same server returns every result and latency."""
    import torch
    print("serve run after failure", flush=True)
    hts, reqs, Ys, dY, outs = _setup(n=120, rate=1e9, qmax=40, seed=331)
    rng = np.random.default_rng(7)
    Xb = rng.standard_normal((300000, 4)).astype(np.float32)
    bad = HashTables()
    assert bad.reset(300000, 4, 4, 1, 1e6, np.zeros((4, 1, 4), np.float32),
                     np.zeros((4, 1), np.float32))
    assert bad.index(Xb)
    srv = Server([hts[0], bad], batching=True, slots_per_index=3, hbm_fraction=0.1)
    yb = torch.from_numpy(Xb[:8].copy()).cuda()
    ib = torch.empty(64, dtype=torch.int32, device="cuda")
    db = torch.empty(64, dtype=torch.float32, device="cuda")
    ids, ds = _buffers(outs)
    a_reqs = [r for r in reqs if r[1] == 0]
    failing = a_reqs[:60] + [(a_reqs[59][0], 1, 2, 5, 1, 0, 0)] + a_reqs[60:]
    assert srv.run(failing, [dY[0], yb], [ids[0], ib], [ds[0], db]) is None
    assert "candidates" in srv.last_error
    small = a_reqs[:10]
    ids, ds = _buffers(outs)
    r = srv.run(small, [dY[0], yb], [ids[0], ib], [ds[0], db])
    assert r is not None, srv.last_error
    lat, stats = r
    assert stats["requests"] == 10 and lat.shape == (10,) and (lat > 0).all()
    _check_alone([hts[0], bad], small, Ys, ids, ds)
