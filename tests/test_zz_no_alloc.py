"""The allocation-free device search (culsh_st_search_device_noalloc) against
the allocating one (culsh_st_search_device), through the C ABI: bit-identical
ids and distances on two calls in a row, the second with a smaller query
set; a workload larger than the fixed scratch is refused before any launch;
a query beyond the fixed candidate caps sets the caller's status word."""
import numpy as np
import pytest

from culsh_amd import HashTables, SearchTables
from culsh_amd import workload as wl

pytestmark = pytest.mark.gpu


def _index(N, D, L, M, W, seed, centers=30, sigma=0.2):
    X = wl.mixture(N, D, centers, sigma, seed)
    A, b = wl.projection(L, D, M, W, seed + 1)
    ht = HashTables()
    assert ht.reset(N, D, L, M, W, A, b), ht.last_error
    assert ht.index(X), ht.last_error
    return X, ht


def _alloc_search(ht, Y, K, T):
    import torch
    st = SearchTables()
    assert st.reset(ht, K, T)
    dY = torch.from_numpy(Y).cuda()
    ids = torch.full((Y.shape[0], K), -7, dtype=torch.int32, device="cuda")
    d = torch.full((Y.shape[0], K), -7.0, dtype=torch.float32, device="cuda")
    assert st.search_device(dY, ids, d), st.last_error
    assert st.wait(), st.last_error
    return ids.cpu().numpy(), d.cpu().numpy()


@pytest.mark.parametrize("T", [1, 16, 64])
def test_noalloc_matches_the_allocating_search(T):
    import torch
    print(f"noalloc T={T}", flush=True)
    X, ht = _index(30000, 32, 8, 10, 2.0, 1701)
    st = SearchTables()
    K = 20
    assert st.reset(ht, K, T)
    status = torch.full((1,), -5, dtype=torch.int32, device="cuda")
    for Q, seed in ((1500, 1702), (300, 1703)):  # second call smaller
        Y = wl.queries_near(X, Q, 0.05, seed)
        ref_i, ref_d = _alloc_search(ht, Y, K, T)
        dY = torch.from_numpy(Y).cuda()
        ids = torch.full((Q, K), -7, dtype=torch.int32, device="cuda")
        d = torch.full((Q, K), -7.0, dtype=torch.float32, device="cuda")
        assert st.search_device_noalloc(dY, ids, d, status), st.last_error
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        np.testing.assert_array_equal(ids.cpu().numpy(), ref_i)
        np.testing.assert_array_equal(d.cpu().numpy(), ref_d)


def test_noalloc_wide_buckets_take_its_overflow_kernel():
    """Wide buckets: queries with thousands of unique candidates run the
    fixed-size overflow kernel on the scratch array; results still equal the
    allocating search on two calls in a row."""
    import torch
    print("noalloc overflow kernel", flush=True)
    X = wl.mixture(30000, 16, 3, 0.3, 81)
    A, b = wl.projection(3, 16, 3, 8.0, 83)
    ht = HashTables()
    assert ht.reset(30000, 16, 3, 3, 8.0, A, b)
    assert ht.index(X)
    st = SearchTables()
    assert st.reset(ht, 10, 4)
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    for Q in (40, 17):
        Y = wl.queries_near(X, Q, 0.05, 82)
        ref_i, ref_d = _alloc_search(ht, Y, 10, 4)
        dY = torch.from_numpy(Y).cuda()
        ids = torch.empty((Q, 10), dtype=torch.int32, device="cuda")
        d = torch.empty((Q, 10), dtype=torch.float32, device="cuda")
        assert st.search_device_noalloc(dY, ids, d, status), st.last_error
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        np.testing.assert_array_equal(ids.cpu().numpy(), ref_i)
        np.testing.assert_array_equal(d.cpu().numpy(), ref_d)


def test_noalloc_limits():
    import torch
    print("noalloc limits", flush=True)
    X, ht = _index(20000, 32, 16, 10, 2.0, 1711)
    st = SearchTables()
    assert st.reset(ht, 10, 64)
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    # 16 tables x 64 probes x 4 B = 4 KiB per query: 20,000 queries do not fit.
    Q = 20000
    dY = torch.zeros((Q, 32), dtype=torch.float32, device="cuda")
    ids = torch.empty((Q, 10), dtype=torch.int32, device="cuda")
    d = torch.empty((Q, 10), dtype=torch.float32, device="cuda")
    assert not st.search_device_noalloc(dY, ids, d, status)
    assert "fixed scratch" in st.last_error
    assert st.set_form(2)
    assert not st.search_device_noalloc(dY[:10], ids[:10], d[:10], status)
    assert "grouped" in st.last_error
    # Every point in one bucket per table: 4 x 300,000 candidates > 2^17.
    rng = np.random.default_rng(7)
    Xb = rng.standard_normal((300000, 4)).astype(np.float32)
    hb = HashTables()
    assert hb.reset(300000, 4, 4, 1, 1e6, np.zeros((4, 1, 4), np.float32),
                    np.zeros((4, 1), np.float32))
    assert hb.index(Xb)
    sb = SearchTables()
    assert sb.reset(hb, 5, 1)
    qy = torch.from_numpy(Xb[:2].copy()).cuda()
    i2 = torch.empty((2, 5), dtype=torch.int32, device="cuda")
    d2 = torch.empty((2, 5), dtype=torch.float32, device="cuda")
    assert sb.search_device_noalloc(qy, i2, d2, status), sb.last_error
    torch.cuda.synchronize()
    assert int(status.item()) != 0
    # The status is cleared by the next call that succeeds.
    X2, h2 = _index(5000, 32, 4, 8, 2.0, 1721)
    s2 = SearchTables()
    assert s2.reset(h2, 5, 8)
    Y2 = wl.queries_near(X2, 50, 0.05, 1722)
    ref_i, _ = _alloc_search(h2, Y2, 5, 8)
    i3 = torch.empty((50, 5), dtype=torch.int32, device="cuda")
    d3 = torch.empty((50, 5), dtype=torch.float32, device="cuda")
    assert s2.search_device_noalloc(torch.from_numpy(Y2).cuda(), i3, d3, status), s2.last_error
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    np.testing.assert_array_equal(i3.cpu().numpy(), ref_i)
