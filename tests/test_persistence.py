"""Index persistence (HashTables::save / load / reset(filename),
cuLSH_HashTables.cu:243-249,395-483) and the CLI driver (cuLSH_main.cu).

The file format is the reference's: a writer restating its layout from the
oracle's buckets must produce the same bytes as save(), and load() of such a
file must search exactly like the oracle."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

from culsh_amd import HashTables, SearchTables, _lib
from culsh_amd import workload as wl
from lshtest import write_reference_format

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_load_missing_file_is_io_error(tmp_path):
    ht = HashTables()
    assert not ht.load(str(tmp_path / "absent.lsh"))
    assert "Could not open file" in ht.last_error


def test_load_truncated_header_is_io_error(tmp_path):
    p = tmp_path / "trunc.lsh"
    p.write_bytes(struct.pack("<ii", 10, 4))
    ht = HashTables()
    assert not ht.reset(str(p))
    assert "truncated" in ht.last_error


def test_save_without_index_is_state_error(tmp_path):
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.culsh_ht_create(ctypes.byref(h)) == 0
    try:
        assert L.culsh_ht_save(h, str(tmp_path / "x.lsh").encode(), None) == 2
    finally:
        L.culsh_ht_destroy(h)


@pytest.mark.gpu
def test_save_is_byte_identical_to_reference_layout_and_load_round_trips(tmp_path):
    import oracle
    print("save/load round trip", flush=True)
    N, D, L, M, W, Q, K, T = 3000, 16, 4, 6, 1.0, 60, 7, 10
    X = wl.mixture(N, D, 20, 0.2, 201)
    Y = wl.queries_near(X, Q, 0.05, 202)
    A, b = wl.projection(L, D, M, W, 203)
    ix = oracle.Index(X, L, M, W, A, b)
    ref_path = tmp_path / "ref.lsh"
    write_reference_format(ref_path, ix, A, b, N, D, L, M, W)

    ht = HashTables()
    assert ht.reset(N, D, L, M, W, A, b) and ht.index(X)
    ours = tmp_path / "ours.lsh"
    assert ht.save(str(ours)), ht.last_error
    assert ours.read_bytes() == ref_path.read_bytes()

    want = ix.search(Y, X, K, T, want_buckets=True)
    ht2 = HashTables()
    assert ht2.reset(str(ref_path)), ht2.last_error
    assert (ht2.getN(), ht2.getD(), ht2.getL(), ht2.getM()) == (N, D, L, M)
    for t in range(L):
        np.testing.assert_array_equal(ht2.getBuckets_codes(t), ix.table(t)[3])
    st = SearchTables()
    assert st.reset(ht2, K, T)
    assert st.search(Y, Q, X), st.last_error
    np.testing.assert_array_equal(st.getKnnIds(), want["ids"])
    np.testing.assert_array_equal(st.getKnnDistances(), want["dists"])
    for t in range(L):
        np.testing.assert_array_equal(st.getQueryBuckets(t), want["query_buckets"][t])
    again = tmp_path / "again.lsh"
    assert ht2.save(str(again))
    assert again.read_bytes() == ref_path.read_bytes()


@pytest.mark.gpu
def test_load_rejects_inconsistent_buckets(tmp_path):
    import oracle
    N, D, L, M, W = 500, 8, 2, 4, 1.0
    X = wl.mixture(N, D, 5, 0.2, 211)
    A, b = wl.projection(L, D, M, W, 212)
    ix = oracle.Index(X, L, M, W, A, b)
    p = tmp_path / "bad.lsh"
    write_reference_format(p, ix, A, b, N, D, L, M, W)
    raw = bytearray(p.read_bytes())
    # corrupt table 0's first member id (offset after header, projections, B[L])
    off = 20 + L * (D * M + M) * 4 + L * 4
    raw[off:off + 4] = struct.pack("<I", N + 5)
    p.write_bytes(bytes(raw))
    ht = HashTables()
    assert not ht.load(str(p))
    assert "out of range" in ht.last_error


@pytest.mark.gpu
def test_cli_runs_reference_arguments(tmp_path):
    print("cli", flush=True)
    X = wl.mixture(4000, 32, 10, 0.2, 221)
    f = tmp_path / "data.f32"
    f.write_bytes(X.tobytes())
    exe = os.path.join(ROOT, "culsh_amd", "lib", "culsh_cli")
    out = subprocess.run([exe, str(f), "32", "4000", "6", "8", "2.0", "100", "5", "8"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "seconds to search" in out.stdout
    # the first query is data vector 0: its nearest neighbour is itself
    assert "query 0 neighbours: 0(0)" in out.stdout
