"""Pins of the FP64 oracle (oracle/oracle.cpp) against closed forms, brute
force and hand-computed known answers.  The reference ships no tests or
golden vectors, so each oracle function is checked here by a method that
does not share its code path:

  projection+floor  exact rational arithmetic (fractions)        Indexing.cu:27-66
  bucket build      group-by + lexicographic sort in numpy       Indexing.cu:71-144
  multiprobe        enumeration of all subsets sorted by score,  Kernels.cu:188-289
                    plus a hand-traced known answer of the queue
                    including its re-pop of the last position
  lookup/candidates dict-based brute force                       Querying.cu:124-165,
                                                                 SearchTables.cu:249-364
  distance          exact integer / rational sums                Kernels.cu:301-348
  top-K             textbook K-smallest on distinct distances;   Kernels.cu:356-456
                    a closed-form tie rule derived independently
                    of the heap, on tie-heavy inputs
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

FLT_MAX = float(np.finfo(np.float32).max)


# ---------------------------------------------------------------- projection
def test_projection_hand_known_answer():
    # D=2, M=1: x=(1,2), a=(3,4), b=1, W=2 -> (3 + 8 - 1) / 2 = 5
    X = np.array([[1, 2]], np.float32)
    A = np.array([[3, 4]], np.float32)       # column-major [D x M] -> (M, D)
    v = oracle.project(X, A, np.array([1.0], np.float32), 2.0)
    assert v[0, 0] == 5.0
    # negative side: x=(-1, 0.5), a=(2, -1), b=0.25, W=0.5 -> (-2 - 0.5 - 0.25)/0.5 = -5.5
    X = np.array([[-1, 0.5]], np.float32)
    A = np.array([[2, -1]], np.float32)
    v = oracle.project(X, A, np.array([0.25], np.float32), 0.5)
    assert v[0, 0] == -5.5
    assert math.floor(v[0, 0]) == -6


@pytest.mark.parametrize("W", [1.0, 2.0, 3.0, 0.75, 400.0])
def test_projection_matches_exact_rationals(W):
    rng = np.random.default_rng(int(W * 4))
    N, D, M = 300, 9, 5
    X = rng.integers(-7, 8, size=(N, D)).astype(np.float32)
    A = (rng.integers(-16, 17, size=(M, D)) / 8).astype(np.float32)
    b = (rng.integers(0, int(W * 8), size=M) / 8).astype(np.float32)
    v = oracle.project(X, A, b, W)
    Wf = Fraction(float(np.float32(W)))
    for n in range(0, N, 7):
        for m in range(M):
            s = sum(Fraction(float(X[n, d])) * Fraction(float(A[m, d])) for d in range(D))
            exact = (s - Fraction(float(b[m]))) / Wf
            assert math.floor(exact) == math.floor(v[n, m]), (n, m)
            assert abs(Fraction(v[n, m]) - exact) <= abs(exact) * Fraction(1, 2 ** 52) + Fraction(
                1, 2 ** 1000)


def test_floor_invariant_on_random_data():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(500, 16)).astype(np.float32)
    A = rng.normal(size=(6, 16)).astype(np.float32)
    b = rng.uniform(0, 2, size=6).astype(np.float32)
    v = oracle.project(X, A, b, 2.0)
    # a dropped term or a wrong sign moves v well outside this band
    ref = (X.astype(np.float64) @ A.T.astype(np.float64) - b) / 2.0
    assert np.max(np.abs(v - ref)) < 1e-12


# -------------------------------------------------------------- bucket build
def _codes(X, A_t, b_t, W):
    return np.floor(oracle.project(X, A_t, b_t, W)).astype(np.int64)


def test_buckets_match_groupby():
    rng = np.random.default_rng(7)
    N, D, L, M, W = 1500, 6, 3, 4, 1.5
    X = rng.normal(size=(N, D)).astype(np.float32)
    A = rng.normal(size=(L, M, D)).astype(np.float32)
    b = rng.uniform(0, W, size=(L, M)).astype(np.float32)
    ix = oracle.Index(X, L, M, W, A, b)
    for t in range(L):
        codes = _codes(X, A[t], b[t], W)
        groups = {}
        for n in range(N):
            groups.setdefault(tuple(codes[n]), []).append(n)
        keys = sorted(groups)                       # lexicographic, dim 0 most significant
        total, start, size, bcodes = ix.table(t)
        assert ix.buckets(t) == len(keys)
        assert size.sum() == N
        np.testing.assert_array_equal(start, np.concatenate([[0], np.cumsum(size)[:-1]]))
        for k, key in enumerate(keys):
            np.testing.assert_array_equal(bcodes[k], np.array(key, np.float32))
            members = total[start[k]:start[k] + size[k]]
            assert list(members) == groups[key]      # ascending ids inside a bucket
        assert sorted(total.tolist()) == list(range(N))


# ---------------------------------------------------------------- multiprobe
def test_probe_sequence_T1_is_base_code():
    v = np.array([0.3, -1.2, 4.99])
    codes, masks, _ = oracle.probe_sequence(v, 1)
    np.testing.assert_array_equal(codes[0], [0, -2, 4])


def test_probe_sequence_known_answer_with_repop():
    """M=1, v=0.3: deltas dn=0.3 (index 0, '-1'), up=0.7 (index 1, '+1').
    Hand trace of the queue (cuLSH_Kernels.cu:216-255): [0], {0}->-1, then
    {1}->+1 whose largest position is maxT-1=1, so it is never replaced and
    is popped again on every later step."""
    codes, masks, perm = oracle.probe_sequence(np.array([0.3]), 6)
    assert list(perm) == [0, 1]
    assert codes[:, 0].tolist() == [0, -1, 1, 1, 1, 1]
    assert masks.tolist() == [0, 1, 2, 2, 2, 2]


def _bruteforce_sequence(v, T):
    M = len(v)
    base = [math.floor(x) for x in v]
    d = [x - math.floor(x) for x in v] + [1 - (x - math.floor(x)) for x in v]
    maxT = min(2 * M, 64)
    perm = sorted(range(2 * M), key=lambda i: d[i])[:maxT]
    subsets = []
    for r in range(1, maxT + 1):
        for S in itertools.combinations(range(maxT), r):
            subsets.append((math.fsum(d[perm[i]] for i in S), S))
    subsets.sort()
    # Compared prefix: up to the first set holding position maxT-1, after
    # which the reference re-emits that set.
    end = next((k for k, (_, S) in enumerate(subsets) if S[-1] == maxT - 1), len(subsets))
    end = min(end + 1, T - 1)
    seq = [tuple(base)]
    sets = [()]
    for score, S in subsets[:end]:
        c = list(base)
        for i in S:
            c[perm[i] % M] += 1 if perm[i] >= M else -1
        seq.append(tuple(c))
        sets.append(S)
    gaps = np.diff([s for s, _ in subsets[:end + 1]]) if end > 0 else [1.0]
    return seq, sets, perm, (min(gaps) if len(gaps) else 1.0)


@pytest.mark.parametrize("M", [2, 3, 4, 5, 6])
def test_probe_sequence_equals_score_ordered_subsets(M):
    rng = np.random.default_rng(M)
    checked = 0
    for trial in range(300):
        v = rng.uniform(-3, 3, size=M)
        maxT = 2 * M
        T = min(2 ** maxT, 40)
        seq, sets, perm, gap = _bruteforce_sequence(v, T)
        if gap < 1e-9:
            continue      # near-equal scores: order decided by rounding
        codes, masks, operm = oracle.probe_sequence(v, T)
        assert list(operm) == perm
        for t in range(len(seq)):
            S = sets[t]
            assert tuple(codes[t]) == seq[t], (trial, t)
            want_mask = sum(1 << i for i in S)
            assert int(masks[t]) == want_mask
            if S and S[-1] == maxT - 1:
                break     # after this pop the reference re-emits it (A.9)
        checked += 1
    assert checked > 200


# -------------------------------------------------- lookup, candidates, search
def _bruteforce_search(X, Y, A, b, W, K, T):
    N = X.shape[0]
    L = A.shape[0]
    tables = []
    for t in range(L):
        codes = _codes(X, A[t], b[t], W)
        groups = {}
        for n in range(N):
            groups.setdefault(tuple(codes[n]), []).append(n)
        tables.append(groups)
    out = []
    for q in range(Y.shape[0]):
        cand, cnu = set(), 0
        for t in range(L):
            v = oracle.project(Y[q:q + 1], A[t], b[t], W)[0]
            pcodes, _, _ = oracle.probe_sequence(v, T)
            for p in range(T):
                members = tables[t].get(tuple(pcodes[p]), [])
                cnu += len(members)
                cand.update(members)
        out.append((cnu, sorted(cand)))
    return out


def test_search_candidates_match_bruteforce():
    rng = np.random.default_rng(11)
    N, D, L, M, W, Q, K, T = 800, 5, 3, 3, 1.0, 40, 4, 6
    X = rng.normal(size=(N, D)).astype(np.float32)
    Y = (X[rng.integers(0, N, Q)] + rng.normal(scale=0.1, size=(Q, D))).astype(np.float32)
    A = rng.normal(size=(L, M, D)).astype(np.float32)
    b = rng.uniform(0, W, size=(L, M)).astype(np.float32)
    ix = oracle.Index(X, L, M, W, A, b)
    r = ix.search(Y, X, K, T, want_buckets=True)
    bf = _bruteforce_search(X, Y, A, b, W, K, T)
    for q in range(Q):
        cnu, cand = bf[q]
        assert int(r["cnu"][q]) == cnu
        assert int(r["cu"][q]) == len(cand)
        d = np.array([oracle.dist(Y[q], X[i]) for i in cand])
        order = np.lexsort((np.array(cand), d))
        want = [cand[i] for i in order[:K]]
        if len(set(d.tolist())) == len(d):       # distinct distances: K smallest
            assert r["ids"][q, :len(want)].tolist() == want
    # query buckets: index of the matching bucket (lexicographic numbering)
    for t in range(L):
        total, start, size, codes = ix.table(t)
        lut = {tuple(codes[k].astype(int)): k for k in range(ix.buckets(t))}
        for q in range(0, Q, 5):
            v = oracle.project(Y[q:q + 1], A[t], b[t], W)[0]
            pcodes, _, _ = oracle.probe_sequence(v, T)
            for p in range(T):
                assert r["query_buckets"][t, q, p] == lut.get(tuple(pcodes[p]), -1)


# ------------------------------------------------------------------ distance
def test_distance_exact_on_integers():
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.integers(0, 256, size=128).astype(np.float32)
        x = rng.integers(0, 256, size=128).astype(np.float32)
        exact = sum((int(a) - int(c)) ** 2 for a, c in zip(q, x))
        assert oracle.dist(q, x) == exact


def test_distance_close_to_rational():
    rng = np.random.default_rng(6)
    q = rng.normal(size=96).astype(np.float32)
    x = rng.normal(size=96).astype(np.float32)
    exact = sum((Fraction(float(a)) - Fraction(float(c))) ** 2 for a, c in zip(q, x))
    assert abs(Fraction(oracle.dist(q, x)) - exact) <= exact * Fraction(96, 2 ** 52)


# --------------------------------------------------------------------- top-K
def topk_closed_form(dist, ids, K):
    """Top-K the reference's heap returns (derived independently of the heap):
    all candidates below the K-th distance d_K, then the r smallest ids among
    the d_K-ties whose id is >= id*, id* being the K-th largest id among the
    candidates at or below d_K (the heap scans ids in descending order and
    admits strictly smaller distances only)."""
    pairs = sorted(zip(dist, ids))
    pairs = [p for p in pairs if p[0] < FLT_MAX]
    if len(pairs) <= K:
        return pairs + [(FLT_MAX, -1)] * (K - len(pairs))
    dK = pairs[K - 1][0]
    lt = [p for p in pairs if p[0] < dK]
    r = K - len(lt)
    le_desc = sorted((p[1] for p in pairs if p[0] <= dK), reverse=True)
    idstar = le_desc[K - 1]
    ties = sorted(p[1] for p in pairs if p[0] == dK and p[1] >= idstar)
    return lt + [(dK, i) for i in ties[:r]]


def test_topk_distinct_is_k_smallest():
    rng = np.random.default_rng(8)
    for _ in range(200):
        C = int(rng.integers(1, 60))
        ids = np.sort(rng.choice(100000, size=C, replace=False)).astype(np.uint32)
        d = rng.uniform(0, 10, size=C)
        K = int(rng.integers(1, 20))
        oi, od = oracle.topk_heap(d, ids, K)
        order = np.argsort(d)[:K]
        assert oi[:len(order)].tolist() == ids[order].astype(int).tolist()
        assert (oi[len(order):] == -1).all() and (od[len(order):] == FLT_MAX).all()


def test_topk_ties_follow_closed_form():
    rng = np.random.default_rng(9)
    for trial in range(3000):
        C = int(rng.integers(1, 40))
        ids = np.sort(rng.choice(500, size=C, replace=False)).astype(np.uint32)
        d = rng.integers(0, 4, size=C).astype(np.float64)
        K = int(rng.integers(1, 12))
        oi, od = oracle.topk_heap(d, ids, K)
        want = topk_closed_form(d.tolist(), ids.astype(int).tolist(), K)
        assert list(zip(od.tolist(), oi.tolist())) == want, trial


def test_topk_single_slot_takes_largest_tied_id():
    oi, od = oracle.topk_heap(np.array([1.0, 1.0, 1.0]), np.array([3, 7, 9], np.uint32), 1)
    assert oi.tolist() == [9]


# ------------------------------------------------------- recall ground truth
def test_bruteforce_knn_matches_numpy_argsort():
    """orc_bruteforce_knn (the exact kNN that recall@10 is measured against)
    equals a numpy lexsort on (exact distance, id), including exact ties
    (small integer data) and K beyond N (padding with -1)."""
    rng = np.random.default_rng(314)
    for N, D, Q, K, ints in [(500, 7, 30, 10, False), (400, 3, 40, 25, True), (12, 5, 4, 20, False)]:
        if ints:
            X = rng.integers(0, 3, size=(N, D)).astype(np.float32)
            Y = rng.integers(0, 3, size=(Q, D)).astype(np.float32)
        else:
            X = rng.standard_normal((N, D)).astype(np.float32)
            Y = rng.standard_normal((Q, D)).astype(np.float32)
        ids, d = oracle.bruteforce_knn(X, Y, K)
        for q in range(Q):
            # Exact rational distances (float inputs are dyadic rationals).
            exact = [sum((Fraction(float(Y[q, j])) - Fraction(float(X[n, j]))) ** 2
                         for j in range(D)) for n in range(N)]
            order = sorted(range(N), key=lambda n: (exact[n], n))[:K]
            k = min(K, N)
            np.testing.assert_array_equal(ids[q, :k], order[:k])
            ref = np.array([float(exact[n]) for n in order[:k]])
            # The oracle sums with one FMA rounding per term: relative error
            # at most D * 2^-53.
            np.testing.assert_allclose(d[q, :k], ref, rtol=D * 2.0 ** -53, atol=0)
            if K > N:
                assert (ids[q, N:] == -1).all()


def test_recall_counts_ties_at_the_kth_distance():
    """recall_at_k counts a returned id when its exact distance is <= the
    exact k-th distance, so an id tied with the k-th neighbour but not the one
    brute force lists still counts; an id beyond it and -1 do not."""
    X = np.array([[0.0], [1.0], [-1.0], [2.0], [5.0]], np.float32)
    Y = np.array([[0.0]], np.float32)
    # Exact: 0 (id 0), 1 (ids 1, 2 tied), 4 (id 3), 25 (id 4).
    assert oracle.recall_at_k(X, Y, np.array([[0, 2]]), k=2) == 1.0
    assert oracle.recall_at_k(X, Y, np.array([[0, 3]]), k=2) == 0.5
    assert oracle.recall_at_k(X, Y, np.array([[-1, 1]]), k=2) == 0.5
