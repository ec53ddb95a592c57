"""Beam search over pruned indexes on the GPU against the oracle's restatement of
get_neighbors (common.cu:65-127): identical ids, distances and evaluation counts for
several pool sizes and index kinds, exact top-k on a complete graph, the recall judge for
external queries against numpy, bitmap reset between calls (including the overflow path
that clears the whole bitmap), and argument checks."""
import numpy as np
import pytest

import oracle
from tagore_amd import State, TagoreError
from tests.conftest import gmm_int, gmm_unit

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pruned(X, K, kind, fr, rr, th, FD=32, TOPM=64):
    G, GD = oracle.knn_bruteforce(X, K)
    s = State(*X.shape)
    s.set_data(X)
    s.set_knn(G, GD)
    s.prune(kind, fr, rr, FD, TOPM, th)
    deg, ids, ep = s.get_pruned()
    return s, deg, ids, ep


@pytest.mark.parametrize("name,kind,fr,rr,th", [
    ("NSG", 0, 0, 0, 1.0), ("NSSG", 1, 1, 1, 0.5), ("CAGRA", 2, 0, 2, 0.0),
])
def test_search_parity(name, kind, fr, rr, th):
    Xall = gmm_int(3300, 128, 24, 55)
    X, Q = Xall[:3000], Xall[3000:]
    s, deg, ids, ep = _pruned(X, 48, kind, fr, rr, th)
    with s:
        for L in (10, 64, 100, 200):
            _say(f"search parity {name} L={L} nq={Q.shape[0]}")
            gi, gd, st = s.search(Q, k=10, L=L, seed=0x123)
            oi, od, oc = oracle.search(X, deg, ids, ep, Q, L, 10, 0x123)
            assert np.array_equal(gi, oi)
            assert np.array_equal(gd, od)
            assert st["evals"] == oc["evals"] and st["expansions"] == oc["expansions"]


def test_search_parity_unit_d96_and_entry_point():
    Xall = gmm_unit(2600, 96, 16, 77)
    X, Q = Xall[:2500], Xall[2500:]
    s, deg, ids, ep = _pruned(X, 40, 0, 0, 0, 1.2, FD=24)
    with s:
        for e in (ep, 17):
            gi, gd, st = s.search(Q, k=8, L=48, seed=5, ep=e)
            oi, od, oc = oracle.search(X, deg, ids, e, Q, 48, 8, 5)
            assert np.array_equal(gi, oi) and np.array_equal(gd, od)
            assert st["evals"] == oc["evals"]


def test_search_complete_graph_exact():
    rng = np.random.default_rng(4)
    N, D = 120, 32
    X = rng.integers(0, 30, (N, D)).astype(np.float32)
    Q = rng.integers(0, 30, (40, D)).astype(np.float32)
    # CAGRA with FD = K = N-1 keeps every neighbour: a complete graph, so the first
    # expansion evaluates every vector and the result is the exact top-k
    s, deg, ids, ep = _pruned(X, N - 1, 2, 0, 2, 0.0, FD=N - 1)
    with s:
        gi, gd, _ = s.search(Q, k=10, L=64, seed=9)
        oi, od, _ = oracle.search(X, deg, ids, ep, Q, 64, 10, 9)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od)
        for q in range(Q.shape[0]):
            d = np.array([oracle.dist(Q[q], x) for x in X], np.float32)
            want = sorted(range(N), key=lambda i: (d[i], i))[:10]
            assert gi[q].tolist() == want


def test_search_recall_judge_matches_numpy():
    Xall = gmm_int(5200, 64, 32, 91)
    X, Q = Xall[:5000], Xall[5000:]
    s, deg, ids, ep = _pruned(X, 32, 0, 0, 0, 1.0, FD=24)
    with s:
        gi, _, _ = s.search(Q, k=10, L=40, seed=3)
        r = s.search_recall(Q, gi, 10)
        X64, Q64 = X.astype(np.float64), Q.astype(np.float64)
        hits = 0
        for q in range(Q.shape[0]):
            d = ((X64 - Q64[q]) ** 2).sum(1)
            kth = np.sort(d)[9]
            hits += int((d[gi[q].astype(np.int64)] <= kth).sum())
        assert abs(r - hits / (10 * Q.shape[0])) < 1e-12
        assert r > 0.5


def test_search_bitmap_reset_and_overflow():
    """A query that visits more ids than the per-query list holds clears the whole bitmap;
    repeated searches must give identical results."""
    rng = np.random.default_rng(8)
    N, D = 30000, 16
    X = rng.integers(0, 64, (N, D)).astype(np.float32)
    Q = rng.integers(0, 64, (8, D)).astype(np.float32)
    s = State(N, D)
    with s:
        s.set_data(X)
        s.knn_build(128, 4, 0x77, "bf16")
        s.prune(2, 0, 2, 128, 64, 0.0)
        deg, ids, ep = s.get_pruned()
        a = s.search(Q, k=16, L=256, seed=1)
        b = s.search(Q, k=16, L=256, seed=1)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        # about 8k visited ids per query at this size: several queries overflow the 8192-id
        # list and clear their whole bitmap
        assert a[2]["evals"] > 7000 * Q.shape[0]
        oi, od, oc = oracle.search(X, deg, ids, ep, Q, 256, 16, 1)
        assert np.array_equal(a[0], oi) and np.array_equal(a[1], od)
        assert a[2]["evals"] == oc["evals"]
        _say(f"evals per query {a[2]['evals'] / Q.shape[0]:.0f}")


def test_search_rejects_bad_arguments():
    X = gmm_int(500, 32, 4, 1)
    s, deg, ids, ep = _pruned(X, 16, 0, 0, 0, 1.0, FD=16)
    with s:
        Q = X[:4]
        with pytest.raises(TagoreError):
            s.search(Q, k=10, L=300)
        with pytest.raises(TagoreError):
            s.search(Q, k=20, L=10)
        with pytest.raises(TagoreError):
            s.search(Q, k=5, L=10, ep=500)
        with pytest.raises(TagoreError):
            s.search_recall(Q, np.zeros((4, 20), np.uint32), 20)


@pytest.mark.parametrize("index_type,kind,fr,rr,th", [
    ("NSG", 0, 0, 0, 1.0), ("Vamana", 0, 0, 0, 1.2), ("NSSG", 1, 1, 1, 0.5),
    ("CAGRA", 2, 0, 2, 0.0), ("DPG", 3, 0, 2, 0.0),
])
def test_stored_index_loads_back_and_searches_alike(tmp_path, index_type, kind, fr, rr, th):
    """tg_store_index -> tg_read_index -> tg_set_pruned gives the same graph, and the search
    over the loaded index equals the search over the pruned one (same entry point)."""
    Xall = gmm_int(2200, 64, 16, 31)
    X, Q = Xall[:2000], Xall[2000:]
    path = str(tmp_path / f"idx.{index_type}")
    s, deg, ids, ep = _pruned(X, 32, kind, fr, rr, th, FD=24)
    with s:
        s.store_index(path, index_type)
        a = s.search(Q, k=10, L=48, seed=2, ep=ep)
    with State(*X.shape) as t:
        t.set_data(X)
        eps = t.load_index(path, index_type)
        d2, i2, ep2 = t.get_pruned()
        assert np.array_equal(d2, deg)
        for u in range(X.shape[0]):
            assert np.array_equal(i2[u, :deg[u]], ids[u, :deg[u]])
        if index_type in ("NSG", "Vamana"):
            assert eps.tolist() == [ep] and ep2 == ep
        b = t.search(Q, k=10, L=48, seed=2, ep=ep)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("name,th", [("NSG", 1.0), ("Vamana", 1.2)])
def test_index_search_recall_regression(name, th):
    """The pruned indexes are searchable: on 20k x 32 unit-norm Gaussian-mixture vectors
    with an exact kNN graph (K=48), NSG / Vamana (FD=32, m=64) searched with L=64 from the
    index's entry point reach recall@10 >= 0.95 on 1000 held-out queries (0.98-0.99 when
    this test was written), judged by the FP64 brute force.  Parity with the oracle cannot
    catch a rule that is wrong on both sides; this can."""
    Xall = gmm_unit(21000, 32, 24, 909)
    X, Q = Xall[:20000], Xall[20000:]
    G, GD = oracle.knn_bruteforce(X, 48)
    with State(*X.shape) as s:
        s.set_data(X)
        s.set_knn(G, GD)
        s.prune("path", "dist", "dist", 32, 64, th)
        ids, _, _ = s.search(Q, k=10, L=64, seed=0x5EA)
        r = s.search_recall(Q, ids, 10)
    _say(f"{name}: search recall@10 {r:.4f}")
    assert r >= 0.95
