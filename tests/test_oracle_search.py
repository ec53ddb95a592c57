"""Pins of the oracle's beam search (restatement of get_neighbors, common.cu:65-127) by what
the mathematics fixes: on a complete graph the first expansion evaluates every vector, so
the result is the exact (d, id) top-k; a greedy walk along a chain ends at the nearest
vertex; the output is sorted, unique and carries canonical distances; a pool that cannot
fill (N < L) is padded."""
import numpy as np

import oracle

INV = 0xFFFFFFFF


def _topk(X, Q, k):
    out = []
    for q in Q:
        d = np.array([oracle.dist(q, x) for x in X], np.float32)
        order = sorted(range(len(X)), key=lambda i: (d[i], i))[:k]
        out.append((order, d[order]))
    return out


def test_complete_graph_gives_exact_topk():
    rng = np.random.default_rng(1)
    N, D = 100, 16
    X = rng.integers(0, 20, (N, D)).astype(np.float32)
    ids = np.array([[j for j in range(N) if j != i] for i in range(N)], np.uint32)
    deg = np.full(N, N - 1, np.uint32)
    Q = rng.integers(0, 20, (12, D)).astype(np.float32)
    for L in (10, 40):
        oi, od, c = oracle.search(X, deg, ids, 3, Q, L, 10, 99)
        for (want_ids, want_d), gi, gd in zip(_topk(X, Q, 10), oi, od):
            assert gi.tolist() == want_ids
            assert np.array_equal(gd, want_d)
        assert c["evals"] == 12 * N  # every vector evaluated exactly once per query


def test_greedy_walk_on_a_chain_reaches_the_nearest_vertex():
    # 1-D points 0..49 on a chain; L = 1 and the entry point's single neighbour start the
    # pool, so no random draws are taken and the walk is the greedy descent
    N = 50
    X = np.arange(N, dtype=np.float32).reshape(N, 1).repeat(4, axis=1)
    ids = np.full((N, 2), INV, np.uint32)
    deg = np.zeros(N, np.uint32)
    for i in range(N):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < N]
        ids[i, :len(nb)] = nb
        deg[i] = len(nb)
    for t in (7.3, 31.8, 48.6, 0.2):
        Q = np.full((1, 4), t, np.float32)
        oi, od, _ = oracle.search(X, deg, ids, 0, Q, 1, 1, 5)
        assert oi[0, 0] == int(round(t))
        assert od[0, 0] == oracle.dist(Q[0], X[int(round(t))])


def test_results_sorted_unique_canonical():
    rng = np.random.default_rng(2)
    N, D = 600, 24
    X = rng.standard_normal((N, D)).astype(np.float32)
    G, GD = oracle.knn_bruteforce(X, 20)
    deg = np.full(N, 20, np.uint32)
    Q = rng.standard_normal((30, D)).astype(np.float32)
    oi, od, c = oracle.search(X, deg, G, 0, Q, 32, 16, 7)
    for q in range(30):
        keys = list(zip(od[q].tolist(), oi[q].tolist()))
        assert keys == sorted(keys)
        assert len(set(oi[q].tolist())) == 16
        for i, d in zip(oi[q], od[q]):
            assert d == oracle.dist(Q[q], X[i])
    assert c["expansions"] >= 30


def test_pool_smaller_than_L_is_padded():
    rng = np.random.default_rng(3)
    N, D = 20, 8
    X = rng.standard_normal((N, D)).astype(np.float32)
    G, _ = oracle.knn_bruteforce(X, 4)
    deg = np.full(N, 4, np.uint32)
    Q = rng.standard_normal((3, D)).astype(np.float32)
    oi, od, _ = oracle.search(X, deg, G, 0, Q, 32, 32, 1)
    for q in range(3):
        assert sorted(oi[q, :N].tolist()) == list(range(N))  # every vector drawn or reached
        assert (oi[q, N:] == INV).all() and np.isinf(od[q, N:]).all()
