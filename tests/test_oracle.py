"""Pins for the CPU oracle (oracle/tagore_oracle.cpp).

Each test checks the oracle against something it does not share code with: a closed form,
numpy FP64 arithmetic, a textbook special case of the rule, or an independent brute force on
tiny inputs.  A plausible slip in the restatement (a dropped term, a wrong comparison
direction, a transposed index, a sequential instead of tree reduction) fails at least one.
"""
import struct

import numpy as np
import pytest

import oracle
from tests.conftest import gmm_int

INV = oracle.INV


# --------------------------------------------------------------------- canonical distance
@pytest.mark.parametrize("D", [96, 128, 960])
def test_dist_matches_fp64(D):
    rng = np.random.default_rng(D)
    for scale in (1.0, 255.0):
        a = (rng.random(D) * scale).astype(np.float32)
        b = (rng.random(D) * scale).astype(np.float32)
        d = float(oracle.dist(a, b))
        ref = float(np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2))
        assert abs(d - ref) <= 1e-5 * ref
        assert oracle.dist(b, a) == oracle.dist(a, b)  # bitwise symmetric


def test_dist_integer_closed_form():
    a = np.arange(1, 129, dtype=np.float32)
    assert float(oracle.dist(a, np.zeros(128, np.float32))) == 128 * 129 * 257 / 6


def test_dist_is_a_tree_not_a_running_sum():
    # t0^2 = 2^24 and 63 ones: a running FP32 sum stays at 2^24 (each +1 rounds away),
    # the exact sum is 2^24+63, the 64-partial tree gives 2^24+62.
    a = np.ones(64, np.float32)
    a[0] = 4096.0
    assert float(oracle.dist(a, np.zeros(64, np.float32))) == 2.0 ** 24 + 62


def test_dist_sums_each_lane_run_first():
    # p = [2^24, 1, 0, 1, 0...]: (p0 + p1) rounds to 2^24 (tie to even) and (p2 + p3) = 1,
    # so the in-run order gives 2^24 + 1 -> 2^24; pairing p0 with p2 and p1 with p3 (a tree
    # over s = 32..1) would give 2^24 + 2, and an exact sum 2^24 + 2 as well.
    a = np.zeros(64, np.float32)
    a[0], a[1], a[3] = 4096.0, 1.0, 1.0
    assert float(oracle.dist(a, np.zeros(64, np.float32))) == 2.0 ** 24
    # with ones elsewhere: q[0] = (2^24 + 1 -> 2^24) + 2, q[t] = 4 for t >= 1, and the
    # tree over the 16 runs adds the remaining 60 exactly
    b = np.ones(64, np.float32)
    b[0] = 4096.0
    assert float(oracle.dist(b, np.zeros(64, np.float32))) == 2.0 ** 24 + 62


# --------------------------------------------------------------------------- brute force
def test_knn_bruteforce_matches_numpy():
    rng = np.random.default_rng(3)
    X = rng.integers(0, 16, (300, 16)).astype(np.float32)  # distances exact in FP32
    K = 12
    ids, d = oracle.knn_bruteforce(X, K)
    X64 = X.astype(np.float64)
    D2 = ((X64[:, None, :] - X64[None, :, :]) ** 2).sum(-1)
    for u in range(X.shape[0]):
        order = sorted((D2[u, v], v) for v in range(X.shape[0]) if v != u)[:K]
        assert [v for _, v in order] == ids[u].tolist()
        assert np.array_equal(np.array([x for x, _ in order], np.float32), d[u])


def test_centroid_exact_on_integer_data():
    X = gmm_int(5000, 32, 8, 1)
    c = oracle.centroid(X)
    assert np.array_equal(c, X.astype(np.float64).mean(0).astype(np.float32))


def test_recall_tie_rule():
    # 4 points: q=0 at 0, points 1 and 2 tie at distance 1, point 3 farther.
    X = np.array([[0, 0], [1, 0], [0, 1], [3, 0]], np.float32)
    q = np.array([0], np.uint32)
    ids, kth = oracle.knn_f64(X, q, 1)
    assert kth[0] == 1.0
    G = np.array([[2, 1, 3]] * 4, np.uint32)  # node 0 returns the tied id 2 first
    assert oracle.recall(X, G, q, 1, kth) == 1.0
    G[0, 0] = 3
    assert oracle.recall(X, G, q, 1, kth) == 0.0


# ---------------------------------------------------------------------------- Filter rule
def _pool_for(X, u):
    others = np.array([v for v in range(X.shape[0]) if v != u], np.uint32)
    d = oracle.dist_pairs(X, np.full(others.shape, u, np.uint32), others)
    # one trailing invalid entry so that no real candidate sits at the last index
    return np.append(others, INV), np.append(d, np.float32(np.inf))


def test_filter_collinear_keeps_only_nearest():
    X = np.zeros((5, 4), np.float32)
    X[:, 0] = [0, 1, 2, 3, 4]
    pid, pd = _pool_for(X, 0)
    ids, d = oracle.filter_pool(X, pid, pd, 0, 10, oracle.RULE_DIST, 1.0)
    assert ids.tolist() == [1] and d.tolist() == [1.0]


def _textbook_select(X, u, th, rule):
    """Sequential RNG / alpha / angle selection written from the published rules:
    visit candidates in (d(u,c), c) order and keep c unless a kept w occludes it."""
    X64 = X.astype(np.float64)
    cands = sorted((float(oracle.dist(X[u], X[v])), v) for v in range(len(X)) if v != u)
    kept = []
    for duc, c in cands:
        ok = True
        for w in kept:
            if rule == "dist":
                dwc = np.float32(oracle.dist(X[w], X[c]))
                if np.float32(np.float32(th) * dwc) < np.float32(duc):
                    ok = False
            else:
                a = X64[w] - X64[u]
                b = X64[c] - X64[u]
                cosv = a @ b / np.sqrt((a @ a) * (b @ b))
                if cosv > th:
                    ok = False
            if not ok:
                break
        if ok:
            kept.append(c)
    return kept


@pytest.mark.parametrize("th", [1.0, 1.2])
def test_filter_reduces_to_rng_and_alpha_rule(th):
    rng = np.random.default_rng(5)
    X = rng.integers(0, 50, (150, 8)).astype(np.float32)
    for u in (0, 17, 99):
        pid, pd = _pool_for(X, u)
        ids, _ = oracle.filter_pool(X, pid, pd, u, 1000, oracle.RULE_DIST, th)
        assert ids.tolist() == _textbook_select(X, u, th, "dist")


def _angle_safe_data(N, D, us):
    """First seed whose cosines at the nodes `us` stay 1e-4 away from 0.5 (60 degrees)."""
    for seed in range(100):
        X = np.random.default_rng(seed).standard_normal((N, D)).astype(np.float32)
        X64 = X.astype(np.float64)
        ok = True
        for u in us:
            diff = np.delete(X64, u, 0) - X64[u]
            nrm = np.sqrt((diff ** 2).sum(1))
            C = (diff @ diff.T) / np.outer(nrm, nrm)
            np.fill_diagonal(C, 0.0)
            ok &= bool(np.min(np.abs(C - 0.5)) > 1e-4)
        if ok:
            return X
    raise AssertionError("no boundary-free seed")


def test_filter_reduces_to_nssg_angle_rule():
    us = (0, 13, 30)
    X = _angle_safe_data(40, 6, us)
    for u in us:
        pid, pd = _pool_for(X, u)
        ids, _ = oracle.filter_pool(X, pid, pd, u, 1000, oracle.RULE_ANGLE, 0.5)
        assert ids.tolist() == _textbook_select(X, u, 0.5, "angle")


def test_filter_final_degree_and_drop_last():
    X = np.eye(5, dtype=np.float32) * 3
    X[0] = 0  # u at the origin, 4 mutually orthogonal candidates: none occludes another
    others = np.arange(1, 5, dtype=np.uint32)
    d = np.full(4, 9.0, np.float32)
    # pool exhausted by a selection at the last index: that selection is dropped
    ids, _ = oracle.filter_pool(X, others, d, 0, 100, oracle.RULE_DIST, 1.0)
    assert ids.tolist() == [1, 2, 3]
    # trailing invalid slot: the pool runs dry instead and every selection is kept
    ids, _ = oracle.filter_pool(X, np.append(others, INV), np.append(d, np.inf), 0, 100,
                                oracle.RULE_DIST, 1.0)
    assert ids.tolist() == [1, 2, 3, 4]
    # final degree reached: FD-1 kept
    ids, _ = oracle.filter_pool(X, np.append(others, INV), np.append(d, np.inf), 0, 3,
                                oracle.RULE_DIST, 1.0)
    assert ids.tolist() == [1, 2]
    # u itself and invalid ids are skipped for the first pick; repeats are removed
    ids, _ = oracle.filter_pool(X, np.array([0, 1, 1, 2, INV], np.uint32),
                                np.array([0, 9, 9, 9, np.inf], np.float32), 0, 100,
                                oracle.RULE_DIST, 1.0)
    assert ids.tolist() == [1, 2]
    # nothing valid: degree 0
    ids, _ = oracle.filter_pool(X, np.array([0, INV], np.uint32),
                                np.array([0, np.inf], np.float32), 0, 100, oracle.RULE_DIST,
                                1.0)
    assert ids.size == 0


# ---------------------------------------------------------------------- Collect / ep
def _complete_graph(X):
    N = X.shape[0]
    return oracle.knn_bruteforce(X, N - 1)


def test_cal_ep_on_complete_graph_is_nearest_to_centroid():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((60, 10)).astype(np.float32)
    G, _ = _complete_graph(X)
    c = oracle.centroid(X)
    d64 = ((X.astype(np.float64) - c.astype(np.float64)) ** 2).sum(1)
    assert oracle.cal_ep(X, G, 16) == int(np.argmin(d64))


def test_collect_path_on_complete_graph_sees_every_node():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((40, 10)).astype(np.float32)
    G, GD = _complete_graph(X)
    pid, pd, c = oracle.collect_path(X, G, GD, 16, ep=7, u=5)
    # start list = 39 nodes (all but ep); the first expansion adds ep itself; nothing else
    assert c == 40
    valid = pid[pid != INV]
    assert sorted(valid.tolist()) == list(range(40))
    d_ref = oracle.dist_pairs(X, np.full(40, 5, np.uint32), np.arange(40, dtype=np.uint32))
    order = sorted((d_ref[v], v) for v in range(40))
    assert valid.tolist() == [v for _, v in order]


def test_collect_path_ring_wraps_at_480():
    X = gmm_int(2000, 16, 4, 8)
    G, GD = oracle.knn_bruteforce(X, 64)
    pid, pd, c = oracle.collect_path(X, G, GD, 64, ep=0, u=1234)
    assert 64 <= c <= 64 + 10 * 64
    assert pid.shape[0] == 64 + 480
    valid = pid[pid != INV]
    assert len(set(valid.tolist())) == valid.size
    assert np.all(np.diff(pd[pid != INV]) >= 0)


# ------------------------------------------------------------------ one-hop rank rules
def _brute_rank(X, G, GD, d, kind):
    """CAGRA / DPG detour ranks, reverse lists and the fixed-degree merge, written from the
    rule text: detour[j] counts r<j with G[u][j] in G[G[u][r]] (CAGRA) or with
    d(G[u][r],G[u][j]) < D[u][j] (DPG)."""
    N, K = G.shape
    T = np.zeros((N, d), np.uint32)
    R = [[] for _ in range(N)]
    for u in range(N):
        det = []
        for j in range(K):
            cnt = 0
            for r in range(j):
                if kind == "cagra":
                    cnt += int(G[u, j] in set(G[G[u, r]].tolist()))
                else:
                    a, b = X[G[u, r]].astype(np.float64), X[G[u, j]].astype(np.float64)
                    cnt += int(((a - b) ** 2).sum() < GD[u, j])
            det.append(cnt)
        order = sorted(range(K), key=lambda j: (det[j], G[u, j]))[:d]
        T[u] = G[u, order]
        for j in order:
            R[G[u, j]].append((det[j], u))
    out = np.zeros((N, d), np.uint32)
    h = d // 2
    for u in range(N):
        Rv = sorted(R[u])[:d]
        out[u, :h] = T[u, :h]
        r = len(Rv)
        if r < h:
            A = [v for _, v in Rv if v not in T[u, :d - r].tolist()]
        else:
            A = [v for _, v in Rv if v not in T[u, :h].tolist()][:h]
        a = len(A)
        out[u, h:h + a] = A
        out[u, h + a:] = T[u, h:d - a]
    return out


@pytest.mark.parametrize("kind", ["cagra", "dpg"])
def test_rank_rules_match_brute_force(kind):
    rng = np.random.default_rng(12)
    X = rng.integers(0, 20, (120, 6)).astype(np.float32)
    G, GD = oracle.knn_bruteforce(X, 16)
    res = oracle.prune(X, G, GD, 8, 64, oracle.KIND_CAGRA if kind == "cagra" else oracle.KIND_DPG,
                       0, oracle.RULE_RANK, 0.0)
    assert np.all(res["deg"] == 8)
    assert np.array_equal(res["ids"], _brute_rank(X, G, GD, 8, kind))


def test_rank_reverse_hand_case():
    # 5 nodes on a line, K=2 (each node's two nearest), d=2: one forward + one reverse slot.
    X = np.array([[0], [1], [3], [6], [10]], np.float32)
    G, GD = oracle.knn_bruteforce(X, 2)
    res = oracle.prune(X, G, GD, 2, 64, oracle.KIND_CAGRA, 0, oracle.RULE_RANK, 0.0,
                       stages=True)
    # forward: detours are all 0 except where G[u][1] is in G[G[u][0]]
    #   u=0: G=[1,2]; 2 in G[1]=[0,2] -> detour 1 -> T=[1,2]
    assert res["fwd_ids"][0].tolist() == [1, 2]
    # node 0's output: T[0]=1 then the best reverse neighbour of 0 not equal to it
    assert res["ids"][0, 0] == 1
    assert res["ids"][0, 1] in (1, 2)
    assert np.all(res["deg"] == 2)


# -------------------------------------------------------------- whole pruning invariants
@pytest.mark.parametrize("kind,fr,rr,th", [(0, 0, 0, 1.0), (0, 0, 0, 1.2), (1, 1, 1, 0.5),
                                           (0, 0, 1, 0.5), (1, 1, 0, 1.0)])
def test_prune_invariants(small_int_data, kind, fr, rr, th):
    X = small_int_data[:1500]
    G, GD = oracle.knn_bruteforce(X, 48)
    FD = 24
    r = oracle.prune(X, G, GD, FD, 64, kind, fr, rr, th, stages=True)
    N = X.shape[0]
    assert np.all(r["deg"] <= FD) and np.all(r["fwd_deg"] <= FD - 1)
    for u in range(0, N, 37):
        fw = r["fwd_ids"][u, :r["fwd_deg"][u]]
        fd = r["fwd_d"][u, :r["fwd_deg"][u]]
        assert len(set(fw.tolist())) == fw.size and np.all(fw < N)
        assert np.all(np.diff(fd) >= 0)
        # every kept pair obeys the forward rule
        for i in range(fw.size):
            for j in range(i + 1, fw.size):
                res = np.float32(oracle.dist(X[fw[i]], X[fw[j]]))
                if fr == 0:
                    assert not (np.float32(np.float32(th) * res) - fd[j] < 0)
        out = r["ids"][u, :r["deg"][u]]
        assert np.all(out < N)
        # final rows: unique, no self, drawn from the forward row and the reverse list
        assert len(set(out.tolist())) == out.size and u not in out.tolist()
        rv = (r["rev_keys"][u, :r["rev_cnt"][u]] & 0xFFFFFFFF).tolist()
        assert set(out.tolist()) <= set(fw.tolist()) | set(rv)
        if fw.size + len(rv) <= FD:
            # no second filter: the forward row is kept whole, in order, at the front
            assert out[:fw.size].tolist() == fw.tolist()
            assert set(out.tolist()) == set(fw.tolist()) | set(rv)
    if kind == 0:
        assert r["ep"] == oracle.cal_ep(X, G, 64)


def test_prune_rejects_mixed_layouts(small_int_data):
    X = small_int_data[:200]
    G, GD = oracle.knn_bruteforce(X, 16)
    with pytest.raises(ValueError):
        oracle.prune(X, G, GD, 8, 64, oracle.KIND_CAGRA, 0, oracle.RULE_DIST, 1.0)
    with pytest.raises(ValueError):
        oracle.prune(X, G, GD, 8, 64, oracle.KIND_PATH, 0, oracle.RULE_RANK, 1.0)


# ------------------------------------------------------------------------------- writers
def test_writers_layout(tmp_path):
    deg = np.array([2, 0, 3], np.uint32)
    ids = np.array([[1, 2, INV], [INV] * 3, [0, 1, 2]], np.uint32)
    p = str(tmp_path / "i")
    oracle.store(p, "NSG", deg, ids, ep=2)
    b = open(p, "rb").read()
    assert struct.unpack("<2I", b[:8]) == (3, 2)
    assert struct.unpack("<3I", b[8:20]) == (2, 1, 2)
    assert struct.unpack("<I", b[20:24]) == (0,) and len(b) == 8 + 12 + 4 + 16
    oracle.store(p, "Vamana", deg, ids, ep=1)
    b = open(p, "rb").read()
    size, maxd, ep, frozen = struct.unpack("<QIIQ", b[:24])
    assert (size, maxd, ep, frozen) == (24 + 4 * (3 + 1 + 4), 3, 1, 0) and len(b) == size
    eps = oracle.nssg_eps(3, 99)
    oracle.store(p, "NSSG", deg, ids, eps=eps)
    b = open(p, "rb").read()
    assert struct.unpack("<2I", b[:8]) == (3, 10) and len(b) == 8 + 40 + 12 + 4 + 16
    assert sorted(struct.unpack("<10I", b[8:48])[:3]) == [0, 1, 2]
    full = np.array([[1, 2], [0, 2], [0, 1]], np.uint32)
    d2 = np.full(3, 2, np.uint32)
    oracle.store(p, "CAGRA", d2, full)
    assert open(p, "rb").read() == np.array([[2, 1, 2], [2, 0, 2], [2, 0, 1]], "<u4").tobytes()
    oracle.store(p, "DPG", d2, full)
    b = open(p, "rb").read()
    assert b[:8] == b"KNNGRAPH" and struct.unpack("<3I", b[8:20]) == (2, 0, 3)
    assert struct.unpack("<4I", b[20:36]) == (2, 2, 1, 2)


# ------------------------------------------------------------------------- GNN-Descent
def test_gnn_descent_invariants_and_recall():
    X = gmm_int(4000, 64, 16, 21)
    K = 32
    G, GD, ev = oracle.gnn_descent(X, K, 10, 77)
    N = X.shape[0]
    for u in range(0, N, 97):
        assert u not in G[u]
        assert len(set(G[u].tolist())) == K
        keys = list(zip(GD[u].tolist(), G[u].tolist()))
        assert keys == sorted(keys)
        assert np.array_equal(GD[u], oracle.dist_pairs(X, np.full(K, u, np.uint32), G[u]))
    q = np.arange(0, N, 7, dtype=np.uint32)
    _, kth = oracle.knn_f64(X, q, 10)
    assert oracle.recall(X, G, q, 10, kth) >= 0.97
    # fewer iterations must not be better than more
    G2, _, _ = oracle.gnn_descent(X, K, 2, 77)
    assert oracle.recall(X, G2, q, 10, kth) <= oracle.recall(X, G, q, 10, kth)


def test_gnn_descent_deterministic_across_thread_counts():
    X = gmm_int(1500, 32, 8, 5)
    n0 = oracle.num_threads()
    try:
        oracle.set_threads(1)
        G1, D1, _ = oracle.gnn_descent(X, 16, 6, 3)
        oracle.set_threads(max(2, n0))
        G2, D2, _ = oracle.gnn_descent(X, 16, 6, 3)
    finally:
        oracle.set_threads(n0)
    assert np.array_equal(G1, G2) and np.array_equal(D1, D2)
