"""Hand-worked pins of the oracle's Collect and reverse rules (CPU).

Each case is built so that its answer follows from the reference's rule text by reasoning,
and the expected ids are written out literally:

* path Collect (Tagore_src.cu:1261-1339): exactly 10 expansions, and a 480-slot ring filled
  in evaluation order that keeps the last 480 evaluations when it wraps;
* two-hop Collect (Tagore_src.cu:1357-1359): slot i holds G[G[u][i / K]][i % K], i.e. the
  whole lists of u's first neighbours, in list order;
* filter_reverse (Tagore_src.cu:1561-1649, reverse cap :1465-1475): the filter runs only when
  |C| > FD, otherwise the row is the forward row followed by the reverse ids not already in
  it, in (d, id) order; the reverse list holds the K nearest pushers.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import gmm_int

INV = 0xFFFFFFFF


# ------------------------------------------------------------------------ path Collect
def _layer_graph():
    """ep = node 0 at x = 0.  Layer l (l = 0..11) holds ids 1 + 60l .. 60 + 60l at
    x = 100(l+1) + k (k = 0..59).  Node 0's list is layer 0; every node of layer l lists
    layer l+1 (layer 11 lists layer 0).  u = 721 sits at x = 2000, so every node of layer
    l+1 is nearer u than any node of layer l, and its own list is layer 0.  K = 60."""
    K, L = 60, 12
    N = 1 + 60 * L + 1
    u = N - 1
    X = np.zeros((N, 4), np.float32)
    layer = lambda l: np.arange(1 + 60 * l, 61 + 60 * l, dtype=np.uint32)
    for l in range(L):
        X[layer(l), 0] = 100 * (l + 1) + np.arange(60)
    X[u, 0] = 2000
    G = np.zeros((N, K), np.uint32)
    G[0] = layer(0)
    for l in range(L):
        G[layer(l)] = layer((l + 1) % L)
    G[u] = layer(0)
    GD = np.zeros((N, K), np.float32)
    GD[u] = (2000.0 - X[layer(0), 0]) ** 2
    return X, G, GD, u, layer


def test_path_collect_ten_expansions_ring_keeps_last_480():
    X, G, GD, u, layer = _layer_graph()
    pid, pd, c = oracle.collect_path(X, G, GD, 64, ep=0, u=u)
    # node 0's list (layer 0) is 60 evaluations; expansion e (1..10) takes the unexplored
    # entry nearest u, the top node of layer e-1, and evaluates its list, layer e
    assert c == 60 + 10 * 60
    # evaluations 0..659 went to slots e % 480: the ring holds evaluations 180..659, i.e.
    # layers 3..10; the pool adds u's own list (layer 0)
    want = set(layer(0).tolist())
    for l in range(3, 11):
        want |= set(layer(l).tolist())
    valid = pid[pid != INV]
    assert set(valid.tolist()) == want
    assert valid.size == len(want)
    # sorted by distance to u: the top of layer 10 first
    assert valid[0] == layer(10)[-1]
    assert np.all(np.diff(pd[pid != INV]) >= 0)


def test_path_collect_stops_when_nothing_is_left_to_expand():
    # a complete 6-node graph: after node 0's list every node is known, the expansions
    # evaluate nothing new and the beam runs out of unexplored finite entries
    X = np.array([[0, 0], [1, 0], [3, 0], [6, 0], [10, 0], [15, 0]], np.float32)
    X = np.pad(X, ((0, 0), (0, 2)))
    G, GD = oracle.knn_bruteforce(X, 5)
    pid, pd, c = oracle.collect_path(X, G, GD, 8, ep=0, u=5)
    # start list of node 0: 5 evaluations; expanding any node then finds only node 0 new
    assert c == 6
    assert sorted(pid[pid != INV].tolist()) == [0, 1, 2, 3, 4, 5]


# ---------------------------------------------------------------------- two-hop Collect
def _two_hop_graph():
    """K = 24; u = 0 lists ids 1..24; neighbour r (id 1 + r) lists 100 + 24r .. 123 + 24r.
    Ids 100..675 list 1..24.  Points on a line: id i >= 100 at x = i, ids 1..24 at x = -i,
    u at x = 1000, so the ids nearest u are the largest."""
    K = 24
    N = 100 + 24 * K
    X = np.zeros((N, 4), np.float32)
    X[1:25, 0] = -np.arange(1, 25)
    X[100:, 0] = np.arange(100, N)
    X[0, 0] = 1000
    G = np.zeros((N, K), np.uint32)
    G[0] = np.arange(1, 25)
    for r in range(K):
        G[1 + r] = 100 + 24 * r + np.arange(24)
    G[25:] = np.arange(1, 25)
    G[100:] = np.arange(1, 25)
    GD = np.zeros((N, K), np.float32)
    GD[0] = (1000.0 + np.arange(1, 25)) ** 2
    for r in range(K):  # any sorted distances will do for the other rows
        GD[1 + r] = np.arange(24)
    return X, G, GD


def test_two_hop_pool_is_whole_lists_of_the_first_neighbours():
    X, G, GD = _two_hop_graph()
    pid, pd = oracle.collect_2hop(X, G, GD, 0)
    # 480 slots = 20 whole lists: neighbours r = 0..19, ids 100..579; plus u's own list
    want = set(range(1, 25)) | set(range(100, 580))
    valid = pid[pid != INV]
    assert set(valid.tolist()) == want and valid.size == len(want)
    assert valid[0] == 579  # the nearest two-hop id


def test_two_hop_forward_row_through_prune():
    X, G, GD = _two_hop_graph()
    # dist rule with a huge threshold prunes nothing: the forward row is the FD-1 nearest
    # pool entries
    r = oracle.prune(X, G, GD, 8, 64, oracle.KIND_2HOP, oracle.RULE_DIST, oracle.RULE_DIST,
                     1e6, stages=True)
    assert r["fwd_ids"][0, :r["fwd_deg"][0]].tolist() == [579, 578, 577, 576, 575, 574, 573]
    assert r["counters"]["collect_evals"] == 480 * X.shape[0]


# ------------------------------------------------------------------------ reverse phase
def _star():
    """u = node 0 at the origin; nodes 1..4 on the four axes at distances 1, 2, 3, 4
    (mutually orthogonal: none occludes another under the dist rule); node 5 at (5,0,0,0),
    behind node 1 (d(1,5) = 16 < d(0,5) = 25); node 6 at (0,0,0,-6), opposite node 4."""
    X = np.zeros((7, 4), np.float32)
    X[1, 0], X[2, 1], X[3, 2], X[4, 3] = 1, 2, 3, 4
    X[5, 0] = 5
    X[6, 3] = -6
    return X


def _rev(X, K, FD, fwd, pushes, rule=oracle.RULE_DIST, th=1.0):
    N = X.shape[0]
    fdeg = np.zeros(N, np.uint32)
    fids = np.full((N, FD), INV, np.uint32)
    fdd = np.zeros((N, FD), np.float32)
    fdeg[0] = len(fwd)
    for k, v in enumerate(fwd):
        fids[0, k] = v
        fdd[0, k] = oracle.dist(X[0], X[v])
    P = [(0, s, float(oracle.dist(X[0], X[s]))) for s in pushes]
    deg, ids, ev = oracle.reverse(X, K, FD, rule, th, fdeg, fids, fdd, P)
    return ids[0, :deg[0]].tolist(), ev


def test_reverse_union_when_candidates_equal_final_degree():
    X = _star()
    # |C| = 2 forward + 2 reverse = FD: no filter, the union keeps all four
    row, ev = _rev(X, 8, 4, [1, 2], [3, 4])
    assert row == [1, 2, 3, 4] and ev == 0


def test_reverse_union_order_and_dedup():
    X = _star()
    # forward row first, in its order; then the reverse ids by (d, id), skipping ids
    # already in the forward row
    row, _ = _rev(X, 8, 4, [1, 3], [4, 2])
    assert row == [1, 3, 2, 4]
    row, _ = _rev(X, 8, 4, [1, 2], [2, 3])
    assert row == [1, 2, 3]


def test_reverse_filter_when_candidates_exceed_final_degree():
    X = _star()
    # C = {1} + {2, 3, 5, 6} = 5 > FD = 4: sorted 1, 2, 3, 5, 6.  Pick 1; 5 is occluded by
    # 1; pick 2, 3, then 6 (FD picks reached); the last pick is dropped.
    row, ev = _rev(X, 8, 4, [1], [5, 2, 3, 6])
    assert row == [1, 2, 3]
    # rounds: p=1 checks 2,3,5,6; p=2 checks 3,6; p=3 checks 6
    assert ev == 4 + 2 + 1


def test_reverse_list_holds_the_k_nearest_pushers():
    X = _star()
    # K = 4: of the pushers 1,2,3,4,6 only the 4 nearest enter C = [1,2,3,4] (4 > FD = 3),
    # so the filter runs: picks 1, 2, 3, drops the last
    row, _ = _rev(X, 4, 3, [], [6, 4, 3, 2, 1])
    assert row == [1, 2]
    # K = 4, FD = 4: C = [1,2,3,4] is not above FD, so the union keeps all four; node 6 (the
    # fifth pusher) never enters
    row, _ = _rev(X, 4, 4, [], [6, 4, 3, 2, 1])
    assert row == [1, 2, 3, 4]


def test_final_rows_unique_and_within_forward_and_reverse():
    X = gmm_int(1200, 32, 12, 21)
    G, GD = oracle.knn_bruteforce(X, 32)
    N, FD, K = X.shape[0], 16, 32
    for kind, fr, rr, th in [(0, 0, 0, 1.0), (0, 0, 0, 1.2), (1, 1, 1, 0.5)]:
        r = oracle.prune(X, G, GD, FD, 64, kind, fr, rr, th, stages=True)
        for u in range(N):
            out = r["ids"][u, :r["deg"][u]].tolist()
            fw = r["fwd_ids"][u, :r["fwd_deg"][u]].tolist()
            rv = [int(k) & 0xFFFFFFFF for k in r["rev_keys"][u, :r["rev_cnt"][u]]]
            assert len(set(out)) == len(out) and u not in out
            assert set(out) <= set(fw) | set(rv)
            assert r["rev_cnt"][u] <= K
            if len(fw) + len(rv) <= FD:  # union: the forward row is kept as it is
                assert out[:len(fw)] == fw and len(out) == len(set(fw) | set(rv))
            else:
                assert len(out) <= FD - 1
