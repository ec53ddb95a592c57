"""Pins of the largeIndex host restatement (oracle/large_index.py) on hand-worked cases."""
import struct

import numpy as np

from oracle import large_index as L


def _rows(body):
    a = np.frombuffer(body, np.uint32).tolist()
    out, i = [], 0
    while i < len(a):
        d = a[i]
        out.append(a[i + 1:i + 1 + d])
        i += 1 + d
    return out


def test_split_of_oversized_cluster():
    # 6 points, cluster 0 holds all of them (first centre), cluster 1 half as second centre
    cen = [0, 0, 0, 0, 0, 0]
    cen2 = [1, 1, 1, 2, 2, 2]
    clusters, c1, c2, C = L.clusters_from_centres(cen, cen2, 3, 4)
    # cluster 0 (6 > 4) splits into 2 parts round-robin by member position: new id 3
    assert C == 4
    assert c1 == [0, 3, 0, 3, 0, 3]
    assert clusters[0] == [0, 2, 4] and clusters[3] == [1, 3, 5]
    assert clusters[1] == [0, 1, 2] and clusters[2] == [3, 4, 5]


def test_reorder_and_cached_merge():
    # both clusters hold all 4 points; everything is merged while both are buffered
    cen, cen2 = [0, 0, 1, 1], [1, 1, 0, 0]
    clusters, cen, cen2, C = L.clusters_from_centres(cen, cen2, 2, 10)
    order, pos = L.reorder(cen, cen2, C, 2)
    assert order == [0, 1] and pos == [0, 1]
    g0 = [[1, 2], [0], [3], [2, 0]]
    g1 = [[3], [2, 3], [0, 1], [1]]
    spills, idx = L.merge(clusters, cen, cen2, order, pos, [g0, g1], [7, 9], 2, 2)
    # node 0: first centre 0 (processed first) -> its list first: [1,2] + [3] -> [1,2]
    # node 2: first centre 1 (the current cluster) -> current list first: [0,1] + [3]
    assert spills[0] == [[1, 2], [0, 2], [0, 1], [1, 2]]
    assert spills[1] == [[], [], [], []]
    size, md, ep, z = struct.unpack("<QIIQ", idx[:24])
    assert (md, ep, z) == (2, 9, 0) and size == len(idx)
    assert _rows(idx[24:]) == [[1, 2], [0, 2], [0, 1], [1, 2]]


def test_eviction_leaves_uncached_points():
    cen = [0, 0, 1, 0, 1, 0]
    cen2 = [1, 2, 2, 1, 2, 2]
    clusters, cen, cen2, C = L.clusters_from_centres(cen, cen2, 3, 10)
    assert clusters == [[0, 1, 3, 5], [0, 2, 3, 4], [1, 2, 4, 5]]
    order, pos = L.reorder(cen, cen2, C, 2)
    # ties keep the first maximum; cluster 2 evicts slot 0 (cluster 0)
    assert order == [0, 1, 2] and pos == [0, 1, 0]
    g = {0: [[10], [11], [12], [13]], 1: [[20], [21], [22], [23]], 2: [[30], [31], [32], [33]]}
    spills, idx = L.merge(clusters, cen, cen2, order, pos, [g[0], g[1], g[2]], [0, 0, 5], 2, 3)
    # cluster 0 is spilled after cluster 1 merged points 0 and 3 into it
    assert spills[0] == [[10, 20], [11], [12, 22], [13]]
    # cluster 2 merges points 2 and 4 into cluster 1; points 1 and 5 meet an evicted cluster
    assert spills[1] == [[], [21, 31], [], [23, 32]]
    assert spills[2] == [[30], [], [], [33]]
    # final: cached points take their merged record, uncached ones the union of both
    assert _rows(idx[24:]) == [[10, 20], [11, 30], [21, 31], [12, 22], [23, 32], [13, 33]]
    assert struct.unpack("<QIIQ", idx[:24])[2] == 5
