"""Python restatement of the reference's largeIndex host side (TEST INFRASTRUCTURE ONLY).

Follows pybind.cu:297-391 and large_index.cu:9-522: clusters from the two-centre file with
the round-robin split of oversized clusters, the Algorithm-4 dispatch order, the streaming
union merge through buffer_size cluster graphs with spill files, and the merged index in the
DiskANN layout.  Where the reference reads undefined values (reorder's uninitialised
max_id / max_evict when no candidate overlaps), the smallest unvisited cluster is taken and
slot 0 is evicted; buffer_size is clamped to the cluster count.

Cluster graphs are given per cluster as lists of rows (global ids) in member order.
"""
import struct

import numpy as np


def read_centres(path, n):
    a = np.fromfile(path, dtype=np.uint32, count=2 * n)
    assert a.size == 2 * n
    return a[0::2].tolist(), a[1::2].tolist()


def clusters_from_centres(cen, cen2, cluster_num, max_points):
    """pybind.cu:326-355.  Returns (clusters, cen, cen2, C) after the split."""
    cen, cen2 = list(cen), list(cen2)
    clusters = [[] for _ in range(cluster_num)]
    for i in range(len(cen)):
        clusters[cen[i]].append(i)
        clusters[cen2[i]].append(i)
    extra = 0
    for i in range(cluster_num):
        if len(clusters[i]) <= max_points:
            continue
        div = 2
        while len(clusters[i]) // div > max_points:
            div += 1
        cand = [i] + [cluster_num + extra + j for j in range(div - 1)]
        for j, node in enumerate(clusters[i]):
            if cen[node] == i:
                cen[node] = cand[j % div]
            else:
                cen2[node] = cand[j % div]
        extra += div - 1
    C = cluster_num + extra
    clusters = [[] for _ in range(C)]
    for i in range(len(cen)):
        clusters[cen[i]].append(i)
        clusters[cen2[i]].append(i)
    return clusters, cen, cen2, C


def reorder(cen, cen2, C, B):
    """Algorithm 4 (large_index.cu:9-106).  Returns (order, position)."""
    inter = [dict() for _ in range(C)]
    for a, b in zip(cen, cen2):
        inter[a][b] = inter[a].get(b, 0) + 1
        inter[b][a] = inter[b].get(a, 0) + 1
    visited = [False] * C
    result, buffer = [0], [0]
    visited[0] = True
    position = [0] * C
    for i in range(1, B):
        best, best_id = 0, None
        for j in range(C):
            if visited[j]:
                continue
            t = sum(inter[j].get(r, 0) for r in result)
            if t > best:
                best, best_id = t, j
        if best_id is None:
            best_id = visited.index(False)
        result.append(best_id)
        buffer.append(best_id)
        position[best_id] = i
        visited[best_id] = True
    for _ in range(C - B):
        best, best_id, best_ev = 0, None, 0
        for j in range(C):
            if visited[j]:
                continue
            lsum, lev = 0, 0
            for k in range(B):
                t = sum(inter[j].get(buffer[l % B], 0) for l in range(k + 1, k + B))
                if t > lsum:
                    lsum, lev = t, k
            if lsum > best:
                best, best_id, best_ev = lsum, j, lev
        if best_id is None:
            best_id, best_ev = visited.index(False), 0
        buffer[best_ev] = best_id
        result.append(best_id)
        visited[best_id] = True
        position[best_id] = best_ev
    return result, position


def merge(clusters, cen, cen2, order, position, graphs, metroids, B, max_degree):
    """order_merge_new (large_index.cu:287-521).  graphs[c]: rows of cluster c (global ids).
    Returns (spills: {c: list of rows}, index file bytes)."""
    N = len(cen)
    C = len(clusters)
    rows = {c: [list(r) for r in graphs[c]] for c in range(C)}
    visited = [False] * N
    pointers = [None] * N
    cached = [False] * N
    buffer = [None] * B
    spills = {}

    def spill(clu):
        spills[clu] = [list(r) for r in rows[clu]]
        for node in clusters[clu]:
            visited[node] = True
            pointers[node] = None

    for j, node in enumerate(clusters[order[0]]):
        visited[node] = True
        pointers[node] = j
    buffer[0] = order[0]
    for i in range(1, C):
        c = order[i]
        for j, node in enumerate(clusters[c]):
            if not visited[node]:
                visited[node] = True
                pointers[node] = j
                continue
            if pointers[node] is None:
                continue
            ano = cen2[node] if cen[node] == c else cen[node]
            ra, rc = rows[ano][pointers[node]], rows[c][j]
            first, second = (ra, rc) if cen[node] != c else (rc, ra)
            nh, seen = list(first), set(first)
            for v in second:
                if v not in seen:
                    seen.add(v)
                    nh.append(v)
            rows[ano][pointers[node]] = nh[:max_degree]
            rows[c][j] = []
            cached[node] = True
        buffer[position[c]] = c
        if i >= B - 1 and i + 1 < C:
            clu = buffer[position[order[i + 1]]]
            if clu is not None:
                spill(clu)
    for clu in buffer:
        if clu is not None:
            spill(clu)
    cur = {c: 0 for c in spills}

    def nxt(c):
        r = spills[c][cur[c]]
        cur[c] += 1
        return r

    body = []
    for i in range(N):
        if cached[i]:
            r = nxt(cen[i])
            if len(r) == 0:
                r = nxt(cen2[i])
            else:
                assert len(nxt(cen2[i])) == 0
        else:
            a, b = nxt(cen[i]), nxt(cen2[i])
            r, seen = list(a), set(a)
            for v in b:
                if v not in seen:
                    seen.add(v)
                    r.append(v)
            r = r[:max_degree]
        body.append(len(r))
        body.extend(r)
    size = 24 + 4 * len(body)
    head = struct.pack("<QIIQ", size, max_degree, metroids[-1], 0)
    return spills, head + np.asarray(body, np.uint32).tobytes()


def spill_bytes(rows):
    out = []
    for r in rows:
        out.append(len(r))
        out.extend(r)
    return np.asarray(out, np.uint32).tobytes()
