"""GPU: segmented handles (many clusters in the same launches) against per-segment builds.

Each segment of a segmented build must equal a standalone build of its rows with the segment
seed (ids shifted by the segment start), and that standalone build must equal the oracle.
"""
import numpy as np
import pytest

import oracle
import tagore_amd
from tagore_amd import State
from tests.conftest import gmm_int, gmm_unit

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("precision", ["exact", "bf16"])
def test_segmented_build_equals_standalone(precision):
    sizes = [700, 1500, 260, 2300, 90 + 64]
    D, K, iters, seed = 96, 32, 6, 0x4242
    parts = [gmm_unit(n, D, 6, 10 + i) if i % 2 else gmm_int(n, D, 6, 10 + i) / 255.0
             for i, n in enumerate(sizes)]
    parts = [p.astype(np.float32) for p in parts]
    X = np.concatenate(parts)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    _say(f"segmented {precision}: sizes {sizes}")
    with State(*X.shape) as s:
        s.set_data(X)
        s.set_segments(starts)
        s.knn_build(K, iters, seed, precision)
        G, GD = s.get_knn()
        s.prune("path", "dist", "dist", 24, 48, 1.2)
        deg, ids, _ = s.get_pruned()
        eps = s.entry_points()
    for c, n in enumerate(sizes):
        lo, hi = starts[c], starts[c + 1]
        Xc = X[lo:hi]
        sc = tagore_amd.segment_seed(seed, c)
        Gr, GDr, _ = oracle.gnn_descent(Xc, K, iters, sc)
        assert np.array_equal(G[lo:hi] - lo, Gr), f"segment {c} kNN ids"
        assert np.array_equal(GD[lo:hi], GDr), f"segment {c} kNN distances"
        ref = oracle.prune(Xc, Gr, GDr, 24, 48, 0, 0, 0, 1.2)
        assert eps[c] - lo == ref["ep"], f"segment {c} entry point"
        assert np.array_equal(deg[lo:hi], ref["deg"]), f"segment {c} degrees"
        for u in range(n):
            d = deg[lo + u]
            assert np.array_equal(ids[lo + u, :d] - lo, ref["ids"][u, :d]), f"segment {c} row {u}"


def test_segment_rules():
    X = gmm_int(1000, 64, 4, 1)
    with State(*X.shape) as s:
        s.set_data(X)
        with pytest.raises(tagore_amd.TagoreError):
            s.set_segments([0, 500, 999])       # does not end at N
        with pytest.raises(tagore_amd.TagoreError):
            s.set_segments([0, 500, 500, 1000])  # empty segment
        s.set_segments([0, 20, 1000])
        with pytest.raises(tagore_amd.TagoreError):
            s.knn_build(32, 2, 1, "exact")      # a segment with <= K rows
        s.set_segments([0, 400, 1000])
        with pytest.raises(tagore_amd.TagoreError):
            s.knn_build(32, 2, 1, "exact", init="learned")
        s.knn_build(32, 2, 1, "exact")
        with pytest.raises(tagore_amd.TagoreError):
            s.recall(10, 10, 1)
        s.set_segments(None)
        s.knn_build(32, 2, 1, "exact")
        assert 0.0 <= s.recall(10, 10, 1) <= 1.0
