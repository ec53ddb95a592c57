"""GNN-Descent with the iteration sequence captured as a HIP graph (TG_FLAG_GRAPH_ITERS):
every replay gives the graph of individual launches, bit for bit, with other builds (other
seeds, other iteration counts, eager launches) and host synchronisations between the
replays; on a plain handle (odd and even expansion counts, so the list buffers alternate),
after the learned initial lists, and on a segmented handle (several clusters in one graph)."""
import numpy as np
import pytest

import oracle
from tagore_amd import State
from tests.conftest import gmm_int, gmm_unit

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _build(s, K, it, seed, prec, **kw):
    s.knn_build(K, it, seed, prec, **kw)
    return s.get_knn()


@pytest.mark.parametrize("iters", [2, 4, 5])
def test_graph_iterations_replay_equals_eager(iters):
    X = gmm_int(3000, 128, 24, 31)
    K, seed = 32, 0x6A
    with State(*X.shape) as s:
        s.set_data(X)
        G0, D0 = _build(s, K, iters, seed, "bf16")
        for rep in range(3):
            _say(f"iters={iters} replay {rep}")
            G, D = _build(s, K, iters, seed, "bf16", graph_iters=True)
            assert np.array_equal(G, G0) and np.array_equal(D, D0), f"replay {rep}"
            # other work between the replays: another seed, eagerly, and a prune
            _build(s, K, iters, seed + 1 + rep, "bf16")
            s.prune("path", "dist", "dist", 24, 64, 1.0)
    Gr, Dr, _ = oracle.gnn_descent(X, K, iters, seed)
    assert np.array_equal(G0, Gr) and np.array_equal(D0, Dr)


def test_graph_iterations_after_learned_init():
    X = gmm_unit(4000, 96, 16, 32)
    K, seed = 32, 0x6B
    with State(*X.shape) as s:
        s.set_data(X)
        G0, D0 = _build(s, K, 2, seed, "bf16", init="learned", sample=2000)
        for rep in range(3):
            G, D = _build(s, K, 2, seed, "bf16", init="learned", sample=2000, graph_iters=True)
            assert np.array_equal(G, G0) and np.array_equal(D, D0), f"replay {rep}"
            _build(s, K, 3, seed, "exact")


def test_graph_iterations_segmented():
    sizes = [700, 1500, 450, 1100]
    X = gmm_int(sum(sizes), 64, 12, 34)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    K, seed = 24, 0x6C
    with State(*X.shape) as s:
        s.set_data(X)
        s.set_segments(starts)
        G0, D0 = _build(s, K, 6, seed, "fp32")
        for rep in range(3):
            G, D = _build(s, K, 6, seed, "fp32", graph_iters=True)
            assert np.array_equal(G, G0) and np.array_equal(D, D0), f"replay {rep}"
            _build(s, K, 4, seed + rep, "exact")
        # new segment boundaries on the same handle: captured again, not replayed stale
        starts2 = np.array([0, 2000, int(starts[-1])], np.uint32)
        s.set_segments(starts2)
        E0, F0 = _build(s, K, 6, seed, "fp32")
        E, F = _build(s, K, 6, seed, "fp32", graph_iters=True)
        assert np.array_equal(E, E0) and np.array_equal(F, F0)
