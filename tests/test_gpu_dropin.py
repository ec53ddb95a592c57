"""GPU: every branch of the drop-in module's dispatch (pybind.cu:177-222, 264-284) against
the oracle run with the reference's mapping of names to collect kinds and rules.

Pruning(index_type): NSG = path collect, dist rules, threshold fixed at 1.0 whatever `thre`
says (pybind.cu:181-183); Vamana = path, dist, `thre`; NSSG = 2-hop, angle, `thre`;
CAGRA = 1-hop rank; DPG = 1-hop DPG rank.  The index files must equal the oracle's writers'
bytes.  Pruning_with_CFS(mode, metric): mode picks the collect and the forward rule (path:
dist, 2-hop: angle, 1-hop: CAGRA rank), metric the reverse rule; the pruned rows must equal
the oracle's, and combinations that mix the two row layouts are rejected.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import gmm_int

pytestmark = pytest.mark.gpu

N, D, K, IT, SEED = 3000, 128, 48, 6, 0x0D1
FD, TOPM = 24, 64


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    from tagore_amd import Tagore

    X = gmm_int(N, D, 24, 77)
    p = str(tmp_path_factory.mktemp("dropin") / "base.fvecs")
    rec = np.empty((N, D + 1), np.float32)
    rec[:, 0] = np.frombuffer(np.uint32(D).tobytes(), np.float32)[0]
    rec[:, 1:] = X
    rec.tofile(p)
    ptrs = Tagore.GNN_descent(K, N, D, IT, p, seed=SEED)
    G, GD, _ = oracle.gnn_descent(X, K, IT, SEED)
    return Tagore, ptrs, X, G, GD


@pytest.mark.parametrize("index_type,thre,kind,fr,rr,th", [
    ("NSG", 1.3, 0, 0, 0, 1.0),       # NSG ignores thre: 1.0 is hard-wired
    ("Vamana", 1.2, 0, 0, 0, 1.2),
    ("NSSG", 0.5, 1, 1, 1, 0.5),
    ("CAGRA", 1.0, 2, 0, 2, 0.0),
    ("DPG", 1.0, 3, 0, 2, 0.0),
])
def test_pruning_index_types(built, tmp_path, index_type, thre, kind, fr, rr, th):
    Tagore, ptrs, X, G, GD = built
    _say(f"Tagore.Pruning {index_type} thre={thre}")
    path = str(tmp_path / f"{index_type}.index")
    Tagore.Pruning(K, N, D, FD, TOPM, ptrs, index_type, thre, path)
    ref = oracle.prune(X, G, GD, FD, TOPM, kind, fr, rr, th)
    p_ref = str(tmp_path / "ref.index")
    oracle.store(p_ref, index_type, ref["deg"], ref["ids"], ref["ep"],
                 oracle.nssg_eps(N, 0x5EED))
    assert open(path, "rb").read() == open(p_ref, "rb").read()
    deg, ids, ep = Tagore.pruned_graph(K, N, D, ptrs)
    assert np.array_equal(deg, ref["deg"])
    if kind == 0:
        assert ep == ref["ep"]


@pytest.mark.parametrize("mode,metric,thre,kind,fr,rr", [
    ("path", "dist", 1.2, 0, 0, 0),
    ("path", "angle", 0.5, 0, 0, 1),
    ("2-hop", "dist", 1.0, 1, 1, 0),
    ("2-hop", "angle", 0.5, 1, 1, 1),
    ("1-hop", "rank", 0.0, 2, 0, 2),
])
def test_pruning_with_cfs_modes(built, mode, metric, thre, kind, fr, rr):
    Tagore, ptrs, X, G, GD = built
    _say(f"Tagore.Pruning_with_CFS {mode}/{metric} thre={thre}")
    assert Tagore.Pruning_with_CFS(K, N, D, FD, TOPM, ptrs, mode, metric, thre) is None
    deg, ids, ep = Tagore.pruned_graph(K, N, D, ptrs)
    ref = oracle.prune(X, G, GD, FD, TOPM, kind, fr, rr, thre)
    assert np.array_equal(deg, ref["deg"])
    for u in range(N):
        assert np.array_equal(ids[u, :deg[u]], ref["ids"][u, :deg[u]]), f"row {u}"
    if kind == 0:
        assert ep == ref["ep"]


def test_dispatch_rejections(built):
    Tagore, ptrs, X, G, GD = built
    for mode, metric in [("1-hop", "dist"), ("1-hop", "angle"), ("path", "rank"),
                         ("2-hop", "rank"), ("3-hop", "dist"), ("path", "cosine")]:
        with pytest.raises(ValueError):
            Tagore.Pruning_with_CFS(K, N, D, FD, TOPM, ptrs, mode, metric, 1.0)
    with pytest.raises(ValueError):
        Tagore.Pruning(K, N, D, FD, TOPM, ptrs, "HNSW")
    with pytest.raises(ValueError):  # shape of the graph in ptrs
        Tagore.Pruning(K + 1, N, D, FD, TOPM, ptrs, "NSG")
    with pytest.raises(ValueError):  # 1-hop rows cannot exceed K
        Tagore.Pruning(K, N, D, K + 1, TOPM, ptrs, "CAGRA")
    # the graph is intact after all of the above: one more prune still matches
    Tagore.Pruning_with_CFS(K, N, D, FD, TOPM, ptrs, "path", "dist", 1.0)
    deg, _, _ = Tagore.pruned_graph(K, N, D, ptrs)
    assert np.array_equal(deg, oracle.prune(X, G, GD, FD, TOPM, 0, 0, 0, 1.0)["deg"])
