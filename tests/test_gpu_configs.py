"""GPU parity at the shapes of the five workloads, against the CPU oracle, bit for bit.

* 10k x 128 integer mixture, K=32, 10 iterations, NSG FD=32 m=64: in full, also through the
  drop-in module (GNN_descent -> Pruning('NSG') index file);
* the 1M x 128 K=96 headline shape at 20k vectors: K=96 in every precision and from the
  learned lists, then NSG, Vamana (alpha 1.2) and CAGRA with FD=32, m=64 (the beam's and
  the lists' two-register-slot paths, K > 64);
* K=128, the upper limit of a list;
* 96-dimensional unit vectors, K=64, NSSG (2-hop, angle 60 degrees);
* the clustered shape through the drop-in largeIndex([0], ...): D=96, K=64, Vamana alpha 1.2,
  max_degree 22, byte for byte against the oracle pipeline;
* dimension 960 with K=64 in the FP8 and BF16 screens;
* the headline at its full size (1M): recall, row invariants, and sampled rows of the kNN
  lists, the entry point, the forward rows and the final rows recomputed by the oracle.
"""
import os

import numpy as np
import pytest

import oracle
import tagore_amd
from oracle import large_index as LI
from tagore_amd import State
from tests.conftest import gmm_unit
from tests.test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

INV = 0xFFFFFFFF


def _say(msg):
    print(msg, flush=True)


def _gen(N, D, kind, ncomp, seed):
    with State(N, D) as s:
        s.gen_data(kind, ncomp, seed, 20.0)
        return s.get_data()


def _knn_equal(G, GD, Gr, GDr, what):
    bad = np.nonzero((G != Gr).any(1))[0]
    assert bad.size == 0, f"{what}: {bad.size} rows differ, first {bad[:5]}"
    assert np.array_equal(GD, GDr), f"{what}: distances differ"


def _write_fvecs(path, X):
    rec = np.empty((X.shape[0], X.shape[1] + 1), np.float32)
    rec[:, 0] = np.frombuffer(np.uint32(X.shape[1]).tobytes(), np.float32)[0]
    rec[:, 1:] = X
    rec.tofile(path)


# --------------------------------------------------------------- 10k x 128, K=32 (in full)
def test_small_config_in_full(tmp_path):
    N, D, K, it, seed = 10000, 128, 32, 10, 0x7A600001
    X = _gen(N, D, 0, 64, seed)
    assert X.min() >= 0 and X.max() <= 255 and np.array_equal(X, np.rint(X))
    Gr, GDr, ev = oracle.gnn_descent(X, K, it, seed)
    for prec in ("exact", "fp32", "bf16"):
        _say(f"10k x 128 K=32 10 iterations, {prec}")
        with State(N, D) as s:
            s.set_data(X)
            s.knn_build(K, it, seed, prec)
            G, GD = s.get_knn()
            st = s.stats()
            r = s.recall(2000, 10, 0x5A11)
        _knn_equal(G, GD, Gr, GDr, prec)
        assert st["knn_evals"] == ev
    _say(f"  recall@10 {r:.4f}")
    assert r >= 0.95
    for rows in ("fp32", "bf16"):
        _compare(X, Gr, GDr, 0, 0, 0, 32, 64, 1.0, rows)
    # the drop-in module: GNN_descent from an .fvecs file, then Pruning('NSG') writes the index
    p = str(tmp_path / "base.fvecs")
    _write_fvecs(p, X)
    from tagore_amd import Tagore
    ptrs = Tagore.GNN_descent(K, N, D, it, p, seed=seed)
    idx = str(tmp_path / "nsg.index")
    Tagore.Pruning(K, N, D, 32, 64, ptrs, "NSG", index_path=idx)
    ref = oracle.prune(X, Gr, GDr, 32, 64, 0, 0, 0, 1.0)
    p_ref = str(tmp_path / "ref.index")
    oracle.store(p_ref, "NSG", ref["deg"], ref["ids"], ref["ep"])
    assert open(idx, "rb").read() == open(p_ref, "rb").read()


# ------------------------------------------------- the 1M x 128, K=96 shape at 20k vectors
@pytest.fixture(scope="module")
def k96_case():
    N, D, K, seed = 20000, 128, 96, 0x7A600002
    X = _gen(N, D, 0, 1024, seed)
    Gr, GDr, ev = oracle.gnn_descent(X, K, 10, seed)
    return X, Gr, GDr, ev, seed


@pytest.mark.parametrize("prec", ["exact", "fp32", "bf16"])
def test_k96_knn_bit_identical(k96_case, prec):
    X, Gr, GDr, ev, seed = k96_case
    _say(f"20k x 128 K=96 10 iterations, {prec}")
    with State(*X.shape) as s:
        s.set_data(X)
        s.knn_build(96, 10, seed, prec)
        G, GD = s.get_knn()
        st = s.stats()
    _knn_equal(G, GD, Gr, GDr, prec)
    assert st["knn_evals"] == ev


def test_k96_learned_two_iterations_and_prunes(k96_case):
    """The headline pipeline at 20k: learned lists, 2 BF16 iterations, then NSG, Vamana and
    CAGRA (FD=32, m=64) on that graph, with FP32 and u8 rows."""
    X, _, _, _, seed = k96_case
    N, D = X.shape
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(96, 0, seed, "bf16", init="learned")
        G0, GD0 = s.get_knn()
        s.knn_build(96, 2, seed, "bf16", init="learned")
        G, GD = s.get_knn()
        r = s.recall(2000, 10, 0x5A11)
    _say(f"learned start + 2 iterations: recall@10 {r:.4f}")
    Gr, GDr, _ = oracle.gnn_descent(X, 96, 2, seed, init=(G0, GD0))
    _knn_equal(G, GD, Gr, GDr, "learned + 2")
    for name, kind, fr, rr, th in [("NSG", 0, 0, 0, 1.0), ("Vamana", 0, 0, 0, 1.2),
                                   ("CAGRA", 2, 0, 2, 0.0)]:
        for rows in ("fp32", "bf16"):
            _compare(X, G, GD, kind, fr, rr, 32, 64, th, rows)


def test_k96_prunes_on_the_ten_iteration_graph(k96_case):
    X, Gr, GDr, _, _ = k96_case
    _compare(X, Gr, GDr, 0, 0, 0, 32, 64, 1.0)
    _compare(X, Gr, GDr, 0, 0, 0, 32, 100, 1.2, "bf16")  # two beam register slots


# ------------------------------------------------------------------------- K=128 limit
@pytest.mark.parametrize("prec", ["exact", "bf16"])
def test_k128_upper_limit(prec):
    N, D, K, it, seed = 5000, 128, 128, 6, 0x128
    X = _gen(N, D, 0, 32, seed)
    _say(f"5k x 128 K=128, {prec}")
    Gr, GDr, ev = oracle.gnn_descent(X, K, it, seed)
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, it, seed, prec)
        G, GD = s.get_knn()
        assert s.stats()["knn_evals"] == ev
    _knn_equal(G, GD, Gr, GDr, prec)
    if prec == "exact":
        _compare(X, Gr, GDr, 0, 0, 0, 32, 128, 1.0)
        _compare(X, Gr, GDr, 0, 0, 0, 128, 64, 1.2, "bf16")
        _compare(X, Gr, GDr, 1, 1, 1, 64, 64, 0.5)
        _compare(X, Gr, GDr, 2, 0, 2, 64, 64, 0.0)


# ----------------------------------------------------- 96-dim unit vectors, K=64, NSSG
def test_unit_k64_nssg():
    N, D, K, seed = 20000, 96, 64, 0x7A600003
    X = _gen(N, D, 1, 512, seed)
    assert np.allclose(np.linalg.norm(X.astype(np.float64), axis=1), 1.0, atol=1e-6)
    _say("20k x 96 unit K=64: random start 6 iterations (bf16), learned start 4 (bf16)")
    Gr, GDr, ev = oracle.gnn_descent(X, K, 6, seed)
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, 6, seed, "bf16")
        G, GD = s.get_knn()
        assert s.stats()["knn_evals"] == ev
        _knn_equal(G, GD, Gr, GDr, "random 6")
        s.knn_build(K, 0, seed, "bf16", init="learned")
        G0, GD0 = s.get_knn()
        s.knn_build(K, 4, seed, "bf16", init="learned")
        GL, GDL = s.get_knn()
        # stored FP32 distances of 2000 judged rows within 1e-5 relative of FP64
        assert s.dist_check(2000, 0x5A11) <= 1e-5
    GLr, GDLr, _ = oracle.gnn_descent(X, K, 4, seed, init=(G0, GD0))
    _knn_equal(GL, GDL, GLr, GDLr, "learned 4")
    for rows in ("fp32", "bf16"):
        _compare(X, GL, GDL, 1, 1, 1, 32, 64, 0.5, rows)  # NSSG: 2-hop, angle 60 degrees
    _compare(X, GL, GDL, 1, 1, 0, 32, 64, 1.0)  # CFS 2-hop / dist


# ------------------------------------------------------- D=960, K=64 in FP8 and BF16
@pytest.mark.parametrize("prec", ["fp8", "bf16"])
def test_d960_k64(prec):
    N, D, K, it, seed = 2000, 960, 64, 4, 0x7A600005
    X = _gen(N, D, 2, 16, seed)
    _say(f"2k x 960 K=64, {prec}")
    Gr, GDr, ev = oracle.gnn_descent(X, K, it, seed)
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, it, seed, prec)
        G, GD = s.get_knn()
        st = s.stats()
        derr = s.dist_check(N, 0x5A11)
    _say(f"  re-scored {st['knn_rescored']} of {st['knn_evals']}, stored distances within "
         f"{derr:.1e} of FP64")
    assert derr <= 1e-5
    _knn_equal(G, GD, Gr, GDr, prec)
    assert st["knn_evals"] == ev
    _compare(X, Gr, GDr, 3, 0, 2, 32, 64, 0.0, prec)  # DPG
    _compare(X, Gr, GDr, 2, 0, 2, 32, 64, 0.0)  # CFS 1-hop / rank


# ------------------------------------------- clustered shape through largeIndex([0], ...)
def test_clustered_shape_through_dropin_large_index(tmp_path):
    """Uneven clusters (log-normal component sizes) of 96-dim unit vectors; every point in
    the clusters of its two nearest centres; per cluster K=64, 8+8 iterations, Vamana alpha
    1.2, FD=32, m=64; merge with max_degree 22.  The drop-in module's index and spill files
    equal the oracle pipeline's byte for byte."""
    tmp = str(tmp_path)
    rng = np.random.default_rng(0x7A600004)
    ncl, D = 6, 96
    w = np.exp(rng.normal(0.0, 0.95, ncl))
    sizes = np.maximum((w / w.sum() * 9000).astype(int), 400)
    cent = rng.standard_normal((ncl, D))
    lab = np.repeat(np.arange(ncl), sizes)
    X = cent[lab] + rng.standard_normal((lab.size, D))
    X = (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)
    Npts = X.shape[0]
    cn = cent / np.linalg.norm(cent, axis=1, keepdims=True)
    o = np.argsort(-(X.astype(np.float64) @ cn.T), 1)
    cen, cen2 = o[:, 0].astype(np.uint32), o[:, 1].astype(np.uint32)
    cpath = os.path.join(tmp, "centres.bin")
    np.stack([cen, cen2], 1).astype(np.uint32).tofile(cpath)
    K, FD, TOPM, md, B, thre, maxp = 64, 32, 64, 22, 3, 1.2, 10 ** 6
    clusters, c1, c2, C = LI.clusters_from_centres(cen.tolist(), cen2.tolist(), ncl, maxp)
    _say(f"clusters {[len(c) for c in clusters]}")
    prefix = os.path.join(tmp, "cluster_")
    for c, mem in enumerate(clusters):
        with open(prefix + str(c), "wb") as f:
            np.asarray([len(mem)], np.uint32).tofile(f)
            X[mem].astype(np.float32).tofile(f)
    local = os.path.join(tmp, "local_")
    index = os.path.join(tmp, "index.bin")
    tagore_amd.Tagore.largeIndex([0], D, K, ncl, maxp, prefix, local, index, md, B, Npts,
                                 cpath, thre, FD, TOPM)
    seed = 0x7A600000
    order, pos = LI.reorder(c1, c2, C, min(B, C))
    graphs, eps = [None] * C, {}
    for c in range(C):
        Xc = X[clusters[c]]
        G, GD, _ = oracle.gnn_descent(Xc, K, 16, tagore_amd.segment_seed(seed, c))
        ref = oracle.prune(Xc, G, GD, FD, TOPM, 0, 0, 0, thre)
        graphs[c] = [[clusters[c][t] for t in ref["ids"][j, :ref["deg"][j]]]
                     for j in range(len(clusters[c]))]
        eps[c] = clusters[c][ref["ep"]]
    spills, idx = LI.merge(clusters, c1, c2, order, pos, graphs, [eps[c] for c in order],
                           min(B, C), md)
    assert open(index, "rb").read() == idx, "merged index differs from the oracle pipeline"
    for c, rows in spills.items():
        assert open(local + str(c), "rb").read() == LI.spill_bytes(rows), f"spill file {c}"


# -------------------------------------------------------- the headline at its full size
def test_headline_full_size_properties():
    """1M x 128 integer mixture (1024 components), learned lists + 2 BF16 iterations, NSG
    FD=32 m=64.  Recall@10 >= 0.95 by the FP64 judge; every row valid; sampled kNN rows
    carry canonical distances; the entry point, and the forward and final rows of sampled
    nodes, equal the oracle's recomputation on the GPU's kNN graph."""
    N, D, K, FD, TOPM, seed = 1_000_000, 128, 96, 32, 64, 0x7A600002
    with State(N, D) as s:
        s.gen_data(0, 1024, seed, 20.0)
        s.knn_build(K, 2, seed, "bf16", init="learned")
        rec = s.recall(10000, 10, 0x5A11)
        derr = s.dist_check(10000, 0x5A11)
        _say(f"1M headline: recall@10 {rec:.4f}, stored distances within {derr:.1e} of FP64")
        assert rec >= 0.95
        assert derr == 0.0  # integer rows: the FP32 distances are exact
        X = s.get_data()
        G, GD = s.get_knn()
        s.prune("path", "dist", "dist", FD, TOPM, 1.0)
        deg, ids, ep = s.get_pruned()
        fdeg, fids, fdd = s.get_forward()
    rng = np.random.default_rng(5)
    samp = np.sort(rng.choice(N, 300, replace=False)).astype(np.uint32)
    # kNN rows: sorted by (d, id), no self, no repeats, canonical distances
    Gs, GDs = G[samp], GD[samp]
    assert not (Gs == samp[:, None]).any()
    ss = np.sort(Gs, 1)
    assert not (ss[:, 1:] == ss[:, :-1]).any()
    d0, d1 = GDs[:, :-1], GDs[:, 1:]
    assert ((d0 < d1) | ((d0 == d1) & (Gs[:, :-1] < Gs[:, 1:]))).all()
    ref = oracle.dist_pairs(X, np.repeat(samp, K), Gs.ravel())
    assert np.array_equal(GDs.ravel(), ref)
    # final rows: degree, range, no repeats, no self
    assert deg.max() <= FD
    valid = np.arange(FD)[None, :] < deg[:, None]
    assert (ids[valid] < N).all()
    srt = np.sort(np.where(valid, ids, INV), 1)
    assert not ((srt[:, 1:] == srt[:, :-1]) & (srt[:, 1:] != INV)).any()
    assert not ((ids == np.arange(N, dtype=np.uint32)[:, None]) & valid).any()
    # entry point and sampled forward rows
    assert ep == oracle.cal_ep(X, G, TOPM)
    for u in samp:
        pid, pd, _ = oracle.collect_path(X, G, GD, TOPM, ep, int(u))
        want, _ = oracle.filter_pool(X, pid, pd, int(u), FD, oracle.RULE_DIST, 1.0)
        assert fids[u, :fdeg[u]].tolist() == want.tolist(), f"forward row {u}"
    # sampled final rows: the reverse phase over the pushes the GPU's forward rows address
    # to the sampled nodes
    fvalid = np.arange(FD)[None, :] < fdeg[:, None]
    hit = np.isin(fids, samp) & fvalid
    src, k = np.nonzero(hit)
    pushes = zip(fids[src, k].tolist(), src.tolist(), fdd[src, k].tolist())
    rdeg, rids, _ = oracle.reverse(X, K, FD, oracle.RULE_DIST, 1.0, fdeg, fids, fdd, pushes)
    for u in samp:
        assert ids[u, :deg[u]].tolist() == rids[u, :rdeg[u]].tolist(), f"final row {u}"
