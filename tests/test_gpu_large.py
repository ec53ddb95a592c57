"""GPU: largeIndex on one GPU against the oracle pipeline (per-cluster restatement builds +
the Python restatement of the host merge), byte for byte, and independent of batching."""
import os

import numpy as np
import pytest

import oracle
import tagore_amd
from oracle import large_index as L

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _make_case(tmp, Npts=3000, D=32, ncl=6, seed=3):
    rng = np.random.default_rng(seed)
    cent = rng.standard_normal((ncl, D)) * 3
    lab = rng.integers(0, ncl, Npts)
    X = (cent[lab] + rng.standard_normal((Npts, D))).astype(np.float32)
    d = ((X[:, None, :] - cent[None]) ** 2).sum(-1)
    o = np.argsort(d, 1)
    cen, cen2 = o[:, 0].astype(np.uint32), o[:, 1].astype(np.uint32)
    cpath = os.path.join(tmp, "centres.bin")
    np.stack([cen, cen2], 1).astype(np.uint32).tofile(cpath)
    return X, cen.tolist(), cen2.tolist(), cpath


def _write_clusters(tmp, X, clusters):
    prefix = os.path.join(tmp, "cluster_")
    for c, mem in enumerate(clusters):
        with open(prefix + str(c), "wb") as f:
            np.asarray([len(mem)], np.uint32).tofile(f)
            X[mem].astype(np.float32).tofile(f)
    return prefix


@pytest.mark.parametrize("precision", ["exact", "bf16"])
def test_large_index_matches_oracle(tmp_path, precision):
    tmp = str(tmp_path)
    K, FD, TOPM, md, B, thre, seed, maxp = 24, 16, 32, 20, 3, 1.2, 0x1D, 1100
    X, cen, cen2, cpath = _make_case(tmp)
    Npts, D = X.shape
    clusters, c1, c2, C = L.clusters_from_centres(cen, cen2, 6, maxp)
    _say(f"largeIndex {precision}: {C} clusters, sizes {[len(c) for c in clusters]}")
    prefix = _write_clusters(tmp, X, clusters)
    outs = {}
    for br in (0, 1):
        local = os.path.join(tmp, f"local{br}_")
        index = os.path.join(tmp, f"index{br}.bin")
        st = tagore_amd.large_index(D, K, 6, maxp, prefix, local, index, md, B, Npts, cpath,
                                    thre, FD, TOPM, seed=seed, precision=precision,
                                    batch_rows=br, tune=(br == 0))
        _say(f"  batch_rows={br}: {st}")
        assert st["clusters"] == C
        assert st["batches"] == (C if br == 1 else 1)
        outs[br] = (local, open(index, "rb").read())
    assert outs[0][1] == outs[1][1], "the index depends on the batching"
    # oracle pipeline
    order, pos = L.reorder(c1, c2, C, min(B, C))
    graphs, eps = [None] * C, {}
    for c in range(C):
        Xc = X[clusters[c]]
        G, GD, _ = oracle.gnn_descent(Xc, K, 16, tagore_amd.segment_seed(seed, c))
        ref = oracle.prune(Xc, G, GD, FD, TOPM, 0, 0, 0, thre)
        graphs[c] = [[clusters[c][t] for t in ref["ids"][j, :ref["deg"][j]]]
                     for j in range(len(clusters[c]))]
        eps[c] = clusters[c][ref["ep"]]
    metroids = [eps[c] for c in order]
    spills, idx = L.merge(clusters, c1, c2, order, pos, graphs, metroids, min(B, C), md)
    assert outs[0][1] == idx, "merged index differs from the oracle pipeline"
    for c, rows in spills.items():
        for br in (0, 1):
            got = open(outs[br][0] + str(c), "rb").read()
            assert got == L.spill_bytes(rows), f"spill file of cluster {c}"


def test_large_index_rejects(tmp_path):
    tmp = str(tmp_path)
    X, cen, cen2, cpath = _make_case(tmp, Npts=600, ncl=3)
    clusters, *_ = L.clusters_from_centres(cen, cen2, 3, 10 ** 6)
    prefix = _write_clusters(tmp, X, clusters)
    args = (32, 24, 3, 10 ** 6, prefix, os.path.join(tmp, "l_"), os.path.join(tmp, "i.bin"),
            20, 2, 600, cpath, 1.2, 16, 32)
    with pytest.raises(tagore_amd.TagoreError):  # max_degree must stay below K
        tagore_amd.large_index(*args[:7], 24, *args[8:])
    with pytest.raises(tagore_amd.TagoreError):  # buffer of one cluster graph
        tagore_amd.large_index(*args[:8], 1, *args[9:])
    bad = os.path.join(tmp, "bad.bin")
    np.zeros(1200, np.uint32).tofile(bad)  # every point with equal centres
    with pytest.raises(tagore_amd.TagoreError):
        tagore_amd.large_index(*args[:10], bad, *args[11:])
    tagore_amd.large_index(*args)  # the valid call works
    assert os.path.getsize(os.path.join(tmp, "i.bin")) > 24


def test_large_index_reports_io_errors_of_overlapped_stages(tmp_path):
    """The file reader (a batch ahead of the builds) and the merger thread (beside them)
    report their I/O errors as status codes, and the library stays usable afterwards."""
    tmp = str(tmp_path)
    X, cen, cen2, cpath = _make_case(tmp, Npts=900, ncl=4, seed=5)
    clusters, *_ = L.clusters_from_centres(cen, cen2, 4, 10 ** 6)
    prefix = _write_clusters(tmp, X, clusters)
    def args(pre, local, index):
        return (32, 24, 4, 10 ** 6, pre, local, index, 20, 2, 900, cpath, 1.2, 16, 32)
    # a later cluster's file is missing: found while earlier batches build
    gone = os.path.join(tmp, "gone_")
    for c in range(len(clusters)):
        if c != 2:
            os.link(prefix + str(c), gone + str(c))
    _say("largeIndex: missing cluster file")
    with pytest.raises(tagore_amd.TagoreError, match="cannot open cluster file"):
        tagore_amd.large_index(*args(gone, os.path.join(tmp, "l1_"), os.path.join(tmp, "i1")),
                               batch_rows=1)
    # a truncated cluster file
    short = os.path.join(tmp, "short_")
    for c in range(len(clusters)):
        data = open(prefix + str(c), "rb").read()
        open(short + str(c), "wb").write(data[:-4] if c == 1 else data)
    _say("largeIndex: truncated cluster file")
    with pytest.raises(tagore_amd.TagoreError, match="holds fewer than its points"):
        tagore_amd.large_index(*args(short, os.path.join(tmp, "l2_"), os.path.join(tmp, "i2")),
                               batch_rows=1)
    # the spill files cannot be written: the merger thread's error
    _say("largeIndex: unwritable spill prefix")
    with pytest.raises(tagore_amd.TagoreError, match="cannot write"):
        tagore_amd.large_index(*args(prefix, os.path.join(tmp, "no_dir", "l3_"),
                                     os.path.join(tmp, "i3")), batch_rows=1)
    # and a valid call afterwards, batched and per cluster, gives one index
    _say("largeIndex: valid calls after the errors")
    tagore_amd.large_index(*args(prefix, os.path.join(tmp, "l4_"), os.path.join(tmp, "i4")))
    tagore_amd.large_index(*args(prefix, os.path.join(tmp, "l5_"), os.path.join(tmp, "i5")),
                           batch_rows=1)
    assert open(os.path.join(tmp, "i4"), "rb").read() == open(os.path.join(tmp, "i5"), "rb").read()


def test_dropin_large_index(tmp_path):
    """The drop-in module's largeIndex (reference argument order) writes the same index."""
    tmp = str(tmp_path)
    X, cen, cen2, cpath = _make_case(tmp, Npts=1500, ncl=4, seed=9)
    clusters, *_ = L.clusters_from_centres(cen, cen2, 4, 10 ** 6)
    prefix = _write_clusters(tmp, X, clusters)
    a = os.path.join(tmp, "a.bin")
    b = os.path.join(tmp, "b.bin")
    tagore_amd.Tagore.largeIndex([0], 32, 24, 4, 10 ** 6, prefix, os.path.join(tmp, "la_"), a,
                                 20, 2, 1500, cpath, 1.2, 16, 32)
    tagore_amd.large_index(32, 24, 4, 10 ** 6, prefix, os.path.join(tmp, "lb_"), b, 20, 2,
                           1500, cpath, 1.2, 16, 32, seed=0x7A600000, precision="bf16")
    assert open(a, "rb").read() == open(b, "rb").read()
    with pytest.raises(ValueError):
        tagore_amd.Tagore.largeIndex([0, 1], 32, 24, 4, 10 ** 6, prefix, os.path.join(tmp, "l_"),
                                     a, 20, 2, 1500, cpath, 1.2, 16, 32)
