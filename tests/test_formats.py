"""Vector and index file readers (CPU only): .bvecs / .fbin round trips, and every index
layout the reference's Store functions write (common.cu:151-248, written here by the
oracle's restatement) read back to the same degrees, ids and entry points; malformed files
raise instead of being read."""
import numpy as np
import pytest

import oracle
import tagore_amd
from tagore_amd import TagoreError

INV = 0xFFFFFFFF


def test_bvecs_and_fbin_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    X = rng.integers(0, 256, (37, 20)).astype(np.uint8)
    p = tmp_path / "a.bvecs"
    with open(p, "wb") as f:
        for r in X:
            f.write(np.uint32(20).tobytes())
            f.write(r.tobytes())
    assert np.array_equal(tagore_amd.read_bvecs(str(p)), X.astype(np.float32))
    Y = rng.standard_normal((11, 7)).astype(np.float32)
    q = tmp_path / "a.fbin"
    with open(q, "wb") as f:
        f.write(np.array([11, 7], np.uint32).tobytes())
        f.write(Y.tobytes())
    assert np.array_equal(tagore_amd.read_fbin(str(q)), Y)
    with open(q, "ab") as f:
        f.write(bytes(4))
    with pytest.raises(TagoreError):
        tagore_amd.read_fbin(str(q))
    with open(p, "ab") as f:
        f.write(bytes(1))
    with pytest.raises(TagoreError):
        tagore_amd.read_bvecs(str(p))


def _graph(N, FD, seed, full=False):
    rng = np.random.default_rng(seed)
    deg = np.full(N, FD, np.uint32) if full else rng.integers(1, FD + 1, N).astype(np.uint32)
    ids = np.full((N, FD), INV, np.uint32)
    for u in range(N):
        ids[u, :deg[u]] = rng.choice(N, deg[u], replace=False)
    return deg, ids


@pytest.mark.parametrize("kind", ["NSG", "Vamana", "NSSG", "DPG", "CAGRA"])
def test_index_layouts_read_back(tmp_path, kind):
    N, FD = 300, 16
    deg, ids = _graph(N, FD, 3, full=(kind == "CAGRA"))
    path = str(tmp_path / f"g.{kind}")
    eps = oracle.nssg_eps(N, 0x5EED)
    oracle.store(path, kind, deg, ids, ep=17, eps=eps)
    rdeg, rids, reps = tagore_amd.read_index(path, kind, N)
    w = rids.shape[1]
    if kind in ("NSG", "NSSG"):
        assert w == FD
    else:
        assert w == int(deg.max())  # Vamana header max degree, DPG / CAGRA largest row
    assert np.array_equal(rdeg, deg)
    for u in range(N):
        assert np.array_equal(rids[u, :deg[u]], ids[u, :deg[u]])
        assert (rids[u, deg[u]:] == INV).all()
    want = {"NSG": [17], "Vamana": [17], "NSSG": eps.tolist(), "DPG": [], "CAGRA": []}[kind]
    assert reps.tolist() == want


def test_index_reader_rejects_malformed(tmp_path):
    N, FD = 50, 8
    deg, ids = _graph(N, FD, 4)
    path = str(tmp_path / "g.nsg")
    oracle.store(path, "NSG", deg, ids, ep=3)
    with pytest.raises(TagoreError):
        tagore_amd.read_index(path, "NSG", N + 1)  # more rows expected than stored
    with pytest.raises(TagoreError):
        tagore_amd.read_index(path, "NSG", N - 1)  # trailing rows
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-4])
    with pytest.raises(TagoreError):
        tagore_amd.read_index(path, "NSG", N)  # truncated
    bad = np.frombuffer(raw, np.uint32).copy()
    bad[3] = N + 5  # first id of row 0 out of range
    open(path, "wb").write(bad.tobytes())
    with pytest.raises(TagoreError):
        tagore_amd.read_index(path, "NSG", N)
    with pytest.raises(TagoreError):
        tagore_amd.read_index(path, "HNSW", N)
