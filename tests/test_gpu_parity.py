"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Pruning must be bit-identical (ids, order, degrees, entry point, evaluation counts and
index-file bytes); the exact-mode kNN build must be bit-identical to the oracle's
GNN-Descent restatement; recall is judged against FP64 brute force.
"""
import os

import numpy as np
import pytest

import oracle
import tagore_amd
from tagore_amd import State, TagoreError
from tests.conftest import gmm_int, gmm_unit

pytestmark = pytest.mark.gpu

INV = 0xFFFFFFFF


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def int_case():
    X = gmm_int(4000, 128, 32, 101)
    G, GD = oracle.knn_bruteforce(X, 64)
    return X, G, GD


@pytest.fixture(scope="module")
def unit_case():
    X = gmm_unit(3000, 96, 16, 202)
    G, GD = oracle.knn_bruteforce(X, 48)
    return X, G, GD


def _gpu_prune(X, G, GD, kind, fr, rr, FD, TOPM, th, precision="fp32"):
    with State(*X.shape) as s:
        s.set_data(X)
        s.set_knn(G, GD)
        s.prune(kind, fr, rr, FD, TOPM, th, precision)
        deg, ids, ep = s.get_pruned()
        fdeg, fids, fd = s.get_forward()
        st = s.stats()
    return deg, ids, ep, fdeg, fids, fd, st


def _compare(X, G, GD, kind, fr, rr, FD, TOPM, th, precision="fp32"):
    _say(f"prune parity kind={kind} fwd={fr} rev={rr} FD={FD} TOPM={TOPM} th={th} "
         f"N={X.shape[0]} D={X.shape[1]} K={G.shape[1]} rows={precision}")
    ref = oracle.prune(X, G, GD, FD, TOPM, kind, fr, rr, th, stages=True)
    deg, ids, ep, fdeg, fids, fd, st = _gpu_prune(X, G, GD, kind, fr, rr, FD, TOPM, th,
                                                  precision)
    N = X.shape[0]
    assert np.array_equal(fdeg, ref["fwd_deg"])
    for u in range(N):
        n = fdeg[u]
        assert np.array_equal(fids[u, :n], ref["fwd_ids"][u, :n]), f"forward row {u}"
        if kind < 2:
            assert np.array_equal(fd[u, :n], ref["fwd_d"][u, :n]), f"forward dists {u}"
    assert np.array_equal(deg, ref["deg"])
    for u in range(N):
        assert np.array_equal(ids[u, :deg[u]], ref["ids"][u, :deg[u]]), f"row {u}"
    if kind == 0:
        assert ep == ref["ep"]
    c = ref["counters"]
    pc = st["prune_counters"]
    assert pc["collect"] == c["collect_evals"]
    assert pc["filter"] == c["filter_evals"]
    assert pc["reverse"] == c["reverse_evals"]
    assert pc["adj_rows"] == c["adj_rows"]
    if kind == 0:
        assert pc["ep"] == c["ep_evals"]
    return ref


@pytest.mark.parametrize("name,kind,fr,rr,th", [
    ("NSG", 0, 0, 0, 1.0),
    ("Vamana", 0, 0, 0, 1.2),
    ("CFS path/angle", 0, 0, 1, 0.5),
    ("CFS 2-hop/dist", 1, 1, 0, 1.0),
    ("CAGRA", 2, 0, 2, 0.0),
    ("DPG", 3, 0, 2, 0.0),
])
def test_prune_parity_int(int_case, name, kind, fr, rr, th):
    X, G, GD = int_case
    _compare(X, G, GD, kind, fr, rr, 32, 64, th)


@pytest.mark.parametrize("name,kind,fr,rr,th", [
    ("NSSG", 1, 1, 1, 0.5),
    ("NSG", 0, 0, 0, 1.0),
    ("Vamana", 0, 0, 0, 1.2),
])
def test_prune_parity_unit(unit_case, name, kind, fr, rr, th):
    X, G, GD = unit_case
    _compare(X, G, GD, kind, fr, rr, 32, 64, th)


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
@pytest.mark.parametrize("name,kind,fr,rr,th", [
    ("NSG", 0, 0, 0, 1.0),
    ("NSSG", 1, 1, 1, 0.5),
    ("DPG", 3, 0, 2, 0.0),
])
def test_prune_parity_bf16_rows(int_case, name, kind, fr, rr, th, prec):
    """BF16 rows are read only for data sets that BF16 holds exactly: BF16-rounded unit
    vectors take the BF16 path, integer data in [0,255] the u8 copy in the forward and
    reverse passes (BF16 in the others), plain unit vectors fall back to FP32 rows.  DPG
    screens its pair distances on an MFMA Gram (BF16 or FP8).  All must equal the oracle
    bit for bit."""
    X, G, GD = int_case
    _compare(X, G, GD, kind, fr, rr, 32, 64, th, prec)
    U = gmm_unit(2000, 96, 16, 303)
    Ub = (U.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)  # exact in BF16
    for Y in (U, Ub):
        GY, GDY = oracle.knn_bruteforce(Y, 48)
        _compare(Y, GY, GDY, kind, fr, rr, 32, 64, th, prec)


@pytest.mark.parametrize("D,lo,hi", [(128, 0, 255), (96, 0, 255), (200, 0, 200), (200, -60, 60),
                                    (64, -256, 256)])
def test_prune_parity_integer_bf16_rows(D, lo, hi):
    """Integer data under precision "bf16": values in [0, 255] take the u8 copy (exact dot4
    distances, one or two 64-byte chunks per row), signed values the BF16 copy, which holds
    them exactly; D not a multiple of 32 or 64, values at +-256.  Every rule, and a beam of
    two register slots (TOPM > 64), must equal the oracle bit for bit."""
    rng = np.random.default_rng(D + hi)
    C = rng.integers(lo, hi + 1, (24, D))
    X = np.clip(C[rng.integers(0, 24, 2500)] + rng.integers(-12, 13, (2500, D)), lo, hi)
    X = X.astype(np.float32)
    G, GD = oracle.knn_bruteforce(X, 48)
    for name, kind, fr, rr, th in [("NSG", 0, 0, 0, 1.0), ("Vamana", 0, 0, 0, 1.2),
                                   ("CFS path/angle", 0, 0, 1, 0.5), ("NSSG", 1, 1, 1, 0.5),
                                   ("CFS 2-hop/dist", 1, 1, 0, 1.0)]:
        _compare(X, G, GD, kind, fr, rr, 32, 64, th, "bf16")
    _compare(X, G, GD, 0, 0, 0, 24, 100, 1.0, "bf16")  # two beam slots (TOPM > 64)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_dpg_parity_d960(prec):
    """DPG and CAGRA (1-hop/rank) at dimension 960, the C5 shape, in every precision."""
    rng = np.random.default_rng(961)
    C = rng.standard_normal((8, 960))
    X = np.abs(C[rng.integers(0, 8, 1200)] + rng.standard_normal((1200, 960)))
    X = (X / X.max()).astype(np.float32)
    G, GD = oracle.knn_bruteforce(X, 64)
    _compare(X, G, GD, 3, 0, 2, 32, 64, 0.0, prec)
    _compare(X, G, GD, 2, 0, 2, 32, 64, 0.0, prec)


def test_prune_parity_small_k_and_odd_sizes():
    # K < TOPM (beam padding), FD > K/2, a degree cap that is reached often
    X = gmm_int(1500, 64, 8, 7)
    G, GD = oracle.knn_bruteforce(X, 32)
    _compare(X, G, GD, 0, 0, 0, 24, 64, 1.0)
    _compare(X, G, GD, 0, 0, 0, 6, 16, 1.3)
    _compare(X, G, GD, 1, 1, 1, 30, 64, 0.3)


@pytest.mark.parametrize("case", ["int", "unit"])
def test_prune_parity_noncanonical_list_distances(int_case, unit_case, case):
    """kNN lists whose stored distances are not the canonical ones (every row shifted by a
    constant, so the rows stay sorted): the pool then holds one id under two keys, and the
    filter's rule that invalidates a repeat of the selected id without counting it decides
    the rows and the evaluation counts (u8 rows with the two-round passes for the integer
    data, FP32 rows for the unit data)."""
    X, G, GD = int_case if case == "int" else unit_case
    GDs = (GD + (np.float32(3.0) if case == "int" else np.float32(0.25))).astype(np.float32)
    _compare(X, G, GDs, 0, 0, 0, 32, 64, 1.0)
    _compare(X, G, GDs, 0, 0, 0, 32, 64, 1.2)
    _compare(X, G, GDs, 1, 1, 1, 32, 64, 0.5)


def test_index_files_are_byte_identical(int_case, tmp_path):
    X, G, GD = int_case
    with State(*X.shape) as s:
        s.set_data(X)
        s.set_knn(G, GD)
        for name, (kind, fr, rr, fixed) in tagore_amd.INDEX_TYPES.items():
            _say(f"index file {name}")
            th = fixed if fixed is not None else (0.5 if kind == 1 else 1.2)
            s.prune(kind, fr, rr, 32, 64, th)
            p_gpu = str(tmp_path / f"gpu.{name}")
            s.store_index(p_gpu, name, 0x5EED)
            ref = oracle.prune(X, G, GD, 32, 64, kind, fr, rr, th)
            p_ref = str(tmp_path / f"ref.{name}")
            oracle.store(p_ref, name, ref["deg"], ref["ids"], ref["ep"],
                         oracle.nssg_eps(X.shape[0], 0x5EED))
            assert open(p_gpu, "rb").read() == open(p_ref, "rb").read(), name


@pytest.mark.parametrize("precision", ["exact", "fp32", "bf16", "fp8"])
@pytest.mark.parametrize("N,D,K,iters,kind", [(3000, 128, 32, 6, "int"), (2500, 96, 48, 8, "unit"),
                                              (1200, 64, 64, 4, "int"), (2000, 96, 32, 6, "unitb"),
                                              (1500, 128, 32, 6, "sint")])
def test_knn_bit_identical(N, D, K, iters, kind, precision):
    """Every precision gives the oracle's graph: the screens only decide what to re-score.

    int: non-negative integers (exact MFMA screen, no re-score); sint: signed integers
    (exact screen under the tighter signed bound); unitb: unit vectors truncated to BF16
    (re-scores read the BF16 copy); unit: general FP32 (re-scores read FP32 rows)."""
    _say(f"kNN {precision} N={N} D={D} K={K} iters={iters} data={kind}")
    if kind == "int":
        X = gmm_int(N, D, 16, 5)
    elif kind == "sint":
        X = gmm_int(N, D, 16, 5) - np.float32(128)
    else:
        X = gmm_unit(N, D, 16, 5)
        if kind == "unitb":
            X = (X.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)
    seed = 0x1234
    Gr, GDr, ev = oracle.gnn_descent(X, K, iters, seed)
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, iters, seed, precision)
        G, GD = s.get_knn()
        st = s.stats()
    _say(f"  re-scored {st['knn_rescored']} of {st['knn_evals']} evaluations")
    bad = np.nonzero((G != Gr).any(1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    assert np.array_equal(GD, GDr)
    assert st["knn_evals"] == ev


@pytest.mark.parametrize("precision", ["exact", "bf16"])
def test_knn_bit_identical_hub(precision):
    """A hub: 1500 integer points within +-3 of one centre, plus 500 points of distant
    clusters.  The centre enters nearly every list near it, so its reverse run and its push
    list hold far more than 256 keys (the chunked top-R path) -- still the oracle's graph."""
    rng = np.random.default_rng(404)
    N, D, K = 2000, 64, 32
    X = np.empty((N, D), np.float32)
    X[0] = 128.0
    X[1:1500] = 128.0 + rng.integers(-3, 4, (1499, D))
    X[1500:] = gmm_int(500, D, 8, 404)
    seed = 0x4B
    Gr, GDr, ev = oracle.gnn_descent(X, K, 4, seed)
    indeg = np.bincount(Gr.ravel(), minlength=N)
    _say(f"hub {precision}: centre in-degree {indeg[0]}, max {indeg.max()}")
    assert indeg.max() > 256
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, 4, seed, precision)
        G, GD = s.get_knn()
        st = s.stats()
    bad = np.nonzero((G != Gr).any(1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    assert np.array_equal(GD, GDr)
    assert st["knn_evals"] == ev
    if precision == "bf16":  # the prune's reverse pass sees the hub too
        _compare(X, Gr, GDr, 0, 0, 0, 32, 64, 1.0, "bf16")


def test_recall_judge_matches_oracle():
    X = gmm_unit(3000, 96, 16, 9)
    with State(*X.shape) as s:
        s.set_data(X)
        s.knn_build(32, 4, 77, "exact")
        G, _ = s.get_knn()
        r_gpu = s.recall(500, 10, 0x5A11)
    perm = np.arange(X.shape[0])
    # same query selection as tg_recall: SplitMix64 Fisher-Yates prefix
    M = (1 << 64) - 1
    x = 0x5A11
    for i in range(500):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        j = i + z % (X.shape[0] - i)
        perm[i], perm[j] = perm[j], perm[i]
    q = perm[:500].astype(np.uint32)
    _, kth = oracle.knn_f64(X, q, 10)
    r_ref = oracle.recall(X, G, q, 10, kth)
    assert abs(r_gpu - r_ref) < 1e-12
    assert 0.5 < r_gpu <= 1.0


def test_distance_check_matches_numpy():
    """tg_knn_dist_check: the largest |d32 - d64| / d64 over the kNN rows of the judged nodes
    (tg_recall's selection), against the same figure from numpy FP64; and its rejections."""
    X = gmm_unit(3000, 96, 16, 9)
    with State(*X.shape) as s:
        s.set_data(X)
        with pytest.raises(RuntimeError, match="no kNN graph"):
            s.dist_check(100)
        s.knn_build(32, 4, 77, "bf16")
        G, GD = s.get_knn()
        e_gpu = s.dist_check(500, 0x5A11)
        with pytest.raises(RuntimeError, match="nq must be"):
            s.dist_check(0)
    perm = np.arange(X.shape[0])
    M = (1 << 64) - 1
    x = 0x5A11
    for i in range(500):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        j = i + z % (X.shape[0] - i)
        perm[i], perm[j] = perm[j], perm[i]
    q = perm[:500]
    X64 = X.astype(np.float64)
    d64 = ((X64[q][:, None, :] - X64[G[q].astype(np.int64)]) ** 2).sum(-1)
    e_ref = float((np.abs(GD[q].astype(np.float64) - d64) / d64).max())
    assert abs(e_gpu - e_ref) <= 1e-12
    assert 0.0 < e_gpu <= 1e-5


def test_generated_data_properties():
    with State(20000, 128) as s:
        s.gen_data(0, 64, 42, 20.0)
        X = s.get_data()
        s.gen_data(0, 64, 42, 20.0)
        assert np.array_equal(X, s.get_data())  # deterministic
    assert X.min() >= 0 and X.max() <= 255 and np.array_equal(X, np.rint(X))
    with State(5000, 96) as s:
        s.gen_data(1, 32, 3)
        U = s.get_data()
    assert np.allclose(np.linalg.norm(U.astype(np.float64), axis=1), 1.0, atol=1e-6)
    with State(5000, 64) as s:
        s.gen_data(2, 32, 3)
        A = s.get_data()
    assert A.min() >= 0 and A.max() == 1.0


def test_dropin_module_end_to_end(tmp_path):
    from tagore_amd import Tagore

    X = gmm_int(2000, 128, 16, 31)
    p = tmp_path / "base.fvecs"
    rec = np.empty((X.shape[0], X.shape[1] + 1), np.float32)
    rec[:, 0] = np.frombuffer(np.uint32(128).tobytes(), np.float32)[0]
    rec[:, 1:] = X
    rec.tofile(p)
    ptrs = Tagore.GNN_descent(32, 2000, 128, 6, str(p), seed=99)
    assert len(ptrs) == 8
    idx = str(tmp_path / "nsg.index")
    Tagore.Pruning(32, 2000, 128, 24, 64, ptrs, "NSG", index_path=idx)
    Tagore.Pruning_with_CFS(32, 2000, 128, 24, 64, ptrs, "path", "dist", 1.2)
    with pytest.raises(ValueError):
        Tagore.Pruning_with_CFS(32, 2000, 128, 24, 64, ptrs, "1-hop", "dist", 1.0)
    G, GD, _ = oracle.gnn_descent(X, 32, 6, 99)
    ref = oracle.prune(X, G, GD, 24, 64, 0, 0, 0, 1.0)
    p_ref = str(tmp_path / "ref.index")
    oracle.store(p_ref, "NSG", ref["deg"], ref["ids"], ref["ep"])
    assert open(idx, "rb").read() == open(p_ref, "rb").read()


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_knn_tiles_and_autotune(precision):
    """Both MFMA staging depths give the oracle's graph; the autotuner times every variant."""
    N, D, K, iters, seed = 2500, 96, 32, 6, 0x99
    X = gmm_unit(N, D, 16, 6)
    Gr, GDr, _ = oracle.gnn_descent(X, K, iters, seed)
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, iters, seed, precision, tile=1)
        G, GD = s.get_knn()
        assert np.array_equal(G, Gr) and np.array_equal(GD, GDr)
        (bp, bt), cands = s.autotune(K, seed)
        _say(f"autotune: best {bp} tile {bt}; {cands}")
        assert len(cands) == 7 and (bp, bt) in [(c[0], c[1]) for c in cands]
        assert min(c[2] for c in cands) == [c[2] for c in cands if (c[0], c[1]) == (bp, bt)][0]


@pytest.mark.parametrize("precision", ["exact", "fp8", "bf16"])
def test_knn_bit_identical_d960(precision):
    """The C5 shape (dimension 960, non-negative data scaled to [0,1]) in every screen."""
    N, D, K, iters, seed = 1500, 960, 32, 4, 0x960
    rng = np.random.default_rng(960)
    C = rng.standard_normal((8, D))
    X = np.abs(C[rng.integers(0, 8, N)] + rng.standard_normal((N, D)))
    X = (X / X.max()).astype(np.float32)
    _say(f"kNN D=960 {precision}")
    Gr, GDr, ev = oracle.gnn_descent(X, K, iters, seed)
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, iters, seed, precision)
        G, GD = s.get_knn()
        st = s.stats()
    _say(f"  re-scored {st['knn_rescored']} of {st['knn_evals']} evaluations")
    assert np.array_equal(G, Gr) and np.array_equal(GD, GDr)
    assert st["knn_evals"] == ev


def test_set_knn_rejects_unsorted_or_negative_rows(int_case):
    """tg_set_knn takes rows sorted by (distance, id) with distances >= 0 (the prune kernels
    merge them as sorted lists); others are rejected before anything is copied."""
    X, G, GD = int_case
    with State(*X.shape) as s:
        s.set_data(X)
        G2, D2 = G.copy(), GD.copy()
        G2[7, [3, 4]] = G2[7, [4, 3]]
        D2[7, [3, 4]] = D2[7, [4, 3]]
        if D2[7, 3] == D2[7, 4]:  # a tie: swapped ids are out of id order
            assert G2[7, 3] > G2[7, 4]
        with pytest.raises(TagoreError, match="not sorted"):
            s.set_knn(G2, D2)
        D3 = GD.copy()
        D3[5, 0] = -1.0
        with pytest.raises(TagoreError, match=">= 0"):
            s.set_knn(G, D3)
        s.set_knn(G, GD)  # the sorted rows are accepted


@pytest.mark.parametrize("D", [104, 128, 200, 248])
def test_knn_bit_identical_u8_rows(D):
    """Integer data in [0, 255] under the BF16 screen: the local join's products on int8
    MFMA from the u8 copy (rows staged whole, 16-byte pieces, partial pieces when D is not a
    multiple of 16, one or four K-steps), exact, so the graph is the oracle's bit for bit."""
    rng = np.random.default_rng(D)
    C = rng.integers(0, 256, (20, D))
    X = np.clip(C[rng.integers(0, 20, 2500)] + rng.integers(-20, 21, (2500, D)), 0, 255)
    X = X.astype(np.float32)
    K, it, seed = 32, 6, 0x0B
    Gr, GDr, ev = oracle.gnn_descent(X, K, it, seed)
    with State(*X.shape) as s:
        s.set_data(X)
        s.knn_build(K, it, seed, "bf16")
        G, GD = s.get_knn()
        assert s.stats()["knn_evals"] == ev
    bad = np.nonzero((G != Gr).any(1))[0]
    assert bad.size == 0, f"D={D}: {bad.size} rows differ, first {bad[:5]}"
    assert np.array_equal(GD, GDr)
