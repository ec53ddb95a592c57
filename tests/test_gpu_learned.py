"""GPU checks of the learned initial lists (encoder kernels, seeding, iterations, gain).

* encoder forward/backward (tg_mlp_eval) against the oracle's FP64 restatement with the same
  BF16 rounding points; Adam (tg_mlp_adam) against the FP64 closed form;
* the seeded lists are valid kNN lists (sorted by (d, id), canonical distances, no self, no
  repeats) and deterministic; GNN-Descent started from them is bit-identical to the oracle
  started from the same lists;
* the gain: after 2 iterations (1 local join + 1 expansion) the learned start reaches a far
  higher recall than the random start.
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


def _params(D, seed):
    rng = np.random.default_rng(seed)
    H, C = tagore_amd.MLP_HID, tagore_amd.MLP_CODE
    p = np.zeros(tagore_amd.mlp_nparams(D), np.float32)
    p[: H * (D + 1)] = (rng.uniform(-1, 1, H * (D + 1)) / np.sqrt(D)).astype(np.float32)
    p[H * (D + 1):] = (rng.uniform(-1, 1, C * (H + 1)) / np.sqrt(H)).astype(np.float32)
    return p


@pytest.mark.parametrize("D,B,P", [(128, 64, 4), (96, 32, 2), (64, 128, 1)])
def test_encoder_matches_oracle(D, B, P):
    _say(f"encoder D={D} B={B} P={P}")
    rng = np.random.default_rng(D + B)
    X = np.rint(rng.uniform(0, 255, (B * (1 + P), D))).astype(np.float32)
    for b in range(B):
        for j in range(P):
            X[B + b * P + j] = np.clip(X[b] + np.rint(rng.normal(0, 8, D)), 0, 255)
    mu = X.mean(0).astype(np.float32)
    scale = np.float32(1.0 / X.std())
    prm = _params(D, 3)
    loss, g, z = tagore_amd.mlp_eval(X, B, P, mu, float(scale), prm, 1.0)
    rl, rg, rz = oracle.mlp_eval(X, B, P, mu, float(scale), prm.astype(np.float64), 1.0, True)
    assert abs(loss - rl) <= 2e-3 * abs(rl) + 1e-6, (loss, rl)
    assert np.abs(z - rz).max() <= 2e-3 * np.abs(rz).max() + 1e-6
    err = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    _say(f"  loss {loss:.6f} vs {rl:.6f}, relative gradient error {err:.2e}")
    assert err < 1e-2
    # each block separately (a transposed or misplaced block would show here)
    n1 = tagore_amd.MLP_HID * (D + 1)
    for a, b in ((0, n1), (n1, g.size)):
        assert np.linalg.norm(g[a:b] - rg[a:b]) < 1e-2 * np.linalg.norm(rg[a:b])


def test_adam_matches_oracle():
    D = 128
    n = tagore_amd.mlp_nparams(D)
    rng = np.random.default_rng(11)
    p = rng.standard_normal(n).astype(np.float32)
    m = (0.1 * rng.standard_normal(n)).astype(np.float32)
    v = (0.01 * rng.random(n)).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    gp, gm, gv = tagore_amd.mlp_adam(D, 3, 1e-3, g, p, m, v)
    rp, rm, rv = p.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    oracle.mlp_adam(3, float(np.float32(1e-3)), g.astype(np.float64), rp, rm, rv)
    np.testing.assert_allclose(gm, rm, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(gv, rv, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(gp, rp, rtol=1e-5, atol=1e-6)


def _check_lists(X, G, GD):
    N, K = G.shape
    assert (G < N).all()
    assert not (G == np.arange(N, dtype=np.uint32)[:, None]).any(), "self in a list"
    s = np.sort(G, 1)
    assert not (s[:, 1:] == s[:, :-1]).any(), "repeated id in a list"
    ref = oracle.dist_pairs(X, np.repeat(np.arange(N, dtype=np.uint32), K), G.ravel())
    assert np.array_equal(GD.ravel(), ref), "distances are not the canonical FP32 distances"
    d0, d1, i0, i1 = GD[:, :-1], GD[:, 1:], G[:, :-1], G[:, 1:]
    assert ((d0 < d1) | ((d0 == d1) & (i0 < i1))).all(), "rows not sorted by (d, id)"


@pytest.mark.parametrize("kind,N,D,K", [("int", 20000, 128, 32), ("unit", 12000, 96, 24)])
def test_learned_lists_and_parity(kind, N, D, K):
    _say(f"learned init {kind} N={N} D={D} K={K}")
    X = gmm_int(N, D, 20, 5) if kind == "int" else gmm_unit(N, D, 12, 5)
    seed = 0x77
    with State(N, D) as s:
        s.set_data(X)
        s.knn_build(K, 0, seed, "exact", init="learned", sample=5000)
        G0, GD0 = s.get_knn()
        st = s.stats()
        _say(f"  loss {st['learn_loss']}, buckets {st['learn_buckets']}, steps {st['learn_steps']}")
        assert st["learn_steps"] == 2 * (5000 // 2048)  # 2 epochs of 2048-anchor batches
        assert st["learn_loss"][1] < st["learn_loss"][0]
        _check_lists(X, G0, GD0)
        codes = s.codes()
        assert codes.shape == (N, 32) and np.isfinite(codes).all()
        # deterministic
        s.knn_build(K, 0, seed, "exact", init="learned", sample=5000)
        G0b, GD0b = s.get_knn()
        assert np.array_equal(G0, G0b) and np.array_equal(GD0, GD0b)
        # iterations from the learned lists: GPU (every precision) == oracle from the same lists
        ref, _, _ = oracle.gnn_descent(X, K, 4, seed, init=(G0, GD0))
        for prec in ("exact", "bf16"):
            s.knn_build(K, 4, seed, prec, init="learned", sample=5000)
            G4, GD4 = s.get_knn()
            bad = np.nonzero((G4 != ref).any(1))[0]
            assert bad.size == 0, f"{prec}: {bad.size} rows differ from the oracle"
        # and from the given lists: the same graph
        s.set_knn(G0, GD0)
        s.knn_build(K, 4, seed, "exact", init="given")
        G4g, _ = s.get_knn()
        assert np.array_equal(G4g, ref)


@pytest.mark.parametrize("kind,N,D,K,ncomp", [(0, 200000, 128, 96, 200), (1, 200000, 96, 64, 100)])
def test_learned_init_gain(kind, N, D, K, ncomp):
    """The network's gain: recall@10 after 2 iterations, learned start vs random start."""
    _say(f"gain data kind {kind} N={N} D={D} K={K}")
    with State(N, D) as s:
        s.gen_data(kind, ncomp, 0x7A600003, 20.0)
        s.knn_build(K, 2, 0x51, "bf16", init="random")
        r_rand = s.recall(2000, 10, 0x5A11)
        s.knn_build(K, 2, 0x51, "bf16", init="learned")
        r_learn = s.recall(2000, 10, 0x5A11)
    _say(f"  recall@10 after 2 iterations: random {r_rand:.4f}, learned {r_learn:.4f}")
    assert r_learn >= 0.85
    assert r_learn >= r_rand + 0.2


def test_training_graph_replays_match_eager():
    """The encoder's training loop runs as a captured HIP graph, replayed by later builds;
    every replay (with other device work in between) equals the eager launches exactly."""
    X = gmm_int(12000, 128, 20, 8)
    with State(*X.shape) as s:
        s.set_data(X)
        s.knn_build(32, 0, 0x77, "exact", init="learned", sample=5000)
        G_e, GD_e = s.get_knn()
        Z_e = s.codes()
        loss_e = s.stats()["learn_loss"]
        for rep in range(3):
            _say(f"graph build {rep}")
            s.knn_build(32, 0, 0x77, "exact", init="learned", sample=5000, graph=True)
            G_g, GD_g = s.get_knn()
            assert np.array_equal(Z_g := s.codes(), Z_e), f"codes differ on replay {rep}"
            assert np.array_equal(G_g, G_e) and np.array_equal(GD_g, GD_e), f"lists, replay {rep}"
            assert s.stats()["learn_loss"] == loss_e
            s.knn_build(32, 2, 1 + rep, "bf16")  # unrelated device work between replays
