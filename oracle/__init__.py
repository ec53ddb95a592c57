"""CPU oracle for Tagore's GNN-Descent kNN construction and CFS pruning.

TEST INFRASTRUCTURE ONLY.  This package is the checker.  It may be imported by tests/,
by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product package
(tagore_amd) never imports it.

The arithmetic lives in tagore_oracle.cpp (plain C++17 + OpenMP, built by the Makefile next
to it).  This module is a thin ctypes wrapper that takes and returns numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

INV = 0xFFFFFFFF

# index kinds / rules, as in or_prune
KIND_PATH, KIND_2HOP, KIND_CAGRA, KIND_DPG = 0, 1, 2, 3
RULE_DIST, RULE_ANGLE, RULE_RANK = 0, 1, 2

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.or_dist.restype = ctypes.c_float
        L.or_dist.argtypes = [f32p, f32p, ctypes.c_int]
        L.or_dist_pairs.restype = None
        L.or_dist_pairs.argtypes = [f32p, ctypes.c_int, u32p, u32p, ctypes.c_int64, f32p]
        L.or_centroid.restype = None
        L.or_centroid.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, f32p]
        L.or_knn_bruteforce.restype = ctypes.c_int
        L.or_knn_bruteforce.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, u32p, f32p]
        L.or_knn_f64.restype = ctypes.c_int
        L.or_knn_f64.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, u32p, ctypes.c_int64,
                                 ctypes.c_int, u32p, f64p]
        L.or_recall.restype = ctypes.c_double
        L.or_recall.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, u32p, ctypes.c_int, u32p,
                                ctypes.c_int64, ctypes.c_int, f64p]
        L.or_cal_ep.restype = ctypes.c_uint32
        L.or_cal_ep.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, u32p, ctypes.c_int,
                                ctypes.c_int, f32p, vp]
        L.or_prune.restype = ctypes.c_int
        L.or_prune.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, u32p, f32p, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_float, u32p, u32p, vp, vp, vp, vp,
                               vp, vp, vp]
        L.or_store.restype = ctypes.c_int
        L.or_store.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                               u32p, u32p, ctypes.c_uint32, u32p]
        L.or_nssg_eps.restype = None
        L.or_nssg_eps.argtypes = [ctypes.c_int64, ctypes.c_uint64, u32p]
        L.or_gnn_descent.restype = ctypes.c_int
        L.or_gnn_descent.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_uint64, u32p, f32p, vp]
        L.or_gnn_descent_from.restype = ctypes.c_int
        L.or_gnn_descent_from.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_uint64, u32p, f32p, u32p,
                                          f32p, vp]
        L.or_mlp_eval.restype = ctypes.c_int
        L.or_mlp_eval.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  f32p, ctypes.c_float, f64p, ctypes.c_double, ctypes.c_int,
                                  f64p, f64p, f64p]
        L.or_mlp_adam.restype = None
        L.or_mlp_adam.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_double, f64p, f64p,
                                  f64p, f64p]
        L.or_filter_pool.restype = ctypes.c_int
        L.or_filter_pool.argtypes = [f32p, ctypes.c_int, u32p, f32p, ctypes.c_int,
                                     ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float, u32p, f32p]
        L.or_collect_path.restype = ctypes.c_uint64
        L.or_collect_path.argtypes = [f32p, ctypes.c_int, u32p, f32p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                      f32p]
        L.or_collect_2hop.restype = None
        L.or_collect_2hop.argtypes = [f32p, ctypes.c_int, u32p, f32p, ctypes.c_int,
                                      ctypes.c_uint32, u32p, f32p]
        L.or_reverse.restype = ctypes.c_uint64
        L.or_reverse.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_float, u32p, u32p, f32p,
                                 ctypes.c_int64, u32p, u32p, f32p, u32p, u32p]
        L.or_search.restype = ctypes.c_int
        L.or_search.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, u32p, u32p, ctypes.c_int,
                                ctypes.c_uint32, f32p, ctypes.c_int64, ctypes.c_int,
                                ctypes.c_int, ctypes.c_uint64, u32p, f32p, u64p]
        L.or_num_threads.restype = ctypes.c_int
        L.or_set_threads.restype = None
        L.or_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def _u32(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def num_threads() -> int:
    return int(lib().or_num_threads())


def dist(a, b) -> np.float32:
    a, b = _f32(a), _f32(b)
    return np.float32(lib().or_dist(a, b, a.shape[0]))


def dist_pairs(X, a, b):
    X = _f32(X)
    a, b = _u32(a), _u32(b)
    out = np.empty(a.shape[0], np.float32)
    lib().or_dist_pairs(X, X.shape[1], a, b, a.shape[0], out)
    return out


def centroid(X):
    X = _f32(X)
    c = np.empty(X.shape[1], np.float32)
    lib().or_centroid(X, X.shape[0], X.shape[1], c)
    return c


def knn_bruteforce(X, K: int):
    """Exact canonical-FP32 kNN graph: (ids u32[N,K], dists f32[N,K]), rows by (d,id)."""
    X = _f32(X)
    N = X.shape[0]
    ids = np.empty((N, K), np.uint32)
    d = np.empty((N, K), np.float32)
    if lib().or_knn_bruteforce(X, N, X.shape[1], K, ids, d) != 0:
        raise ValueError("K must be < N")
    return ids, d


def knn_f64(X, queries, k: int):
    X = _f32(X)
    q = _u32(queries)
    ids = np.empty((q.shape[0], k), np.uint32)
    kth = np.empty(q.shape[0], np.float64)
    if lib().or_knn_f64(X, X.shape[0], X.shape[1], q, q.shape[0], k, ids, kth) != 0:
        raise ValueError("k must be < N")
    return ids, kth


def recall(X, G, queries, k: int, kth):
    X = _f32(X)
    G = _u32(G)
    q = _u32(queries)
    return float(lib().or_recall(X, X.shape[0], X.shape[1], G, G.shape[1], q, q.shape[0], k,
                                 np.ascontiguousarray(kth, np.float64)))


def cal_ep(X, G, TOPM: int, c=None):
    X, G = _f32(X), _u32(G)
    if c is None:
        c = centroid(X)
    n = np.zeros(1, np.uint64)
    ep = lib().or_cal_ep(X, X.shape[0], X.shape[1], G, G.shape[1], TOPM, _f32(c), _ptr(n))
    return int(ep)


def prune(X, G, GD, FD: int, TOPM: int, kind: int, fwd_rule: int, rev_rule: int,
          th: float, stages: bool = False):
    """Run the CFS rules.  Returns dict(deg, ids, ep, counters[, fwd_deg, fwd_ids, fwd_d,
    rev_cnt, rev_keys])."""
    X, G, GD = _f32(X), _u32(G), _f32(GD)
    N, K = G.shape
    deg = np.zeros(N, np.uint32)
    ids = np.full((N, FD), INV, np.uint32)
    ep = np.zeros(1, np.uint32)
    cnt = np.zeros(8, np.uint64)
    st = {}
    if stages:
        cap = FD if kind >= 2 else K
        st = dict(fwd_deg=np.zeros(N, np.uint32), fwd_ids=np.zeros((N, FD), np.uint32),
                  fwd_d=np.zeros((N, FD), np.float32), rev_cnt=np.zeros(N, np.uint32),
                  rev_keys=np.zeros((N, cap), np.uint64))
    rc = lib().or_prune(X, N, X.shape[1], G, GD, K, FD, TOPM, kind, fwd_rule, rev_rule,
                        float(th), deg, ids, _ptr(ep), _ptr(st.get("fwd_deg")),
                        _ptr(st.get("fwd_ids")), _ptr(st.get("fwd_d")),
                        _ptr(st.get("rev_cnt")), _ptr(st.get("rev_keys")), _ptr(cnt))
    if rc != 0:
        raise ValueError(f"or_prune rejected the arguments (code {rc})")
    out = dict(deg=deg, ids=ids, ep=int(ep[0]), counters=dict(
        collect_evals=int(cnt[0]), filter_evals=int(cnt[1]), reverse_evals=int(cnt[2]),
        adj_rows=int(cnt[3]), ep_evals=int(cnt[4])))
    out.update(st)
    return out


def search(X, deg, ids, ep: int, Q, L: int, k: int, seed: int):
    """Beam search (get_neighbors restatement) of queries Q over the pruned rows
    (deg, ids).  Returns (ids [nq, k], dists [nq, k], dict(evals, expansions))."""
    X, Q = _f32(X), _f32(Q)
    deg, ids = _u32(deg), _u32(ids)
    nq = Q.shape[0]
    oi = np.zeros((nq, k), np.uint32)
    od = np.zeros((nq, k), np.float32)
    cnt = np.zeros(2, np.uint64)
    rc = lib().or_search(X, X.shape[0], X.shape[1], deg, ids, ids.shape[1], ep, Q, nq, L, k,
                         seed & 0xFFFFFFFFFFFFFFFF, oi, od, cnt)
    if rc != 0:
        raise ValueError(f"or_search rejected the arguments (code {rc})")
    return oi, od, dict(evals=int(cnt[0]), expansions=int(cnt[1]))


STORE_KIND = {"NSG": 0, "Vamana": 1, "NSSG": 2, "DPG": 3, "CAGRA": 4}


def nssg_eps(N: int, seed: int):
    eps = np.zeros(10, np.uint32)
    lib().or_nssg_eps(N, seed, eps)
    return eps


def store(path: str, index_type: str, deg, ids, ep: int = 0, eps=None):
    deg, ids = _u32(deg), _u32(ids)
    eps = _u32(eps if eps is not None else np.zeros(10, np.uint32))
    rc = lib().or_store(path.encode(), STORE_KIND[index_type], ids.shape[0], ids.shape[1],
                        deg, ids, ep, eps)
    if rc != 0:
        raise IOError(f"or_store failed for {path}")


def gnn_descent(X, K: int, iters: int, seed: int, init=None):
    """GNN-Descent restatement; `init` = (G0, GD0) starts from given sorted lists instead of
    the seeded random ones.  iters == 0 returns the initial lists."""
    X = _f32(X)
    N = X.shape[0]
    G = np.empty((N, K), np.uint32)
    GD = np.empty((N, K), np.float32)
    cnt = np.zeros(4, np.uint64)
    if init is None:
        rc = lib().or_gnn_descent(X, N, X.shape[1], K, iters, seed, G, GD, _ptr(cnt))
    else:
        G0, GD0 = _u32(init[0]), _f32(init[1])
        assert G0.shape == (N, K) and GD0.shape == (N, K)
        rc = lib().or_gnn_descent_from(X, N, X.shape[1], K, iters, seed, G0, GD0, G, GD,
                                       _ptr(cnt))
    if rc != 0:
        raise ValueError("invalid GNN-Descent arguments")
    return G, GD, int(cnt[0])


def filter_pool(X, pool_ids, pool_d, u: int, FD: int, rule: int, th: float):
    """Sort + dedup the pool, run the Filter step; returns (ids, d) kept."""
    X = _f32(X)
    pid, pd = _u32(pool_ids), _f32(pool_d)
    oid = np.zeros(max(FD, 1), np.uint32)
    od = np.zeros(max(FD, 1), np.float32)
    n = lib().or_filter_pool(X, X.shape[1], pid, pd, pid.shape[0], u, FD, rule, float(th),
                             oid, od)
    return oid[:n].copy(), od[:n].copy()


def collect_path(X, G, GD, TOPM: int, ep: int, u: int):
    """Sorted, deduplicated path pool of node u (length K + 480) and the ring counter."""
    X, G, GD = _f32(X), _u32(G), _f32(GD)
    K = G.shape[1]
    pid = np.zeros(K + 480, np.uint32)
    pd = np.zeros(K + 480, np.float32)
    c = lib().or_collect_path(X, X.shape[1], G, GD, K, TOPM, ep, u, pid, pd)
    return pid, pd, int(c)


def collect_2hop(X, G, GD, u: int):
    """Sorted, deduplicated two-hop pool of node u (length K + 480)."""
    X, G, GD = _f32(X), _u32(G), _f32(GD)
    K = G.shape[1]
    pid = np.zeros(K + 480, np.uint32)
    pd = np.zeros(K + 480, np.float32)
    lib().or_collect_2hop(X, X.shape[1], G, GD, K, u, pid, pd)
    return pid, pd


def reverse(X, K: int, FD: int, rule: int, th: float, fwd_deg, fwd_ids, fwd_d, pushes):
    """The reverse phase on given forward rows (fwd_deg [N], fwd_ids/fwd_d [N, FD]) and Store
    pushes, a sequence of (dst, src, d): final rows (deg [N], ids [N, FD]) and the rule
    evaluations."""
    X = _f32(X)
    N = X.shape[0]
    P = list(pushes)
    dst = _u32([p[0] for p in P]) if P else np.zeros(0, np.uint32)
    src = _u32([p[1] for p in P]) if P else np.zeros(0, np.uint32)
    dd = _f32([p[2] for p in P]) if P else np.zeros(0, np.float32)
    deg = np.zeros(N, np.uint32)
    ids = np.full((N, FD), INV, np.uint32)
    ev = lib().or_reverse(X, N, X.shape[1], K, FD, rule, float(th), _u32(fwd_deg),
                          _u32(fwd_ids), _f32(fwd_d), len(P), dst, src, dd, deg, ids)
    return deg, ids, int(ev)


MLP_HID, MLP_CODE = 256, 32


def mlp_nparams(D: int) -> int:
    return MLP_HID * (D + 1) + MLP_CODE * (MLP_HID + 1)


def mlp_eval(X, B: int, P: int, mu, scale: float, params, margin: float = 1.0,
             round_bf16: bool = True):
    """Encoder forward/backward in FP64 (optionally with the HIP path's BF16 roundings).
    Returns (loss, grad[nparams], codes[M, 32])."""
    X = _f32(X)
    M, D = X.shape
    prm = np.ascontiguousarray(params, np.float64)
    assert prm.shape == (mlp_nparams(D),)
    loss = np.zeros(1, np.float64)
    grad = np.zeros_like(prm)
    codes = np.zeros((M, MLP_CODE), np.float64)
    if lib().or_mlp_eval(X, M, D, B, P, _f32(mu), float(scale), prm, float(margin),
                         1 if round_bf16 else 0, loss, grad, codes) != 0:
        raise ValueError("bad encoder batch shape")
    return float(loss[0]), grad, codes


def mlp_adam(t: int, lr: float, g, p, m, v):
    """One FP64 Adam step in place on p, m, v (float64 arrays)."""
    lib().or_mlp_adam(p.shape[0], t, lr, np.ascontiguousarray(g, np.float64), p, m, v)
