"""tagore_amd -- MI355X-native GNN-Descent kNN construction and CFS pruning.

The product is the HIP library libtagore.so (C ABI in include/tagore.h) plus the drop-in
pybind11 module `Tagore` built next to it.  This package loads both:

    from tagore_amd import Tagore          # the reference's Python surface
    from tagore_amd import State           # ctypes wrapper of the C ABI (tests, bench)

Nothing here falls back to CPU code: if the HIP library has not been built, loading raises.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import sysconfig

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtagore.so")
INCLUDE = os.path.join(os.path.dirname(_HERE), "include", "tagore.h")

TG_OK, TG_ERR_ARG, TG_ERR_IO, TG_ERR_HIP, TG_ERR_STATE, TG_ERR_CAPACITY = range(6)
PREC = {"exact": 0, "fp32": 1, "bf16": 2, "fp8": 3}
KIND = {"path": 0, "2-hop": 1, "cagra": 2, "dpg": 3}
RULE = {"dist": 0, "angle": 1, "rank": 2}

# index types of Pruning(): (kind, fwd_rule, rev_rule, fixed threshold or None)
INDEX_TYPES = {
    "NSG": (0, 0, 0, 1.0),
    "Vamana": (0, 0, 0, None),
    "NSSG": (1, 1, 1, None),
    "CAGRA": (2, 0, 2, None),
    "DPG": (3, 0, 2, None),
}


class TagoreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[tagore status {code}] {msg}")
        self.code = code


INIT = {"random": 0, "learned": 1, "given": 2}
MLP_HID, MLP_CODE = 256, 32


class KnnParams(ctypes.Structure):
    _fields_ = [("K", ctypes.c_uint32), ("iters", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("precision", ctypes.c_int32), ("init", ctypes.c_int32),
                ("sample", ctypes.c_uint32), ("epochs", ctypes.c_uint32),
                ("tile", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class PruneParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("fwd_rule", ctypes.c_int32),
                ("rev_rule", ctypes.c_int32), ("FD", ctypes.c_uint32),
                ("TOPM", ctypes.c_uint32), ("thre", ctypes.c_float),
                ("precision", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("knn_ms", ctypes.c_double), ("prune_ms", ctypes.c_double),
                ("knn_kernel_ms", ctypes.c_double * 8), ("prune_kernel_ms", ctypes.c_double * 4),
                ("knn_evals", ctypes.c_uint64), ("knn_rescored", ctypes.c_uint64),
                ("prune_counters", ctypes.c_uint64 * 15),
                ("learn_ms", ctypes.c_double * 5), ("learn_loss", ctypes.c_float * 2),
                ("learn_buckets", ctypes.c_uint32), ("learn_steps", ctypes.c_uint32)]


class LargeParams(ctypes.Structure):
    _fields_ = [("devicelist", ctypes.POINTER(ctypes.c_int32)), ("ndevices", ctypes.c_uint32),
                ("DIM", ctypes.c_uint32), ("K", ctypes.c_uint32), ("cluster_num", ctypes.c_uint32),
                ("max_points", ctypes.c_uint32), ("file_prefix", ctypes.c_char_p),
                ("local_index_path_prefix", ctypes.c_char_p), ("index_store_path", ctypes.c_char_p),
                ("max_degree", ctypes.c_uint32), ("buffer_size", ctypes.c_uint32),
                ("all_node_num", ctypes.c_uint32), ("centers_path", ctypes.c_char_p),
                ("thre", ctypes.c_float), ("FINAL_DEGREE", ctypes.c_uint32),
                ("TOPM", ctypes.c_uint32), ("iters", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("precision", ctypes.c_int32), ("batch_rows", ctypes.c_uint64),
                ("verbose", ctypes.c_int32), ("tune", ctypes.c_int32), ("tile", ctypes.c_uint32)]


class LargeStats(ctypes.Structure):
    _fields_ = [("clusters", ctypes.c_uint32), ("batches", ctypes.c_uint32),
                ("build_ms", ctypes.c_double), ("merge_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("precision", ctypes.c_int32),
                ("tile", ctypes.c_uint32), ("device_ms", ctypes.c_double)]


class SearchParams(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("k", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("ep", ctypes.c_uint32)]


class SearchStats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double), ("evals", ctypes.c_uint64),
                ("expansions", ctypes.c_uint64)]


class TuneResult(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("tile", ctypes.c_uint32),
                ("ncand", ctypes.c_uint32), ("cand_precision", ctypes.c_int32 * 8),
                ("cand_tile", ctypes.c_uint32 * 8), ("cand_ms", ctypes.c_double * 8)]


KNN_SECTIONS = ["init", "sample", "rsample", "localjoin", "merge", "reverse", "phase2", "final"]
PRUNE_SECTIONS = ["ep", "forward", "reverse", "group"]

_lib = None


def exported_symbols():
    """Function names declared in include/tagore.h."""
    import re

    txt = open(INCLUDE).read()
    return sorted(set(re.findall(r"\b(tg_[a-z_0-9]+)\s*\(", txt)))


def lib():
    """The loaded libtagore.so with argument types set.  Raises if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tagore_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P, vp, c = ctypes.POINTER, ctypes.c_void_p, ctypes
    L.tg_last_error.restype = c.c_char_p
    L.tg_create.argtypes = [c.c_uint32, c.c_uint32, c.c_int, P(vp)]
    L.tg_free.argtypes = [vp]
    L.tg_set_data.argtypes = [vp, vp]
    L.tg_gen_data.argtypes = [vp, c.c_int32, c.c_uint32, c.c_uint64, c.c_float]
    L.tg_get_data.argtypes = [vp, vp, c.c_uint32, c.c_uint32]
    L.tg_read_fvecs.argtypes = [c.c_char_p, P(vp), P(c.c_uint32), P(c.c_uint32)]
    L.tg_read_bvecs.argtypes = [c.c_char_p, P(vp), P(c.c_uint32), P(c.c_uint32)]
    L.tg_read_fbin.argtypes = [c.c_char_p, P(vp), P(c.c_uint32), P(c.c_uint32)]
    L.tg_read_index.argtypes = [c.c_char_p, c.c_char_p, c.c_uint32, P(c.c_uint32), P(vp),
                                P(vp), vp, P(c.c_uint32)]
    L.tg_set_pruned.argtypes = [vp, c.c_uint32, vp, vp, c.c_uint32]
    L.tg_host_free.argtypes = [vp]
    L.tg_host_free.restype = None
    L.tg_knn_build.argtypes = [vp, P(KnnParams)]
    L.tg_set_knn.argtypes = [vp, c.c_uint32, vp, vp]
    L.tg_get_knn.argtypes = [vp, vp, vp]
    L.tg_prune.argtypes = [vp, P(PruneParams)]
    L.tg_get_pruned.argtypes = [vp, vp, vp, P(c.c_uint32)]
    L.tg_get_forward.argtypes = [vp, vp, vp, vp]
    L.tg_pruned_width.argtypes = [vp, P(c.c_uint32)]
    L.tg_store_index.argtypes = [vp, c.c_char_p, c.c_char_p, c.c_uint64]
    L.tg_recall.argtypes = [vp, c.c_uint32, c.c_uint64, c.c_uint32, P(c.c_double)]
    L.tg_knn_dist_check.argtypes = [vp, c.c_uint32, c.c_uint64, P(c.c_double)]
    L.tg_buffer.argtypes = [vp, c.c_int32, P(vp), P(c.c_size_t)]
    L.tg_get_stats.argtypes = [vp, P(Stats)]
    L.tg_set_profiling.argtypes = [vp, c.c_int]
    L.tg_get_codes.argtypes = [vp, vp]
    L.tg_set_segments.argtypes = [vp, c.c_uint32, vp, vp]
    L.tg_large_index.argtypes = [P(LargeParams), P(LargeStats)]
    L.tg_autotune_knn.argtypes = [vp, c.c_uint32, c.c_uint64, P(TuneResult)]
    L.tg_get_entry_points.argtypes = [vp, vp]
    L.tg_mlp_eval.argtypes = [c.c_int, c.c_uint32, c.c_uint32, c.c_uint32, vp, vp, c.c_float, vp,
                              c.c_float, P(c.c_float), vp, vp]
    L.tg_mlp_adam.argtypes = [c.c_int, c.c_uint32, c.c_uint32, c.c_float, vp, vp, vp, vp]
    L.tg_selftest_sort.argtypes = [c.c_int, c.c_uint32, c.c_uint32, c.c_int32, vp, vp]
    L.tg_search.argtypes = [vp, c.c_uint32, vp, P(SearchParams), vp, vp, P(SearchStats)]
    L.tg_search_noalloc.argtypes = [vp, c.c_uint32, vp, P(SearchParams), vp, vp, vp,
                                    P(SearchStats)]
    L.tg_search_recall.argtypes = [vp, c.c_uint32, vp, c.c_uint32, vp, P(c.c_double)]
    for name in exported_symbols():
        f = getattr(L, name)
        if name not in ("tg_last_error", "tg_host_free"):
            f.restype = c.c_int
    _lib = L
    return L


def _check(rc):
    if rc != TG_OK:
        raise TagoreError(rc, lib().tg_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class State:
    """Owner of one device handle (tg_state) -- N vectors of dimension D on one GPU."""

    def __init__(self, N: int, D: int, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(lib().tg_create(N, D, device, ctypes.byref(self._h)))
        self.N, self.D = N, D
        self.K = None
        self.FD = None

    def close(self):
        if self._h:
            lib().tg_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # data
    def set_data(self, X):
        X = np.ascontiguousarray(X, np.float32)
        assert X.shape == (self.N, self.D)
        _check(lib().tg_set_data(self._h, _ptr(X)))

    def gen_data(self, kind: int, ncomp: int, seed: int, sigma: float = 20.0):
        _check(lib().tg_gen_data(self._h, kind, ncomp, seed, sigma))

    def get_data(self, first: int = 0, count: int | None = None):
        count = self.N - first if count is None else count
        out = np.empty((count, self.D), np.float32)
        _check(lib().tg_get_data(self._h, _ptr(out), first, count))
        return out

    # kNN
    def knn_build(self, K: int, iters: int, seed: int = 0x7A600000, precision: str = "exact",
                  init: str = "random", sample: int = 0, epochs: int = 0, tile: int = 0,
                  graph: bool = False, graph_iters: bool = False):
        """GNN-Descent.  graph: capture the encoder's training loop as a HIP graph;
        graph_iters: capture the iteration sequence as a HIP graph (replayed by later builds
        with the same launch arguments)."""
        p = KnnParams(K, iters, seed, PREC[precision], INIT[init], sample, epochs, tile,
                      (1 if graph else 0) | (2 if graph_iters else 0))
        _check(lib().tg_knn_build(self._h, ctypes.byref(p)))
        self.K = K

    def set_segments(self, starts, keys=None):
        """Rows [starts[c], starts[c+1]) become independent data sets seeded by keys[c]
        (default c); None: one data set."""
        if starts is None:
            _check(lib().tg_set_segments(self._h, 0, None, None))
            self.nseg = 0
            return
        st = np.ascontiguousarray(starts, np.uint32)
        ky = None if keys is None else np.ascontiguousarray(keys, np.uint32)
        _check(lib().tg_set_segments(self._h, st.shape[0] - 1, _ptr(st), _ptr(ky)))
        self.nseg = st.shape[0] - 1

    def entry_points(self):
        out = np.empty(max(1, getattr(self, "nseg", 0)), np.uint32)
        _check(lib().tg_get_entry_points(self._h, _ptr(out)))
        return out

    def autotune(self, K: int, seed: int = 0x7A600000):
        """Fastest (precision, tile) for a kNN build of this data, and every candidate's ms."""
        r = TuneResult()
        _check(lib().tg_autotune_knn(self._h, K, seed, ctypes.byref(r)))
        inv = {v: k for k, v in PREC.items()}
        cands = [(inv[r.cand_precision[i]], int(r.cand_tile[i]), float(r.cand_ms[i]))
                 for i in range(r.ncand)]
        return (inv[r.precision], int(r.tile)), cands

    def codes(self):
        out = np.empty((self.N, MLP_CODE), np.float32)
        _check(lib().tg_get_codes(self._h, _ptr(out)))
        return out

    def set_knn(self, ids, dists):
        ids = np.ascontiguousarray(ids, np.uint32)
        dists = np.ascontiguousarray(dists, np.float32)
        _check(lib().tg_set_knn(self._h, ids.shape[1], _ptr(ids), _ptr(dists)))
        self.K = ids.shape[1]

    def get_knn(self):
        ids = np.empty((self.N, self.K), np.uint32)
        d = np.empty((self.N, self.K), np.float32)
        _check(lib().tg_get_knn(self._h, _ptr(ids), _ptr(d)))
        return ids, d

    # pruning
    def prune(self, kind, fwd_rule, rev_rule, FD: int, TOPM: int = 64, thre: float = 1.0,
              precision: str = "fp32"):
        kind = KIND[kind] if isinstance(kind, str) else kind
        fwd_rule = RULE[fwd_rule] if isinstance(fwd_rule, str) else fwd_rule
        rev_rule = RULE[rev_rule] if isinstance(rev_rule, str) else rev_rule
        p = PruneParams(kind, fwd_rule, rev_rule, FD, TOPM, thre, PREC[precision])
        _check(lib().tg_prune(self._h, ctypes.byref(p)))
        self.FD = FD

    def get_pruned(self):
        deg = np.empty(self.N, np.uint32)
        ids = np.empty((self.N, self.FD), np.uint32)
        ep = ctypes.c_uint32()
        _check(lib().tg_get_pruned(self._h, _ptr(deg), _ptr(ids), ctypes.byref(ep)))
        return deg, ids, int(ep.value)

    def get_forward(self):
        deg = np.empty(self.N, np.uint32)
        ids = np.empty((self.N, self.FD), np.uint32)
        d = np.empty((self.N, self.FD), np.float32)
        _check(lib().tg_get_forward(self._h, _ptr(deg), _ptr(ids), _ptr(d)))
        return deg, ids, d

    def store_index(self, path: str, index_type: str, nssg_seed: int = 0x5EED):
        _check(lib().tg_store_index(self._h, path.encode(), index_type.encode(), nssg_seed))

    def set_pruned(self, deg, ids, ep: int = 0):
        """Installs a pruned graph (deg [N], ids [N, width]) as the current one."""
        deg = np.ascontiguousarray(deg, np.uint32)
        ids = np.ascontiguousarray(ids, np.uint32)
        _check(lib().tg_set_pruned(self._h, ids.shape[1], _ptr(deg), _ptr(ids), ep))
        self.FD = ids.shape[1]

    def load_index(self, path: str, index_type: str):
        """Reads an index file (tg_read_index) and installs it; returns its entry points."""
        deg, ids, eps = read_index(path, index_type, self.N)
        self.set_pruned(deg, ids, int(eps[0]) if eps.size else 0)
        return eps

    def search(self, queries, k: int = 10, L: int = 64, seed: int = 0x5EA4C4, ep=None):
        """Beam search of host queries [nq, D] over the last pruned graph.  Returns
        (ids [nq, k], dists [nq, k], dict(ms, evals, expansions))."""
        Q = np.ascontiguousarray(queries, np.float32)
        nq = Q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        p = SearchParams(L, k, seed & 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF if ep is None else ep)
        st = SearchStats()
        _check(lib().tg_search(self._h, nq, _ptr(Q), ctypes.byref(p), _ptr(ids), _ptr(d),
                               ctypes.byref(st)))
        return ids, d, dict(ms=st.ms, evals=int(st.evals), expansions=int(st.expansions))

    def search_recall(self, queries, ids, k: int = 10):
        """recall@k of result rows ids [nq, >=k] for host queries (FP64 brute force)."""
        Q = np.ascontiguousarray(queries, np.float32)
        I = np.ascontiguousarray(np.asarray(ids)[:, :k], np.uint32)
        r = ctypes.c_double()
        _check(lib().tg_search_recall(self._h, Q.shape[0], _ptr(Q), k, _ptr(I), ctypes.byref(r)))
        return r.value

    def recall(self, nq: int, k: int = 10, seed: int = 0x5A11):
        r = ctypes.c_double()
        _check(lib().tg_recall(self._h, nq, seed, k, ctypes.byref(r)))
        return r.value

    def dist_check(self, nq: int, seed: int = 0x5A11):
        """Largest relative difference of the stored FP32 kNN distances of nq judged nodes
        (those of recall() with the same seed) from their FP64 distances."""
        r = ctypes.c_double()
        _check(lib().tg_knn_dist_check(self._h, nq, seed, ctypes.byref(r)))
        return r.value

    def stats(self):
        s = Stats()
        _check(lib().tg_get_stats(self._h, ctypes.byref(s)))
        return dict(knn_ms=s.knn_ms, prune_ms=s.prune_ms,
                    knn_kernel_ms=dict(zip(KNN_SECTIONS, list(s.knn_kernel_ms))),
                    prune_kernel_ms=dict(zip(PRUNE_SECTIONS, list(s.prune_kernel_ms))),
                    knn_evals=int(s.knn_evals), knn_rescored=int(s.knn_rescored),
                    prune_counters=dict(zip(["collect", "filter", "reverse", "adj_rows", "ep",
                                             "rescored", "collect_cycles", "filter_cycles",
                                             "pool_cycles", "eval_cycles", "beam_pre_cycles",
                                             "beam_sort_cycles", "beam_merge_cycles",
                                             "beam_hash_cycles", "pool_sort_cycles"],
                                            list(s.prune_counters))),
                    learn_ms=dict(zip(["sample_graph", "train", "encode", "buckets", "seed"],
                                      list(s.learn_ms))),
                    learn_loss=list(s.learn_loss), learn_buckets=int(s.learn_buckets),
                    learn_steps=int(s.learn_steps))

    def set_profiling(self, on):
        """False/0: device windows only; True/1: per-kernel event times; 2: also the prune
        forward kernel's shader-clock phase counters (adds work to that kernel)."""
        _check(lib().tg_set_profiling(self._h, int(on)))

    def buffer(self, idx: int):
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _check(lib().tg_buffer(self._h, idx, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value


def large_index(DIM, K, cluster_num, max_points, file_prefix, local_index_path_prefix,
                index_store_path, max_degree, buffer_size, all_node_num, centers_path, thre,
                FINAL_DEGREE, TOPM, device=0, seed=0x7A600000, precision="bf16", batch_rows=0,
                iters=16, verbose=False, tune=False, tile=0):
    """tg_large_index through ctypes; returns the stats as a dict."""
    dev = (ctypes.c_int32 * 1)(device)
    p = LargeParams(dev, 1, DIM, K, cluster_num, max_points, file_prefix.encode(),
                    local_index_path_prefix.encode(), index_store_path.encode(), max_degree,
                    buffer_size, all_node_num, centers_path.encode(), thre, FINAL_DEGREE, TOPM,
                    iters, seed, PREC[precision], batch_rows, 1 if verbose else 0,
                    1 if tune else 0, tile)
    st = LargeStats()
    _check(lib().tg_large_index(ctypes.byref(p), ctypes.byref(st)))
    inv = {v: k for k, v in PREC.items()}
    return dict(clusters=st.clusters, batches=st.batches, build_ms=st.build_ms,
                merge_ms=st.merge_ms, total_ms=st.total_ms, precision=inv[st.precision],
                tile=st.tile, device_ms=st.device_ms)


def segment_seed(seed: int, c: int) -> int:
    """Seed of segment c of a segmented handle."""
    return (seed ^ ((c + 1) * 0xA0761D6478BD642F)) & 0xFFFFFFFFFFFFFFFF


def mlp_nparams(D: int) -> int:
    return MLP_HID * (D + 1) + MLP_CODE * (MLP_HID + 1)


def mlp_eval(X, B: int, P: int, mu, scale: float, params, margin: float = 1.0, device: int = 0):
    """Encoder forward/backward on the GPU: (loss, grad[nparams], codes[M, 32])."""
    X = np.ascontiguousarray(X, np.float32)
    M, D = X.shape
    assert M == B * (1 + P)
    prm = np.ascontiguousarray(params, np.float32)
    assert prm.shape == (mlp_nparams(D),)
    mu = np.ascontiguousarray(mu, np.float32)
    grad = np.empty_like(prm)
    codes = np.empty((M, MLP_CODE), np.float32)
    loss = ctypes.c_float()
    _check(lib().tg_mlp_eval(device, D, B, P, _ptr(X), _ptr(mu), scale, _ptr(prm), margin,
                             ctypes.byref(loss), _ptr(grad), _ptr(codes)))
    return loss.value, grad, codes


def mlp_adam(D: int, t: int, lr: float, grad, params, m, v, device: int = 0):
    """One Adam step on the GPU; returns the new (params, m, v) as float32 arrays."""
    g = np.ascontiguousarray(grad, np.float32)
    p = np.array(params, np.float32)
    m = np.array(m, np.float32)
    v = np.array(v, np.float32)
    _check(lib().tg_mlp_adam(device, D, t, lr, _ptr(g), _ptr(p), _ptr(m), _ptr(v)))
    return p, m, v


def selftest_sort(keys, npl: int, mode: int = 0, device: int = 0):
    """Runs the prune kernels' register sorts on keys [groups, 64*npl] (u64) on the GPU."""
    k = np.ascontiguousarray(keys, np.uint64)
    assert k.ndim == 2 and k.shape[1] == 64 * npl
    out = np.empty_like(k)
    _check(lib().tg_selftest_sort(device, npl, k.shape[0], mode, _ptr(k), _ptr(out)))
    return out


def _read_rows(fn, path: str):
    rows = ctypes.c_void_p()
    n, d = ctypes.c_uint32(), ctypes.c_uint32()
    _check(fn(path.encode(), ctypes.byref(rows), ctypes.byref(n), ctypes.byref(d)))
    try:
        buf = (ctypes.c_float * (n.value * d.value)).from_address(rows.value)
        return np.array(buf, np.float32).reshape(n.value, d.value)
    finally:
        lib().tg_host_free(rows)


def read_fvecs(path: str):
    """.fvecs vectors as float32 [N, D]."""
    return _read_rows(lib().tg_read_fvecs, path)


def read_bvecs(path: str):
    """.bvecs (u8) vectors as float32 [N, D]."""
    return _read_rows(lib().tg_read_bvecs, path)


def read_fbin(path: str):
    """.fbin (u32 N, u32 D, f32 rows) vectors as float32 [N, D]."""
    return _read_rows(lib().tg_read_fbin, path)


def read_index(path: str, index_type: str, N: int):
    """An index file in one of the five layouts (NSG, Vamana, NSSG, DPG, CAGRA).  Returns
    (deg [N], ids [N, width] INV-padded, entry points)."""
    w = ctypes.c_uint32()
    dp, ip = ctypes.c_void_p(), ctypes.c_void_p()
    eps = np.zeros(10, np.uint32)
    ne = ctypes.c_uint32()
    _check(lib().tg_read_index(path.encode(), index_type.encode(), N, ctypes.byref(w),
                               ctypes.byref(dp), ctypes.byref(ip), _ptr(eps), ctypes.byref(ne)))
    try:
        deg = np.array((ctypes.c_uint32 * N).from_address(dp.value), np.uint32)
        ids = np.array((ctypes.c_uint32 * (N * w.value)).from_address(ip.value),
                       np.uint32).reshape(N, w.value)
    finally:
        lib().tg_host_free(dp)
        lib().tg_host_free(ip)
    return deg, ids, eps[:ne.value].copy()


def nssg_entry_points(N: int, seed: int):
    """The 10 NSSG entry ids tg_store_index writes (SplitMix64 Fisher-Yates prefix)."""
    perm = list(range(N))
    x = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for i in range(min(10, N)):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        j = i + z % (N - i)
        perm[i], perm[j] = perm[j], perm[i]
    return (perm[:10] + [0] * 10)[:10]


def _load_module():
    path = os.path.join(_HERE, "Tagore" + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m tagore_amd.build`")
    lib()  # the module links libtagore.so; load it first so a missing build fails here
    spec = importlib.util.spec_from_file_location("Tagore", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def __getattr__(name):
    if name == "Tagore":
        mod = _load_module()
        globals()["Tagore"] = mod
        return mod
    raise AttributeError(name)
