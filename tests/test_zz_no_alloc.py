"""The allocation-free search entry point (tg_search_noalloc) against tg_search, through the
C ABI: same ids, distances and evaluation counts, bit for bit, on two calls in a row (the
second with fewer queries and a smaller pool), with device buffers the test allocates, on
the handle's stream and on a stream of the caller's.  At 200k vectors the static scratch
holds 1,161 workgroups' bitmaps, fewer than the first call's 3,000 queries, so the
workgroups loop over queries there.

The device buffers come from the HIP runtime libtagore.so itself links (the process must not
load a second one, so no torch here)."""
import ctypes

import numpy as np
import pytest

import tagore_amd
from tagore_amd import SearchParams, SearchStats, State

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _say(msg):
    print(msg, flush=True)


def _hip():
    tagore_amd.lib()  # loads libamdhip64.so.7 as libtagore.so's dependency
    h = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipStreamDestroy.argtypes = [vp]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _Dev:
    """A device buffer of `nbytes` from hipMalloc, freed on exit."""

    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.hip.hipFree(self.p)


def _noalloc(s, hip, Q, k, L, seed, stream=None):
    Q = np.ascontiguousarray(Q, np.float32)
    nq = Q.shape[0]
    ids = np.full((nq, k), 0xFFFFFFFF, np.uint32)
    d = np.full((nq, k), -1.0, np.float32)
    with _Dev(hip, Q.nbytes) as q_dev, _Dev(hip, ids.nbytes) as i_dev, \
            _Dev(hip, d.nbytes) as d_dev:
        assert hip.hipMemcpy(q_dev.p, Q.ctypes.data, Q.nbytes, H2D) == 0
        assert hip.hipMemcpy(i_dev.p, ids.ctypes.data, ids.nbytes, H2D) == 0
        p = SearchParams(L, k, seed, 0xFFFFFFFF)
        st = SearchStats()
        rc = tagore_amd.lib().tg_search_noalloc(s._h, nq, q_dev.p, ctypes.byref(p), i_dev.p,
                                                d_dev.p, stream, ctypes.byref(st))
        assert rc == 0, tagore_amd.lib().tg_last_error()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(ids.ctypes.data, i_dev.p, ids.nbytes, D2H) == 0
        assert hip.hipMemcpy(d.ctypes.data, d_dev.p, d.nbytes, D2H) == 0
    return ids, d, dict(evals=int(st.evals), expansions=int(st.expansions))


def test_search_noalloc_matches_search():
    hip = _hip()
    N, D, nq = 200_000, 128, 3000
    with State(N + nq, D) as g:
        g.gen_data(0, 256, 0x0A11, 20.0)
        Xall = g.get_data()
    X, Q = Xall[:N], Xall[N:]
    side = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(side)) == 0
    try:
        with State(N, D) as s:
            s.set_data(X)
            s.knn_build(32, 4, 0x0A11, "bf16")
            s.prune("path", "dist", "dist", 32, 64, 1.0)
            for n, k, L, stream in ((nq, 10, 64, None), (400, 5, 32, side)):
                _say(f"no-alloc search: {n} queries, k={k}, L={L}, "
                     f"{'caller' if stream else 'handle'} stream")
                ri, rd, rs = s.search(Q[:n], k=k, L=L, seed=0x77)
                gi, gd, gs = _noalloc(s, hip, Q[:n], k, L, 0x77, stream)
                assert np.array_equal(gi, ri), f"ids differ in {(gi != ri).any(1).sum()} rows"
                assert np.array_equal(gd, rd), "distances differ"
                assert gs["evals"] == rs["evals"] and gs["expansions"] == rs["expansions"]
    finally:
        hip.hipStreamDestroy(side)


def test_search_noalloc_rejects_bad_arguments():
    hip = _hip()
    N, D = 3000, 32
    with State(N, D) as s, _Dev(hip, 64 * D * 4) as buf:
        s.gen_data(1, 8, 3, 1.0)
        L = tagore_amd.lib()
        ptr = buf.p
        p = SearchParams(16, 4, 1, 0xFFFFFFFF)
        # no pruned graph yet
        assert L.tg_search_noalloc(s._h, 4, ptr, ctypes.byref(p), ptr, ptr, None,
                                   None) == tagore_amd.TG_ERR_STATE
        s.knn_build(16, 2, 3, "exact")
        s.prune("path", "dist", "dist", 16, 32, 1.0)
        assert L.tg_search_noalloc(s._h, 4, None, ctypes.byref(p), ptr, ptr, None,
                                   None) == tagore_amd.TG_ERR_ARG
        bad = SearchParams(300, 4, 1, 0xFFFFFFFF)
        assert L.tg_search_noalloc(s._h, 4, ptr, ctypes.byref(bad), ptr, ptr, None,
                                   None) == tagore_amd.TG_ERR_ARG
        bad = SearchParams(16, 4, 1, N)
        assert L.tg_search_noalloc(s._h, 4, ptr, ctypes.byref(bad), ptr, ptr, None,
                                   None) == tagore_amd.TG_ERR_ARG
