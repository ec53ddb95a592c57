"""Host-side checks of the C ABI and the drop-in module (no GPU needed)."""
import ctypes
import os

import numpy as np
import pytest

import tagore_amd
from tagore_amd import Tagore, TagoreError


def test_library_exports_every_declared_symbol():
    L = tagore_amd.lib()
    names = tagore_amd.exported_symbols()
    assert len(names) >= 19
    for n in names:
        assert hasattr(L, n), n


def test_module_surface_matches_reference_names():
    for n in ("GNN_descent", "Pruning", "Pruning_with_CFS"):
        assert callable(getattr(Tagore, n))
    assert Tagore.__doc__ == "Tagore GPU Indexing"
    doc = Tagore.Pruning.__doc__
    for kw in ("K", "POINTS", "DIM", "FINAL_DEGREE", "TOPM", "ptrs", "index_type", "thre",
               "index_path"):
        assert kw in doc


def test_errors_are_status_codes_not_exits(tmp_path):
    L = tagore_amd.lib()
    h = ctypes.c_void_p()
    # invalid dimension is rejected before any device call
    assert L.tg_create(100, 3, 0, ctypes.byref(h)) == tagore_amd.TG_ERR_ARG
    assert b"multiple of 4" in L.tg_last_error()
    # missing file
    rows, n, d = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
    rc = L.tg_read_fvecs(str(tmp_path / "nope.fvecs").encode(), ctypes.byref(rows),
                         ctypes.byref(n), ctypes.byref(d))
    assert rc == tagore_amd.TG_ERR_IO


def _write_fvecs(path, X):
    X = np.asarray(X, np.float32)
    rec = np.empty((X.shape[0], X.shape[1] + 1), np.float32)
    rec[:, 0] = np.frombuffer(np.uint32(X.shape[1]).tobytes() * 1, np.float32)[0]
    rec[:, 1:] = X
    rec.tofile(path)


def test_fvecs_reader_roundtrip(tmp_path):
    X = np.random.default_rng(1).random((37, 12)).astype(np.float32)
    p = tmp_path / "x.fvecs"
    _write_fvecs(p, X)
    L = tagore_amd.lib()
    rows, n, d = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
    assert L.tg_read_fvecs(str(p).encode(), ctypes.byref(rows), ctypes.byref(n),
                           ctypes.byref(d)) == 0
    assert (n.value, d.value) == (37, 12)
    got = np.ctypeslib.as_array(ctypes.cast(rows, ctypes.POINTER(ctypes.c_float)),
                                shape=(37, 12)).copy()
    L.tg_host_free(rows)
    assert np.array_equal(got, X)
    # truncated file
    b = p.read_bytes()
    p.write_bytes(b[:-3])
    assert L.tg_read_fvecs(str(p).encode(), ctypes.byref(rows), ctypes.byref(n),
                           ctypes.byref(d)) == tagore_amd.TG_ERR_IO


def test_dropin_rejects_bad_inputs_before_the_gpu(tmp_path):
    p = tmp_path / "x.fvecs"
    _write_fvecs(p, np.zeros((10, 8), np.float32))
    with pytest.raises(ValueError):
        Tagore.GNN_descent(4, 11, 8, 2, str(p))  # POINTS does not match the file
    with pytest.raises(ValueError):
        Tagore.Pruning(4, 10, 8, 4, 64, (1, 2, 3), "NSG")  # not our capsules


def test_nssg_entry_points_match_oracle():
    import oracle

    for N, seed in ((5, 1), (1000, 0x5EED), (12, 99)):
        assert tagore_amd.nssg_entry_points(N, seed)[:min(10, N)] == \
            oracle.nssg_eps(N, seed).tolist()[:min(10, N)]
