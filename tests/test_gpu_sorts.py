"""Register sorts of the prune kernels (DPP / swizzle / bpermute lane exchanges) against
numpy on the CPU: full sorts of 64, 128 and 512 keys, the 64-key bitonic merge, and the
top-64 beam update, with duplicate keys and +inf padding keys mixed in."""
import numpy as np
import pytest

import tagore_amd

pytestmark = pytest.mark.gpu

INF_KEY = np.uint64(0xFFFFFFFFFFFFFFFF)


def _keys(rng, groups, n):
    # (float bits << 32 | id) with small ranges so that equal distances and equal keys occur
    d = rng.integers(0, 40, (groups, n)).astype(np.float32)
    ids = rng.integers(0, 50, (groups, n)).astype(np.uint64)
    k = (d.view(np.uint32).astype(np.uint64) << np.uint64(32)) | ids
    k[rng.random((groups, n)) < 0.1] = INF_KEY
    return k


@pytest.mark.parametrize("npl", [1, 2, 8])
def test_register_bitonic_sorts(npl):
    rng = np.random.default_rng(npl)
    k = _keys(rng, 300, 64 * npl)
    out = tagore_amd.selftest_sort(k, npl, 0)
    assert np.array_equal(out, np.sort(k, axis=1))


def test_register_bitonic_sorts_full_range_keys():
    rng = np.random.default_rng(11)
    k = rng.integers(0, 2**63, (200, 128), dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, (200, 128), dtype=np.uint64)
    out = tagore_amd.selftest_sort(k, 2, 0)
    assert np.array_equal(out, np.sort(k, axis=1))


def test_bitonic_merge64():
    rng = np.random.default_rng(5)
    k = _keys(rng, 400, 64)
    split = rng.integers(0, 65, 400)
    bit = np.empty_like(k)
    for g in range(400):
        s = np.sort(k[g])
        a, b = s[:split[g]], s[split[g]:]
        bit[g] = np.concatenate([a, b[::-1]])  # ascending then descending
    out = tagore_amd.selftest_sort(bit, 1, 1)
    assert np.array_equal(out, np.sort(k, axis=1))


def test_top64_beam_update():
    rng = np.random.default_rng(9)
    a = np.sort(_keys(rng, 300, 64), axis=1)
    b = np.sort(_keys(rng, 300, 64), axis=1)
    out = tagore_amd.selftest_sort(np.concatenate([a, b], axis=1), 2, 2)
    want = np.sort(np.concatenate([a, b], axis=1), axis=1)[:, :64]
    assert np.array_equal(out[:, :64], want)
    assert np.array_equal(out[:, 64:], b)
