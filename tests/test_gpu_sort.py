"""The broadphase's device sort (k_sort.hip, a stable LSD radix sort of
64-bit keys) against numpy's stable sort: the keys, and the values carried
in the same permutation, on sizes around its 4096-key tile, all-equal and
few-distinct keys (stability), full-range 64-bit keys, and the
broadphase's own key shapes."""
import numpy as np
import pytest

from cubez_amd import World

pytestmark = pytest.mark.gpu


def _check(w, keys, with_vals=True):
    vals = np.arange(len(keys), dtype=np.int32)[::-1].copy() if with_vals else None
    ko, vo = w.selftest_sort(keys, vals)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(ko, keys[perm])
    if with_vals:
        assert np.array_equal(vo, vals[perm])


def test_radix_sort_against_numpy_stable_sort():
    rng = np.random.default_rng(9)
    with World() as w:
        for n in (1, 2, 63, 64, 65, 4095, 4096, 4097, 8192 + 77, 100_003):
            _check(w, rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2)
                   + rng.integers(0, 2, n).astype(np.uint64))
            _check(w, rng.integers(0, 5, n).astype(np.uint64))           # few distinct keys: stability
        _check(w, np.full(50_000, 7, np.uint64))                         # all equal
        _check(w, np.arange(30_000, dtype=np.uint64)[::-1].copy())       # reversed
        # the broadphase's keys: world << 32 | ordered float bound, and
        # body << 32 | slot in its row
        world = rng.integers(0, 4096, 200_000).astype(np.uint64)
        bound = rng.integers(0, 2**32, 200_000).astype(np.uint64)
        _check(w, (world << np.uint64(32)) | bound)
        _check(w, (rng.integers(0, 500_000, 300_000).astype(np.uint64) << np.uint64(32))
               | rng.integers(0, 300, 300_000).astype(np.uint64), with_vals=False)
        ko, vo = w.selftest_sort(np.zeros(0, np.uint64))
        assert len(ko) == 0
