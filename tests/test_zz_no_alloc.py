"""ds_sort_pairs_u64_noalloc (the pair sort on a fixed device scratch array,
for callers that cannot allocate) against ds_sort_pairs_u64: two calls in a
row, the second with a smaller input, each bit-exact with the allocating
entry point and with numpy's stable argsort; sizes past the scratch's limit
are rejected before any GPU work."""
from __future__ import annotations

import numpy as np
import pytest

import diffsplat_amd as ds

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 20


def sorted_by(entry, keys, vals, bits, stream=None):
    n = len(keys)
    dk = ds.DeviceArray.from_numpy(keys)
    dv = ds.DeviceArray.from_numpy(vals)
    ko = ds.DeviceArray(8 * n)
    vo = ds.DeviceArray(4 * n)
    ds.check(getattr(ds.lib(), entry)(dk.ptr, dv.ptr, ko.ptr, vo.ptr, n, bits, stream), entry)
    ds.check(ds.lib().ds_synchronize(), "sync")
    return ko.download(np.uint64, n), vo.download(np.uint32, n)


def pairs(n, bits, seed):
    rng = np.random.default_rng(seed)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(2 ** 64 - 1)
    keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) & mask
    keys[: n // 4] = keys[0]  # ties: stability
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals


def test_noalloc_sort_matches_allocating_sort_on_two_calls():
    # the first call fills the whole scratch; the second, smaller one sees the
    # first call's leftovers in it
    for n, bits, seed in ((MAX_PAIRS, 64, 1), (70001, 44, 2)):
        keys, vals = pairs(n, bits, seed)
        k0, v0 = sorted_by("ds_sort_pairs_u64", keys, vals, bits)
        k1, v1 = sorted_by("ds_sort_pairs_u64_noalloc", keys, vals, bits)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k1, k0) and np.array_equal(v1, v0), n
        assert np.array_equal(k1, keys[order]) and np.array_equal(v1, vals[order]), n


def test_noalloc_sort_rejects_sizes_beyond_its_scratch():
    d = ds.DeviceArray(16)
    st = ds.lib().ds_sort_pairs_u64_noalloc(d.ptr, d.ptr, d.ptr, d.ptr, MAX_PAIRS + 1, 64, None)
    assert st == 4  # DS_ERR_CAPACITY, before any launch
    assert b"at most 1048576" in ds.lib().ds_last_error()
    # the library still works afterwards
    keys, vals = pairs(1000, 20, 3)
    k1, v1 = sorted_by("ds_sort_pairs_u64_noalloc", keys, vals, 20)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k1, keys[order]) and np.array_equal(v1, vals[order])
