"""Oracle primitives pinned against closed forms (numpy stable sort, cumsum,
searchsorted) and the reference's own known-answer cases."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n", [0, 1, 7, 256, 1000, 4097, 65536])
@pytest.mark.parametrize("bits", [8, 21, 32])
def test_radix_sort_matches_stable_sort(n, bits):
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = O.radix_sort_kv(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def test_radix_sort_duplicates_keep_input_order():
    # AdlTest radixSortTest: random 8-bit keys, value = index (UnitTests.h:174)
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 256, 5000).astype(np.uint32)
    k, v = O.radix_sort_kv(keys, np.arange(5000, dtype=np.uint32), 8)
    for key in np.unique(k):
        vv = v[k == key]
        assert np.all(np.diff(vv.astype(np.int64)) > 0)


def test_radix_sort_c_api_known_answer():
    # opencl/c_api/main.c:44-98: four keys in reverse order come back sorted
    k, v = O.radix_sort_kv(np.array([3, 2, 1, 0], np.uint32), np.array([0, 1, 2, 3], np.uint32))
    assert list(k) == [0, 1, 2, 3] and list(v) == [3, 2, 1, 0]


def test_radix_sort_padding_keys_sort_last():
    keys = np.array([5, 0xffffffff, 1, 0xffffffff, 3], np.uint32)
    k, v = O.radix_sort_kv(keys, np.arange(5, dtype=np.uint32))
    assert list(k) == [1, 3, 5, 0xffffffff, 0xffffffff] and list(v) == [2, 4, 0, 1, 3]


def test_radix_sort_high_bits_ignored_beyond_sort_bits():
    keys = np.array([0x100, 0x001, 0x200], np.uint32)
    k, v = O.radix_sort_kv(keys, np.arange(3, dtype=np.uint32), 8)
    # only the low 8 bits order: 0x100 and 0x200 both have digit 0
    assert list(v) == [0, 2, 1]


@pytest.mark.parametrize("n", [0, 1, 128, 1000, 262144])
def test_exclusive_scan(n):
    # AdlTest scanTest (UnitTests.h:132): all-ones and random inputs
    x = np.ones(n, np.uint32)
    out, tot = O.exclusive_scan(x)
    assert np.array_equal(out, np.arange(n, dtype=np.uint32)) and tot == n
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, n).astype(np.uint32)
    out, tot = O.exclusive_scan(x)
    ref = np.concatenate([[0], np.cumsum(x.astype(np.uint64))[:-1]]).astype(np.uint32) if n else x
    assert np.array_equal(out, ref) and tot == int(x.sum())


def test_bound_search_lower_upper_count():
    # AdlTest boundSearchTest (UnitTests.h:237): key == i on [lower, upper)
    rng = np.random.default_rng(1)
    keys = np.sort(rng.integers(0, 200, 3000)).astype(np.uint32)
    n_dst = 256
    lo = O.bound_search(keys, n_dst, "lower")
    hi = O.bound_search(keys, n_dst, "upper")
    cnt = O.bound_search(keys, n_dst, "count")
    for k in range(n_dst):
        c = int(np.sum(keys == k))
        assert cnt[k] == c
        if c:
            assert lo[k] == np.searchsorted(keys, k, "left")
            assert hi[k] == np.searchsorted(keys, k, "right")
            assert np.all(keys[lo[k]:hi[k]] == k)
