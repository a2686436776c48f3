"""The radix sort entry point that allocates nothing
(mi_rb_radix_sort_kv_noalloc, workspace in a file-scope device array) gives
the same pairs as mi_rb_radix_sort_kv with its handle's workspace, on two
calls in a row (the second smaller), and rejects sizes above its maximum."""
import ctypes

import numpy as np
import pytest

from experiments_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MAX_N = 1 << 20


def _pairs(n, bits, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    if n > 8:
        keys[::5] = keys[1]  # duplicates: the order of equal keys is checked
    vals = rng.permutation(n).astype(np.uint32)
    return np.stack([keys, vals], 1)


def _sorted_with_handle(kv, bits):
    h = ctypes.c_void_p()
    L.call("mi_rb_radix_sort_create", max(len(kv), 1), ctypes.byref(h))
    try:
        d = torch.from_numpy(kv.view(np.int32).copy()).cuda()
        L.call("mi_rb_radix_sort_kv", h, ctypes.c_void_p(d.data_ptr()), len(kv), bits, None)
        torch.cuda.synchronize()
        return d.cpu().numpy().view(np.uint32)
    finally:
        L.lib().mi_rb_radix_sort_destroy(h)


def _sorted_noalloc(kv, bits, stream=None):
    d = torch.from_numpy(kv.view(np.int32).copy()).cuda()
    L.call("mi_rb_radix_sort_kv_noalloc", ctypes.c_void_p(d.data_ptr()), len(kv), bits, stream)
    torch.cuda.synchronize()
    return d.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("sizes,bits", [((MAX_N, 65537), 32), ((300000, 4097), 19),
                                        ((4096, 255), 8)])
def test_noalloc_sort_equals_handle_sort_twice(sizes, bits):
    for k, n in enumerate(sizes):
        kv = _pairs(n, bits, seed=n + 31 * k + bits)
        want = _sorted_with_handle(kv, bits)
        got = _sorted_noalloc(kv, bits)
        assert np.array_equal(got, want), (n, bits)
        # stable: equal keys keep their input order (the values say where from)
        order = np.argsort(kv[:, 0], kind="stable")
        assert np.array_equal(got[:, 1], kv[order, 1])


def test_noalloc_sort_on_a_stream():
    s = torch.cuda.Stream()
    kv = _pairs(70000, 26, seed=5)
    want = _sorted_with_handle(kv, 26)
    d = torch.from_numpy(kv.view(np.int32).copy()).cuda()
    torch.cuda.synchronize()
    L.call("mi_rb_radix_sort_kv_noalloc", ctypes.c_void_p(d.data_ptr()), len(kv), 26,
           ctypes.c_void_p(s.cuda_stream))
    s.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint32), want)


def test_noalloc_sort_rejects_sizes_above_its_maximum():
    d = torch.zeros((8, 2), dtype=torch.int32, device="cuda")
    st = L.lib().mi_rb_radix_sort_kv_noalloc(ctypes.c_void_p(d.data_ptr()), MAX_N + 1, 32, None)
    assert st == L.CAPACITY
    st = L.lib().mi_rb_radix_sort_kv_noalloc(None, 8, 32, None)
    assert st == L.INVALID_ARG
