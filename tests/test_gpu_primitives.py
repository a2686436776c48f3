"""Radix sort, scan and bound search on the GPU, through the C ABI, against
the oracle (bit-exact)."""
import ctypes

import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def sorter():
    h = ctypes.c_void_p()
    L.call("mi_rb_radix_sort_create", 1 << 21, ctypes.byref(h))
    yield h
    L.lib().mi_rb_radix_sort_destroy(h)


def test_radix_sort_c_api_known_answer(sorter):
    # opencl/c_api/main.c:44-98: four keys in reverse order come back sorted
    kv = _dev(np.array([[3, 0], [2, 1], [1, 2], [0, 3]], np.int32))
    L.call("mi_rb_radix_sort_kv", sorter, _p(kv), 4, 32, None)
    torch.cuda.synchronize()
    assert kv.cpu().numpy().tolist() == [[0, 3], [1, 2], [2, 1], [3, 0]]


@pytest.mark.parametrize("n", [1, 2, 255, 4096, 4097, 65537, 1 << 20])
@pytest.mark.parametrize("bits", [8, 19, 32])
def test_radix_sort_kv(sorter, n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    if n > 8:
        keys[::7] = keys[0]  # many duplicates: stability matters
    vals = rng.permutation(n).astype(np.uint32)
    kv = _dev(np.stack([keys, vals], 1).view(np.int32))
    L.call("mi_rb_radix_sort_kv", sorter, _p(kv), n, bits, None)
    torch.cuda.synchronize()
    got = kv.cpu().numpy().view(np.uint32)
    rk, rv = O.radix_sort_kv(keys, vals, bits)
    assert np.array_equal(got[:, 0], rk)
    assert np.array_equal(got[:, 1], rv)


@pytest.mark.parametrize("n", [0, 1, 100, 4096, 4097, 100000, 1 << 20])
def test_exclusive_scan(sorter, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 50, max(n, 1)).astype(np.uint32)[:n]
    din = _dev(x.view(np.int32)) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
    dout = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    tot = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    L.call("mi_rb_exclusive_scan_u32", sorter, _p(din), _p(dout), n, _p(tot), None)
    torch.cuda.synchronize()
    ref, rt = O.exclusive_scan(x)
    assert np.array_equal(dout.cpu().numpy()[:n].view(np.uint32), ref)
    assert int(tot.cpu().numpy().view(np.uint32)[0]) == rt


@pytest.mark.parametrize("mode", ["lower", "upper", "count"])
def test_bound_search(mode):
    rng = np.random.default_rng(3)
    keys = np.sort(rng.integers(0, 300, 20000)).astype(np.uint32)
    kv = _dev(np.stack([keys, np.arange(len(keys), dtype=np.uint32)], 1).view(np.int32))
    n_dst = 256
    ref = O.bound_search(keys, n_dst, mode)
    init = np.zeros(n_dst, np.uint32) if mode == "count" else ref.copy()
    # lower/upper only write present keys: start from the oracle's values for absent ones
    present = np.isin(np.arange(n_dst), keys)
    init[present] = 0xDEADBEEF if mode != "count" else 0
    dst = _dev(init.view(np.int32))
    L.call("mi_rb_bound_search", _p(kv), len(keys), _p(dst), n_dst,
           {"lower": 0, "upper": 1, "count": 2}[mode], None)
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy().view(np.uint32), ref)
