"""The stand-alone hierarchical grid (mi_rb_grid_*) on the GPU: the sorted
(x > y) pair list equals the brute-force overlap set and the oracle's
single-level grid, bit for bit, for bodies of very different sizes."""
import ctypes

import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def run_grid(mn, mx, cell, dims, levels, max_pairs):
    desc = L.GridDesc(cell, (ctypes.c_int * 3)(*dims), max_pairs, 1, levels)
    g = ctypes.c_void_p()
    L.call("mi_rb_grid_create", ctypes.byref(desc), len(mn), ctypes.byref(g))
    dmn = torch.from_numpy(mn).cuda()
    dmx = torch.from_numpy(mx).cuda()
    pairs = torch.zeros((max_pairs, 2), dtype=torch.int32, device="cuda")
    num = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    L.call("mi_rb_grid_find_pairs", g, P(dmn), P(dmx), None, len(mn), P(pairs), P(num), None)
    torch.cuda.synchronize()
    flags = ctypes.c_int()
    L.call("mi_rb_grid_overflow", g, ctypes.byref(flags))
    ms = np.zeros(4, np.float32)
    L.call("mi_rb_grid_last_ms", g, ms.ctypes.data_as(ctypes.c_void_p))
    L.lib().mi_rb_grid_destroy(g)
    k = int(num.cpu()[0])
    return pairs[:k].cpu().numpy(), flags.value, ms


@pytest.mark.parametrize("n,box", [(1, 10.0), (2000, 20.0), (20000, 60.0)])
def test_mixed_sizes_equal_brute_force(n, box):
    mn, mx, _, _ = scenes.broadphase_aabbs(n, box=box, seed=7 + n)
    got, flags, _ = run_grid(mn, mx, 1.414, (32, 32, 32), 4, 1 << 22)
    ref = O.find_pairs_brute(mn, mx)
    assert flags == 0
    assert np.array_equal(got, ref)


def test_top_level_overflow_is_reported():
    mn, mx, _, _ = scenes.broadphase_aabbs(100, box=10.0, seed=3)
    mx[5, :3] = mn[5, :3] + 50.0   # larger than the top level's cell (1.414 * 2)
    got, flags, _ = run_grid(mn, mx, 1.414, (16, 16, 16), 2, 1 << 16)
    assert flags & L.OVERFLOW_LEVEL


def test_pair_capacity_overflow_is_reported():
    mn, mx, _, _ = scenes.broadphase_aabbs(5000, box=15.0, seed=4)
    got, flags, _ = run_grid(mn, mx, 1.414, (16, 16, 16), 4, 100)
    assert flags & L.OVERFLOW_PAIRS and len(got) == 100


def test_million_bodies_equal_single_level_oracle():
    """1,048,576 spheres and boxes, radii log-uniform in [0.25, 2] in a 400^3
    domain, level-0 cell = 2 x the median radius.  The oracle's single-level
    grid with a cell large enough for every body (no overflow list) finds
    the same set, in the same order."""
    mn, mx, _, r = scenes.broadphase_aabbs()
    cell = float(2 * np.median(r))
    got, flags, ms = run_grid(mn, mx, cell, (256, 256, 256), 4, 1 << 24)
    assert flags == 0
    big = 8.0   # >= the largest AABB extent (2 * 2 * sqrt(3))
    dims = (64, 64, 64)
    kv = O.calc_hash(mn, mx, np.float32(1 / big), dims)
    k, v = O.radix_sort_kv(kv[:, 0], kv[:, 1], 18)
    ref = O.find_pairs_grid(np.stack([k, v], 1), mn, mx, np.float32(1 / big), dims,
                            max_pairs=1 << 24)
    assert len(ref) > 100000
    assert np.array_equal(got, ref)
    print("1M grid phase ms (hash, sort, pairs, pair order):", ms)
