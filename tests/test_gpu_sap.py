"""Sweep-and-prune pair search (mi_rb_sap_*) on the GPU: the pair list of the
reference's 5,876 SAP AABBs and of random mixed-size sets equals the oracle's
sweep (orc_sap_pairs) and the brute-force overlap set, bit for bit."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def run_sap(mn, mx, axis, max_pairs, is_large=None):
    mn4 = np.zeros((len(mn), 4), np.float32)
    mx4 = np.zeros((len(mx), 4), np.float32)
    mn4[:, :3], mx4[:, :3] = mn[:, :3], mx[:, :3]
    dmn = torch.from_numpy(mn4).cuda()
    dmx = torch.from_numpy(mx4).cuda()
    desc = L.SapDesc(axis, max_pairs)
    h = ctypes.c_void_p()
    L.call("mi_rb_sap_create", ctypes.byref(desc), max(len(mn), 1), ctypes.byref(h))
    pairs = torch.zeros((max_pairs, 2), dtype=torch.int32, device="cuda")
    num = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    dl = None if is_large is None else torch.from_numpy(np.ascontiguousarray(is_large, np.int32)).cuda()
    L.call("mi_rb_sap_find_pairs", h, P(dmn), P(dmx), None if dl is None else P(dl), len(mn), P(pairs),
           P(num), None)
    torch.cuda.synchronize()
    flags = ctypes.c_int()
    L.call("mi_rb_sap_overflow", h, ctypes.byref(flags))
    L.lib().mi_rb_sap_destroy(h)
    k = int(num.cpu()[0])
    return pairs[:k].cpu().numpy(), flags.value


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_reference_sap_aabbs(axis):
    g = np.load(os.path.join(GOLD, "sap_aabbs.npz"))
    mn, mx = g["aabb_min"], g["aabb_max"]
    got, flags = run_sap(mn, mx, axis, 1 << 20)
    assert flags == 0
    ref = O.find_pairs_brute(mn, mx)
    assert len(ref) > 1000
    assert np.array_equal(got, ref)
    assert np.array_equal(got, O.sap_pairs(mn, mx, axis))


@pytest.mark.parametrize("n,box", [(1, 5.0), (5000, 30.0), (60000, 90.0)])
def test_mixed_sizes(n, box):
    mn, mx, _, _ = scenes.broadphase_aabbs(n, box=box, seed=21 + n)
    got, flags = run_sap(mn, mx, 0, 1 << 22)
    assert flags == 0
    assert np.array_equal(got, O.find_pairs_brute(mn, mx))


def test_pair_capacity_overflow_is_reported():
    mn, mx, _, _ = scenes.broadphase_aabbs(4000, box=12.0, seed=5)
    got, flags = run_sap(mn, mx, 0, 100)
    assert flags & L.OVERFLOW_PAIRS and len(got) == 100


def test_large_aabbs_and_long_pair_ranges():
    """Large AABBs (a ground under 12,000 bodies, a wall, two overlapping
    slabs) are tested against every AABB instead of sweeping, and a pair of
    two large ones is listed once.  The ground has the highest index, so its
    12,000 pairs form one range, ordered through the bitmap path."""
    mn, mx, _, _ = scenes.broadphase_aabbs(12000, box=40.0, seed=9)
    mn, mx = np.asarray(mn, np.float32)[:, :3], np.asarray(mx, np.float32)[:, :3]
    big_mn = np.array([[-5, -5, 10], [-5, -5, -5], [-3, -3, -3], [-50, -50, -50]], np.float32)
    big_mx = np.array([[45, 45, 12], [2, 45, 45], [30, 3, 30], [50, 50, 50]], np.float32)
    mn = np.concatenate([mn, big_mn])
    mx = np.concatenate([mx, big_mx])
    large = np.zeros(len(mn), np.int32)
    large[-4:] = 1
    ref = O.find_pairs_brute(mn, mx)
    assert np.sum(ref[:, 0] == len(mn) - 1) == len(mn) - 1   # the last box touches all
    for axis in (0, 2):
        got, flags = run_sap(mn, mx, axis, 1 << 22, is_large=large)
        assert flags == 0
        assert np.array_equal(got, ref)
    got, flags = run_sap(mn, mx, 1, 1 << 22)   # no large list: the same set
    assert flags == 0 and np.array_equal(got, ref)
