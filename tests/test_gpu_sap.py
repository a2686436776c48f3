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


def run_sap(mn, mx, axis, max_pairs):
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
    L.call("mi_rb_sap_find_pairs", h, P(dmn), P(dmx), len(mn), P(pairs), P(num), None)
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
