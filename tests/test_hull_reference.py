"""Hull features at registration against the reference's own
(btConvexHullComputer + btConvexUtility::initializePolyhedralFeatures,
opencl/gpu_rigidbody_pipeline/btConvexUtility.cpp:31-510): the committed
fixture tests/golden/hull_features.npz (written by make_hull_features.py from
oracle/_ref/libref_hull.so, the reference code compiled where it lies) and,
when oracle/_ref is built, the reference code itself on the whole hull pool of
the 65,536-hull scene.

What must agree: the vertices (the reference snaps points to its integer
grid; the library uses the same grid) bit for bit as a set, and the faces as
sets of vertices with planes within float rounding.  The order of vertices,
faces and unique edges differs (the reference walks its hull's edge graph),
and so may the grouping of near-coplanar triangles (normal dot > 0.999) into
merged faces, which depends on the order faces are visited
(btConvexUtility.cpp:117-145); where it differs, the differing faces must
cover the same vertices."""
import ctypes
import os

import numpy as np
import pytest

from experiments_amd import scenes
from test_abi import hull

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "hull_features.npz")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_hull.so")


def load_fixture():
    z = np.load(GOLD)
    out = []
    for k in range(int(z["n"])):
        out.append((z["points_%d" % k], z["verts_%d" % k], z["planes_%d" % k],
                    z["face_index_%d" % k], z["indices_%d" % k], z["edges_%d" % k]))
    return out


def faces_of(v, pl, fi, ind):
    return {frozenset(map(tuple, v[ind[s:s + c]])): pl[f] for f, (s, c) in enumerate(fi)}


def compare(ref, mine):
    """Returns True if the faces agree as sets, False if they differ only in
    how near-coplanar triangles were grouped; asserts everything else."""
    rv, rpl, rfi, rind, red = ref
    v, pl, fi, ind, ed = mine
    assert sorted(map(tuple, rv)) == sorted(map(tuple, v))  # bit for bit
    fr, fm = faces_of(rv, rpl, rfi, rind), faces_of(v, pl, fi, ind)
    common = set(fr) & set(fm)
    weld = np.sin(np.arccos(0.999)) * 2.0
    for f in common:
        if len(f) == 3 or np.allclose(fr[f], fm[f], atol=2e-6):
            assert np.allclose(fr[f], fm[f], atol=2e-6), (fr[f], fm[f])
        else:
            # a merged face keeps the plane of the triangle its group started
            # from (btConvexUtility.cpp:181-183), which depends on the order
            # the faces are visited: both planes are within the weld angle
            assert np.dot(fr[f][:3], fm[f][:3]) > np.cos(2 * np.arccos(0.999))
            pts = np.array(sorted(f))
            for n in (fr[f], fm[f]):
                assert np.abs(pts @ n[:3] + n[3]).max() < weld
    if len(common) == len(fr) == len(fm):
        # unique edge directions: the same set up to sign (directions within
        # 1e-6 of each other count once, and which one stays depends on the
        # order, btConvexUtility.cpp:369-385)
        assert len(red) == len(ed)
        for e in red:
            d = np.minimum(np.abs(ed - e).max(1), np.abs(ed + e).max(1))
            assert d.min() <= 2e-6, e
        return True
    only_r, only_m = set(fr) - common, set(fm) - common
    assert frozenset().union(*only_r) == frozenset().union(*only_m)
    # every differing face is a plane of the hull: its vertices are coplanar
    # within the weld angle
    for f in only_r | only_m:
        pts = np.array(sorted(f))
        n = (fr if f in fr else fm)[f]
        assert np.abs(pts @ n[:3] + n[3]).max() < weld
    return False


def test_fixture_hulls_match_reference_features():
    fx = load_fixture()
    assert len(fx) == 7
    same = 0
    for pts, *ref in fx:
        same += compare(ref, hull(pts))
    # the cube and these pool shapes group their faces alike
    assert same == len(fx)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")
def test_fixture_is_what_the_reference_code_gives():
    lib = ctypes.CDLL(REF)
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mhf", os.path.join(HERE, "golden", "make_hull_features.py"))
    mhf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mhf)
    for pts, *ref in load_fixture():
        live = mhf.ref_features(lib, pts)
        for a, b in zip(ref, live):
            assert np.array_equal(a, b)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")
def test_pool_hulls_match_reference_features():
    lib = ctypes.CDLL(REF)
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mhf", os.path.join(HERE, "golden", "make_hull_features.py"))
    mhf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mhf)
    p8, p32 = scenes.hull_pool()
    pool = [scenes.cube_points(0.5), scenes.box_points(3.0, 0.25, 1.5)] + p8 + p32
    same = sum(compare(mhf.ref_features(lib, p), hull(p)) for p in pool)
    # the merge grouping differs for a few percent of the random hulls
    assert same >= 0.95 * len(pool), (same, len(pool))
    print("%d of %d hulls with the same faces" % (same, len(pool)))
