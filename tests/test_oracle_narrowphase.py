"""Oracle SAT, clipping and manifold reduction pinned against the reference's
golden clip fixture (testdata/clipFacesAndContactReductionKernel.7z) and
closed-form cube cases."""
import os

import numpy as np
import pytest

import oracle as O
from experiments_amd import scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]], np.float32)


def _dot3(a, b):
    f = np.float32
    return f(f(f(a[0] * b[0]) + f(a[1] * b[1])) + f(a[2] * b[2]))


def test_clip_and_reduce_match_reference_gold():
    """380 pairs: clip the incident polygon against the reference face's side
    planes, keep depth <= 0.02, reduce to 4 -- must reproduce the reference
    kernel's contact points and their order (within 1e-5; the reference ran
    with fused multiply-adds)."""
    g = np.load(os.path.join(GOLD, "clip_golden.npz"))
    faces, fa, na, pb, sep = g["faces"], g["face_a"], g["normal_a"], g["poly_b"], g["sep"]
    cp, cn = g["contact_pos"], g["contact_normal"]
    exact = 0
    for p in range(len(faces)):
        nA, nB = faces[p, 2], faces[p, 3]
        vA, poly, nrm = fa[p, :nA], pb[p, :nB].copy(), na[p]
        for e0 in range(nA):
            pn = -_cross(vA[e0] - vA[(e0 + 1) % nA], nrm)
            poly = O.clip_face(poly, pn, -_dot3(vA[e0], pn))
        eq = -_dot3(nrm, vA[0])
        pts = []
        for v in poly:
            d = np.float32(_dot3(nrm, v) + eq)
            if d <= 0.02:
                pts.append([v[0], v[1], v[2], d])
        pts = np.array(pts, np.float32)
        idx = O.reduce_manifold(pts, -sep[p])
        assert len(idx) == int(cn[p, 3]), p
        got, exp = pts[idx], cp[p, :len(idx)]
        assert np.allclose(got, exp, rtol=1e-5, atol=1e-5), p
        exact += int(np.array_equal(got, exp))
    assert exact >= 360  # most are bit-identical


def cube_shapes():
    verts = np.array([[x, y, z, 0] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)],
                     np.float32)
    # faces wound counter-clockwise seen from outside
    faces = [[0, 1, 3, 2], [4, 6, 7, 5], [0, 4, 5, 1], [2, 3, 7, 6], [0, 2, 6, 4], [1, 5, 7, 3]]
    planes, fi, ind = [], [], []
    for f in faces:
        p = verts[f, :3]
        n = np.cross(p[1] - p[0], p[2] - p[1])
        n /= np.linalg.norm(n)
        planes.append([n[0], n[1], n[2], -float(np.dot(n, p[0]))])
        fi.append([len(ind), 4])
        ind += f
    H = np.zeros(1, O.HULL_DTYPE)
    H["num_faces"], H["num_vertices"], H["num_edges"], H["num_indices"] = 6, 8, 3, 24
    H["aabb_min"][0, :3], H["aabb_max"][0, :3] = -.5, .5
    edges = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
    return O.Shapes(H, verts, np.array(planes, np.float32), np.array(fi), np.array(ind), edges)


@pytest.mark.parametrize("pen", [0.0, 0.01, 0.05])
def test_stacked_cubes_closed_form(pen):
    sh = cube_shapes()
    pos = np.array([[0, 0, 0], [0, 1 - pen, 0]], np.float32)
    q = np.tile(np.array([0, 0, 0, 1], np.float32), (2, 1))
    slots, has = O.narrowphase(np.array([[1, 0]]), pos, q, np.array([1, 1], np.float32),
                               np.array([0, 0]), sh)
    assert has[0] == 1
    c = slots[0]
    assert np.allclose(c["normal"][:3], [0, -1, 0])
    assert c["normal"][3] == 4
    pts = c["pos"][:4]
    assert np.allclose(np.sort(pts[:, 0]), [-.5, -.5, .5, .5])
    assert np.allclose(np.sort(pts[:, 2]), [-.5, -.5, .5, .5])
    assert np.allclose(pts[:, 1], 0.5, atol=1e-6)
    assert np.allclose(pts[:, 3], -pen, atol=1e-6)
    assert c["body_a"] == 1 and c["body_b"] == 0 and c["pair_index"] == 0


def test_separated_cubes_no_contact():
    sh = cube_shapes()
    pos = np.array([[0, 0, 0], [0, 1.05, 0]], np.float32)
    q = np.tile(np.array([0, 0, 0, 1], np.float32), (2, 1))
    r, sep = O.sat(sh, 0, 0, pos[1], q[1], pos[0], q[0])
    assert r == 0
    _, has = O.narrowphase(np.array([[1, 0]]), pos, q, np.ones(2, np.float32), np.zeros(2, int), sh)
    assert has[0] == 0


def test_static_static_pair_skipped():
    sh = cube_shapes()
    pos = np.array([[0, 0, 0], [0, 0.9, 0]], np.float32)
    q = np.tile(np.array([0, 0, 0, 1], np.float32), (2, 1))
    _, has = O.narrowphase(np.array([[1, 0]]), pos, q, np.zeros(2, np.float32), np.zeros(2, int), sh)
    assert has[0] == 0


def test_sat_axis_unit_and_oriented():
    rng = np.random.default_rng(2)
    sh = cube_shapes()
    for _ in range(50):
        qa = rng.normal(size=4); qa /= np.linalg.norm(qa)
        qb = rng.normal(size=4); qb /= np.linalg.norm(qb)
        pa = rng.uniform(-0.6, 0.6, 3).astype(np.float32)
        pb = np.zeros(3, np.float32)
        r, sep = O.sat(sh, 0, 0, pa, qa.astype(np.float32), pb, qb.astype(np.float32))
        if r:
            assert abs(np.linalg.norm(sep[:3]) - 1) < 1e-5
            assert np.dot(sep[:3], pa - pb) >= -1e-6


def test_reduce_identity_below_four_points():
    pts = np.array([[0, 0, 0, -0.1], [1, 0, 0, -0.1], [0, 0, 1, -0.2]], np.float32)
    assert list(O.reduce_manifold(pts, [0, 1, 0, 0])) == [0, 1, 2]


def test_reduce_regular_polygon_extremes():
    # 8 points on a circle in the xz plane, normal +y: u = n x (p0 - c), so the
    # +u/-u/+v/-v extremes are 4 of the 8 vertices, 90 degrees apart
    ang = np.arange(8) * (np.pi / 4)
    pts = np.stack([np.cos(ang), np.zeros(8), np.sin(ang), np.zeros(8)], 1).astype(np.float32)
    idx = O.reduce_manifold(pts, [0, 1, 0, 0])
    assert len(set(idx.tolist())) == 4
    a = np.sort(ang[idx])
    assert np.allclose(np.diff(a), np.pi / 2)


# ---- the CPU twin's SAT and the split path's reduction (options) -----------

def _sat64_separated(va, vb):
    """Exact separation of two convex vertex sets in float64: some face
    normal or edge cross product of their hulls separates them."""
    from scipy.spatial import ConvexHull
    axes = []
    for v in (va, vb):
        h = ConvexHull(v)
        axes += [eq[:3] for eq in h.equations]
    ea = [va[j] - va[i] for i in range(len(va)) for j in range(i + 1, len(va))]
    eb = [vb[j] - vb[i] for i in range(len(vb)) for j in range(i + 1, len(vb))]
    for a in ea:
        for b in eb:
            c = np.cross(a, b)
            if np.linalg.norm(c) > 1e-9:
                axes.append(c / np.linalg.norm(c))
    for ax in axes:
        pa, pb = va @ ax, vb @ ax
        if pa.max() < pb.min() - 1e-7 or pb.max() < pa.min() - 1e-7:
            return True
    return False


def test_cpu_twin_sat_rejects_pairs_separated_by_an_edge_axis():
    """findSeparatingAxis of the CPU twin (ConvexHullContact.cpp:326-453) ends
    at an edge axis with disjoint projections; the GPU kernel's never does.
    On cube pairs placed edge to edge the twin's answer is the exact
    (float64) separation test, and the kernel's keeps the pairs that only an
    edge axis separates."""
    from test_abi import hull as build
    v, pl, fi, ind, ed = build(scenes.cube_points(0.5))
    H = np.zeros(1, O.HULL_DTYPE)
    H["num_faces"], H["num_vertices"], H["num_edges"], H["num_indices"] = len(pl), len(v), len(ed), len(ind)
    sh = O.Shapes(H, O.f4(v), pl.astype(np.float32), fi.astype(np.int32), ind.astype(np.int32),
                  O.f4(ed))
    rng = np.random.default_rng(8)
    edge_only = both = 0
    for trial in range(700):
        qa = rng.normal(size=4); qa /= np.linalg.norm(qa)
        qb = rng.normal(size=4); qb /= np.linalg.norm(qb)
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        pa = np.zeros(3, np.float32)
        pb = (d * rng.uniform(0.9, 1.8)).astype(np.float32)
        qa, qb = qa.astype(np.float32), qb.astype(np.float32)
        ok_gpu, _ = O.sat(sh, 0, 0, pa, qa, pb, qb)
        ok_twin, _ = O.sat(sh, 0, 0, pa, qa, pb, qb, edge_reject=True)
        R = lambda q: _rot(q)
        va = v @ R(qa).T + pa
        vb = v @ R(qb).T + pb
        sep = _sat64_separated(va.astype(np.float64), vb.astype(np.float64))
        if ok_twin:
            assert not sep, trial      # the twin keeps only overlapping pairs
            assert ok_gpu
        if ok_gpu and not ok_twin:
            edge_only += 1
            assert sep, trial          # ... and rejects only separated ones
        both += ok_twin
    assert edge_only > 5 and both > 25, (edge_only, both)


def _rot(q):
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_split_path_reduction_known_answers():
    """extractManifoldSequentialGlobal (satClipHullContacts.cl:678-767)."""
    n = np.array([0, 1, 0, 0], np.float32)
    # up to four points pass through
    for k in range(5):
        pts = np.zeros((k, 4), np.float32)
        pts[:, 0] = np.arange(k)
        assert list(O.reduce_manifold(pts, n, split_path=True)) == list(range(k))
    # eight points on a circle in the plane of the normal: u = n x (p0 - c)
    # = -z, v = n x u = -x; the smallest projections on u, -u, v, -v are the
    # points at 90, 270, 0 and 180 degrees
    ang = np.arange(8) * (np.pi / 4)
    pts = np.zeros((8, 4), np.float32)
    pts[:, 0], pts[:, 2], pts[:, 3] = np.cos(ang), np.sin(ang), -0.01
    assert list(O.reduce_manifold(pts, n, split_path=True)) == [2, 6, 0, 4]
    # the deepest point replaces the first when it is none of them
    pts[5, 3] = -0.1
    assert list(O.reduce_manifold(pts, n, split_path=True)) == [5, 6, 0, 4]
    # when it is one of them, nothing is replaced
    pts[5, 3] = -0.01
    pts[6, 3] = -0.1
    assert list(O.reduce_manifold(pts, n, split_path=True)) == [2, 6, 0, 4]


def test_split_path_reduction_may_repeat_points():
    """Unlike extractManifoldSequential, the split path returns four indices
    even when one point is the extreme of two directions."""
    n = np.array([0, 1, 0, 0], np.float32)
    xz = [(0.82, 0.17), (0.33, 0.11), (-1.3, 0.09), (0.91, 0.01), (0.45, 0.16), (-0.54, -0.22)]
    pts = np.array([(x, 0, z, -.01) for x, z in xz], np.float32)
    idx = O.reduce_manifold(pts, n, split_path=True)
    # point 3 has the smallest projection on -u and on v
    assert list(idx) == [0, 3, 3, 2]
    idx1 = O.reduce_manifold(pts, n)
    assert len(idx1) == 4
