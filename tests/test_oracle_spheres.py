"""Oracle contacts of spheres and planes (clipHullHullKernel's first
branches, satClipHullContacts.cl:1055-1118) against closed forms: a sphere
resting in a plane, spheres overlapping or touching, coincident centres,
rotated planes, and the pairs the reference gives no contact (a hull against
a sphere or a plane, a plane against a plane)."""
import numpy as np
import pytest

import oracle as O

from test_oracle_narrowphase import cube_shapes


def shapes_with(sphere_radii=(), planes=()):
    """The cube of cube_shapes (shape 0), then spheres, then planes."""
    cube = cube_shapes()
    H = [cube.hulls[0]]
    pl = [p for p in cube.planes]
    fi = [f for f in cube.face_index]
    for r in sphere_radii:
        h = np.zeros(1, O.HULL_DTYPE)[0]
        h["face_offset"], h["center"][3] = len(pl), r
        h["aabb_min"][:3], h["aabb_max"][:3] = -r, r
        H.append(h)
    for n, c in planes:
        h = np.zeros(1, O.HULL_DTYPE)[0]
        h["face_offset"], h["num_faces"] = len(pl), 1
        h["aabb_min"][:3], h["aabb_max"][:3] = -1e30, 1e30
        pl.append(np.array([n[0], n[1], n[2], c], np.float32))
        fi.append(np.array([len(cube.indices), 0]))
        H.append(h)
    return O.Shapes(np.array(H, O.HULL_DTYPE), cube.verts, np.array(pl, np.float32),
                    np.array(fi), cube.indices, cube.edges)


def quat_axis(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.array([*(a * np.sin(ang / 2)), np.cos(ang / 2)], np.float32)


def run(shapes, shape, pos, quat, inv_mass, pair):
    pos = np.asarray(pos, np.float32)
    quat = np.asarray(quat, np.float32)
    slots, has = O.narrowphase(np.array([pair]), pos, quat, np.asarray(inv_mass, np.float32),
                               np.asarray(shape, np.int32), shapes)
    return slots[0], has[0]


@pytest.mark.parametrize("order", ["plane_first", "sphere_first"])
def test_sphere_in_plane(order):
    sh = shapes_with(sphere_radii=[0.5], planes=[((0, 1, 0), 0.0)])
    shape, pos = [2, 1], [[0, 0, 0], [1.5, 0.45, -2]]   # plane body 0, sphere body 1
    q = [[0, 0, 0, 1], [0, 0, 0, 1]]
    pair = [0, 1] if order == "plane_first" else [1, 0]
    c, has = run(sh, shape, pos, q, [0, 1], pair)
    assert has == 1 and c["normal"][3] == 1
    assert np.allclose(c["normal"][:3], [0, 1, 0])
    assert np.allclose(c["pos"][0], [1.5, -0.05, -2, -0.05], atol=1e-6)
    assert (c["body_a"], c["body_b"]) == (0, 1)  # the plane body first; static: -0 == 0
    # resting exactly on the plane, or above it: no contact (distance < 0 only)
    for y in (0.5, 0.7):
        _, has = run(sh, shape, [[0, 0, 0], [0, y, 0]], q, [0, 1], pair)
        assert has == 0


def test_rotated_offset_plane():
    """Plane n . x = 1 in its body frame; the body turned 40 degrees about z
    and moved: the world plane is n_w . x = 1 + n_w . p."""
    n, c0 = np.array([0, 1, 0], np.float32), 1.0
    sh = shapes_with(sphere_radii=[0.3], planes=[(n, c0)])
    qp = quat_axis([0, 0, 1], np.deg2rad(40))
    pp = np.array([0.2, -0.4, 0.1], np.float32)
    R = np.array([[np.cos(np.deg2rad(40)), -np.sin(np.deg2rad(40)), 0],
                  [np.sin(np.deg2rad(40)), np.cos(np.deg2rad(40)), 0], [0, 0, 1]])
    nw = R @ n
    t = np.cross(nw, [0, 0, 1])
    for depth in (0.01, 0.1):
        # sphere centre at signed distance r - depth from the world plane
        centre = pp + nw * (c0 + 0.3 - depth) + 0.7 * t
        c, has = run(sh, [2, 1], [pp, centre], [qp, [0, 0, 0, 1]], [0, 1], [0, 1])
        assert has == 1
        assert np.allclose(c["normal"][:3], nw, atol=1e-6)
        assert abs(c["pos"][0][3] + depth) < 2e-6
        assert np.allclose(c["pos"][0][:3], centre - 0.3 * nw, atol=2e-6)


def test_spheres():
    sh = shapes_with(sphere_radii=[0.5, 0.25])
    q = [[0, 0, 0, 1], [0, 0, 0, 1]]
    d = np.array([0.6, 0.0, 0.0])
    c, has = run(sh, [1, 2], [[0, 0, 0], d], q, [1, 1], [0, 1])
    assert has == 1
    # normal from A to B, the point on B's surface, depth = distance - (rA + rB)
    assert np.allclose(c["normal"], [1, 0, 0, 1])
    assert np.allclose(c["pos"][0], [0.35, 0, 0, -0.15], atol=1e-6)
    # exactly touching: a contact of depth 0; apart: none
    c, has = run(sh, [1, 2], [[0, 0, 0], [0.75, 0, 0]], q, [1, 1], [0, 1])
    assert has == 1 and c["pos"][0][3] == 0.0
    _, has = run(sh, [1, 2], [[0, 0, 0], [0.76, 0, 0]], q, [1, 1], [0, 1])
    assert has == 0
    # coincident centres: the normal defaults to -(1, 0, 0)
    c, has = run(sh, [1, 2], [[1, 1, 1], [1, 1, 1]], q, [1, 0], [0, 1])
    assert has == 1 and np.allclose(c["normal"], [-1, 0, 0, 1])
    assert c["body_b"] == -1


@pytest.mark.parametrize("shapes_ab", [(0, 1), (1, 0), (0, 2), (2, 0), (2, 3)])
def test_pairs_without_contact(shapes_ab):
    """A cube against a sphere or a plane, a plane against a plane: none,
    even when they overlap."""
    sh = shapes_with(sphere_radii=[0.5], planes=[((0, 1, 0), 0.0), ((1, 0, 0), 0.0)])
    q = [[0, 0, 0, 1], [0, 0, 0, 1]]
    _, has = run(sh, list(shapes_ab), [[0, 0, 0], [0.1, 0.2, 0]], q, [1, 1], [0, 1])
    assert has == 0
