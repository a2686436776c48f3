"""Oracle compound shapes (findCompoundPairsKernel, sat.cl:763-925, and
clipCompoundsHullHullKernel, satClipHullContacts.cl:1184-1310) reduced to
known cases: a compound of one child at the identity gives the contacts of
the plain hull, a compound of two children gives the contacts of two
separate bodies placed where the children are (named after the compound),
and the children's world transforms compose the body's with the child's."""
from types import SimpleNamespace

import numpy as np

import oracle as O
from oracle.step import expand_compound_pairs, oracle_contacts
from test_oracle_narrowphase import cube_shapes


def with_compounds(children):
    """cube_shapes (shape 0: unit cube) plus one compound shape (shape 1) of
    the given (position, orientation) children, all unit cubes."""
    cube = cube_shapes()
    h = np.zeros(1, O.HULL_DTYPE)[0]
    h["num_edges"], h["edge_offset"] = len(children), 0
    ch = np.zeros(len(children), O.CHILD_DTYPE)
    for i, (p, q) in enumerate(children):
        ch[i]["pos"][:3] = p
        ch[i]["orn"] = q
        ch[i]["shape"] = 0
    H = np.array([cube.hulls[0], h], O.HULL_DTYPE)
    return O.Shapes(H, cube.verts, cube.planes, cube.face_index, cube.indices, cube.edges), ch


def snapshot(shapes, children, shape, pos, quat, inv_mass):
    n = len(shape)
    proxies, first = [], np.full(n, -1)
    for i, sid in enumerate(shape):
        h = shapes.hulls[sid]
        if h["num_vertices"] == 0 and h["num_edges"] > 0:
            first[i] = len(proxies)
            proxies += [(i, int(h["edge_offset"]) + c) for c in range(int(h["num_edges"]))]
    pos4 = np.zeros((n, 4), np.float32)
    pos4[:, :3] = pos
    return SimpleNamespace(
        state={"pos": pos4, "quat": np.asarray(quat, np.float32)},
        sh={"hulls": shapes.hulls, "shape": np.asarray(shape, np.int32),
            "inv_mass": np.asarray(inv_mass, np.float32)},
        shapes=shapes, children=children, proxy=np.array(proxies, np.int32).reshape(-1, 2),
        proxy_first=first, cfg=SimpleNamespace(friction=0.7, contact_threshold=0.02))


ID = (0, 0, 0, 1)


def test_one_child_at_identity_is_the_plain_hull():
    shapes, ch = with_compounds([((0, 0, 0), ID)])
    pos = [[0, 0, 0], [0.1, 0.97, -0.05]]
    q = [ID, [0, 0.0998, 0, 0.995]]
    q[1] = list(np.array(q[1]) / np.linalg.norm(q[1]))
    plain = snapshot(shapes, ch, [0, 0], pos, q, [0, 1])
    comp = snapshot(shapes, ch, [0, 1], pos, q, [0, 1])
    pairs = np.array([[0, 1]], np.int32)
    assert np.array_equal(expand_compound_pairs(comp, pairs), [[0, 1], [0, 2]])
    a, b = oracle_contacts(plain, pairs), oracle_contacts(comp, pairs)
    assert len(a) == len(b) == 1
    assert b["pair_index"][0] == 1 and b["body_b"][0] == 1
    for k in ("pos", "normal", "body_a", "body_b", "coeffs"):
        assert np.array_equal(a[k], b[k]), k


def test_two_children_are_two_bodies():
    """A dumbbell of two cubes at x = -0.75 and 0.75 on a static cube: the
    contacts of two separate cubes there, named after the compound."""
    shapes, ch = with_compounds([((-0.75, 0, 0), ID), ((0.75, 0, 0), ID)])
    ground = (0, -0.48, 0)
    comp = snapshot(shapes, ch, [0, 1], [ground, [0, 0.5, 0]], [ID, ID], [0, 1])
    sep = snapshot(shapes, ch, [0, 0, 0], [ground, [-0.75, 0.5, 0], [0.75, 0.5, 0]],
                   [ID, ID, ID], [0, 1, 1])
    c = oracle_contacts(comp, np.array([[0, 1]], np.int32))
    s = oracle_contacts(sep, np.array([[0, 1], [0, 2]], np.int32))
    assert len(c) == len(s) == 2
    assert np.array_equal(c["pos"], s["pos"]) and np.array_equal(c["normal"], s["normal"])
    assert list(c["body_b"]) == [1, 1] and list(s["body_b"]) == [1, 2]
    assert list(c["pair_index"]) == [1, 2]


def test_child_transforms_compose():
    ang = np.deg2rad(90)
    qb = np.array([0, np.sin(ang / 2), 0, np.cos(ang / 2)], np.float32)   # 90 deg about y
    qc = np.array([np.sin(ang / 4), 0, 0, np.cos(ang / 4)], np.float32)   # 45 deg about x
    ch = np.zeros(1, O.CHILD_DTYPE)
    ch[0]["pos"][:3] = (1.0, 0.0, 0.0)
    ch[0]["orn"] = qc
    pos, quat = O.compound_proxies(np.array([[0, 0]]), ch, np.array([[2.0, 1.0, 0.0]]), qb[None])
    # the body's 90-degree turn about y takes the child's offset (1, 0, 0) to (0, 0, -1)
    assert np.allclose(pos[1], [2.0, 1.0, -1.0, 0.0], atol=1e-6)
    # composed rotation: body then child, as a quaternion product
    x1, y1, z1, w1 = qb
    x2, y2, z2, w2 = qc
    ref = [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
           w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2]
    assert np.allclose(quat[1], ref, atol=1e-6)


def test_static_pairs_and_plain_pairs_are_not_expanded():
    shapes, ch = with_compounds([((0, 0, 0), ID), ((1, 0, 0), ID)])
    s = snapshot(shapes, ch, [1, 1, 0], [[0, 0, 0], [0, 1, 0], [3, 0, 0]], [ID] * 3, [0, 0, 1])
    pairs = np.array([[0, 1], [0, 2], [1, 2]], np.int32)
    # (0, 1): two static compounds, nothing; (0, 2): 2 x 1; (1, 2): 2 x 1
    assert np.array_equal(expand_compound_pairs(s, pairs)[3:], [[3, 2], [4, 2], [5, 2], [6, 2]])
