"""mi_rb_register_convex_polyhedron (CLPhysicsDemo::registerConvexPolyhedron(
btConvexUtility*, bool), CLPhysicsDemo.cpp:331-410): the reference's own
polyhedral features (tests/golden/hull_features.npz, from
btConvexUtility/btConvexHullComputer) are registered as they are, in the
reference's order, and a pile of bodies of those shapes steps like the
oracle."""
import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd.world import World, reference_config
from oracle.step import Snapshot, gpu_order, oracle_step
from test_hull_reference import load_fixture

import oracle as O

pytestmark = pytest.mark.gpu


def new_world():
    return World(reference_config(max_bodies=256, max_shapes=16, max_hull_vertices=1024,
                                  max_pairs=8192, max_contacts=4096, max_pairs_per_body=64,
                                  solver_iterations=6))


def test_features_are_stored_as_given():
    fx = load_fixture()
    w = new_world()
    shapes = [w.add_polyhedron(v, pl, fi, ind, ed) for _, v, pl, fi, ind, ed in fx]
    assert shapes == list(range(len(fx)))
    hulls = w.read(L.BUF_HULLS, O.HULL_DTYPE)
    verts = w.read(L.BUF_VERTS, np.float32, (4,))
    planes = w.read(L.BUF_PLANES, np.float32, (4,))
    fidx = w.read(L.BUF_FACE_INDEX, np.int32, (2,))
    ind = w.read(L.BUF_INDICES, np.int32)
    edges = w.read(L.BUF_EDGES, np.float32, (4,))
    for k, (_, v, pl, fi, ii, ed) in enumerate(fx):
        h = hulls[k]
        vo, nv = h["vertex_offset"], h["num_vertices"]
        fo, nf = h["face_offset"], h["num_faces"]
        eo, ne = h["edge_offset"], h["num_edges"]
        assert np.array_equal(verts[vo:vo + nv, :3], v)
        assert np.array_equal(planes[fo:fo + nf], pl)
        assert np.array_equal(fidx[fo:fo + nf, 1], fi[:, 1])
        assert np.array_equal(fidx[fo:fo + nf, 0] - h["index_offset"], fi[:, 0])
        assert np.array_equal(ind[h["index_offset"]:h["index_offset"] + len(ii)], ii)
        assert np.array_equal(edges[eo:eo + ne, :3], ed)
        # centre = vertex average, AABB of the vertices (CLPhysicsDemo.cpp:361-405)
        assert np.allclose(h["center"][:3], v.mean(0), atol=1e-6)
        assert np.array_equal(h["aabb_min"][:3], v.min(0))
        assert np.array_equal(h["aabb_max"][:3], v.max(0))
        assert 0.0 < h["center"][3] < np.abs(v - v.mean(0)).max()
    w.close()


def test_bad_features_are_rejected():
    _, v, pl, fi, ind, ed = load_fixture()[0]
    w = new_world()
    bad = ind.copy()
    bad[3] = len(v)  # index out of range
    with pytest.raises(L.MiRbError) as e:
        w.add_polyhedron(v, pl, fi, bad, ed)
    assert e.value.status == L.INVALID_ARG
    short = fi.copy()
    short[0, 1] = 2  # a face of two vertices
    with pytest.raises(L.MiRbError) as e:
        w.add_polyhedron(v, pl, short, ind, ed)
    assert e.value.status == L.INVALID_ARG
    many = np.tile(v, (9, 1))  # 72 vertices
    with pytest.raises(L.MiRbError) as e:
        w.add_polyhedron(many, pl, fi, ind, ed)
    assert e.value.status == L.CAPACITY
    w.close()


def test_pile_of_reference_shapes_steps_like_the_oracle():
    fx = load_fixture()
    w = new_world()
    shapes = [w.add_polyhedron(v, pl, fi, ind, ed) for _, v, pl, fi, ind, ed in fx]
    ground = w.add_hull(np.array([[x, y, z] for x in (-8, 8) for y in (-1, 0) for z in (-8, 8)],
                                 np.float32))
    rng = np.random.default_rng(3)
    for i in range(60):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        pos = (rng.uniform(-3, 3), 0.8 + 0.9 * (i // 12), rng.uniform(-3, 3))
        w.add_body(1.0 + rng.uniform(0, 2), pos, q, shapes[i % len(shapes)])
    w.add_body(0.0, (0, 0, 0), (0, 0, 0, 1), ground)
    w.upload()
    w.step(40)
    snap = Snapshot(w)
    w.step_partial(L.STAGE_COLOR)
    order = gpu_order(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    assert len(contacts) > 20
    w.set_state(**{k: snap.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    s = w.get_state()
    ref = oracle_step(snap, order, contacts)
    for k in ("pos", "linvel", "angvel"):
        assert np.allclose(s[k][:, :3], ref[k][:, :3], rtol=1e-4, atol=1e-4), k
    assert w.stats()["overflow_flags"] == 0
    w.close()
