"""Oracle concave meshes (registerConcaveMeshShape,
btGpuNarrowPhaseAndSolver.cpp:264-333; the per-triangle polyhedron and the
triangle AABB cull of findConcaveSeparatingAxisKernel, sat.cl:970-1191)
reduced to known cases: a cube resting inside one triangle of a flat quad
mesh has the contacts of the same cube on a box whose top face is the mesh
plane; the cull keeps exactly the triangles whose box overlaps the cube's;
a cube clear of every triangle box gets no child pair."""
import numpy as np

import oracle as O
from oracle.step import expand_compound_pairs, is_mesh, oracle_contacts
from test_oracle_compound import snapshot
from test_oracle_narrowphase import cube_shapes


def triangle_hull(v):
    """The polyhedron of one triangle: 3 vertices; the front face (normal
    (v1-v0) x (v2-v0) normalised), the back face, and per edge (v_i, v_prev)
    a 2-index face whose plane holds the edge, perpendicular to the
    triangle; unique edges v1-v0, v2-v1, v0-v2."""
    v = np.asarray(v, np.float64)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    planes = [[*n, -n @ v[0]], [*(-n), n @ v[0]]]
    faces = [[0, 1, 2], [2, 1, 0]]
    prev = 2
    for i in range(3):
        e = np.cross(n, v[prev] - v[i])
        e /= np.linalg.norm(e)
        planes.append([*e, -e @ v[i]])
        faces.append([i, prev])
        prev = i
    edges = [v[1] - v[0], v[2] - v[1], v[0] - v[2]]
    return v, faces, planes, edges


def build(tris, ground_box=None):
    """Shapes: 0 the unit cube, 1 a mesh of the given triangles (or, with
    ground_box=(hx, hy, hz), a box), then the triangle hulls."""
    cube = cube_shapes()
    H = [cube.hulls[0]]
    verts, planes = list(cube.verts), list(cube.planes)
    fi, ind, edges = [list(x) for x in cube.face_index], list(cube.indices), list(cube.edges)

    def add(v, faces, pl, ed, aabb):
        h = np.zeros(1, O.HULL_DTYPE)[0]
        h["vertex_offset"], h["num_vertices"] = len(verts), len(v)
        h["face_offset"], h["num_faces"] = len(planes), len(faces)
        h["edge_offset"], h["num_edges"] = len(edges), len(ed)
        h["index_offset"], h["num_indices"] = len(ind), sum(len(f) for f in faces)
        h["center"][:3] = np.mean(v, axis=0)
        h["aabb_min"][:3], h["aabb_max"][:3] = aabb
        verts.extend([[*p, 0.0] for p in v])
        for f, p in zip(faces, pl):
            fi.append([len(ind), len(f)])
            ind.extend(f)
            planes.append(p)
        edges.extend([[*e, 0.0] for e in ed])
        H.append(h)

    ch = np.zeros(max(len(tris), 1), O.CHILD_DTYPE)
    if ground_box is not None:
        s = np.asarray(ground_box, np.float32)
        add(cube.verts[:, :3] * 2 * s, [list(ind[a:a + c]) for a, c in cube.face_index],
            [[*p[:3], p[3] * 2 * s[np.argmax(np.abs(p[:3]))]] for p in cube.planes],
            [e[:3] for e in cube.edges], (-s, s))
    else:
        m = np.zeros(1, O.HULL_DTYPE)[0]
        m["num_edges"], m["edge_offset"], m["num_indices"] = len(tris), 0, 1
        allv = np.asarray(tris, np.float32).reshape(-1, 3)
        m["aabb_min"][:3], m["aabb_max"][:3] = allv.min(0), allv.max(0)
        H.append(m)
        for t, tri in enumerate(tris):
            v, faces, pl, ed = triangle_hull(tri)
            ch[t]["orn"][3] = 1.0
            ch[t]["shape"] = len(H)
            add(v, faces, pl, ed, (v.min(0), v.max(0)))
    shapes = O.Shapes(np.array(H, O.HULL_DTYPE), np.array(verts, np.float32),
                      np.array(planes, np.float32), np.array(fi), np.array(ind),
                      np.array(edges, np.float32))
    return shapes, ch


def quad_mesh(half=10.0, y=0.0):
    a, b, c, d = (-half, y, -half), (-half, y, half), (half, y, half), (half, y, -half)
    return [[a, b, c], [a, c, d]]   # counter-clockwise seen from +y: normals up


ID = (0, 0, 0, 1)
ROT = (0, float(np.sin(0.15)), 0, float(np.cos(0.15)))   # 0.3 rad about y


def test_triangle_hull_planes():
    v, faces, pl, ed = triangle_hull([(0, 0, 0), (1, 0, 0), (0, 0, -1)])
    pl = np.asarray(pl)
    assert np.allclose(pl[0], [0, 1, 0, 0]) and np.allclose(pl[1], -pl[0])
    centroid = v.mean(0)
    for (i, j), p in zip(faces[2:], pl[2:]):
        assert abs(p[:3] @ v[i] + p[3]) < 1e-12 and abs(p[:3] @ v[j] + p[3]) < 1e-12
        assert abs(p[:3] @ pl[0, :3]) < 1e-12            # perpendicular to the triangle
        assert p[:3] @ centroid + p[3] < 0                # outward


def test_cube_in_one_triangle_is_the_cube_on_a_box():
    pos = [[0, 0, 0], [4.0, 0.49, -3.0]]
    mesh_shapes, ch = build(quad_mesh())
    box_shapes, _ = build([], ground_box=(10.0, 0.5, 10.0))
    mesh = snapshot(mesh_shapes, ch, [1, 0], pos, [ID, ROT], [0, 1])
    box = snapshot(box_shapes, ch[:0], [1, 0], [[0, -0.5, 0], pos[1]], [ID, ROT], [0, 1])
    assert is_mesh(mesh.sh["hulls"][1]) and not is_mesh(box.sh["hulls"][1])
    pairs = np.array([[0, 1]], np.int32)
    # both triangles' boxes are the whole quad: two child pairs (proxies 2 and 3)
    assert np.array_equal(expand_compound_pairs(mesh, pairs), [[0, 1], [2, 1], [3, 1]])
    a, b = oracle_contacts(mesh, pairs), oracle_contacts(box, pairs)
    assert len(a) == len(b) == 1 and a["normal"][0, 3] == b["normal"][0, 3] == 4
    assert a["pair_index"][0] == 2 and a["body_a"][0] == 0 and a["body_b"][0] == 1
    assert np.allclose(a["pos"], b["pos"], atol=1e-6) and np.allclose(a["normal"], b["normal"], atol=1e-6)
    # the depth of each point: 0.01 into the plane y = 0
    assert np.allclose(a["pos"][0, :4, 3], -0.01, atol=1e-5)


def test_triangle_cull():
    # a 4 x 4 grid of 2 x 2 quads (32 triangles) on y = 0, the unit cube over
    # the middle of quad (1, 2): its box overlaps that quad's two triangles only
    tris = []
    for i in range(4):
        for j in range(4):
            a, b = (2 * i, 0, 2 * j), (2 * i, 0, 2 * j + 2)
            c, d = (2 * i + 2, 0, 2 * j + 2), (2 * i + 2, 0, 2 * j)
            tris += [[a, b, c], [a, c, d]]
    shapes, ch = build(tris)
    s = snapshot(shapes, ch, [1, 0], [[0, 0, 0], [3.0, 0.45, 5.0]], [ID, ID], [0, 1])
    got = expand_compound_pairs(s, np.array([[0, 1]], np.int32))[1:]
    quad = 2 * (1 * 4 + 2)
    assert np.array_equal(got, [[2 + quad, 1], [3 + quad, 1]])
    # lifted clear of the plane: no triangle box is reached
    s = snapshot(shapes, ch, [1, 0], [[0, 0, 0], [3.0, 0.6, 5.0]], [ID, ID], [0, 1])
    assert len(expand_compound_pairs(s, np.array([[0, 1]], np.int32))) == 1


def test_static_body_on_a_mesh_gets_no_pairs():
    shapes, ch = build(quad_mesh())
    s = snapshot(shapes, ch, [1, 0], [[0, 0, 0], [0, 0.4, 0]], [ID, ID], [0, 0])
    assert len(expand_compound_pairs(s, np.array([[0, 1]], np.int32))) == 1
