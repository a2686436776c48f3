"""The library's snapshot format ("MIRB", mi_rb_save_state / mi_rb_load_state):
a world saved mid-run and loaded into a fresh world continues bit for bit
like the original uploaded again from the same state (upload derives the
grid and the solver tiles from the bodies' positions, so the comparison is
with the original re-uploaded at the save), for hulls, spheres, planes,
compounds and concave meshes."""
import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd import scenes
from experiments_amd.world import World, reference_config

pytestmark = pytest.mark.gpu


def cfg():
    return reference_config(max_bodies=512, max_shapes=16, max_pairs=1 << 14,
                            max_contacts=1 << 14, max_pairs_per_body=64, solver_iterations=6)


def mixed_world():
    rng = np.random.default_rng(11)
    w = World(cfg())
    cube = w.add_hull(scenes.cube_points(0.5))
    ground = w.add_hull(scenes.box_points(10.0, 1.0, 10.0))
    ball = w.add_sphere(0.4)
    floor = w.add_plane((0, 1, 0), 3.0)   # under the ground box: only the balls reach it
    ident = (0, 0, 0, 1)
    dumbbell = w.add_compound([cube, cube], [(-0.8, 0, 0), (0.8, 0, 0)], [ident, ident])
    w.add_body(0.0, (0, -1, 0), ident, ground)
    w.add_body(0.0, (0, -6, 0), ident, floor)
    for k in range(30):
        q = rng.normal(size=4)
        pos = ((k % 5) * 2.2 - 4.4, 1.0 + (k // 5) * 1.6, 0.0)
        w.add_body(1.0, pos, q / np.linalg.norm(q), [cube, dumbbell][k % 2])
    for k in range(4):
        w.add_body(1.0, (12.0 + k, 2.0, 0.0), ident, ball)
    # a two-triangle mesh away from the rest, a cube dropping on it
    quad = [(-2, 0, -2), (-2, 0, 2), (2, 0, 2), (2, 0, -2)]
    mesh = w.add_concave_mesh(quad, [(0, 1, 2), (0, 2, 3)])
    w.add_body(0.0, (30.0, 0.0, 0.0), ident, mesh)
    w.add_body(1.0, (30.0, 0.8, 0.0), q / np.linalg.norm(q), cube)
    w.upload()
    return w


def test_saved_world_continues_identically(tmp_path):
    a = mixed_world()
    a.step(25)
    path = str(tmp_path / "world.mirb")
    a.save(path)   # (reads the state back to the host first)
    a.upload()     # derive the grid and the tiles from this state, as b will
    b = World(cfg())
    b.load(path)
    b.upload()
    for name, size in (("HULLS", 80), ("VERTS", 16), ("PLANES", 16), ("INV_MASS", 4),
                       ("SHAPE", 4), ("CHILDREN", 48)):
        which = getattr(L, "BUF_" + name)
        ra, rb = a.read(which, np.uint8, (size,)), b.read(which, np.uint8, (size,))
        assert len(ra) > 0 and np.array_equal(ra, rb), name
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.step(15)
    b.step(15)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert a.stats()["num_contacts"] == b.stats()["num_contacts"] > 0
    a.close()
    b.close()


def test_corrupt_snapshot_leaves_the_world_as_it_was(tmp_path):
    a = mixed_world()
    a.step(10)
    good = tmp_path / "world.mirb"
    a.save(str(good))
    raw = good.read_bytes()
    n = a.body_count()
    before = a.get_state()
    # a truncated file: refused, nothing replaced
    bad = tmp_path / "cut.mirb"
    bad.write_bytes(raw[:len(raw) // 2])
    with pytest.raises(L.MiRbError) as e:
        a.load(str(bad))
    assert e.value.status == L.INVALID_ARG
    assert a.body_count() == n
    after = a.get_state()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    a.step(10)
    s = a.get_state()
    assert s["pos"].shape[0] == n and np.all(np.isfinite(s["pos"]))
    a.close()
