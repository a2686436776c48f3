"""A .bullet scene loaded into a world: what is registered equals the
oracle's restatement of the reference's btBulletDataExtractor, and the scene
steps (boxes come to rest on the ground)."""
import os

import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd.world import World, reference_config, shape_arrays
from oracle import bullet_file as B

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def new_world():
    return World(reference_config(max_bodies=4096, max_shapes=16, max_hull_vertices=1 << 12,
                                  max_pairs=1 << 16, max_contacts=1 << 16,
                                  max_pairs_per_body=64))


@pytest.mark.parametrize("name", ["stack_le32.bullet", "stack_be64.bullet"])
@pytest.mark.parametrize("keep_static", [False, True])
def test_registered_scene_equals_oracle(name, keep_static):
    path = os.path.join(GOLD, name)
    _, shapes, bodies = B.read_scene(path)
    hulls, regs = B.registrations(shapes, bodies, keep_static)
    w = new_world()
    summary = w.load_bullet(path, keep_static)
    assert summary["shapes"] == 5 and summary["unsupported_shapes"] == 2
    assert summary["shapes_registered"] == 3
    assert summary["bodies_registered"] + summary["ground_bodies"] == len(regs)
    w.upload()
    st = w.get_state()
    pos = np.array([r[1] for r in regs], np.float32)
    assert np.array_equal(st["pos"][:, :3], pos)
    q = np.array([r[2] for r in regs], np.float32)
    dq = np.minimum(np.abs(st["quat"] - q).max(1), np.abs(st["quat"] + q).max(1))
    assert dq.max() < 1e-6
    inv = w.read(L.BUF_INV_MASS, np.float32)
    mass = np.array([r[0] for r in regs], np.float32)
    expect = np.where(mass > 0, np.float32(1.0) / np.where(mass > 0, mass, 1), 0).astype(np.float32)
    assert np.allclose(inv, expect, rtol=1e-6, atol=0)
    assert np.array_equal(w.read(L.BUF_SHAPE, np.int32), np.array([r[3] for r in regs], np.int32))
    # every hull's vertices are the oracle's scaled points, snapped to the
    # hull computer's grid of 10216 steps per axis of their AABB
    sa = shape_arrays(w)
    assert len(sa["hulls"]) == len(hulls)
    for h, (pts, scale) in zip(sa["hulls"], hulls):
        v = sa["verts"][h["vertex_offset"]:h["vertex_offset"] + h["num_vertices"], :3]
        ref = (pts * scale).astype(np.float32)
        step = (ref.max(0) - ref.min(0)) / 10216.0
        near = (np.abs(v[:, None, :] - ref[None, :, :]) <= step + 1e-7).all(2).any(1)
        assert len(v) >= 4 and near.all()
    w.close()


def test_bullet_scene_steps_to_rest():
    w = new_world()
    w.load_bullet(os.path.join(GOLD, "stack_le32.bullet"), keep_static=True)
    w.upload()
    w.use_graph(True)
    w.step(150)
    st = w.get_state()
    s = w.stats()
    assert s["overflow_flags"] == 0 and s["num_contacts"] > 0
    assert np.all(np.isfinite(st["pos"])) and np.all(np.isfinite(st["quat"]))
    boxes = st["pos"][:36]
    # the lowest layer of the 3 x 3 x 4 stacks rests on the ground top (y = 0)
    low = np.sort(boxes[:, 1])[:9]
    assert np.all(np.abs(low - 0.5) < 0.05), low
    assert np.abs(st["linvel"][:44, :3]).max() < 0.5
    w.close()
