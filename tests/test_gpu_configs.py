"""The benchmark configurations at their full size on the GPU, each advanced
into its contact-rich state and then checked against the oracle on that
exact state: the broadphase bit for bit (world AABBs, sorted cell keys and
bodies, the pair list), the contacts per pair (point counts exact, points
and normals within 1e-5), and one full step within 1e-4 on the GPU's
constraint order.

  * the 16 x 16 x 16 lattice of 4,096 unit cubes on a static ground box,
    after 60 of its 300 steps;
  * the 65,536 random hulls (a third cubes, 8- and 32-vertex hulls) in a
    120^3 box, after 220 of its 500 steps;
  * the 4,096 batched piles (about 500,000 bodies), after 120 steps, the
    step checked scene by scene on each scene's own solver tile.
"""
import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes
from oracle.step import (Snapshot, gpu_order, gpu_tile_orders, oracle_broadphase,
                         oracle_contacts, oracle_step)

pytestmark = pytest.mark.gpu

CONFIGS = {
    "lattice4k": (lambda: scenes.lattice_cubes(16, 16, 16), 60),
    "hulls64k": (lambda: scenes.random_hulls(n=65536, box=120.0, pool=(1024, 1024), steps=500,
                                             iterations=10), 220),
    "batched4k": (lambda: scenes.batched_piles(n_scenes=4096, lo=8, hi=512, iterations=10), 120),
}


@pytest.fixture(scope="module", params=list(CONFIGS))
def settled(request):
    make, steps = CONFIGS[request.param]
    sc = make()
    w = scenes.build_world(sc)
    w.use_graph(True)
    w.step(steps)
    w.use_graph(False)
    print("[config] %s: %d bodies after %d steps" % (request.param, sc.num_bodies, steps),
          flush=True)
    yield request.param, w
    w.close()


def test_broadphase_and_contacts_match_oracle(settled):
    name, w = settled
    s = Snapshot(w)
    w.step_partial(L.STAGE_CONSTRAINTS)
    mn, mx, skv, pairs = oracle_broadphase(s)
    gmn, gmx = w.aabbs()
    assert np.array_equal(gmn, mn) and np.array_equal(gmx, mx)
    assert np.array_equal(w.sorted_kv(), skv)
    assert np.array_equal(w.pairs(), pairs) and len(pairs) > 0
    ref = oracle_contacts(s, pairs)
    got = w.contacts(O.CONTACT_DTYPE)
    assert len(got) == len(ref) > 0
    assert np.array_equal(got["pair_index"], ref["pair_index"])
    assert np.array_equal(got["body_a"], ref["body_a"]) and np.array_equal(got["body_b"], ref["body_b"])
    assert np.array_equal(got["normal"][:, 3], ref["normal"][:, 3])
    n = ref["normal"][:, 3].astype(int)
    pos_ok = np.ones(len(ref), bool)
    for k in range(4):
        m = n > k
        pos_ok[m] &= np.all(np.isclose(got["pos"][m, k], ref["pos"][m, k], rtol=1e-5, atol=1e-6), 1)
    assert pos_ok.all(), np.nonzero(~pos_ok)[0][:10]
    assert np.allclose(got["normal"][:, :3], ref["normal"][:, :3], rtol=1e-5, atol=1e-6)
    exact = np.mean(np.all(got.view(np.uint8).reshape(len(got), -1) ==
                           ref.view(np.uint8).reshape(len(ref), -1), 1))
    print("[config] %s: %d pairs, %d manifolds, %.4f bit-exact" % (name, len(pairs), len(got), exact))
    assert w.stats()["overflow_flags"] == 0
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})


def test_one_step_matches_oracle(settled):
    name, w = settled
    s = Snapshot(w)
    w.step_partial(L.STAGE_COLOR)
    contacts = w.contacts(O.CONTACT_DTYPE)
    scene_tiles = w.params()["tile_mode"] == 1
    order = None if scene_tiles else gpu_order(w)
    segs = gpu_tile_orders(w) if scene_tiles else None
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    ref = oracle_step(s, order, contacts, segments=segs)
    for k in ("linvel", "angvel", "pos", "quat"):
        c = 3 if k != "quat" else 4
        assert np.allclose(g[k][:, :c], ref[k][:, :c], rtol=1e-4, atol=1e-4), k
    gl = w.read(L.BUF_CONSTRAINTS, np.uint8, (176,)).view(O.CONSTRAINT_DTYPE).reshape(-1)
    assert np.allclose(gl["lambda"], ref["cons"]["lambda"], rtol=1e-4, atol=1e-5)
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
