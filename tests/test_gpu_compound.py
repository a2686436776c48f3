"""Compound shapes on the GPU path: L-shapes and dumbbells of cubes with
plain cubes on a static ground.  The expanded pair list (the broadphase
pairs, then one pair per child combination) is bit-exact, contacts match
the oracle per pair, one step matches within 1e-4, and the pile settles."""
import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes
from experiments_amd.world import World, reference_config
from oracle.step import (Snapshot, expand_compound_pairs, gpu_order, oracle_broadphase,
                         oracle_contacts, oracle_step)

pytestmark = pytest.mark.gpu


def quat_axis(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.array([*(a * np.sin(ang / 2)), np.cos(ang / 2)], np.float32)


def compound_pile(seed=4, nx=5, nz=5, layers=3):
    rng = np.random.default_rng(seed)
    w = World(reference_config(max_bodies=1024, max_shapes=32, max_pairs=1 << 16,
                               max_contacts=1 << 16, max_pairs_per_body=64, solver_iterations=8))
    cube = w.add_hull(scenes.cube_points(0.5))
    ground = w.add_hull(scenes.box_points(20.0, 1.0, 20.0))
    ident = (0, 0, 0, 1)
    ell = w.add_compound([cube, cube, cube], [(0, 0, 0), (1.0, 0, 0), (0, 1.0, 0)], [ident] * 3)
    bell = w.add_compound([cube, cube], [(-0.8, 0, 0), (0.8, 0, 0)],
                          [quat_axis((1, 0, 0), 0.4), ident])
    kinds = [ell, bell, cube]
    w.add_body(0.0, (0, -1, 0), ident, ground)
    for ly in range(layers):
        for i in range(nx):
            for j in range(nz):
                q = quat_axis(rng.normal(size=3), rng.uniform(0, 0.5))
                pos = ((i - (nx - 1) / 2) * 3.0, 1.2 + 2.6 * ly, (j - (nz - 1) / 2) * 3.0)
                w.add_body(1.0 + (i + j) % 2, pos, q, kinds[(i + j + ly) % 3])
    w.upload()
    return w


@pytest.fixture(scope="module")
def settled():
    w = compound_pile()
    w.step(50)
    yield w
    w.close()


def test_pairs_contacts_match_oracle(settled):
    w = settled
    s = Snapshot(w)
    assert len(s.proxy) > 0
    w.step_partial(L.STAGE_CONSTRAINTS)
    _, _, _, pairs = oracle_broadphase(s)
    expanded = expand_compound_pairs(s, pairs)
    assert len(expanded) > len(pairs)
    assert np.array_equal(w.pairs(), expanded)
    ref = oracle_contacts(s, pairs)
    got = w.contacts(O.CONTACT_DTYPE)
    assert len(got) == len(ref) > 20
    for k in ("pair_index", "body_a", "body_b"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(got["normal"][:, 3], ref["normal"][:, 3])
    n = ref["normal"][:, 3].astype(int)
    for k in range(len(ref)):
        assert np.allclose(got["pos"][k, :n[k]], ref["pos"][k, :n[k]], rtol=1e-5, atol=1e-6), k
    assert np.allclose(got["normal"][:, :3], ref["normal"][:, :3], rtol=1e-5, atol=1e-6)
    # contacts from child pairs name the compound bodies
    assert (got["pair_index"] >= len(pairs)).any()
    assert np.abs(got["body_b"]).max() < len(s.sh["shape"])


def test_one_step_matches_oracle(settled):
    w = settled
    s = Snapshot(w)
    w.step_partial(L.STAGE_COLOR)
    order = gpu_order(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    ref = oracle_step(s, order, contacts)
    for k in ("linvel", "angvel", "pos", "quat"):
        n = 4 if k == "quat" else 3
        assert np.allclose(g[k][:, :n], ref[k][:, :n], rtol=1e-4, atol=1e-4), k


def test_compound_pile_settles():
    w = compound_pile(seed=9)
    w.use_graph(True)
    ke = []
    for _ in range(8):
        w.step(30)
        v = w.get_state()["linvel"][1:, :3]
        ke.append(0.5 * float((v ** 2).sum()))
    st = w.get_state()
    assert w.stats()["overflow_flags"] == 0
    assert np.all(np.isfinite(st["pos"])) and np.all(np.isfinite(st["quat"]))
    # every body's centre stays above the ground top (y = 0): nothing falls through
    assert st["pos"][1:, 1].min() > 0.3
    assert ke[-1] < 0.05 * max(ke), ke
    w.close()
