"""Spheres and planes on the GPU path: a pile of spheres in a box of static
planes.  The pair list is bit-exact, every contact matches the oracle's
closed forms (1e-5), one full step matches the oracle within 1e-4, and the
spheres come to rest on the floor plane."""
import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd.world import World, reference_config
from oracle.step import Snapshot, gpu_order, oracle_broadphase, oracle_contacts, oracle_step

pytestmark = pytest.mark.gpu


def sphere_box(n_side=6, layers=5, seed=3):
    """A floor plane y = 0 and four wall planes (|x|, |z| <= 4), spheres of
    radius 0.3-0.6 in layers, a sphere-free cube off to the side."""
    rng = np.random.default_rng(seed)
    w = World(reference_config(max_bodies=1024, max_shapes=32, max_pairs=1 << 16,
                               max_contacts=1 << 16, max_pairs_per_body=64, solver_iterations=8))
    floor = w.add_plane((0, 1, 0), 0.0)
    walls = [w.add_plane(n, -4.0) for n in ((1, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, -1))]
    radii = [0.3, 0.4, 0.5, 0.6]
    sph = [w.add_sphere(r) for r in radii]
    cube = w.add_hull(np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)],
                               np.float32))
    ident = (0, 0, 0, 1)
    w.add_body(0.0, (0, 0, 0), ident, floor)
    for s in walls:
        w.add_body(0.0, (0, 0, 0), ident, s)
    k = 0
    for ly in range(layers):
        for i in range(n_side):
            for j in range(n_side):
                x = (i - (n_side - 1) / 2) * 1.25 + rng.uniform(-0.05, 0.05)
                z = (j - (n_side - 1) / 2) * 1.25 + rng.uniform(-0.05, 0.05)
                q = rng.normal(size=4)
                w.add_body(1.0 + k % 3, (x, 0.65 + 1.3 * ly, z), q / np.linalg.norm(q), sph[k % 4])
                k += 1
    w.add_body(1.0, (10.0, 2.0, 0.0), ident, cube)  # falls through the floor: no hull-plane contact
    w.upload()
    return w, radii, k


@pytest.fixture(scope="module")
def settled():
    w, radii, k = sphere_box()
    w.step(45)
    yield w, radii, k
    w.close()


def test_pairs_and_contacts_match_oracle(settled):
    w, _, _ = settled
    s = Snapshot(w)
    w.step_partial(L.STAGE_CONSTRAINTS)
    _, _, skv, pairs = oracle_broadphase(s)
    assert np.array_equal(w.pairs(), pairs)
    ref = oracle_contacts(s, pairs)
    got = w.contacts(O.CONTACT_DTYPE)
    assert len(got) == len(ref) > 50
    for k in ("pair_index", "body_a", "body_b"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.all(got["normal"][:, 3] == 1)
    assert np.allclose(got["pos"][:, 0], ref["pos"][:, 0], rtol=1e-5, atol=1e-6)
    assert np.allclose(got["normal"][:, :3], ref["normal"][:, :3], rtol=1e-5, atol=1e-6)
    # plane contacts (a static body first) and sphere-sphere contacts both occur
    assert (got["body_a"] <= 0).any() and (got["body_a"] > 0).any()


def test_one_step_matches_oracle(settled):
    w, _, _ = settled
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


def test_spheres_rest_on_the_floor():
    w, radii, k = sphere_box(seed=5)
    w.use_graph(True)
    ke = []
    for _ in range(8):
        w.step(30)
        v = w.get_state()["linvel"][5:5 + k, :3]
        ke.append(0.5 * float((v ** 2).sum()))
    st = w.get_state()
    assert w.stats()["overflow_flags"] == 0
    pos = st["pos"][5:5 + k]
    r = np.array([radii[i % 4] for i in range(k)])
    assert np.all(np.isfinite(pos))
    # nothing sinks through the floor or the walls; the bottom layer touches the floor
    assert np.all(pos[:, 1] - r > -0.05)
    assert np.all(np.abs(pos[:, 0]) + r < 4.05) and np.all(np.abs(pos[:, 2]) + r < 4.05)
    assert np.sum(np.abs(pos[:, 1] - r) < 0.05) >= 20
    # the pile settles (the last spheres may still drop through it or roll):
    # kinetic energy falls to a few percent of its peak, nothing flies up
    v = st["linvel"][5:5 + k, :3]
    assert ke[-1] < 0.05 * max(ke), ke
    assert v[:, 1].max() < 1.0
    # the cube has no contact with the planes (as in the reference): it falls
    assert st["pos"][5 + k, 1] < -10.0
    w.close()
