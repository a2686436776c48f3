"""Stage-by-stage parity of the GPU step with the oracle on identical state.

Integer stages (cell keys, sorted key/value pairs, pair list, colouring) are
bit-exact; contacts and constraints are compared per pair (1e-5 relative);
one full step from the same FP32 state within 1e-4."""
import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes
from oracle.step import Snapshot, gpu_order, oracle_broadphase, oracle_contacts, oracle_step

pytestmark = pytest.mark.gpu


def lattice():
    return scenes.lattice_cubes(6, 4, 6, spacing=1.0, gap=0.0)


def hulls_dense():
    return scenes.random_hulls(n=1500, box=18.0, pool=(64, 64), steps=1, iterations=6)


def piles():
    return scenes.batched_piles(n_scenes=24, lo=8, hi=64, pool8=32, iterations=6)


SCENES = {"lattice": lattice, "hulls": hulls_dense, "piles": piles}


def settled(name, steps):
    """World advanced `steps` steps so bodies touch and contacts exist."""
    w = scenes.build_world(SCENES[name](), max_pairs_per_body=64)
    if steps:
        w.step(steps)
    return w


@pytest.fixture(scope="module", params=list(SCENES))
def world(request):
    steps = {"lattice": 30, "hulls": 0, "piles": 40}[request.param]
    w = settled(request.param, steps)
    yield request.param, w
    w.close()


def test_broadphase_bit_exact(world):
    name, w = world
    s = Snapshot(w)
    w.step_partial(L.STAGE_PAIRS)
    mn, mx, skv, pairs = oracle_broadphase(s)
    gmn, gmx = w.aabbs()
    assert np.array_equal(gmn, mn) and np.array_equal(gmx, mx)
    assert np.array_equal(w.sorted_kv(), skv)
    gp = w.pairs()
    assert np.array_equal(gp, pairs)
    # and the pair set is the brute-force overlap set of the same AABBs
    brute = O.find_pairs_brute(mn, mx, scene=s.sh["scene"])
    assert np.array_equal(gp, brute)
    assert len(gp) > 0
    assert w.stats()["overflow_flags"] == 0


@pytest.mark.parametrize("sort_tile,digit_bits", [(1024, 7), (4096, 8), (4096, 6), (1024, 4)])
def test_body_sort_tunings_bit_exact(world, sort_tile, digit_bits):
    """Every tile size and digit width the tuner may pick for the body-key
    sort gives the oracle's sorted pairs and pair list."""
    name, w = world
    s = Snapshot(w)
    w.set_tuning(sort_tile=sort_tile, sort_digit_bits=digit_bits)
    try:
        w.step_partial(L.STAGE_PAIRS)
        _, _, skv, pairs = oracle_broadphase(s)
        assert np.array_equal(w.sorted_kv(), skv)
        assert np.array_equal(w.pairs(), pairs)
    finally:
        w.set_tuning()
    assert w.stats()["overflow_flags"] == 0


def test_narrowphase_contacts(world):
    name, w = world
    s = Snapshot(w)
    w.step_partial(L.STAGE_CONSTRAINTS)
    _, _, _, pairs = oracle_broadphase(s)
    ref = oracle_contacts(s, pairs)
    got = w.contacts(O.CONTACT_DTYPE)
    assert len(got) == len(ref) > 0
    assert np.array_equal(got["pair_index"], ref["pair_index"])
    assert np.array_equal(got["body_a"], ref["body_a"]) and np.array_equal(got["body_b"], ref["body_b"])
    assert np.array_equal(got["normal"][:, 3], ref["normal"][:, 3])  # per-pair point counts
    n = ref["normal"][:, 3].astype(int)
    for k in range(len(ref)):
        assert np.allclose(got["pos"][k, :n[k]], ref["pos"][k, :n[k]], rtol=1e-5, atol=1e-6), k
    assert np.allclose(got["normal"][:, :3], ref["normal"][:, :3], rtol=1e-5, atol=1e-6)
    exact = np.mean([np.array_equal(got[k], ref[k]) for k in range(len(ref))])
    print("bit-exact contact fraction %.4f" % exact)


def test_colouring_valid(world):
    name, w = world
    w.step_partial(L.STAGE_COLOR)
    cons = w.read(L.BUF_CONSTRAINTS, np.uint8, (176,)).view(O.CONSTRAINT_DTYPE).reshape(-1)
    color = w.read(L.BUF_COLORS, np.int32)
    order = w.read(L.BUF_SOLVE_ORDER, np.int32)
    tiles = w.read(L.BUF_PHASES, np.int32, (4,))
    im = w.read(L.BUF_INV_MASS, np.float32)
    assert sorted(order.tolist()) == list(range(len(cons)))
    for first, end, phase, ncol in tiles:
        rows = order[first:end]
        assert np.all(np.diff(color[rows]) >= 0)  # colour-major within the tile
        for c in np.unique(color[rows]):
            rr = rows[color[rows] == c]
            dyn = [b for r in rr for b in {int(cons["body_a"][r]), int(cons["body_b"][r])} if im[b] != 0]
            assert len(dyn) == len(set(dyn))
    # tiles of one phase never share a dynamic body
    for ph in range(9):
        seen = {}
        for t, (first, end, phase, _) in enumerate(tiles):
            if phase != ph:
                continue
            for r in order[first:end]:
                for b in (int(cons["body_a"][r]), int(cons["body_b"][r])):
                    if im[b] != 0:
                        assert seen.setdefault(b, t) == t


def test_one_step_matches_oracle(world):
    name, w = world
    s = Snapshot(w)
    w.step_partial(L.STAGE_COLOR)
    order = gpu_order(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    ref = oracle_step(s, order, contacts)
    for k in ("linvel", "angvel", "pos", "quat"):
        assert np.allclose(g[k][:, :3 if k != "quat" else 4], ref[k][:, :3 if k != "quat" else 4],
                           rtol=1e-4, atol=1e-4), k
    gl = w.read(L.BUF_CONSTRAINTS, np.uint8, (176,)).view(O.CONSTRAINT_DTYPE).reshape(-1)
    assert np.allclose(gl["lambda"], ref["cons"]["lambda"], rtol=1e-4, atol=1e-5)
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})


def test_graph_replay_equals_eager():
    a = scenes.build_world(lattice())
    b = scenes.build_world(lattice())
    b.use_graph(True)
    a.step(12)
    b.step(12)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


def test_long_run_statistics():
    """Chaotic contact dynamics are judged statistically: a 6x4x6 stack on the
    ground stays finite, settles (kinetic energy decays) and keeps penetration
    bounded."""
    w = scenes.build_world(lattice())
    w.step(120)
    s = w.get_state()
    assert np.all(np.isfinite(s["pos"])) and np.all(np.isfinite(s["quat"]))
    ke = 0.5 * np.sum(s["linvel"][1:, :3] ** 2)
    assert ke / (len(s["pos"]) - 1) < 0.05
    assert s["pos"][1:, 1].min() > 0.5 - 0.06   # cubes rest on the ground top (y = 0)
    w.step_partial(L.STAGE_CONSTRAINTS)
    c = w.contacts(O.CONTACT_DTYPE)
    depth = np.concatenate([c["pos"][k, :int(c["normal"][k, 3]), 3] for k in range(len(c))])
    assert depth.min() > -0.05
    w.close()


def test_per_body_pair_overflow_is_reported():
    """A per-body list shorter than the overlaps sets the sticky overflow flag;
    the emitted pairs stay a subset of the true overlap set."""
    w = scenes.build_world(hulls_dense(), max_pairs_per_body=2)
    s = Snapshot(w)
    w.step_partial(L.STAGE_PAIRS)
    assert w.stats()["overflow_flags"] & L.OVERFLOW_BODY_PAIRS
    mn, mx, _, pairs = oracle_broadphase(s)
    gp = w.pairs()
    ref = set(map(tuple, pairs.tolist()))
    assert set(map(tuple, gp.tolist())) <= ref
    assert np.all(np.bincount(gp[:, 0], minlength=len(mn)) <= 2 + s.is_large.sum())
    w.close()


def test_narrowphase_contacts_settled_piles():
    """Per-pair contact counts, pair order and points of the narrowphase on
    1024 batched piles after 200 steps (cubes and 8-vertex hulls resting on
    each other), against the oracle on the same state."""
    w = scenes.build_world(scenes.batched_piles(n_scenes=1024), max_pairs_per_body=64)
    w.use_graph(True)
    w.step(200)
    w.use_graph(False)
    s = Snapshot(w)
    w.step_partial(L.STAGE_CONSTRAINTS)
    _, _, _, pairs = oracle_broadphase(s)
    ref = oracle_contacts(s, pairs)
    got = w.contacts(O.CONTACT_DTYPE)
    w.close()
    gk = dict(zip(got["pair_index"].tolist(), range(len(got))))
    rk = dict(zip(ref["pair_index"].tolist(), range(len(ref))))
    only_gpu = sorted(set(gk) - set(rk))
    only_ref = sorted(set(rk) - set(gk))
    assert not only_gpu and not only_ref, (len(only_gpu), len(only_ref), only_gpu[:5], only_ref[:5])
    n_got, n_ref = got["normal"][:, 3], ref["normal"][:, 3]
    bad = np.nonzero(n_got != n_ref)[0]
    assert len(bad) == 0, [(int(got["pair_index"][k]), got["normal"][k].tolist(),
                            ref["normal"][k].tolist()) for k in bad[:5]]
    assert np.allclose(got["normal"][:, :3], ref["normal"][:, :3], rtol=1e-5, atol=1e-6)
    n = n_ref.astype(int)
    for k in range(len(ref)):
        assert np.allclose(got["pos"][k, :n[k]], ref["pos"][k, :n[k]], rtol=1e-5, atol=1e-6), k


def test_hull_inscribed_radius_is_inside():
    """The ball the narrowphase uses to skip SAT axes (HullInfo center.w about
    the vertex-average centre) lies inside every registered hull: its radius
    is at most the distance from the centre to every facet plane of the
    hull's vertices (scipy's Qhull facets)."""
    from scipy.spatial import ConvexHull
    w = scenes.build_world(scenes.batched_piles(n_scenes=256))
    hulls = w.read(L.BUF_HULLS, np.uint8, (80,)).view(O.HULL_DTYPE).reshape(-1)
    verts = w.read(L.BUF_VERTS, np.float32, (4,))
    w.close()
    assert len(hulls) > 100
    for h in hulls:
        v = verts[h["vertex_offset"]:h["vertex_offset"] + h["num_vertices"], :3].astype(np.float64)
        c = h["center"][:3].astype(np.float64)
        facet = min(-(eq[:3] @ c + eq[3]) for eq in ConvexHull(v).equations)
        assert 0.0 < h["center"][3] <= facet, (h["center"][3], facet)
