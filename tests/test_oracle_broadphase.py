"""Oracle body setup and broadphase pinned against closed forms, brute force
and the reference's 5,876-AABB scene (testdata/m_sapKernelArgs.zip)."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def quat_to_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


def hull_table(aabb_min, aabb_max, center=(0, 0, 0)):
    H = np.zeros(len(aabb_min), O.HULL_DTYPE)
    H["aabb_min"][:, :3] = aabb_min
    H["aabb_max"][:, :3] = aabb_max
    H["center"][:, :3] = center
    return H


def test_world_aabb_closed_form():
    rng = np.random.default_rng(0)
    n = 500
    lo = -rng.uniform(0.1, 2, (n, 3))
    hi = rng.uniform(0.1, 2, (n, 3))
    H = hull_table(lo, hi)
    pos = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    q = random_quats(rng, n)
    mn, mx = O.world_aabbs(pos, q, np.arange(n), H)
    for i in range(n):
        R = quat_to_mat(q[i].astype(np.float64))
        he = (hi[i] - lo[i]) / 2
        lc = (hi[i] + lo[i]) / 2
        wc = R @ lc + pos[i]
        ext = np.abs(R) @ he
        assert np.allclose(mn[i, :3], wc - ext, atol=2e-5)
        assert np.allclose(mx[i, :3], wc + ext, atol=2e-5)
    # identity rotation: exact translation of the local box
    mn, mx = O.world_aabbs(pos, np.tile([0, 0, 0, 1], (n, 1)).astype(np.float32), np.arange(n), H)
    assert np.array_equal(mn[:, :3], (lo.astype(np.float32) * np.float32(1) + pos).astype(np.float32)) or \
        np.allclose(mn[:, :3], lo + pos, atol=1e-5)


def test_cell_hash_closed_form():
    rng = np.random.default_rng(1)
    n = 4000
    c = rng.uniform(-300, 300, (n, 3)).astype(np.float32)
    h = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)
    mn, mx = c - h, c + h
    dims = (64, 32, 16)
    inv_cell = np.float32(1.0 / 1.7)
    kv = O.calc_hash(mn, mx, inv_cell, dims)
    cen = (mn + mx) * np.float32(0.5)
    g = np.floor(cen * inv_cell).astype(np.int64) & (np.array(dims) - 1)
    ref = g[:, 2] * dims[1] * dims[0] + g[:, 1] * dims[0] + g[:, 0]
    assert np.array_equal(kv[:, 0], ref.astype(np.uint32))
    assert np.array_equal(kv[:, 1], np.arange(n, dtype=np.uint32))


def _grid_pairs(mn, mx, cell, dims, is_large=None):
    kv = O.calc_hash(mn, mx, np.float32(1.0 / cell), dims, is_large=is_large)
    k, v = O.radix_sort_kv(kv[:, 0], kv[:, 1], 32)
    skv = np.stack([k, v], 1)
    n = len(mn)
    return O.find_pairs_grid(skv, mn, mx, np.float32(1.0 / cell), dims, is_large=is_large,
                             scene_start=[0] if is_large is not None else None,
                             scene_count=[n] if is_large is not None else None)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_pairs_equal_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    c = rng.uniform(0, 40, (n, 3)).astype(np.float32)
    h = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    mn, mx = c - h, c + h
    cell = 2 * float(np.linalg.norm(h, axis=1).max())
    got = _grid_pairs(mn, mx, cell, (16, 16, 16))
    ref = O.find_pairs_brute(mn, mx)
    assert len(ref) > 100
    assert np.array_equal(got, ref)
    assert np.all(got[:, 0] > got[:, 1])


def test_brute_force_matches_numpy():
    rng = np.random.default_rng(9)
    n = 400
    c = rng.uniform(0, 10, (n, 3)).astype(np.float32)
    h = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    mn, mx = c - h, c + h
    ov = np.all((mn[:, None, :] <= mx[None, :, :]) & (mn[None, :, :] <= mx[:, None, :]), axis=2)
    i, j = np.nonzero(np.tril(ov, -1))
    ref = np.stack([i, j], 1)
    ref = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
    assert np.array_equal(O.find_pairs_brute(mn, mx), ref.astype(np.int32))


def test_sap_fixture_pairs():
    """5,876 AABBs of a reference demo scene: grid pair set (with an overflow
    list for bodies wider than half a cell) == brute force overlap set."""
    g = np.load(os.path.join(GOLD, "sap_aabbs.npz"))
    mn, mx = g["aabb_min"], g["aabb_max"]
    ref = O.find_pairs_brute(mn, mx)
    assert len(ref) > 1000
    h = (mx - mn) / 2
    cell = 4.0
    large = (np.linalg.norm(h, axis=1) > cell / 2).astype(np.int32)
    assert 0 < large.sum() < len(mn)
    got = _grid_pairs(mn, mx, cell, (32, 16, 32), is_large=large)
    assert np.array_equal(got, ref)
    # no overflow list, cell large enough for every body
    cell2 = 2 * float(np.linalg.norm(h, axis=1).max())
    assert np.array_equal(_grid_pairs(mn, mx, cell2, (16, 16, 16)), ref)


def test_scenes_do_not_pair_across():
    rng = np.random.default_rng(3)
    n = 600
    c = rng.uniform(0, 6, (n, 3)).astype(np.float32)
    mn, mx = c - 0.4, c + 0.4
    scene = (np.arange(n) * 3 // n).astype(np.int32)
    ref = O.find_pairs_brute(mn, mx, scene=scene)
    assert np.all(scene[ref[:, 0]] == scene[ref[:, 1]])
    allp = O.find_pairs_brute(mn, mx)
    assert len(allp) > len(ref)


def test_setup_bodies_closed_form():
    rng = np.random.default_rng(4)
    n = 64
    q = random_quats(rng, n)
    inv_i = rng.uniform(0.1, 2, (n, 3)).astype(np.float32)
    im = np.where(np.arange(n) % 5 == 0, 0, 1).astype(np.float32)
    lv = rng.normal(size=(n, 3)).astype(np.float32)
    lv2, world = O.setup_bodies(lv, np.zeros((n, 3)), q, im, inv_i, (0, -9.8, 0), 1 / 60)
    for i in range(n):
        if im[i] == 0:
            assert np.all(world[i] == 0)
            assert np.array_equal(lv2[i, :3], lv[i])
            continue
        A = np.abs(quat_to_mat(q[i].astype(np.float64)))
        ref = A @ np.diag(inv_i[i]) @ A.T
        assert np.allclose(world[i, :, :3], ref, rtol=1e-5, atol=1e-6)
        assert lv2[i, 1] == np.float32(lv[i, 1] + np.float32(-9.8) * np.float32(1 / 60))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sap_oracle_equals_brute_force(axis):
    """Sweep and prune (orc_sap_pairs) on the reference's 5,876 SAP AABBs and
    on random AABBs of mixed sizes, with negative and equal mins (the
    FloatFlip key order): the brute-force overlap set."""
    g = np.load(os.path.join(GOLD, "sap_aabbs.npz"))
    mn, mx = g["aabb_min"], g["aabb_max"]
    assert np.array_equal(O.sap_pairs(mn, mx, axis), O.find_pairs_brute(mn, mx))
    rng = np.random.default_rng(11 + axis)
    c = rng.uniform(-6, 6, (3000, 3)).astype(np.float32)
    h = rng.uniform(0.05, 1.2, (3000, 3)).astype(np.float32)
    mn, mx = c - h, c + h
    mn[::7, axis] = np.float32(0.0)         # equal mins on the sweep axis, both zeros
    mn[1::7, axis] = np.float32(-0.0)
    mx = np.maximum(mx, mn + np.float32(0.1))
    ref = O.find_pairs_brute(mn, mx)
    assert len(ref) > 1000
    assert np.array_equal(O.sap_pairs(mn, mx, axis), ref)
