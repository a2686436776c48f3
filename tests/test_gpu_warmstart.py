"""Warm start on the GPU: previous-frame impulse cache, per-contact
features, the BF16-MFMA network forward, the FP32-MFMA gradients and Adam,
each against the oracle; and one step with each warm start against the
oracle step from the same state."""
import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes
from experiments_amd.warmstart import Mlp
from oracle.step import Snapshot, gpu_order, gpu_tile_orders, oracle_step

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NW, NB = L.MLP_NW, L.MLP_NB


def bf16_round(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f32(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def rand_net(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return bf16_round(rng.uniform(-scale, scale, NW)), rng.uniform(-0.2, 0.2, NB).astype(np.float32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


# Tolerance of the network outputs: FP32 MFMA accumulation differs from the
# oracle's double accumulation in the last bits, and a hidden activation that
# lies at a bf16 rounding midpoint may round the other way (one bf16 ulp,
# 2^-8 relative).  So 99% of rows agree to 1e-4 (1 + |y|) and every row to
# 3e-2 (1 + |y|); a wrong MFMA lane layout gives O(1) errors everywhere.
def assert_outputs_close(got, ref):
    err = np.abs(got - ref) / (1.0 + np.abs(ref))
    tight = np.mean(np.all(err <= 1e-4, axis=1))
    assert tight >= 0.99, tight
    assert err.max() <= 3e-2, err.max()


@pytest.mark.parametrize("n", [1, 15, 16, 1000, 4133])
def test_mlp_forward_matches_oracle(n):
    w, b = rand_net(n)
    x = np.random.default_rng(n + 1).normal(size=(n, 16)).astype(np.float32)
    m = Mlp(seed=3)
    m.set_weights(w, b)
    gw, gb = m.get_weights()
    assert np.array_equal(gw, w) and np.array_equal(gb, b)
    y = m.forward(dev(x)).cpu().numpy()
    _, _, ref = O.mlp_forward(x, w, b)
    assert_outputs_close(y, ref)
    m.close()


@pytest.mark.parametrize("n", [64, 4133, 65536])
def test_mlp_gradients_match_oracle(n):
    w, b = rand_net(10 + n)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    m = Mlp(seed=1)
    m.set_weights(w, b)
    g, loss = m.gradients(dev(x), dev(t))
    g = g.cpu().numpy()
    rgw, rgb, rloss, _ = O.mlp_backward(x, w, b, t)
    ref = np.concatenate([rgw, rgb])
    # the few rows whose activations round differently (see above) perturb
    # the sums by far less than 1e-3 of the gradient norm
    assert np.linalg.norm(g - ref) <= 1e-3 * np.linalg.norm(ref)
    for lo, hi in ((0, 1024), (1024, 5120), (5120, NW), (NW, NW + NB)):
        assert np.abs(g[lo:hi] - ref[lo:hi]).max() <= 5e-3 * np.abs(ref[lo:hi]).max(), (lo, hi)
    assert abs(loss - rloss) <= 1e-4 * rloss
    m.close()


def test_adam_step_matches_numpy():
    """One Adam step on the full batch in row order (seed < 0) from known
    weights equals numpy's Adam on the library's own gradient."""
    w, b = rand_net(77)
    rng = np.random.default_rng(78)
    n = 3000
    x, t = rng.normal(size=(n, 16)).astype(np.float32), rng.normal(size=(n, 3)).astype(np.float32)
    m = Mlp(seed=1)
    m.set_weights(w, b)
    g, _ = m.gradients(dev(x), dev(t))
    g = g.cpu().numpy().astype(np.float64)
    lr = 1e-3
    hist = m.train(dev(x), dev(t), batch=n, steps=1, lr=lr, seed=-1)
    assert hist.shape == (1,) and np.isfinite(hist[0])
    w1, b1 = m.get_weights()
    theta = np.concatenate([bf16_to_f32(w), b]).astype(np.float64)
    mom, vel = 0.1 * g, 0.001 * g * g
    lr_t = lr * np.sqrt(1 - 0.999) / (1 - 0.9)
    ref = theta - lr_t * mom / (np.sqrt(vel) + 1e-8)
    # biases are fp32: compare directly
    assert np.allclose(b1, ref[NW:], rtol=1e-5, atol=1e-7)
    # weights: the bf16 rounding of the fp32 master; at most a few values sit
    # at a rounding midpoint
    rw = bf16_round(ref[:NW].astype(np.float32))
    assert np.mean(rw == w1) >= 0.995
    assert np.all(np.abs(rw.astype(np.int32) - w1.astype(np.int32)) <= 1)
    m.close()


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(5)
    n = 50000
    x = rng.normal(size=(n, 16)).astype(np.float32)
    t = np.stack([x[:, 0] + 0.5 * x[:, 1], np.maximum(x[:, 2], 0), 0.3 * x[:, 3] * x[:, 4]], 1)
    m = Mlp(seed=2)
    X, T = dev(x), dev(t.astype(np.float32))
    hist = m.train(X, T, batch=4096, steps=600, lr=3e-3, seed=1)
    assert np.all(np.isfinite(hist))
    assert hist[-20:].mean() < 0.1 * hist[:5].mean()
    y = m.forward(X).cpu().numpy()
    assert np.mean((y - t) ** 2) < 0.1 * np.mean((t - t.mean(0)) ** 2)
    m.close()


def pile_world(warm, kind="piles"):
    if kind == "piles":   # 24 scenes: one solver tile per scene
        sc = scenes.batched_piles(n_scenes=24, lo=8, hi=64, pool8=32, iterations=6)
    else:                 # one scene: spatial solver tiles in 4 phases
        sc = scenes.lattice_cubes(6, 4, 6, spacing=1.0, gap=0.0)
    return scenes.build_world(sc, max_pairs_per_body=64, warm_start=warm)


@pytest.mark.parametrize("warm,kind", [(L.WARM_PREVIOUS, "piles"), (L.WARM_LEARNED, "piles"),
                                       (L.WARM_PREVIOUS, "lattice"), (L.WARM_LEARNED, "lattice")])
def test_warm_step_matches_oracle(warm, kind):
    w = pile_world(warm, kind)
    if warm == L.WARM_LEARNED:
        wt, bt = rand_net(9, scale=0.1)
        w.set_warmstart_weights(wt, bt)
    w.step(40)
    s = Snapshot(w)
    keys = w.read(L.BUF_PREV_KEYS, np.int32, (2,))
    imp = w.read(L.BUF_PREV_IMPULSES, np.float32, (4,))
    pts = w.read(L.BUF_PREV_POINTS, np.float32, (4, 4))
    assert len(keys) > 0 and len(pts) == len(keys)
    # cache is sorted by pair (the lookup is a binary search)
    kk = keys[:, 0].astype(np.int64) * (1 << 32) + keys[:, 1]
    assert np.all(np.diff(kk) > 0)
    w.step_partial(L.STAGE_COLOR)
    order = gpu_order(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    feats = w.read(L.BUF_FEATURES, np.float32, (16,))
    pred = w.read(L.BUF_PREDICTIONS, np.float32, (3,)) if warm == L.WARM_LEARNED else None
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    ref = oracle_step(s, order, contacts, prev=(keys, imp, pts), warm=warm, prediction=pred,
                      pairs=w.pairs())
    assert np.allclose(feats, ref["features"], rtol=1e-5, atol=1e-6)
    assert np.mean(ref["features"][:, 13]) > 0.5  # most pairs were in contact last step
    if warm == L.WARM_LEARNED:
        _, _, y = O.mlp_forward(feats, wt, bt)
        assert_outputs_close(pred, y)
    for k in ("linvel", "angvel", "pos", "quat"):
        c = 3 if k != "quat" else 4
        assert np.allclose(g[k][:, :c], ref[k][:, :c], rtol=1e-4, atol=1e-4), k
    gl = w.read(L.BUF_CONSTRAINTS, np.uint8, (176,)).view(O.CONSTRAINT_DTYPE).reshape(-1)
    assert np.allclose(gl["lambda"], ref["cons"]["lambda"], rtol=1e-4, atol=1e-5)
    w.close()


def test_training_batch_copies_features_and_impulses():
    w = pile_world(L.WARM_ZERO)
    w.step(30)
    cap = int(w.cfg.max_contacts)
    X = torch.zeros((cap, 16), dtype=torch.float32, device="cuda")
    T = torch.zeros((cap, 3), dtype=torch.float32, device="cuda")
    n = w.copy_training_batch(X.data_ptr(), T.data_ptr(), cap)
    feats = w.read(L.BUF_FEATURES, np.float32, (16,))
    cons = w.read(L.BUF_CONSTRAINTS, np.uint8, (176,)).view(O.CONSTRAINT_DTYPE).reshape(-1)
    assert n == len(feats) == len(cons) > 0
    assert np.array_equal(X[:n].cpu().numpy(), feats)
    t = T[:n].cpu().numpy()
    lam = cons["lambda"].astype(np.float32)
    s = ((lam[:, 0] + lam[:, 1]) + lam[:, 2]) + lam[:, 3]
    assert np.array_equal(t[:, 0], s * np.float32(10))
    assert np.array_equal(t[:, 1:], cons["f_lambda"] * np.float32(10))
    # capacity smaller than the contact count
    assert w.copy_training_batch(X.data_ptr(), T.data_ptr(), 5) == 5
    w.close()


def test_learned_without_weights_is_refused():
    w = pile_world(L.WARM_LEARNED)
    with pytest.raises(L.MiRbError):
        w.step(1)
    w.close()


@pytest.mark.parametrize("warm", [L.WARM_ZERO, L.WARM_PREVIOUS])
def test_scene_tolerance_matches_oracle(warm):
    """Scene tiles with a residual tolerance: every scene stops on its own
    residual inside one launch; the oracle solves each scene's rows alone
    with the same tolerance."""
    sc = scenes.batched_piles(n_scenes=24, lo=8, hi=64, pool8=32, iterations=60)
    w = scenes.build_world(sc, max_pairs_per_body=64, warm_start=warm, residual_tolerance=1e-3)
    w.step(40)
    s = Snapshot(w)
    keys = w.read(L.BUF_PREV_KEYS, np.int32, (2,))
    imp = w.read(L.BUF_PREV_IMPULSES, np.float32, (4,))
    pts = w.read(L.BUF_PREV_POINTS, np.float32, (4, 4))
    w.step_partial(L.STAGE_COLOR)
    segs = gpu_tile_orders(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    st = w.stats()
    ref = oracle_step(s, None, contacts, prev=(keys, imp, pts), warm=warm, segments=segs,
                      pairs=w.pairs())
    for k in ("linvel", "angvel", "pos", "quat"):
        c = 3 if k != "quat" else 4
        assert np.allclose(g[k][:, :c], ref[k][:, :c], rtol=1e-4, atol=1e-4), k
    its = np.array(ref["iters"])
    assert st["normal_iterations"] == its[:, 0].max()
    assert st["friction_iterations"] == its[:, 1].max()
    assert abs(st["mean_normal_iterations"] - its[:, 0].mean()) < 1e-3
    assert its[:, 0].min() < its[:, 0].max()  # scenes really stop at different sweeps
    # the per-scene record: sweeps per phase of every scene with rows, in
    # scene order (the oracle's segments), and final residuals below the
    # tolerance wherever a phase stopped before its cap
    ts = w.read(L.BUF_TILE_STATS, np.int32, (4,))
    act = ts[:, 0] >= 0
    assert np.array_equal(ts[act, 0], its[:, 0]) and np.array_equal(ts[act, 1], its[:, 1])
    res = ts[act, 2:4].copy().view(np.float32)
    assert np.all(res[its[:, 0] < 60, 0] < 1e-3) and np.all(res[its[:, 1] < 60, 1] < 1e-3)
    w.close()


@pytest.mark.parametrize("kind", ["compound", "mesh", "spheres"])
def test_previous_warm_start_with_child_pairs_matches_oracle(kind):
    """Worlds whose contacts come from compound child pairs, mesh triangles or
    sphere/plane pairs: the impulse cache is keyed by each manifold's pair
    entry (unique, sorted for the binary search), so every manifold finds its
    own previous impulses; the features (with the found flag) and one step
    with the previous-frame warm start equal the oracle's."""
    if kind == "compound":
        from test_gpu_compound import compound_pile
        w = compound_pile()
    elif kind == "mesh":
        from test_gpu_mesh import mesh_world
        w = mesh_world()[0]
    else:
        from test_gpu_spheres import sphere_box
        w = sphere_box()[0]
    w.set_solver(8, 0.0, L.WARM_PREVIOUS)
    w.step(45)
    s = Snapshot(w)
    keys = w.read(L.BUF_PREV_KEYS, np.int32, (2,))
    imp = w.read(L.BUF_PREV_IMPULSES, np.float32, (4,))
    pts = w.read(L.BUF_PREV_POINTS, np.float32, (4, 4))
    assert len(keys) > 0
    kk = keys[:, 0].astype(np.int64) * (1 << 32) + keys[:, 1]
    assert np.all(np.diff(kk) > 0)  # unique and sorted
    w.step_partial(L.STAGE_COLOR)
    order = gpu_order(w)
    contacts = w.contacts(O.CONTACT_DTYPE)
    pairs = w.pairs()
    feats = w.read(L.BUF_FEATURES, np.float32, (16,))
    w.set_state(**{k: s.state[k] for k in ("pos", "quat", "linvel", "angvel")})
    w.step(1)
    g = w.get_state()
    ref = oracle_step(s, order, contacts, prev=(keys, imp, pts), warm=L.WARM_PREVIOUS,
                      pairs=pairs)
    assert np.array_equal(feats[:, 13], ref["features"][:, 13])
    assert np.mean(feats[:, 13]) > 0.5
    assert np.allclose(feats, ref["features"], rtol=1e-5, atol=1e-6)
    for k in ("linvel", "angvel", "pos"):
        assert np.allclose(g[k][:, :3], ref[k][:, :3], rtol=1e-4, atol=1e-4), k
    gl = w.read(L.BUF_CONSTRAINTS, np.uint8, (176,)).view(O.CONSTRAINT_DTYPE).reshape(-1)
    assert np.allclose(gl["lambda"], ref["cons"]["lambda"], rtol=1e-4, atol=1e-5)
    w.close()
