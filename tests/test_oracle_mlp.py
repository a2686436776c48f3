"""The oracle's learned-warm-start pieces (MLP forward/backward, per-contact
features, prediction -> impulses), pinned against independent restatements:
numpy matrix-form forward/backward, central finite differences, and closed
forms on hand-built contacts."""
import numpy as np
import pytest

import oracle as O

NW, NB = O.N_W, O.N_B


def bf16_round(x):
    """Round float32 to bfloat16 (nearest even) with numpy integer ops."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_to_f32(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def rand_net(seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-0.4, 0.4, NW).astype(np.float32)
    b = rng.uniform(-0.2, 0.2, NB).astype(np.float32)
    return bf16_round(w), b


def split(wf):
    return wf[:1024].reshape(64, 16), wf[1024:5120].reshape(64, 64), wf[5120:].reshape(3, 64)


def np_forward(x, w_bf16, b, round_acts=True):
    W1, W2, W3 = split(bf16_to_f32(w_bf16).astype(np.float64))
    r = (lambda a: bf16_to_f32(bf16_round(a.astype(np.float32))).astype(np.float64)) \
        if round_acts else (lambda a: a.astype(np.float64))
    xr = r(x)
    h1 = np.maximum(xr @ W1.T + b[:64], 0)
    h2 = np.maximum(r(h1) @ W2.T + b[64:128], 0)
    y = r(h2) @ W3.T + b[128:]
    return xr, h1, h2, y


def np_backward(x, w_bf16, b, t, round_acts=True):
    W1, W2, W3 = split(bf16_to_f32(w_bf16).astype(np.float64))
    xr, h1, h2, y = np_forward(x, w_bf16, b, round_acts)
    r = (lambda a: bf16_to_f32(bf16_round(a.astype(np.float32))).astype(np.float64)) \
        if round_acts else (lambda a: a)
    n = len(x)
    dy = 2 * (y - t) / (3 * n)
    d2 = (dy @ W3) * (h2 > 0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW3, gW2, gW1 = dy.T @ r(h2), d2.T @ r(h1), d1.T @ xr
    g = np.concatenate([gW1.ravel(), gW2.ravel(), gW3.ravel()])
    gb = np.concatenate([d1.sum(0), d2.sum(0), dy.sum(0)])
    loss = ((y - t) ** 2).sum() / (3 * n)
    return g, gb, loss


def test_bf16_rounding_known_values():
    f = O.lib().orc_f32_to_bf16
    assert f(1.0) == 0x3F80
    assert f(np.uint32(0x3F808000).view(np.float32)) == 0x3F80   # tie, even stays
    assert f(np.uint32(0x3F818000).view(np.float32)) == 0x3F82   # tie, odd rounds up
    assert f(np.uint32(0x3F808001).view(np.float32)) == 0x3F81
    assert f(float("inf")) == 0x7F80
    assert (f(float("nan")) & 0x7FFF) > 0x7F80
    vals = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.array_equal(np.array([f(float(v)) for v in vals], np.uint16), bf16_round(vals))


def test_forward_matches_matrix_form():
    w, b = rand_net(1)
    x = np.random.default_rng(2).normal(size=(300, 16)).astype(np.float32)
    h1, h2, y = O.mlp_forward(x, w, b)
    _, rh1, rh2, ry = np_forward(x, w, b)
    np.testing.assert_allclose(h1, rh1, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(h2, rh2, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(y, ry, rtol=1e-6, atol=1e-6)


def test_backward_matches_matrix_form():
    w, b = rand_net(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(257, 16)).astype(np.float32)
    t = rng.normal(size=(257, 3)).astype(np.float32)
    gw, gb, loss, _ = O.mlp_backward(x, w, b, t)
    rg, rgb, rloss = np_backward(x, w, b, t)
    scale = np.abs(rg).max()
    np.testing.assert_allclose(gw, rg, rtol=0, atol=1e-5 * scale)
    np.testing.assert_allclose(gb, rgb, rtol=0, atol=1e-5 * np.abs(rgb).max())
    assert abs(loss - rloss) <= 1e-5 * rloss


def test_matrix_form_backward_matches_finite_differences():
    """Pins the numpy backward (and through the previous test the oracle's)
    on an unrounded float64 network: central differences of the loss."""
    w, b = rand_net(5)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(20, 16))
    t = rng.normal(size=(20, 3))
    W = bf16_to_f32(w).astype(np.float64)

    def loss_of(Wv, bv):
        W1, W2, W3 = Wv[:1024].reshape(64, 16), Wv[1024:5120].reshape(64, 64), Wv[5120:].reshape(3, 64)
        h1 = np.maximum(x @ W1.T + bv[:64], 0)
        h2 = np.maximum(h1 @ W2.T + bv[64:128], 0)
        y = h2 @ W3.T + bv[128:]
        return ((y - t) ** 2).sum() / (3 * len(x))

    # analytic gradient on float64 weights (no rounding anywhere)
    W1, W2, W3 = W[:1024].reshape(64, 16), W[1024:5120].reshape(64, 64), W[5120:].reshape(3, 64)
    bb = b.astype(np.float64)
    h1 = np.maximum(x @ W1.T + bb[:64], 0)
    h2 = np.maximum(h1 @ W2.T + bb[64:128], 0)
    y = h2 @ W3.T + bb[128:]
    dy = 2 * (y - t) / (3 * len(x))
    d2 = (dy @ W3) * (h2 > 0)
    d1 = (d2 @ W2) * (h1 > 0)
    g = np.concatenate([(d1.T @ x).ravel(), (d2.T @ h1).ravel(), (dy.T @ h2).ravel()])
    gb = np.concatenate([d1.sum(0), d2.sum(0), dy.sum(0)])
    eps = 1e-6
    for i in list(rng.choice(1024, 8, replace=False)) + list(1024 + rng.choice(4096, 8, replace=False)) \
            + list(5120 + rng.choice(192, 8, replace=False)):
        Wp, Wm = W.copy(), W.copy()
        Wp[i] += eps
        Wm[i] -= eps
        fd = (loss_of(Wp, bb) - loss_of(Wm, bb)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), i
    for i in rng.choice(NB, 8, replace=False):
        bp, bm = bb.copy(), bb.copy()
        bp[i] += eps
        bm[i] -= eps
        fd = (loss_of(W, bp) - loss_of(W, bm)) / (2 * eps)
        assert abs(fd - gb[i]) <= 1e-6 + 1e-4 * abs(fd), i


def _box_contact(npts=4, depth=-0.01):
    """Dynamic body 1 resting on static body 0 (y up), contact normal -y as
    the narrowphase writes it (from B=ground towards A is +y; normal = -axis)."""
    c = np.zeros(1, O.CONTACT_DTYPE)
    pts = [(0.5, 0.0, 0.5), (-0.5, 0.0, 0.5), (-0.5, 0.0, -0.5), (0.5, 0.0, -0.5)]
    for i in range(4):
        c["pos"][0, i, :3] = pts[i]
        c["pos"][0, i, 3] = depth - 0.001 * i
    c["normal"][0] = (0.0, 1.0, 0.0, float(npts))
    c["body_a"][0] = 1
    c["body_b"][0] = 0
    return c


def _bodies():
    pos = np.array([[0, -1, 0, 0], [0, 0.5, 0, 0]], np.float32)
    lin = np.array([[0, 0, 0, 0], [1.0, -2.0, 0.5, 0]], np.float32)
    ang = np.array([[0, 0, 0, 0], [0.0, 3.0, 0.0, 0]], np.float32)
    im = np.array([0.0, 0.5], np.float32)
    invI = np.zeros((2, 12), np.float32)
    invI[1, [0, 5, 10]] = 6.0
    return pos, lin, ang, im, invI


@pytest.mark.parametrize("npts", [1, 3, 4])
def test_contact_features_closed_form(npts):
    c = _box_contact(npts)
    pos, lin, ang, im, invI = _bodies()
    cons = O.contact_to_constraint(c, pos, lin, ang, im, invI, 1 / 60.0)
    keys = np.array([[1, 0], [2, 1]], np.int32)
    imp = np.array([[0.8, 0.1, -0.2, 3], [5, 5, 5, 4]], np.float32)
    # the previous manifold: 3 points, the nearest to current point i is
    # previous point (i % 3), moved slightly; its impulse is 0.1 * (i % 3 + 1)
    pts = np.zeros((2, 4, 4), np.float32)
    for j in range(3):
        pts[0, j, :3] = c["pos"][0, j, :3] + 0.01
        pts[0, j, 3] = 0.1 * (j + 1)
    pts[0, 0, :3] = c["pos"][0, 0, :3] - 0.01
    f, cw = O.contact_features(c, cons, pos, lin, ang, im, keys, imp, pts, apply_previous=True)
    f = f[0]
    cen = c["pos"][0, :npts, :3].mean(0)
    r = cen - pos[1, :3]
    vr = lin[1, :3] + np.cross(ang[1, :3], r)
    n = np.array([0.0, 1.0, 0.0])
    assert np.allclose(f[:3], n)
    assert np.isclose(f[3], 10 * c["pos"][0, :npts, 3].min())
    assert np.isclose(f[4], vr @ n, atol=1e-6)
    # plane_space of +y: |n.z| <= 0.7071 -> t0 = (-n.y, n.x, 0)/|.| = (-1, 0, 0)
    assert np.isclose(f[5], -vr[0], atol=1e-6)
    assert np.isclose(f[6], vr[2] * 0 + (np.cross(n, [-1, 0, 0]) @ vr), atol=1e-6)
    assert f[7] == np.float32(0.5) and f[8] == 0
    assert np.allclose(f[9:12], 10 * imp[0, :3])
    assert f[12] == npts / 4 and f[13] == 1
    jac = cons["jac_inv"][0, :npts]
    assert np.isclose(f[14], -jac.mean(), rtol=1e-6)
    assert np.isclose(f[15], cons["b"][0, :npts].mean(), rtol=1e-6)
    lam = cw["lambda"][0]
    # point 3 (if present) is nearest to previous point 0 (the corners are 1 apart)
    want = [0.1, 0.2, 0.3, 0.1][:npts]
    assert np.allclose(lam[:npts], want) and np.all(lam[npts:] == 0)
    assert np.allclose(cw["f_lambda"][0], [0.1, -0.2])


def test_contact_features_without_previous():
    c = _box_contact(4)
    pos, lin, ang, im, invI = _bodies()
    cons = O.contact_to_constraint(c, pos, lin, ang, im, invI, 1 / 60.0)
    keys = np.array([[1, 2]], np.int32)   # another pair
    imp = np.array([[9, 9, 9, 4]], np.float32)
    pts = np.ones((1, 4, 4), np.float32)
    f, cw = O.contact_features(c, cons, pos, lin, ang, im, keys, imp, pts, apply_previous=True)
    assert np.all(f[0, 9:12] == 0) and f[0, 13] == 0
    assert np.all(cw["lambda"] == 0) and np.all(cw["f_lambda"] == 0)


def test_warm_from_prediction_clamps():
    c = np.concatenate([_box_contact(4), _box_contact(2), _box_contact(4)])
    pos, lin, ang, im, invI = _bodies()
    cons = O.contact_to_constraint(c, pos, lin, ang, im, invI, 1 / 60.0)
    cons["lambda"][2] = [0.1, 0.3, 0.0, 0.4]                 # matched previous impulses
    cons["f_lambda"][2] = [0.5, -2.0]                        # and previous friction
    y = np.array([[20.0, 30.0, -1.0], [-5.0, 1.0, 1.0], [16.0, 0.0, 0.0]], np.float32)
    out = O.warm_from_prediction(cons, y, 0.7)
    assert np.allclose(out["lambda"][2], [0.2, 0.6, 0.0, 0.8])  # 1.6 in proportion 1:3:0:4
    # a matched manifold keeps its previous friction (not the predicted 0, 0),
    # clamped to the cone 0.7 x 1.6
    assert np.allclose(out["f_lambda"][2], [0.5, -0.7 * 1.6])
    assert np.allclose(out["lambda"][0], 0.5)               # no previous: 2.0 split over 4 points
    assert np.isclose(out["f_lambda"][0, 0], 0.7 * 2.0)     # clamped to the cone
    assert np.isclose(out["f_lambda"][0, 1], -0.1)
    assert np.all(out["lambda"][1] == 0)                    # negative normal -> 0
    assert np.all(out["f_lambda"][1] == 0)                  # cone of zero
