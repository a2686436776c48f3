"""Oracle constraint setup, solve and integration pinned against closed forms."""
import numpy as np
import pytest

import oracle as O


def ground_contact(pen, npts=4, body=1):
    c = np.zeros(1, O.CONTACT_DTYPE)
    corners = [[-.5, 0, -.5], [.5, 0, -.5], [.5, 0, .5], [-.5, 0, .5]]
    for i in range(npts):
        c["pos"][0, i, :3] = corners[i]
        c["pos"][0, i, 3] = -pen
    c["normal"][0] = [0, -1, 0, npts]  # from body A (cube) towards B (ground)
    c["body_a"], c["body_b"] = body, 0
    return c


def world_inertia(n, inv_diag):
    I = np.zeros((n, 3, 4), np.float32)
    for i, d in enumerate(inv_diag):
        I[i, 0, 0], I[i, 1, 1], I[i, 2, 2] = d
    return I


@pytest.mark.parametrize("pen", [0.0, 0.02])
def test_resting_box_converges_to_bias_velocity(pen):
    dt = np.float32(1 / 60)
    pos = np.array([[0, -0.5, 0], [0, 0.5 - pen, 0]], np.float32)
    im = np.array([0, 1], np.float32)
    I = world_inertia(2, [(0, 0, 0), (6, 6, 6)])
    v0 = np.array([[0, 0, 0], [0, -9.8 / 60, 0]], np.float32)
    c = O.contact_to_constraint(ground_contact(pen), pos, v0, np.zeros((2, 3)), im, I, dt)
    # Jacobian inverse of a centred point: -1/(1/m + r x n . I^-1 r x n)
    r = np.array([-.5, 0.5 - 0 - (0.5 - pen) + (0.5 - pen) - 0.5, -.5])
    assert np.all(c["jac_inv"][0] < 0)
    b = (-pen + 0.005) * 0.2 / dt
    assert np.allclose(c["b"][0], b, rtol=1e-6)
    cs, lv, av, it, res = O.solve(c, None, v0, np.zeros((2, 3)), pos, im, I, 200, 0)
    assert abs(lv[1, 1] - (-b)) < 1e-4
    assert np.allclose(av[1, :3], 0, atol=1e-5)
    assert np.all(lv[0] == 0)  # static body stays at rest
    assert np.all(cs["lambda"][0] >= 0)
    # total normal impulse = momentum change of the unit-mass box
    assert abs(cs["lambda"][0].sum() - (-b - v0[1, 1])) < 1e-4


def test_friction_bounded_by_normal_impulse():
    dt = np.float32(1 / 60)
    pos = np.array([[0, -0.5, 0], [0, 0.5, 0]], np.float32)
    im = np.array([0, 1], np.float32)
    I = world_inertia(2, [(0, 0, 0), (6, 6, 6)])
    v0 = np.array([[0, 0, 0], [3.0, -0.5, 0]], np.float32)
    c = O.contact_to_constraint(ground_contact(0.0), pos, v0, np.zeros((2, 3)), im, I, dt)
    cs, lv, av, it, res = O.solve(c, None, v0, np.zeros((2, 3)), pos, im, I, 20, 20, friction=0.6)
    lam_n = cs["lambda"][0].sum()
    assert np.all(np.abs(cs["f_lambda"][0]) <= 0.6 * lam_n + 1e-6)
    assert lv[1, 0] < 3.0  # friction slowed the slide
    assert lv[1, 0] > 3.0 - 0.6 * lam_n - 1e-4


def test_residual_tolerance_stops_early():
    dt = np.float32(1 / 60)
    pos = np.array([[0, -0.5, 0], [0, 0.5, 0]], np.float32)
    im = np.array([0, 1], np.float32)
    I = world_inertia(2, [(0, 0, 0), (6, 6, 6)])
    v0 = np.array([[0, 0, 0], [0, -9.8 / 60, 0]], np.float32)
    c = O.contact_to_constraint(ground_contact(0.0), pos, v0, np.zeros((2, 3)), im, I, dt)
    _, _, _, it, res = O.solve(c, None, v0, np.zeros((2, 3)), pos, im, I, 100, 100, tol=1e-4)
    assert 1 <= it[0] < 100 and res[0] < 1e-4


def test_coloring_is_valid_and_deterministic():
    rng = np.random.default_rng(0)
    n, m = 300, 1200
    cs = np.zeros(m, O.CONSTRAINT_DTYPE)
    cs["body_a"] = rng.integers(0, n, m)
    cs["body_b"] = rng.integers(0, n, m)
    im = np.where(np.arange(n) < 20, 0, 1).astype(np.float32)
    col, nc = O.color_contacts(cs, im)
    assert nc > 0
    for c in range(nc):
        rows = np.nonzero(col == c)[0]
        bodies = np.concatenate([cs["body_a"][rows], cs["body_b"][rows]])
        dyn = bodies[im[bodies] != 0]
        # a row whose two ends are the same body counts once
        same = cs["body_a"][rows] == cs["body_b"][rows]
        dyn_rows = [b for r in rows for b in {int(cs["body_a"][r]), int(cs["body_b"][r])} if im[b] != 0]
        assert len(dyn_rows) == len(set(dyn_rows))
    col2, _ = O.color_contacts(cs, im)
    assert np.array_equal(col, col2)


def test_integrate_constant_velocity_and_rotation():
    n = 3
    pos = np.zeros((n, 3), np.float32)
    q = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    lv = np.array([[1, 2, 3], [0, 0, 0], [0, 0, 0]], np.float32)
    w = np.array([[0, 0, 0], [0, 0, 2.0], [0, 0, 1e-4]], np.float32)
    dt = np.float32(1 / 60)
    p2, q2, w2 = O.integrate(pos, q, lv, w, np.ones(n, np.float32), dt)
    assert np.array_equal(p2[0, :3], (lv[0] * dt).astype(np.float32))
    assert np.allclose(w2[1, :3], w[1] * 0.99)
    ang = 2.0 * 0.99 * float(dt)
    assert np.allclose(q2[1], [0, 0, np.sin(ang / 2), np.cos(ang / 2)], atol=1e-6)
    # small-angle Taylor branch
    ang = 1e-4 * 0.99 * float(dt)
    assert np.allclose(q2[2], [0, 0, np.sin(ang / 2), np.cos(ang / 2)], atol=1e-9)
    # clamp: the angle |w| dt is limited to 0.25 pi but the axis term keeps the
    # unclamped |w| (integrateKernel.cl:131-148), and the result is renormalised
    p3, q3, _ = O.integrate(pos[:1], q[:1], lv[:1] * 0, np.array([[0, 0, 100.0]], np.float32),
                            np.ones(1, np.float32), dt)
    lim = 0.25 * 3.14159254
    wz = 100.0 * 0.99
    fa = lim / float(dt)
    raw = np.array([0, 0, wz * np.sin(0.5 * fa * float(dt)) / fa, np.cos(lim / 2)])
    assert np.allclose(q3[0], raw / np.linalg.norm(raw), atol=1e-5)
