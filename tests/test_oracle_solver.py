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


# ---- closed forms for the constraint rows and the friction solve ----------

def spd_inverse_inertia(rng):
    """A world inverse inertia R diag(d) R^T as 3 float4 rows."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    M = R @ np.diag(rng.uniform(0.5, 8.0, 3)) @ R.T
    out = np.zeros((3, 4), np.float32)
    out[:, :3] = M
    return out


def effective(r, d, im, I):
    """1 / (effective mass) of one body's share along direction d at lever r."""
    a = np.cross(r, d)
    return im + a @ I[:, :3].astype(np.float64) @ a


@pytest.mark.parametrize("n", [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
                               (0, 0, -1), (0.6, 0.0, 0.8), (0.48, -0.6, 0.64)])
def test_plane_space_is_a_right_handed_orthonormal_frame(n):
    t0, t1 = O.plane_space(np.array(n, np.float32))
    n = np.array(n, np.float64)
    assert abs(np.linalg.norm(t0) - 1) < 1e-6 and abs(np.linalg.norm(t1) - 1) < 1e-6
    assert abs(t0 @ n) < 1e-6 and abs(t1 @ n) < 1e-6 and abs(t0 @ t1) < 1e-6
    assert np.allclose(np.cross(n, t0), t1, atol=1e-6)
    # btPlaneSpace1 (solverSetup.cl:491-516) keeps t0 in the plane of the two
    # axes other than the normal's dominant one: z = 0 unless |n.z| > 1/sqrt 2
    assert (t0[0] == 0) if abs(n[2]) > 0.70710678 else (t0[2] == 0)


def test_plane_space_known_answers_for_axis_normals():
    want = {(1, 0, 0): ((0, 1, 0), (0, 0, 1)), (-1, 0, 0): ((0, -1, 0), (0, 0, 1)),
            (0, 1, 0): ((-1, 0, 0), (0, 0, 1)), (0, -1, 0): ((1, 0, 0), (0, 0, 1)),
            (0, 0, 1): ((0, -1, 0), (1, 0, 0)), (0, 0, -1): ((0, 1, 0), (1, 0, 0))}
    for n, (a, b) in want.items():
        t0, t1 = O.plane_space(np.array(n, np.float32))
        assert np.array_equal(t0 + 0.0, np.array(a, np.float32)), n
        assert np.array_equal(t1 + 0.0, np.array(b, np.float32)), n


def test_constraint_rows_match_effective_mass_closed_form():
    rng = np.random.default_rng(12)
    dt = np.float32(1 / 60)
    for trial in range(40):
        pos = rng.uniform(-2, 2, (2, 3)).astype(np.float32)
        im = rng.uniform(0.1, 2.0, 2).astype(np.float32)
        if trial % 5 == 0:
            im[1] = 0  # static partner
        I = np.zeros((2, 3, 4), np.float32)
        I[0] = spd_inverse_inertia(rng)
        I[1] = spd_inverse_inertia(rng) if im[1] else 0
        nrm = rng.normal(size=3)
        nrm = (nrm / np.linalg.norm(nrm)).astype(np.float32)
        npts = 1 + trial % 4
        c = np.zeros(1, O.CONTACT_DTYPE)
        c["pos"][0, :npts, :3] = rng.uniform(-1, 1, (npts, 3))
        c["pos"][0, :npts, 3] = rng.uniform(-0.05, 0.01, npts)
        c["normal"][0, :3] = nrm
        c["normal"][0, 3] = npts
        c["body_a"], c["body_b"] = 0, 1
        v = rng.normal(size=(2, 3)).astype(np.float32)
        w = rng.normal(size=(2, 3)).astype(np.float32)
        cs = O.contact_to_constraint(c, pos, v, w, im, I, dt)[0]
        pA, pB = pos.astype(np.float64)
        n64 = nrm.astype(np.float64)
        for i in range(4):
            if i >= npts:
                assert cs["jac_inv"][i] == 0
                continue
            p = c["pos"][0, i, :3].astype(np.float64)
            k = effective(p - pA, n64, im[0], I[0]) + effective(p - pB, n64, im[1], I[1])
            assert np.isclose(cs["jac_inv"][i], -1.0 / k, rtol=2e-5), (trial, i)
            # Baumgarte bias of the point's depth (restitution 0)
            depth = c["pos"][0, i, 3]
            assert np.isclose(cs["b"][i], (depth + 0.005) * 0.2 / dt, rtol=1e-5, atol=1e-6)
        # friction rows at the points' centre, along the two tangents
        cen = c["pos"][0, :npts, :3].astype(np.float64).mean(0)
        assert np.allclose(cs["center"][:3], cen, atol=1e-6)
        for i, t in enumerate(O.plane_space(nrm)):
            t = t.astype(np.float64)
            k = effective(cen - pA, t, im[0], I[0]) + effective(cen - pB, t, im[1], I[1])
            assert np.isclose(cs["f_jac_inv"][i], -1.0 / k, rtol=2e-5), (trial, i)
        assert np.allclose(cs["linear"][:3], -nrm)


def _ground_rows(pA, vA, wA, lam_n, npts=4, pB=(0, -0.5, 0), imB=0.0, IB=None):
    """Constraint of a cube (body 0) on body 1 with the normal rows made inert
    (jac_inv 0) and the normal impulses preset, so a friction-only solve runs
    against a known cone."""
    pos = np.array([pA, pB], np.float32)
    im = np.array([1.0, imB], np.float32)
    I = np.zeros((2, 3, 4), np.float32)
    I[0, 0, 0] = I[0, 1, 1] = I[0, 2, 2] = 6.0
    if IB is not None:
        I[1] = IB
    c = ground_contact(0.0, npts=npts, body=0)
    c["body_b"] = 1
    c["pos"][0, :npts, 1] += pA[1] - 0.5
    v = np.array([vA, [0, 0, 0]], np.float32)
    w = np.array([wA, [0, 0, 0]], np.float32)
    cs = O.contact_to_constraint(c, pos, v, w, im, I, np.float32(1 / 60))
    cs["jac_inv"][0] = 0
    cs["lambda"][0] = 0
    cs["lambda"][0, 0] = lam_n
    return cs, pos, v, w, im, I


def test_friction_damps_spin_about_the_normal_by_a_tenth_per_sweep():
    # spin about the contact normal with the centre right below the centre of
    # mass: the friction rows see no sliding, only the reference's angular
    # damping (solveFriction.cl:379-389) acts: w.n loses 10% per application
    wy = np.float32(5.0)
    cs, pos, v, w, im, I = _ground_rows((0, 0.5, 0), (0, 0, 0), (0, wy, 0), 10.0)
    for k in (1, 3, 7):
        c, lv, av, _, _ = O.solve(cs.copy(), None, v, w, pos, im, I, 0, k, friction=0.6)
        want = wy
        for _ in range(k):
            want = np.float32(want - np.float32(want * np.float32(0.1)))
        assert av[0, 1] == want, (k, av[0, 1], want)
        assert np.all(c["f_lambda"][0] == 0)
        assert np.all(lv[0] == 0) and av[0, 0] == 0 and av[0, 2] == 0


def test_friction_damping_needs_alignment_or_a_static_partner():
    # both bodies dynamic; with no normal impulse the friction rows cannot act
    # (cone of zero width), so any change of w.n is the damping
    IB = np.zeros((3, 4), np.float32)
    IB[0, 0] = IB[1, 1] = IB[2, 2] = 6.0
    w0 = (0.0, 2.0, 0.0)
    # B right below A: centre on the line between them -> damped
    cs, pos, v, w, im, I = _ground_rows((0, 0.5, 0), (0, 0, 0), w0, 0.0, pB=(0, -0.5, 0),
                                        imB=1.0, IB=IB)
    _, _, av, _, _ = O.solve(cs, None, v, w, pos, im, I, 0, 1, friction=0.6)
    assert av[0, 1] == np.float32(2.0 - 0.2)
    # B far to the side: dot(ab, ac) < 0.95 -> not damped
    cs, pos, v, w, im, I = _ground_rows((0, 0.5, 0), (0, 0, 0), w0, 0.0, pB=(3.0, -0.5, 0),
                                        imB=1.0, IB=IB)
    _, _, av, _, _ = O.solve(cs, None, v, w, pos, im, I, 0, 1, friction=0.6)
    assert av[0, 1] == np.float32(2.0)


def test_friction_stops_sliding_inside_the_cone():
    # a cube sliding on a static ground with a large normal impulse: one
    # friction sweep brings the contact centre's tangential velocity to zero
    # (the rows decouple when the lever arm is along the normal), and the
    # impulses are the closed-form m_eff * v_t
    vA = np.array([1.5, 0.0, -0.8], np.float32)
    cs, pos, v, w, im, I = _ground_rows((0, 0.5, 0), vA, (0, 0, 0), 100.0)
    cs, lv, av, _, _ = O.solve(cs, None, v, w, pos, im, I, 0, 1, friction=0.6)
    r = np.array([0, -0.5, 0])
    vp = lv[0, :3].astype(np.float64) + np.cross(av[0, :3].astype(np.float64), r)
    assert abs(vp[1]) < 1e-6 and np.abs(vp).max() < 1e-5
    n = np.array([0.0, -1.0, 0.0])
    for i, t in enumerate(O.plane_space(n.astype(np.float32))):
        t = t.astype(np.float64)
        meff = 1.0 / effective(r, t, 1.0, I[0])
        # the row's impulse is along -t: lambda = m_eff * (v . t)
        assert np.isclose(cs["f_lambda"][0, i], meff * (vA @ t), rtol=1e-5, atol=1e-7)


def test_friction_is_clamped_to_the_cone():
    vA = np.array([4.0, 0.0, 0.0], np.float32)
    L = np.float32(0.05)
    cs, pos, v, w, im, I = _ground_rows((0, 0.5, 0), vA, (0, 0, 0), L)
    cs, lv, _, _, _ = O.solve(cs, None, v, w, pos, im, I, 0, 1, friction=0.6)
    lim = np.float32(np.float32(0.6) * L)
    f = cs["f_lambda"][0]
    assert np.max(np.abs(f)) == lim
    assert lv[0, 0] > 4.0 - 2 * lim - 1e-6


def test_solve_segments_is_solve_per_segment():
    """orc_solve_segments (the GPU's tile order: independent segments in one
    call) equals orc_solve run on each segment's row order in turn, and a
    segment's rows touch only its own bodies, so the segments' order does
    not matter.  Two stacks of two boxes on their own grounds, sliding."""
    dt = np.float32(1 / 60)
    rng = np.random.default_rng(3)
    # bodies: grounds 0 and 3, stack A = 1, 2 on 0; stack B = 4, 5 on 3
    pos = np.array([[0, -0.5, 0], [0, 0.5, 0], [0, 1.5, 0],
                    [9, -0.5, 0], [9, 0.5, 0], [9, 1.5, 0]], np.float32)
    im = np.array([0, 1, 0.5, 0, 2, 1], np.float32)
    I = world_inertia(6, [(0, 0, 0), (6, 6, 6), (3, 3, 3), (0, 0, 0), (12, 12, 12), (6, 6, 6)])
    v0 = rng.normal(size=(6, 3)).astype(np.float32) * 0.5
    v0[[0, 3]] = 0
    w0 = rng.normal(size=(6, 3)).astype(np.float32) * 0.3
    w0[[0, 3]] = 0
    cs = []
    for a, b, x in ((1, 0, 0), (2, 1, 0), (4, 3, 9), (5, 4, 9)):
        c = ground_contact(0.01, body=a)
        c["body_b"] = b
        c["pos"][0, :, 0] += x
        c["pos"][0, :, 1] += pos[a, 1] - 0.5
        cs.append(O.contact_to_constraint(c, pos, v0, w0, im, I, dt))
    cons = np.concatenate(cs)
    segs = [np.array([1, 0], np.int32), np.array([3, 2], np.int32)]
    got = O.solve_segments(cons, segs, v0, w0, pos, im, I, 7, 5, friction=0.6)
    c1, lv, av = cons, v0, w0
    for seg in segs:
        c1, lv, av, _, _ = O.solve(c1, seg, lv, av, pos, im, I, 7, 5, friction=0.6)
    assert np.array_equal(got[0], c1)
    assert np.array_equal(got[1], lv) and np.array_equal(got[2], av)
    rev = O.solve_segments(cons, segs[::-1], v0, w0, pos, im, I, 7, 5, friction=0.6)
    assert np.array_equal(rev[1], lv) and np.array_equal(rev[0], c1)
