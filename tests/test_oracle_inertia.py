"""The oracle's registration inertia (orc_register_inertia: the world AABB of
CLPhysicsDemo::registerPhysicsInstance, CLPhysicsDemo.cpp:536-552, and the box
inertia of btGpuNarrowphaseAndSolver::registerRigidBody,
btGpuNarrowPhaseAndSolver.cpp:614-640) against closed forms computed here in
float64 from the textbook rotation matrix of a unit quaternion."""
import numpy as np
import pytest

import oracle as O


def rot(q):
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def closed_form(mass, half, q, margin=0.01):
    # world AABB half extents of the box grown by the margin, rotated by q;
    # the reference then takes l = 2 x (AABB size) = 4 x half extent
    e = np.abs(rot(q)) @ (np.asarray(half, np.float64) + margin)
    l = 4.0 * e
    return np.array([12.0 / (mass * (l[1] ** 2 + l[2] ** 2)),
                     12.0 / (mass * (l[0] ** 2 + l[2] ** 2)),
                     12.0 / (mass * (l[0] ** 2 + l[1] ** 2))])


def test_unit_cube_at_identity():
    got = O.register_inertia(1.0, [-.5, -.5, -.5], [.5, .5, .5], [0, 0, 0], [0, 0, 0, 1])
    # l = 2 * 2 * (0.5 + 0.01) = 2.04 on every axis
    assert np.allclose(got[:3], 12.0 / (2 * 2.04 ** 2), rtol=1e-6)
    assert got[3] == 0


def test_static_body_has_no_inertia():
    got = O.register_inertia(0.0, [-1, -1, -1], [1, 1, 1], [0, 0, 0], [0, 0, 0, 1])
    assert np.array_equal(got, np.zeros(4, np.float32))


@pytest.mark.parametrize("q", [[0, 0, 0, 1], [0, np.sin(np.pi / 4), 0, np.cos(np.pi / 4)],
                               [0, np.sin(np.pi / 8), 0, np.cos(np.pi / 8)],
                               [0.1, 0.2, 0.3, 0.9], [-0.5, 0.1, 0.7, -0.2]])
@pytest.mark.parametrize("pos", [[0, 0, 0], [37.5, 61.0, -12.25]])
def test_rotated_boxes_match_closed_form(q, pos):
    half = np.array([1.0, 0.5, 0.25])
    centre = np.array([0.1, -0.2, 0.05])
    m = 3.5
    q = np.asarray(q, np.float32)
    got = O.register_inertia(m, centre - half, centre + half, pos, q)
    # the AABB is formed as centre +- extent in float at the body's position,
    # so the size carries the rounding of that position: 1e-5 relative
    assert np.allclose(got[:3], closed_form(m, half, q), rtol=1e-5)


def test_quarter_turn_about_y_swaps_x_and_z():
    half = [1.0, 0.5, 0.25]
    a = O.register_inertia(2.0, [-h for h in half], half, [0, 0, 0], [0, 0, 0, 1])
    s = np.float32(np.sqrt(0.5))
    b = O.register_inertia(2.0, [-h for h in half], half, [0, 0, 0], [0, s, 0, s])
    assert np.allclose(b[[2, 1, 0]], a[:3], rtol=1e-6)
    assert not np.allclose(a[0], a[2])


def test_quaternion_scale_does_not_matter():
    # setRotation divides by the squared norm (btMatrix3x3.h:206-208)
    q = np.array([0.3, -0.4, 0.2, 0.8], np.float32)
    a = O.register_inertia(1.5, [-.3, -.7, -.2], [.4, .7, .5], [1, 2, 3], q / np.linalg.norm(q))
    b = O.register_inertia(1.5, [-.3, -.7, -.2], [.4, .7, .5], [1, 2, 3], 2 * q)
    assert np.allclose(a, b, rtol=1e-6)
