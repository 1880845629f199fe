"""Pins the oracle's spring forces (orc_apply_springs, after Spring.java:
139-209 and RigidBody.applyForceW) with checks that come from the physics,
not from the oracle's own arithmetic: the elastic force is minus the
gradient of the potential k/2 (|d| - l0)^2 (central differences over a
translation and a rotation of the body), the dashpot term opposes the
closing speed along the spring, a body-body spring exerts equal and opposite
forces along one line (zero net force and zero net moment), the zero-length
spring keeps the field's rest length 0.5 and the 1e-3 cut-off, and the
stiffness/damping multipliers scale the two terms."""
import math

import numpy as np
import pytest

from adaptivemerging_amd.scene import Scene, random_rotation
from oracle.pyoracle import Oracle


def _forces(scene, x=None, R=None, v=None, w=None, **params):
    o = Oracle(scene)
    o.set_params(use_gravity=0, **params)
    o.set_state(x, R, v, w)
    o.clear_bodies()
    o.apply_springs()
    f, t = o.pull_forces()
    return f, t


def _skew_rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _world_spring_scene(rng):
    s = Scene()
    s.add_box([0.5, 0.3, 0.4], [0.0, 2.0, 0.0], density=300.0)
    pB = [0.2, 0.15, -0.1]
    pW = [0.4, 4.0, 0.3]
    s.add_spring(0, pB, pW=pW, k=250.0, d=7.0)
    return s, np.array(pB), np.array(pW)


def test_world_spring_is_minus_potential_gradient():
    """With the dashpot off the force is -dE/dx and the torque -dE/dtheta of
    E = k/2 (|pW - p| - l0)^2, p = R pB + x, by central differences."""
    rng = np.random.default_rng(0)
    s, pB, pW = _world_spring_scene(rng)
    l0 = float(s.spring_array()[0]["l0"])
    k = 250.0
    x = np.array([[0.3, 1.2, -0.2]])
    R = random_rotation(rng)

    def energy(xx, RR):
        L = np.linalg.norm(pW - (RR @ pB + xx))
        return 0.5 * k * (L - l0) ** 2

    f, t = _forces(s, x=x, R=R.reshape(1, 9), spring_damping_mod=0.0)
    h = 1e-6
    grad_x = np.array([(energy(x[0] + h * e, R) - energy(x[0] - h * e, R)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(f[0], -grad_x, rtol=1e-6, atol=1e-6)
    # rotation about the centre of mass by a small angle about each world axis
    grad_r = np.array([(energy(x[0], _skew_rot(e, h) @ R) - energy(x[0], _skew_rot(e, -h) @ R)) / (2 * h)
                       for e in np.eye(3)])
    np.testing.assert_allclose(t[0], -grad_r, rtol=1e-6, atol=1e-6)


def test_world_spring_rest_length_and_direction():
    rng = np.random.default_rng(1)
    s, pB, pW = _world_spring_scene(rng)
    sp = s.spring_array()[0]
    # constructed at rest: l0 is the initial distance, so no force at the start
    assert sp["l0"] == pytest.approx(np.linalg.norm(pW - (pB + [0.0, 2.0, 0.0])), rel=1e-15)
    f, t = _forces(s)
    np.testing.assert_allclose(f[0], 0.0, atol=1e-12)
    np.testing.assert_allclose(t[0], 0.0, atol=1e-12)
    # pulled away from the anchor: the force points at the anchor with magnitude k * stretch
    x = np.array([[0.0, 1.0, 0.0]])
    f, t = _forces(s, x=x)
    p = pB + x[0]
    u = (pW - p) / np.linalg.norm(pW - p)
    assert np.dot(f[0], u) == pytest.approx(250.0 * (np.linalg.norm(pW - p) - sp["l0"]), rel=1e-12)
    np.testing.assert_allclose(np.cross(f[0], u), 0.0, atol=1e-11)
    np.testing.assert_allclose(t[0], np.cross(p - x[0], f[0]), rtol=1e-12, atol=1e-12)


def test_dashpot_opposes_the_closing_speed():
    """At the rest length only the dashpot acts: -d (v_p . u) u, where v_p is
    the attachment point's velocity v + w x r and u points to the anchor."""
    rng = np.random.default_rng(2)
    s, pB, pW = _world_spring_scene(rng)
    v = np.array([[0.3, -1.1, 0.4]])
    w = np.array([[0.5, 0.2, -0.7]])
    f, _ = _forces(s, v=v, w=w)
    p = pB + np.array([0.0, 2.0, 0.0])
    u = (pW - p) / np.linalg.norm(pW - p)
    vp = v[0] + np.cross(w[0], p - np.array([0.0, 2.0, 0.0]))
    np.testing.assert_allclose(f[0], -7.0 * np.dot(vp, u) * u, rtol=1e-12, atol=1e-13)
    # moving away from the anchor (v_p . u < 0) the dashpot pulls towards it
    assert np.dot(vp, u) < 0 and np.dot(f[0], u) > 0


def test_multipliers_and_rest_length_scale():
    rng = np.random.default_rng(3)
    s, pB, pW = _world_spring_scene(rng)
    x = np.array([[0.1, 1.5, 0.2]])
    v = np.array([[0.0, 0.7, 0.0]])
    f1, _ = _forces(s, x=x, v=v, spring_damping_mod=0.0)
    f2, _ = _forces(s, x=x, v=v, spring_damping_mod=0.0, spring_stiffness_mod=3.0)
    np.testing.assert_allclose(f2, 3.0 * f1, rtol=1e-13, atol=1e-13)
    g1, _ = _forces(s, x=x, v=v, spring_stiffness_mod=0.0)
    g2, _ = _forces(s, x=x, v=v, spring_stiffness_mod=0.0, spring_damping_mod=0.5)
    np.testing.assert_allclose(g2, 0.5 * g1, rtol=1e-13, atol=1e-13)
    # ls scales the rest length: ls = 0 makes the spring zero-length (force k * |d|)
    s.springs[0]["ls"] = 0.0
    h, _ = _forces(s, x=x, spring_damping_mod=0.0)
    p = pB + x[0]
    np.testing.assert_allclose(h[0], 250.0 * (pW - p), rtol=1e-12)


def test_zero_spring_rest_length_and_cutoff():
    s = Scene()
    s.add_box([1, 1, 1], [1.0, 0.0, 0.0])
    s.add_spring(0, [0.5, 0.0, 0.0], k=40.0, d=0.0)
    sp = s.spring_array()[0]
    assert sp["type"] == 0 and sp["l0"] == 0.5 and list(sp["pw"]) == [1.5, 0.0, 0.0]
    # at the anchor (distance below 1e-3) the reference skips the spring
    f, t = _forces(s, x=np.array([[1.0 + 5e-4, 0.0, 0.0]]))
    assert not f.any() and not t.any()
    # one unit away: stretch 1 - 0.5 towards the anchor
    f, t = _forces(s, x=np.array([[2.0, 0.0, 0.0]]))
    np.testing.assert_allclose(f[0], [-40.0 * 0.5, 0.0, 0.0], rtol=1e-15)
    # closer than the rest length the zero spring pushes away from its anchor
    f, _ = _forces(s, x=np.array([[1.2, 0.0, 0.0]]))
    assert f[0][0] > 0


def test_body_body_spring_is_equal_and_opposite():
    rng = np.random.default_rng(4)
    s = Scene()
    s.add_box([0.4, 0.4, 0.4], [0.0, 1.0, 0.0])
    s.add_sphere(0.3, [1.5, 1.2, 0.3])
    s.add_spring(0, [0.2, 0.1, 0.0], body2=1, pB2=[0.0, 0.3, 0.0], k=80.0, d=3.0)
    sp = s.spring_array()[0]
    assert sp["type"] == 2 and sp["body2"] == 1
    x = np.array([[0.1, 0.8, -0.1], [2.2, 1.6, 0.1]])
    R = np.stack([random_rotation(rng).reshape(9), random_rotation(rng).reshape(9)])
    v = rng.normal(size=(2, 3))
    w = rng.normal(size=(2, 3))
    f, t = _forces(s, x=x, R=R, v=v, w=w)
    p1 = R[0].reshape(3, 3) @ [0.2, 0.1, 0.0] + x[0]
    p2 = R[1].reshape(3, 3) @ [0.0, 0.3, 0.0] + x[1]
    u = (p2 - p1) / np.linalg.norm(p2 - p1)
    np.testing.assert_allclose(f[0] + f[1], 0.0, atol=1e-12)
    np.testing.assert_allclose(np.cross(f[0], u), 0.0, atol=1e-11)
    # no net moment about the origin: both forces act along the line p1-p2
    m = t[0] + np.cross(x[0], f[0]) + t[1] + np.cross(x[1], f[1])
    np.testing.assert_allclose(m, 0.0, atol=1e-10)
    # the magnitude: stretch plus the relative closing speed along u
    v1 = v[0] + np.cross(w[0], p1 - x[0])
    v2 = v[1] + np.cross(w[1], p2 - x[1])
    fi = 80.0 * (np.linalg.norm(p2 - p1) - sp["l0"]) + 3.0 * np.dot(v2 - v1, u)
    np.testing.assert_allclose(f[0], fi * u, rtol=1e-12, atol=1e-12)


def test_springs_add_after_gravity():
    """applyExternalForces: gravity first, then the springs, accumulated."""
    rng = np.random.default_rng(5)
    s, pB, pW = _world_spring_scene(rng)
    x = np.array([[0.0, 1.0, 0.0]])
    o = Oracle(s)
    o.set_params(gravity_amount=9.81)
    o.set_state(x=x)
    o.clear_bodies(); o.apply_gravity()
    g, _ = o.pull_forces()
    o.apply_springs()
    f, _ = o.pull_forces()
    fs, _ = _forces(s, x=x)
    np.testing.assert_allclose(f[0], g[0] + fs[0], rtol=1e-14)
    m = 0.5 * 0.3 * 0.4 * 300.0
    np.testing.assert_allclose(g[0], [0.0, -9.81 * m, 0.0], rtol=1e-12, atol=1e-12)
