"""Pins composite bodies (XMLParser.createComposite, RigidBodyGeomComposite,
CollisionProcessor.narrowPhase :711-722, Contact.set :166-177) in the
oracle without retyping its arithmetic:
- mass properties: two boxes side by side form one box, so mass, centre and
  inertia must equal the closed form of the union; a box-and-sphere
  composite matches an independent parallel-axis sum;
- a composite of one box at its centre collides exactly like that box (same
  points, normals, depths, Jacobians), its contacts carrying the sub-body in
  the bv id;
- a two-part composite collides like its parts placed as separate bodies at
  the composed poses, part by part in order (points, normals, depths, info);
- composite against composite visits the first body's parts outermost."""
import math

import numpy as np
import pytest

from adaptivemerging_amd._lib import AM_BV_NONE, AM_BV_SUB, AM_GEOM_COMPOSITE
from adaptivemerging_amd.scene import Scene, random_rotation, rot_axis_angle
from oracle.pyoracle import Oracle

GEOM_KEYS = ("point_b1", "normal_b1", "violation", "info")


def test_two_boxes_make_one_box():
    s = Scene()
    s.add_composite([dict(kind="box", size=[1.0, 0.5, 0.4], density=3.0, x=[0.5, 0, 0]),
                     dict(kind="box", size=[1.0, 0.5, 0.4], density=3.0, x=[1.5, 0, 0])], x=[2.0, 1.0, 0.0])
    b = s.bodies[0]
    assert int(b["geom"]) == AM_GEOM_COMPOSITE and int(b["tree_count"]) == 2
    M = 3.0 * 2.0 * 0.5 * 0.4
    assert float(b["mass"]) == pytest.approx(M, rel=1e-15)
    np.testing.assert_allclose(s.x[0], [3.0, 1.0, 0.0], rtol=1e-15)      # x + centre of mass (1, 0, 0)
    want = np.diag([M * (0.25 + 0.16) / 12, M * (4.0 + 0.16) / 12, M * (4.0 + 0.25) / 12])
    np.testing.assert_allclose(np.reshape(b["inertia0"], (3, 3)), want, rtol=1e-13, atol=1e-15)
    bb = np.reshape(b["bbox"], (8, 3))
    np.testing.assert_allclose(bb.min(axis=0), [-1.0, -0.25, -0.2]); np.testing.assert_allclose(bb.max(axis=0), [1.0, 0.25, 0.2])
    # the parts sit at -0.5 and +0.5 from the centre of mass
    np.testing.assert_allclose([t[9:] for t in s.sub_b2c], [[-0.5, 0, 0], [0.5, 0, 0]], atol=1e-15)


def test_box_and_sphere_parallel_axis():
    rng = np.random.default_rng(0)
    Rb = random_rotation(rng)
    s = Scene()
    s.add_composite([dict(kind="box", size=[0.6, 0.2, 0.3], density=2.0, x=[0.1, 0.2, -0.1], R=Rb),
                     dict(kind="sphere", r=0.25, density=5.0, x=[-0.4, 0.0, 0.3])])
    m1 = 0.6 * 0.2 * 0.3 * 2.0
    m2 = 4.0 / 3.0 * math.pi * 0.25 ** 3 * 5.0
    c = (m1 * np.array([0.1, 0.2, -0.1]) + m2 * np.array([-0.4, 0.0, 0.3])) / (m1 + m2)
    I1 = Rb @ np.diag([m1 * (0.04 + 0.09) / 12, m1 * (0.36 + 0.09) / 12, m1 * (0.36 + 0.04) / 12]) @ Rb.T
    I2 = np.eye(3) * 0.4 * m2 * 0.25 ** 2
    I = I1 + I2
    for m, p in ((m1, [0.1, 0.2, -0.1]), (m2, [-0.4, 0.0, 0.3])):
        d = np.array(p) - c
        I += m * (d @ d * np.eye(3) - np.outer(d, d))
    b = s.bodies[0]
    assert float(b["mass"]) == pytest.approx(m1 + m2, rel=1e-14)
    np.testing.assert_allclose(s.x[0], c, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(np.reshape(b["inertia0"], (3, 3)), I, rtol=1e-12, atol=1e-15)


def _contacts_on(scene):
    o = Oracle(scene)
    o.detect()
    return o.contacts()


def test_single_box_composite_collides_like_the_box():
    rng = np.random.default_rng(1)
    R = random_rotation(rng)
    a, b = Scene(), Scene()
    for s in (a, b):
        s.add_plane([0, 0, 0], [0, 1, 0])
    a.add_box([0.5, 0.3, 0.4], [0.0, 0.12, 0.0], R=R)
    a.add_box([0.3, 0.3, 0.3], [0.1, 0.35, 0.05], R=R.T)
    b.add_composite([dict(kind="box", size=[0.5, 0.3, 0.4])], x=[0.0, 0.12, 0.0], R=R)
    b.add_box([0.3, 0.3, 0.3], [0.1, 0.35, 0.05], R=R.T)
    ca, cb = _contacts_on(a), _contacts_on(b)
    assert len(ca["body1"]) == len(cb["body1"]) >= 3
    for k in GEOM_KEYS + ("jacobian", "body1", "body2"):
        np.testing.assert_array_equal(ca[k], cb[k], err_msg=k)
    for side in ("bv1", "bv2"):
        comp = np.where((cb["body1"] if side == "bv1" else cb["body2"]) == 1)[0]
        assert np.all(cb[side][comp] == AM_BV_SUB(0, AM_BV_NONE))
        assert np.all(ca[side][comp] == AM_BV_NONE)


def _compose(R, x, Rbc, t):
    """RigidTransform3D.mult in the oracle's order: (R t + x, R Rbc)."""
    xw = [R[i][0] * t[0] + R[i][1] * t[1] + R[i][2] * t[2] + x[i] for i in range(3)]
    Rw = [[R[i][0] * Rbc[0][j] + R[i][1] * Rbc[1][j] + R[i][2] * Rbc[2][j] for j in range(3)] for i in range(3)]
    return xw, np.array(Rw)


def test_two_part_composite_collides_like_its_parts():
    R = rot_axis_angle([0.3, 1.0, 0.2], 0.7)
    Rp = rot_axis_angle([1.0, 0.0, 0.5], 0.4)
    parts = [dict(kind="box", size=[0.6, 0.2, 0.3], x=[0.3, 0.0, 0.0], R=Rp),
             dict(kind="sphere", r=0.15, x=[-0.35, 0.05, 0.1])]
    comp = Scene()
    comp.add_plane([0, 0, 0], [0, 1, 0])
    ci = comp.add_composite(parts, x=[0.0, 0.1, 0.0], R=R)
    sep = Scene()
    sep.add_plane([0, 0, 0], [0, 1, 0])
    Rm = [list(r) for r in np.asarray(R)]
    for k, pt in enumerate(parts):
        t = comp.sub_b2c[k]
        xw, Rw = _compose(Rm, comp.x[ci], np.reshape(t[:9], (3, 3)).tolist(), t[9:])
        if pt["kind"] == "box":
            sep.add_box(pt["size"], xw, R=Rw)
        else:
            sep.add_sphere(pt["r"], xw, R=Rw)
    cc, cs = _contacts_on(comp), _contacts_on(sep)
    assert len(cc["body1"]) == len(cs["body1"]) > 1
    for k in ("violation", "info"):
        np.testing.assert_array_equal(cc[k], cs[k], err_msg=k)

    def world(scene, c, j):
        b = int(c["body1"][j])
        R = np.asarray(scene.R[b]).reshape(3, 3)
        return R @ c["point_b1"][j] + scene.x[b], R @ c["normal_b1"][j]

    for j in range(len(cc["body1"])):
        pc, nc = world(comp, cc, j)
        ps, ns = world(sep, cs, j)
        np.testing.assert_allclose(pc, ps, atol=1e-15)
        np.testing.assert_allclose(nc, ns, atol=1e-15)
        # the composite is one side of every contact; its bv is the part and the part's own id
        side = "1" if int(cc["body1"][j]) == ci else "2"
        part = int(cs["body" + side][j]) - 1
        assert cc["bv" + side][j] == AM_BV_SUB(part, int(cs["bv" + side][j]))


def test_composite_against_composite_order():
    """Two two-box composites overlapping: the contacts come part of the first
    by part of the second, like the separate boxes' pairs in that order."""
    parts = [dict(kind="box", size=[0.4, 0.4, 0.4], x=[-0.3, 0, 0]), dict(kind="box", size=[0.4, 0.4, 0.4], x=[0.3, 0, 0])]
    comp = Scene()
    comp.add_composite(parts, x=[0.0, 0.0, 0.0])
    comp.add_composite(parts, x=[0.05, 0.35, 0.02], R=rot_axis_angle([0, 0, 1], 0.2))
    o = Oracle(comp)
    o.detect()
    c = o.contacts()
    assert len(c["body1"]) > 0
    order = [((-3 - int(b1)) >> 16, (-3 - int(b2)) >> 16) for b1, b2 in zip(c["bv1"], c["bv2"])]
    assert order == sorted(order)
    assert all(int(b) == 0 for b in c["body1"]) and all(int(b) == 1 for b in c["body2"])


def _warm_pair(comp_first: bool):
    """A box and a composite (part 0 a box, part 1 a sphere) with three old
    contacts and three new ones: two on the box part whose info numbers swap
    between the steps (their points do not move), one on the sphere part
    whose old point lies 0.2 m from the new one."""
    s = Scene()
    s.add_box([2.0, 0.4, 2.0], [0.0, 0.0, 0.0])
    s.add_composite([dict(kind="box", size=[0.4, 0.2, 0.4], x=[0.0, 0.0, 0.0]),
                     dict(kind="sphere", r=0.1, x=[0.6, 0.0, 0.0])], x=[0.0, 0.3, 0.0])
    box, comp = 0, 1
    o = Oracle(s)
    b1, b2 = (comp, box) if comp_first else (box, comp)
    part = lambda q: AM_BV_SUB(q, AM_BV_NONE if q == 0 else 0)
    bvs = [part(0), part(0), part(1)]
    bv1 = [bv if comp_first else AM_BV_NONE for bv in bvs]
    bv2 = [AM_BV_NONE if comp_first else bv for bv in bvs]
    n = np.tile([0.0, -1.0 if comp_first else 1.0, 0.0], (3, 1))
    old_p = np.array([[-0.2, 0.2, -0.2], [0.2, 0.2, 0.2], [0.5, 0.2, 0.0]])
    o.push_contacts([b1] * 3, [b2] * 3, old_p, n, [-1e-3] * 3, info=[0, 1, 0], bv1=bv1, bv2=bv2)
    o.push_lambdas(np.array([[1.0, 0.1, 0.0], [2.0, 0.0, 0.2], [3.0, 0.0, 0.0]]))
    o.update_contacts_map()
    new_p = np.array([[0.2, 0.2, 0.2], [-0.2, 0.2, -0.2], [0.7, 0.2, 0.0]])
    o.push_contacts([b1] * 3, [b2] * 3, new_p, n, [-1e-3] * 3, info=[0, 1, 0], bv1=bv1, bv2=bv2)
    bad, rep = o.warm_start()
    return o.contacts(), bad, rep


def test_composite_box_part_warm_start_repairs_info_swap():
    """CollisionProcessor.warmStart (:489-503): a box-or-composite pair takes
    the box-box repair on a box part (the info numbers of the two box-part
    contacts swapped, each takes the lambda of the old contact at its own
    point); a sphere part of body 1 takes the plain hash match whatever its
    distance."""
    c, bad, rep = _warm_pair(comp_first=True)
    np.testing.assert_array_equal(c["lambda"][0], [2.0, 0.0, 0.2])
    np.testing.assert_array_equal(c["lambda"][1], [1.0, 0.1, 0.0])
    np.testing.assert_array_equal(c["lambda"][2], [3.0, 0.0, 0.0])
    assert list(c["new_this_step"]) == [0, 0, 0]
    assert (bad, rep) == (2, 2)


def test_composite_sphere_part_of_body2_falls_through():
    """A sphere part of body 2 takes the plain hash match and then goes on
    through the box-box test (:501-503): its old point is 0.2 m away and no
    other info matches better, so it keeps the plain match's lambda but is
    marked new."""
    c, bad, rep = _warm_pair(comp_first=False)
    np.testing.assert_array_equal(c["lambda"][0], [2.0, 0.0, 0.2])
    np.testing.assert_array_equal(c["lambda"][1], [1.0, 0.1, 0.0])
    np.testing.assert_array_equal(c["lambda"][2], [3.0, 0.0, 0.0])
    assert list(c["new_this_step"]) == [0, 0, 1]
    assert bad == 3
