"""Scene description: bodies, sphere trees and initial state as numpy arrays.

Body construction follows XMLParser.createBox / createSphere / createPlane
(src/mergingBodies3D/XMLParser.java:360-447) and PlaneRigidBody
(PlaneRigidBody.java:55-62); the .sph sphere-tree format follows
XMLParser.createMesh (:500-535).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import (AM_BODY_KINEMATIC, AM_BODY_PINNED, AM_GEOM_BOX, AM_GEOM_COLLECTION, AM_GEOM_COMPOSITE, AM_GEOM_PLANE, AM_GEOM_TREE, AM_SPRING_BODYBODY,
                   AM_SPRING_WORLD, AM_SPRING_ZERO, BODY_DTYPE, NODE_DTYPE, SPRING_DTYPE, KINEMATIC_DTYPE)


def rot_axis_angle(axis, angle) -> np.ndarray:
    """Row-major rotation matrix for a unit axis and an angle (radians)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    x, y, z = a
    c, s = math.cos(angle), math.sin(angle)
    C1 = 1 - c
    return np.array(
        [
            [c + x * x * C1, x * y * C1 - z * s, x * z * C1 + y * s],
            [y * x * C1 + z * s, c + y * y * C1, y * z * C1 - x * s],
            [z * x * C1 - y * s, z * y * C1 + x * s, c + z * z * C1],
        ]
    )


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


@dataclass
class Scene:
    bodies: list = field(default_factory=list)       # list of BODY_DTYPE records
    nodes: list = field(default_factory=list)        # list of NODE_DTYPE records
    x: list = field(default_factory=list)
    R: list = field(default_factory=list)
    v: list = field(default_factory=list)
    w: list = field(default_factory=list)
    names: list = field(default_factory=list)
    springs: list = field(default_factory=list)      # list of SPRING_DTYPE records
    subs: list = field(default_factory=list)         # sub-bodies of composites (BODY_DTYPE records)
    sub_b2c: list = field(default_factory=list)      # their transforms into the composite frame (12 each)
    spinners: dict = field(default_factory=dict)     # body -> RigidBody.spinner (added to omega by the app's keys)
    kinematics: list = field(default_factory=list)   # prescribed motions (KINEMATIC_DTYPE records)

    # -- body builders -------------------------------------------------
    def _append(self, rec, x, R, v, w, name):
        self.bodies.append(rec)
        self.x.append(np.asarray(x, dtype=np.float64).reshape(3))
        self.R.append(np.eye(3) if R is None else np.asarray(R, dtype=np.float64).reshape(3, 3))
        self.v.append(np.zeros(3) if v is None else np.asarray(v, dtype=np.float64).reshape(3))
        self.w.append(np.zeros(3) if w is None else np.asarray(w, dtype=np.float64).reshape(3))
        self.names.append(name or f"body{len(self.bodies) - 1}")
        return len(self.bodies) - 1

    def add_box(self, size, x, R=None, v=None, w=None, density=1.0, friction=0.8, restitution=0.0,
                pinned=False, scene=0, name=None) -> int:
        """XMLParser.createBox (XMLParser.java:360-396)."""
        s = np.asarray(size, dtype=np.float64)
        rec = np.zeros((), dtype=BODY_DTYPE)
        rec["geom"] = AM_GEOM_BOX
        rec["scene"] = scene
        rec["size"] = s
        m = s[0] * s[1] * s[2] * density
        I = np.array([1.0 / 12 * m * (s[1] * s[1] + s[2] * s[2]),
                      1.0 / 12 * m * (s[0] * s[0] + s[2] * s[2]),
                      1.0 / 12 * m * (s[0] * s[0] + s[1] * s[1])])
        h = 0.5 * s
        bb = []
        for sx in (-1, 1):
            for sy in (-1, 1):
                for sz in (-1, 1):
                    bb += [sx * h[0], sy * h[1], sz * h[2]]
        rec["bbox"] = bb
        rec["radius"] = math.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
        rec["friction"] = friction
        rec["restitution"] = restitution
        if pinned:
            rec["flags"] = AM_BODY_PINNED
        else:
            rec["mass"] = m
            rec["minv"] = 1.0 / m
            rec["jinv0"] = [1.0 / I[0], 0, 0, 0, 1.0 / I[1], 0, 0, 0, 1.0 / I[2]]
            rec["inertia0"] = [I[0], 0, 0, 0, I[1], 0, 0, 0, I[2]]
        return self._append(rec, x, R, v, w, name)

    def add_plane(self, p, n, friction=0.8, restitution=0.0, scene=0, name=None) -> int:
        """PlaneRigidBody(p, n) (PlaneRigidBody.java:55-62): pinned, d = -(p.n)."""
        n = np.asarray(n, dtype=np.float64)
        n = n * (1.0 / math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]))
        p = np.asarray(p, dtype=np.float64)
        rec = np.zeros((), dtype=BODY_DTYPE)
        rec["geom"] = AM_GEOM_PLANE
        rec["flags"] = AM_BODY_PINNED
        rec["scene"] = scene
        rec["plane_n"] = n
        rec["plane_p"] = p
        rec["plane_d"] = -(p[0] * n[0] + p[1] * n[1] + p[2] * n[2])
        rec["friction"] = friction
        rec["restitution"] = restitution
        return self._append(rec, np.zeros(3), None, None, None, name)

    def add_sphere(self, r, x, R=None, v=None, w=None, density=1.0, friction=0.8, restitution=0.0, scene=0,
                   name=None) -> int:
        """XMLParser.createSphere (XMLParser.java:420-447): a one-leaf sphere
        tree; RigidBody.radius stays 0 as in the reference."""
        m = 4.0 / 3 * math.pi * r * r * r * density
        Ic = 2.0 / 5 * m * r * r
        rec = np.zeros((), dtype=BODY_DTYPE)
        rec["geom"] = AM_GEOM_TREE
        rec["scene"] = scene
        rec["mass"] = m
        rec["minv"] = 1.0 / m
        rec["jinv0"] = [1.0 / Ic, 0, 0, 0, 1.0 / Ic, 0, 0, 0, 1.0 / Ic]
        rec["inertia0"] = [Ic, 0, 0, 0, Ic, 0, 0, 0, Ic]
        rec["friction"] = friction
        rec["restitution"] = restitution
        bb = []
        for sx in (-1, 1):
            for sy in (-1, 1):
                for sz in (-1, 1):
                    bb += [sx * r, sy * r, sz * r]
        rec["bbox"] = bb
        rec["tree_first"] = len(self.nodes)
        rec["tree_count"] = 1
        node = np.zeros((), dtype=NODE_DTYPE)
        node["r"] = r
        node["first_child"] = -1
        self.nodes.append(node)
        return self._append(rec, x, R, v, w, name)

    def add_tree(self, centres, radii, children, x, R=None, v=None, w=None, mass=1.0, inertia=(1, 1, 1),
                 bbox=None, friction=0.8, restitution=0.0, scene=0, name=None, pinned=False) -> int:
        """A sphere-tree body.  children[i] lists node i's children; node 0 is
        the root.  Nodes are re-indexed breadth first so every node's children
        are contiguous (the traversal visits them in the given order); nodes
        the root does not reach are dropped, as the reference never visits
        them.  inertia is the diagonal (3 values) or the full rest-pose tensor
        (9 values, row-major)."""
        order, q = [0], 0
        while q < len(order):
            order += list(children[order[q]])
            q += 1
        n = len(order)
        assert len(set(order)) == n, "a sphere tree node has two parents"
        new_index = {old: k for k, old in enumerate(order)}
        first = len(self.nodes)
        for old in order:
            node = np.zeros((), dtype=NODE_DTYPE)
            node["c"] = centres[old]
            node["r"] = radii[old]
            ch = list(children[old])
            if ch:
                node["first_child"] = new_index[ch[0]]
                node["n_children"] = len(ch)
                assert [new_index[c] for c in ch] == list(range(new_index[ch[0]], new_index[ch[0]] + len(ch)))
            else:
                node["first_child"] = -1
                node["n_children"] = 0
            self.nodes.append(node)
        rec = np.zeros((), dtype=BODY_DTYPE)
        rec["geom"] = AM_GEOM_TREE
        rec["scene"] = scene
        if pinned:
            # RigidBody with pinned set afterwards (XMLParser.java:577-583)
            rec["flags"] = AM_BODY_PINNED
        else:
            rec["mass"] = mass
            rec["minv"] = 1.0 / mass
            if len(inertia) == 3:
                rec["jinv0"] = [1.0 / inertia[0], 0, 0, 0, 1.0 / inertia[1], 0, 0, 0, 1.0 / inertia[2]]
                rec["inertia0"] = [inertia[0], 0, 0, 0, inertia[1], 0, 0, 0, inertia[2]]
            else:
                I = np.asarray(inertia, dtype=np.float64).reshape(3, 3)
                rec["inertia0"] = I.reshape(9)
                rec["jinv0"] = np.linalg.inv(I).reshape(9)
        rec["friction"] = friction
        rec["restitution"] = restitution
        if bbox is None:
            c = np.asarray([centres[i] for i in order])
            r = np.asarray([radii[i] for i in order])
            lo = (c - r[:, None]).min(axis=0)
            hi = (c + r[:, None]).max(axis=0)
            bbox = [[lx, ly, lz] for lx in (lo[0], hi[0]) for ly in (lo[1], hi[1]) for lz in (lo[2], hi[2])]
        rec["bbox"] = np.asarray(bbox, dtype=np.float64).reshape(24)
        rec["tree_first"] = first
        rec["tree_count"] = n
        return self._append(rec, x, R, v, w, name)

    def add_mesh(self, obj_path: str, sph_path: str, x, scale: float = 1.0, density: float = 1.0, R=None, v=None,
                 w=None, friction=0.8, restitution=0.0, pinned=False, scene=0, name=None) -> int:
        """XMLParser.createMesh (XMLParser.java:449-538): mass properties and
        bounding box from the OBJ mesh (mesh.mesh_body), collision geometry
        from the .sph sphere tree, both scaled and moved to the centre of
        mass."""
        from .mesh import mesh_body

        mb = mesh_body(obj_path, scale, density)
        centres, radii, children = load_sph(sph_path, scale, mb["com"])
        return self.add_tree(centres, radii, children, x, R=R, v=v, w=w, mass=mb["mass"], inertia=mb["inertia"],
                             bbox=mb["bbox"], friction=friction, restitution=restitution, scene=scene, name=name,
                             pinned=pinned)

    def add_composite(self, parts, x=(0.0, 0.0, 0.0), R=None, v=None, w=None, friction=0.8, restitution=0.0,
                      pinned=False, scene=0, name=None) -> int:
        """XMLParser.createComposite (XMLParser.java:180-353): a rigid body
        made of boxes and spheres.  parts: dicts with kind "box" (size) or
        "sphere" (r), density (default 1) and the part's pose in the
        composite's file frame, x and R.  The composite's mass is the sum, its
        centre of mass the mass-weighted mean, its inertia the sum of the
        parts' world-aligned inertias plus m (|d|^2 I - d d^T) for each part's
        offset d; its bounding box spans the parts' boxes (upper corner
        starting from Double.MIN_VALUE, as in the reference) and its pose is
        x + centre of mass.  Each part keeps its transform relative to the
        centre of mass."""
        if not parts:
            raise ValueError("a composite needs at least one part")
        built = []
        for pt in parts:
            dens = float(pt.get("density", 1.0))
            Rb = np.eye(3) if pt.get("R") is None else np.asarray(pt["R"], dtype=np.float64).reshape(3, 3)
            xb = [float(t) for t in pt.get("x", (0.0, 0.0, 0.0))]
            rec = np.zeros((), dtype=BODY_DTYPE)
            if pt["kind"] == "box":
                sz = np.asarray(pt["size"], dtype=np.float64)
                m = sz[0] * sz[1] * sz[2] * dens
                I0 = [1.0 / 12 * m * (sz[1] * sz[1] + sz[2] * sz[2]), 1.0 / 12 * m * (sz[0] * sz[0] + sz[2] * sz[2]),
                      1.0 / 12 * m * (sz[0] * sz[0] + sz[1] * sz[1])]
                h = 0.5 * sz
                bb = [[sx * h[0], sy * h[1], sz_ * h[2]] for sx in (-1, 1) for sy in (-1, 1) for sz_ in (-1, 1)]
                rec["geom"] = AM_GEOM_BOX
                rec["size"] = sz
                rec["radius"] = math.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
            elif pt["kind"] == "sphere":
                r = float(pt["r"])
                m = 4.0 / 3 * math.pi * r * r * r * dens
                I0 = [2.0 / 5 * m * r * r] * 3
                bb = [[sx * r, sy * r, sz_ * r] for sx in (-1, 1) for sy in (-1, 1) for sz_ in (-1, 1)]
                rec["geom"] = AM_GEOM_TREE
                rec["tree_first"] = len(self.nodes)
                rec["tree_count"] = 1
                rec["radius"] = r
                node = np.zeros((), dtype=NODE_DTYPE)
                node["r"] = r
                node["first_child"] = -1
                self.nodes.append(node)
            else:
                raise ValueError(f"composite parts are boxes or spheres, not {pt['kind']!r}")
            rec["bbox"] = np.asarray(bb).reshape(24)
            rec["scene"] = scene
            # massAngular as updateRotationalInertiaFromTransformation leaves it: R I0 R^T
            Iw = Rb @ np.diag(I0) @ Rb.T
            built.append((rec, m, xb, Rb, Iw, bb))
        mass = 0.0
        com = [0.0, 0.0, 0.0]
        for _, m, xb, _, _, _ in built:
            mass += m
            com = [m * xb[a] + com[a] for a in range(3)]
        inv = 1.0 / mass
        com = [com[a] * inv for a in range(3)]
        I = np.zeros((3, 3))
        for _, m, xb, _, Iw, _ in built:
            I = I + Iw
            d = [xb[0] - com[0], xb[1] - com[1], xb[2] - com[2]]
            op = np.array([[d[1] * d[1] + d[2] * d[2], -d[0] * d[1], -d[0] * d[2]],
                           [-d[1] * d[0], d[0] * d[0] + d[2] * d[2], -d[1] * d[2]],
                           [-d[2] * d[0], -d[2] * d[1], d[0] * d[0] + d[1] * d[1]]]) * m
            I = I + op
        ll = [1.7976931348623157e308] * 3
        ur = [5e-324] * 3
        for _, _, xb, Rb, _, bb in built:
            for q in bb:
                pw = [Rb[i, 0] * q[0] + Rb[i, 1] * q[1] + Rb[i, 2] * q[2] + xb[i] for i in range(3)]
                for a in range(3):
                    ll[a] = min(pw[a], ll[a])
                    ur[a] = max(pw[a], ur[a])
        ll = [ll[a] - com[a] for a in range(3)]
        ur = [ur[a] - com[a] for a in range(3)]
        bbox = [ll, [ll[0], ll[1], ur[2]], [ll[0], ur[1], ll[2]], [ll[0], ur[1], ur[2]], ur, [ur[0], ll[1], ll[2]],
                [ur[0], ll[1], ur[2]], [ur[0], ur[1], ll[2]]]
        first = len(self.subs)
        radius = 0.0
        for rec, m, xb, Rb, _, _ in built:
            t = [xb[0] - com[0], xb[1] - com[1], xb[2] - com[2]]
            self.subs.append(rec)
            self.sub_b2c.append(list(Rb.reshape(9)) + t)
            radius = max(radius, math.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) + float(rec["radius"]))
        crec = np.zeros((), dtype=BODY_DTYPE)
        crec["geom"] = AM_GEOM_COMPOSITE
        crec["scene"] = scene
        crec["tree_first"] = first
        crec["tree_count"] = len(built)
        # enclosing radius about the centre of mass: the broad phase's bound
        crec["radius"] = radius * (1 + 1e-12)
        crec["bbox"] = np.asarray(bbox).reshape(24)
        crec["friction"] = friction
        crec["restitution"] = restitution
        if pinned:
            crec["flags"] = AM_BODY_PINNED
        else:
            crec["mass"] = mass
            crec["minv"] = 1.0 / mass
            crec["inertia0"] = I.reshape(9)
            crec["jinv0"] = np.linalg.inv(I).reshape(9)
        pos = [float(t) for t in x]
        return self._append(crec, [pos[a] + com[a] for a in range(3)], R, v, w, name)

    def sub_arrays(self):
        """(sub-bodies, their transforms (n, 12)) for am_scene_upload_ex."""
        s = np.array(self.subs, dtype=BODY_DTYPE) if self.subs else np.zeros(0, dtype=BODY_DTYPE)
        t = np.ascontiguousarray(np.array(self.sub_b2c, dtype=np.float64).reshape(len(self.subs), 12))
        return s, t

    # -- springs (Spring.java:50-104) -----------------------------------
    def body_point_world(self, body: int, pB, x=None, R=None):
        """RigidTransform3D.transform: R pB + x, with the Matrix3d.transform
        evaluation order; x and R default to the body's current pose."""
        R = self.R[body] if R is None else np.asarray(R, dtype=np.float64).reshape(3, 3)
        x = self.x[body] if x is None else np.asarray(x, dtype=np.float64).reshape(3)
        p = [float(t) for t in pB]
        return [R[i, 0] * p[0] + R[i, 1] * p[1] + R[i, 2] * p[2] + x[i] for i in range(3)]

    def add_spring(self, body1: int, pB, pW=None, body2: int | None = None, pB2=None, k: float = 100.0,
                   d: float = 10.0, ls: float = 1.0, pose1=None) -> int:
        """Spring(pB, body), Spring(pB, body, pW) or Spring(pB1, body1, pB2,
        body2) (Spring.java:50-104).  The attachment points' world positions,
        hence pw and l0, come from the poses at construction: pose1 = (x, R)
        overrides body1's current pose (a spring parsed before a body's <x> or
        <R> sees the pose it had then)."""
        rec = np.zeros((), dtype=SPRING_DTYPE)
        rec["body1"] = body1
        rec["body2"] = -1
        rec["pb1"] = pB
        rec["k"], rec["d"], rec["ls"] = k, d, ls
        x1, R1 = (None, None) if pose1 is None else pose1
        p1w = self.body_point_world(body1, pB, x1, R1)
        if body2 is not None:
            rec["type"] = AM_SPRING_BODYBODY
            rec["body2"] = body2
            rec["pb2"] = pB2
            p2w = self.body_point_world(body2, pB2)
            disp = [p1w[i] - p2w[i] for i in range(3)]
            rec["l0"] = math.sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2])
        elif pW is not None:
            rec["type"] = AM_SPRING_WORLD
            rec["pw"] = pW
            disp = [float(pW[i]) - p1w[i] for i in range(3)]
            rec["l0"] = math.sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2])
        else:
            rec["type"] = AM_SPRING_ZERO
            rec["pw"] = p1w
            rec["l0"] = 0.5   # the field's initial value; the constructor leaves it
        self.springs.append(rec)
        return len(self.springs) - 1

    # -- prescribed motion (include/am.h am_kinematic) --------------------
    def set_kinematic(self, body: int, amplitude: float, omega: float, phase: float, direction=(0.0, 1.0, 0.0)):
        """Makes `body` kinematic: pinned, and at time t placed at
        x0 + (amplitude sin(omega t + phase)) direction with velocity
        (amplitude omega cos(omega t + phase)) direction, x0 its current
        position."""
        rec = self.bodies[body]
        rec["flags"] = int(rec["flags"]) | AM_BODY_PINNED | AM_BODY_KINEMATIC
        rec["mass"] = 0.0
        rec["minv"] = 0.0
        rec["jinv0"] = 0.0
        rec["inertia0"] = 0.0
        k = np.zeros((), dtype=KINEMATIC_DTYPE)
        k["body"] = body
        k["x0"] = self.x[body]
        k["dir"] = direction
        k["amplitude"], k["omega"], k["phase"] = amplitude, omega, phase
        self.kinematics.append(k)
        return len(self.kinematics) - 1

    def kinematic_array(self):
        if not self.kinematics:
            return np.zeros(0, dtype=KINEMATIC_DTYPE)
        return np.array(self.kinematics, dtype=KINEMATIC_DTYPE)

    def spring_array(self):
        return np.array(self.springs, dtype=SPRING_DTYPE) if self.springs else np.zeros(0, dtype=SPRING_DTYPE)

    # -- arrays ---------------------------------------------------------
    def arrays(self):
        """(bodies, nodes, x, R, v, w) as contiguous numpy arrays."""
        b = np.array(self.bodies, dtype=BODY_DTYPE) if self.bodies else np.zeros(0, dtype=BODY_DTYPE)
        nd = np.array(self.nodes, dtype=NODE_DTYPE) if self.nodes else np.zeros(0, dtype=NODE_DTYPE)
        n = len(self.bodies)
        x = np.ascontiguousarray(np.array(self.x).reshape(n, 3))
        R = np.ascontiguousarray(np.array(self.R).reshape(n, 9))
        v = np.ascontiguousarray(np.array(self.v).reshape(n, 3))
        w = np.ascontiguousarray(np.array(self.w).reshape(n, 3))
        return b, nd, x, R, v, w

    @property
    def n(self) -> int:
        return len(self.bodies)


def load_sph(path: str, scale: float = 1.0, com=(0.0, 0.0, 0.0)):
    """Reads a .sph sphere tree (XMLParser.java:500-535): first line N, then N
    lines 'x y z r', then lines 'parent count child...' (1-based).  Returns
    (centres, radii, children) with centres scaled and shifted by -com."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    n = int(lines[0][0])
    centres, radii = [], []
    for i in range(n):
        x, y, z, r = (float(t) for t in lines[1 + i][:4])
        centres.append([x * scale - com[0], y * scale - com[1], z * scale - com[2]])
        radii.append(r * scale)
    children = [[] for _ in range(n)]
    for ln in lines[1 + n:]:
        p, cnt = int(ln[0]) - 1, int(ln[1])
        children[p] = [int(t) - 1 for t in ln[2:2 + cnt]]
    return centres, radii, children
