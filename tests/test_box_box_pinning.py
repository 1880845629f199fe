"""Pins the oracle's box-box routine (BoxBox.dBoxBox, BoxBox.java:433-859)
where the closed forms of test_oracle.py leave it open: the branch where box
2's face is the reference face (codes 4-6, BoxBox.java:656-858), the
incident-face clipping (intersectRectQuad, :140-209) at oblique poses, and the
edge-edge contact (:600-651).

Every expectation here is computed independently of the oracle's expressions:
separating axes and overlaps from world-frame projections, the axis choice
from the reference's rule (faces replace on s > best + tinyOffset, edges on
s * fudgeFactor > best, :291-330), the contact polygon by clipping the
incident face against the reference face's four half-planes, points lifted
back onto the incident face by a line-plane intersection, and the edge-edge
point as the midpoint of the two edge lines' closest points.  Cases whose axis
choice sits within a margin of a tie are skipped, and each test asserts how
many cases of each branch it checked, so none of them can pass vacuously.

test_planted_mistakes_are_caught builds copies of the oracle with one
plausible mistake each (most of them in the box-2-reference branch) and
requires every copy to fail at least one of these checks."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from adaptivemerging_amd import Scene
from adaptivemerging_amd.scene import random_rotation, rot_axis_angle
from oracle.pyoracle import Oracle

TINY, FUDGE = 1e-5, 1.05  # BoxBox.java:50-51 (tinyOffset, fudgeFactor)
LCA_EPS = 1e-4  # BoxBox.java:52, dLineClosestApproach's parallel guard
MARGIN = 1e-7  # decisions closer than this to a tie are skipped
TOL = 1e-11


# ------------------------------------------------------------ independent geometry
def support(h, R, u):
    """Half-extent of the box (half sizes h, rotation R) along unit u."""
    return float(np.sum(h * np.abs(R.T @ u)))


def candidates(x1, R1, h1, x2, R2, h2):
    """(code, s, unit axis, |cross|) in the reference's test order; s < 0 is
    overlap along the axis."""
    p = x2 - x1
    out = []
    for k in range(3):
        u = R1[:, k]
        out.append((k + 1, abs(p @ u) - support(h1, R1, u) - support(h2, R2, u), u, 1.0))
    for k in range(3):
        u = R2[:, k]
        out.append((k + 4, abs(p @ u) - support(h1, R1, u) - support(h2, R2, u), u, 1.0))
    for i in range(3):
        for j in range(3):
            u = np.cross(R1[:, i], R2[:, j])
            ln = float(np.linalg.norm(u))
            if ln < 1e-12:
                out.append((7 + 3 * i + j, -math.inf, u, ln))
                continue
            u = u / ln
            out.append((7 + 3 * i + j, abs(p @ u) - support(h1, R1, u) - support(h2, R2, u), u, ln))
    return out


def choose(x1, R1, h1, x2, R2, h2):
    """The reference's choice of axis: None when an axis separates the boxes,
    else (code, s, normal from box 1 towards box 2).  The flag says whether
    some decision sat within MARGIN of a tie."""
    p = x2 - x1
    best, pick, amb = -math.inf, None, False
    for code, s, u, ln in candidates(x1, R1, h1, x2, R2, h2):
        if abs(s) < MARGIN:
            amb = True
        if s > 0:
            return None, amb
        if code <= 6:
            if best > -math.inf and abs(s - (best + TINY)) < MARGIN:
                amb = True
            if s > best + TINY:
                best, pick = s, (code, s, u)
        else:
            if ln < 0.05:
                # near-parallel edges: the normalisation amplifies rounding,
                # so such an axis must stay far from being chosen
                if s * FUDGE > best - 1e-3:
                    amb = True
                continue
            if abs(s * FUDGE - best) < MARGIN:
                amb = True
            if s * FUDGE > best:
                best, pick = s, (code, s, u)
    code, s, u = pick
    n = u if p @ u >= 0 else -u
    if abs(p @ u) < MARGIN:
        amb = True
    return (code, s, n), amb


def clip(poly, axis, sign, h):
    """Sutherland-Hodgman clip of a 2D polygon to sign * X[axis] <= h."""
    out = []
    for i in range(len(poly)):
        P, Q = poly[i], poly[(i + 1) % len(poly)]
        inP, inQ = sign * P[axis] <= h, sign * Q[axis] <= h
        if inP:
            out.append(P)
        if inP != inQ:
            t = (sign * h - P[axis]) / (Q[axis] - P[axis])
            out.append(P + t * (Q - P))
    return out


def expected_face(x1, R1, h1, x2, R2, h2, code, n12):
    """Face contact: (points, depths).  The reference face belongs to box 1
    for codes 1-3 and to box 2 for codes 4-6; the incident face is the other
    box's face most opposed to it."""
    if code <= 3:
        xa, Ra, ha, xb, Rb, hb, kN, n2 = x1, R1, h1, x2, R2, h2, code - 1, n12
    else:
        xa, Ra, ha, xb, Rb, hb, kN, n2 = x2, R2, h2, x1, R1, h1, code - 4, -n12
    o0, o1 = [k for k in range(3) if k != kN]
    nr = Rb.T @ n2
    lanr = int(np.argmax(np.abs(nr)))
    sgn = -1.0 if nr[lanr] > 0 else 1.0
    nb = Rb[:, lanr]
    fc = xb + sgn * hb[lanr] * nb
    a1, a2 = [k for k in range(3) if k != lanr]
    corners = [fc + s1 * hb[a1] * Rb[:, a1] + s2 * hb[a2] * Rb[:, a2]
               for s1, s2 in ((-1, -1), (-1, 1), (1, 1), (1, -1))]
    poly = [np.array([(q - xa) @ Ra[:, o0], (q - xa) @ Ra[:, o1]]) for q in corners]
    for axis, lim in ((0, ha[o0]), (1, ha[o1])):
        for sign in (-1.0, 1.0):
            poly = clip(poly, axis, sign, lim)
    pts, deps = [], []
    eN = Ra[:, kN]
    for X, Y in poly:
        base = xa + X * Ra[:, o0] + Y * Ra[:, o1]
        t = ((fc - base) @ nb) / (eN @ nb)
        q = base + t * eN
        d = ha[kN] - n2 @ (q - xa)
        if d >= 0:
            pts.append(q)
            deps.append(d)
    return np.array(pts).reshape(-1, 3), np.array(deps)


def expected_edge(x1, R1, h1, x2, R2, h2, code, n12, s):
    """Edge-edge contact: the midpoint of the closest points of box 1's edge
    farthest along the normal and box 2's edge farthest against it."""
    i, j = (code - 7) // 3, (code - 7) % 3
    pa = x1 + sum(np.sign(n12 @ R1[:, k]) * h1[k] * R1[:, k] for k in range(3))
    pb = x2 - sum(np.sign(n12 @ R2[:, k]) * h2[k] * R2[:, k] for k in range(3))
    ua, ub = R1[:, i], R2[:, j]
    M = np.array([[ua @ ua, -(ua @ ub)], [ua @ ub, -(ub @ ub)]])
    rhs = np.array([(pb - pa) @ ua, (pb - pa) @ ub])
    t, u = np.linalg.solve(M, rhs)
    return 0.5 * ((pa + t * ua) + (pb + u * ub)), -s


# ------------------------------------------------------------ oracle access
def pose_scene(h1, x1, R1, h2, x2, R2):
    s = Scene()
    s.add_box(2 * np.asarray(h1), x1, R1)
    s.add_box(2 * np.asarray(h2), x2, R2)
    return s


def run_oracle(scene, i, j, lib_path=None):
    """(world points, world normals, depths, info) of oracle box_box(i, j)."""
    o = Oracle(scene, lib_path=lib_path)
    n = o.box_box(i, j)
    c = o.contacts()
    x, R, _, _ = o.pull_state()
    b1 = c["body1"]
    pts = np.array([R[b1[k]].reshape(3, 3) @ c["point_b1"][k] + x[b1[k]] for k in range(n)]).reshape(-1, 3)
    nrm = np.array([R[b1[k]].reshape(3, 3) @ c["normal_b1"][k] for k in range(n)]).reshape(-1, 3)
    return pts, nrm, -c["violation"][:n], c["info"][:n]


def same_point_sets(a, b, tol):
    if len(a) != len(b):
        return False
    used = np.zeros(len(b), bool)
    for p in a:
        d = np.linalg.norm(b - p, axis=1)
        d[used] = np.inf
        k = int(np.argmin(d))
        if d[k] > tol:
            return False
        used[k] = True
    return True


# ------------------------------------------------------------ cases
def generic_case(rng):
    h1, h2 = rng.uniform(0.1, 0.3, 3), rng.uniform(0.1, 0.3, 3)
    return h1, np.zeros(3), random_rotation(rng), h2, rng.uniform(-0.35, 0.35, 3), random_rotation(rng)


def resting_case(rng):
    """A small box resting, tilted and shifted, on a wider one's face
    (penetration 1-20 mm); half of them with the wide box as body 2, so its
    face is the reference face of codes 4-6."""
    k = int(rng.integers(3))
    up = np.eye(3)[k]
    hw = rng.uniform(0.3, 0.6, 3)
    hw[k] = rng.uniform(0.05, 0.2)
    hs = rng.uniform(0.08, 0.25, 3)
    hs[k] = rng.uniform(0.05, 0.15)
    Rw = np.eye(3)
    Rs = rot_axis_angle(rng.normal(size=3), rng.uniform(0.0, 0.35))
    lat = rng.uniform(-1.0, 1.0, 3) * hw
    lat[k] = 0.0
    xs = lat.copy()
    corners = np.array([Rs @ (hs * np.array(c)) for c in np.array(np.meshgrid([-1, 1], [-1, 1], [-1, 1])).T.reshape(-1, 3)])
    xs[k] = hw[k] - corners[:, k].min() - rng.uniform(0.001, 0.02)
    if rng.uniform() < 0.5:
        return hs, xs, Rs, hw, np.zeros(3), Rw
    return hw, np.zeros(3), Rw, hs, xs, Rs


def case_list(n_generic=250, n_resting=250, seed=11):
    rng = np.random.default_rng(seed)
    return [generic_case(rng) for _ in range(n_generic)] + [resting_case(rng) for _ in range(n_resting)]


# ------------------------------------------------------------ the checks
def check_axis_points_depths(cases, lib_path=None):
    """Oracle contacts equal the independently computed ones; returns the
    number of cases checked per branch."""
    seen = {"face_box1": 0, "face_box2": 0, "edge": 0, "separated": 0}
    for h1, x1, R1, h2, x2, R2 in cases:
        pick, amb = choose(x1, R1, h1, x2, R2, h2)
        if amb:
            continue
        pts, nrm, dep, info = run_oracle(pose_scene(h1, x1, R1, h2, x2, R2), 0, 1, lib_path)
        if pick is None:
            assert len(pts) == 0, "separated boxes gave contacts"
            seen["separated"] += 1
            continue
        code, s, n12 = pick
        if code > 6:
            p, d = expected_edge(x1, R1, h1, x2, R2, h2, code, n12, s)
            if 1.0 - (R1[:, (code - 7) // 3] @ R2[:, (code - 7) % 3]) ** 2 <= 2 * LCA_EPS:
                continue
            assert len(pts) == 1, f"edge-edge code {code}: {len(pts)} contacts"
            np.testing.assert_allclose(pts[0], p, rtol=0, atol=TOL)
            np.testing.assert_allclose(dep[0], d, rtol=0, atol=TOL)
            seen["edge"] += 1
        else:
            p, d = expected_face(x1, R1, h1, x2, R2, h2, code, n12)
            assert len(pts) == len(p), f"face code {code}: {len(pts)} contacts, expected {len(p)}"
            assert same_point_sets(pts, p, TOL), f"face code {code}: contact points differ"
            order = [int(np.argmin(np.linalg.norm(p - q, axis=1))) for q in pts]
            np.testing.assert_allclose(dep, d[order], rtol=0, atol=TOL)
            seen["face_box1" if code <= 3 else "face_box2"] += 1
        assert list(info) == list(range(len(pts))), "info is not the emission order"
        for k in range(len(pts)):
            np.testing.assert_allclose(nrm[k], n12, rtol=0, atol=TOL)
    return seen


def check_swap_symmetry(cases, lib_path=None):
    """box_box(0, 1) against box_box(1, 0) where both orders pick the same
    physical axis: same points, negated normals, equal depths."""
    seen = {"face_box1": 0, "face_box2": 0, "edge": 0}
    for h1, x1, R1, h2, x2, R2 in cases:
        a, amb_a = choose(x1, R1, h1, x2, R2, h2)
        b, amb_b = choose(x2, R2, h2, x1, R1, h1)
        if amb_a or amb_b or a is None or b is None:
            continue
        ca, cb = a[0], b[0]
        if ca <= 3:
            same = cb == ca + 3
        elif ca <= 6:
            same = cb == ca - 3
        else:
            i, j = (ca - 7) // 3, (ca - 7) % 3
            same = cb == 7 + 3 * j + i
        if not same:
            continue
        sc = pose_scene(h1, x1, R1, h2, x2, R2)
        p1, n1, d1, _ = run_oracle(sc, 0, 1, lib_path)
        p2, n2, d2, _ = run_oracle(sc, 1, 0, lib_path)
        assert len(p1) == len(p2), f"code {ca}/{cb}: {len(p1)} vs {len(p2)} contacts after the swap"
        assert same_point_sets(p1, p2, TOL), f"code {ca}/{cb}: points change under the swap"
        order = [int(np.argmin(np.linalg.norm(p2 - q, axis=1))) for q in p1]
        np.testing.assert_allclose(d1, d2[order], rtol=0, atol=TOL)
        np.testing.assert_allclose(n1, -n2[order], rtol=0, atol=TOL)
        seen["face_box1" if ca <= 3 else ("face_box2" if ca <= 6 else "edge")] += 1
    return seen


def resting_on_box2(k, tilt_axis, tilt, depth):
    """Box 1 (small) rests on box 2's face +e_k, tilted by `tilt` rad about
    `tilt_axis`, its lowest corner `depth` below the face."""
    hw = np.array([0.8, 0.8, 0.8])
    hw[k] = 0.1
    hs = np.array([0.2, 0.15, 0.12])
    hs[k] = 0.07
    Rs = rot_axis_angle(tilt_axis, tilt)
    corners = np.array([Rs @ (hs * np.array(c)) for c in np.array(np.meshgrid([-1, 1], [-1, 1], [-1, 1])).T.reshape(-1, 3)])
    xs = np.array([0.03, -0.02, 0.05])
    xs[k] = hw[k] - corners[:, k].min() - depth
    return hs, xs, Rs, hw, np.zeros(3), np.eye(3), corners + xs


def check_resting_on_box2_closed_form(lib_path=None):
    """Box 2's face is the reference face (code 4 + k): the contacts are box
    1's corners below that face, depth = how far below, normal -e_k."""
    for k in range(3):
        other = np.eye(3)[(k + 1) % 3]
        for tilt, n_expect in ((0.004, 4), (0.15, 2)):
            h1, x1, R1, h2, x2, R2, corners = resting_on_box2(k, other, tilt, 0.01)
            pick, amb = choose(x1, R1, h1, x2, R2, h2)
            assert not amb and pick[0] == 4 + k
            pts, nrm, dep, info = run_oracle(pose_scene(h1, x1, R1, h2, x2, R2), 0, 1, lib_path)
            below = corners[corners[:, k] < h2[k]]
            assert len(below) == n_expect
            assert len(pts) == n_expect, f"axis {k} tilt {tilt}: {len(pts)} contacts, expected {n_expect}"
            assert same_point_sets(pts, below, TOL)
            order = [int(np.argmin(np.linalg.norm(below - q, axis=1))) for q in pts]
            np.testing.assert_allclose(dep, h2[k] - below[order, k], rtol=0, atol=TOL)
            np.testing.assert_allclose(nrm, np.tile(-np.eye(3)[k], (n_expect, 1)), rtol=0, atol=1e-15)


# ------------------------------------------------------------ tests
def test_box_box_against_independent_geometry():
    seen = check_axis_points_depths(case_list())
    assert seen["face_box1"] >= 40 and seen["face_box2"] >= 40 and seen["edge"] >= 20 and seen["separated"] >= 10, seen


def test_box_box_swap_symmetry_strict():
    seen = check_swap_symmetry(case_list())
    assert seen["face_box1"] >= 40 and seen["face_box2"] >= 40 and seen["edge"] >= 20, seen


def test_box_box_box2_reference_face_closed_form():
    check_resting_on_box2_closed_form()


ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "am_oracle.c")

# (name, text in am_oracle.c, replacement): each a plausible slip in the
# restatement of BoxBox.java:656-858 and :140-209
PLANTED = [
    ("normal2 not negated for box 2's face",
     "else { normal2[0] = -1.0 * normal[0]; normal2[1] = -1.0 * normal[1]; normal2[2] = -1.0 * normal[2]; }",
     "else { normal2[0] = normal[0]; normal2[1] = normal[1]; normal2[2] = normal[2]; }"),
    ("sizes not swapped for box 2's face",
     "else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }",
     "else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = A; Sb = B; }"),
    ("centres not swapped for box 2's face",
     "else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }",
     "else { Ra = R2; Rb = R1; pa = p1; pb = p2; Sa = B; Sb = A; }"),
    ("rotations not swapped for box 2's face",
     "else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }",
     "else { Ra = R1; Rb = R2; pa = p2; pb = p1; Sa = B; Sb = A; }"),
    ("reference axis off by one for box 2's face",
     "if (t.code <= 3) codeN = t.code - 1; else codeN = t.code - 4;",
     "if (t.code <= 3) codeN = t.code - 1; else codeN = (t.code - 3) % 3;"),
    ("incident face on the wrong side",
     "if (nr[lanr] < 0) {",
     "if (nr[lanr] > 0) {"),
    ("clip interpolation towards the wrong edge",
     "((double)sign * h[dir] - pq[dir]);",
     "(h[dir] - pq[dir]);"),
    ("depth filter keeps separated points",
     "if (dep[cnum] >= 0) {",
     "if (dep[cnum] >= -1.0) {"),
]


def _build_mutant(tmp, name, old, new):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"planted mistake '{name}': its text is not found once in am_oracle.c"
    d = tmp / name.replace(" ", "_").replace("'", "")
    d.mkdir()
    shutil.copy(os.path.join(os.path.dirname(ORACLE_SRC), "am_oracle.h"), d / "am_oracle.h")
    (d / "am_oracle.c").write_text(src.replace(old, new))
    inc = os.path.join(os.path.dirname(os.path.dirname(ORACLE_SRC)), "include")
    lib = d / "liboracle_mut.so"
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-D_DEFAULT_SOURCE", f"-I{inc}",
                    "-shared", "-o", str(lib), str(d / "am_oracle.c"), "-lm"], check=True, capture_output=True)
    return str(lib)


@pytest.mark.parametrize("name,old,new", PLANTED, ids=[p[0] for p in PLANTED])
def test_planted_mistakes_are_caught(tmp_path, name, old, new):
    lib = _build_mutant(tmp_path, name, old, new)
    cases = case_list(80, 120, seed=5)
    failures = []
    for check in (lambda: check_resting_on_box2_closed_form(lib), lambda: check_axis_points_depths(cases, lib),
                  lambda: check_swap_symmetry(cases, lib)):
        try:
            check()
        except AssertionError as e:
            failures.append(str(e)[:80])
    assert failures, f"planted mistake '{name}' passed every box-box check"
