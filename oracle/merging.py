"""CPU restatement of adaptive merging over the C oracle.

TEST INFRASTRUCTURE ONLY: used by the test suite to check libam's merging
phases; never imported by the product package.

The phases follow src/mergingBodies3D/ of the reference:
  RigidBodySystem.advanceTime            RigidBodySystem.java:102-185
  CollisionProcessor.updateInCollections CollisionProcessor.java:232-303
  BodyPairContact (histories, tests)     BodyPairContact.java:83-266
  Merging.merge / unmerge                Merging.java:73-393
  RigidCollection (mass properties)      RigidCollection.java:55-937
  Sleeping.wake / sleep / wakeAll        Sleeping.java:49-160
Detection, warm start, the PGS sweeps and the integration of top-level
entities run in the C oracle; the body-pair bookkeeping, the collection
physics and the graph edits are restated here in plain Python floats (IEEE
double, the reference's operation order).  Where the reference iterates a
HashSet this restatement uses the order libam documents: pairs in contact-list
order, internal pairs in creation order, components depth first from the
first member.  vecmath's Matrix3d.invert is replaced by a cofactor inverse
(parity with the JVM unpinned there).
"""
from __future__ import annotations

import math

import numpy as np

from adaptivemerging_amd._lib import (AM_BODY_KINEMATIC, AM_BODY_PINNED, AM_BODY_SLEEPING, AM_GEOM_COLLECTION,
                                      AM_GEOM_PLANE,
                                      AM_SPRING_BODYBODY, AM_STATE_BROKEN, AM_STATE_CLEAR, AM_STATE_ONEDGE,
                                      BODY_DTYPE)

from .pyoracle import Oracle

MAXD = 1.7976931348623157e308


# ---- javax.vecmath operations ---------------------------------------------
def mulv(A, v):
    return [A[0] * v[0] + A[1] * v[1] + A[2] * v[2], A[3] * v[0] + A[4] * v[1] + A[5] * v[2],
            A[6] * v[0] + A[7] * v[1] + A[8] * v[2]]


def mmul(A, B):
    return [A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j] for i in range(3) for j in range(3)]


def mmul_bt(A, B):  # A B^T
    return [A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2]
            for i in range(3) for j in range(3)]


def mmul_at(A, B):  # A^T B
    return [A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j] for i in range(3) for j in range(3)]


def transpose(A):
    return [A[j * 3 + i] for i in range(3) for j in range(3)]


def rm0rt(R, M):
    return mmul_bt(mmul(R, M), R)


def rtmr(R, M):
    return mmul(mmul_at(R, M), R)


def invert3(a):
    c00 = a[4] * a[8] - a[5] * a[7]
    c01 = a[5] * a[6] - a[3] * a[8]
    c02 = a[3] * a[7] - a[4] * a[6]
    det = a[0] * c00 + a[1] * c01 + a[2] * c02
    s = 1.0 / det
    return [c00 * s, (a[2] * a[7] - a[1] * a[8]) * s, (a[1] * a[5] - a[2] * a[4]) * s,
            c01 * s, (a[0] * a[8] - a[2] * a[6]) * s, (a[2] * a[3] - a[0] * a[5]) * s,
            c02 * s, (a[1] * a[6] - a[0] * a[7]) * s, (a[0] * a[4] - a[1] * a[3]) * s]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def jmax(a, b):
    if a != a:
        return a
    if a == 0.0 and b == 0.0:
        return b if math.copysign(1.0, a) < 0 else a
    return a if a >= b else b


def jmin(a, b):
    if a != a:
        return a
    if a == 0.0 and b == 0.0:
        return a if math.copysign(1.0, a) < 0 else b
    return a if a <= b else b


IDENT = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
ZERO = [0.0] * 9


def get_op(bx, mass, com):
    """RigidCollection.getOp (:808-830)."""
    x, y, z = bx[0] - com[0], bx[1] - com[1], bx[2] - com[2]
    x2, y2, z2 = x * x, y * y, z * z
    op = [y2 + z2, -x * y, -x * z, -y * x, x2 + z2, -y * z, -z * x, -z * y, x2 + y2]
    return [e * mass for e in op]


class Ent:
    """One entity's state and mass properties (RigidBody fields)."""

    def __init__(self, **kw):
        self.sleeping = False
        self.__dict__.update(kw)

    def copy(self):
        return Ent(**{k: (list(v) if isinstance(v, list) else v) for k, v in self.__dict__.items()})


def bbox_corners(mn, mx):
    return [mx[0], mx[1], mx[2], mx[0], mn[1], mn[2], mn[0], mx[1], mn[2], mn[0], mn[1], mx[2],
            mn[0], mn[1], mn[2], mn[0], mx[1], mx[2], mx[0], mn[1], mx[2], mx[0], mx[1], mn[2]]


# ---- BodyPairContact tests (pure functions of the histories) --------------
def pair_state(not_on_edge: int, n_active: int) -> int:
    """The state accumulateForMerging appends (BodyPairContact.java:100-116)."""
    if not_on_edge == 0:
        return AM_STATE_ONEDGE
    if not_on_edge >= 2:
        return AM_STATE_CLEAR
    if n_active == 1:
        return AM_STATE_CLEAR
    return AM_STATE_ONEDGE


def merge_condition(P, hist, states, breath: bool, pinned1: bool, pinned2: bool, kinematic: bool = False) -> bool:
    """checkMergeCondition (:151-177) without the same-collection test:
    breathing, pinned bodies, checkMotionMetricForMerging (:187-198) and
    areContactsStable (:212-224).  A pair with a body on a prescribed motion
    (no reference counterpart) never merges."""
    accum = P.step_accum_merging
    if P.enable_merge_let_it_breathe and breath:
        return False
    if not P.enable_merge_pinned and (pinned1 or pinned2):
        return False
    if kinematic:
        return False
    if len(hist) != accum or any(m > P.threshold_merge for m in hist):
        return False
    if P.enable_merge_stable_contact and (len(states) != accum or AM_STATE_ONEDGE in states):
        return False
    return True


def organized_layers(body1, body2, is_new, n_ext: int):
    """CollisionProcessor.getOrganizedContacts / getNextLayer (:346-441) on a
    combined contact list (external contacts [0, n_ext), then internal ones):
    the breadth-first layer of each contact's body pair.  Layer 0 holds the
    external pairs with a contact new this step; layer k+1 the pairs not yet
    reached that share a body with a pair of layer k (getNextLayer walks
    body.bodyPairContacts, pinned bodies included); then come the external
    pairs never reached, then the internal ones.  The reference visits the
    pairs of one layer in HashSet order; only the layer is reproducible.
    Returns the layer of every contact."""
    n = len(body1)
    pair = [(min(int(a), int(b)), max(int(a), int(b))) for a, b in zip(body1, body2)]
    pairs, ext_pair, new_pair = [], {}, {}
    for k in range(n):
        p = pair[k]
        if p not in ext_pair:
            pairs.append(p)
            ext_pair[p] = k < n_ext
            new_pair[p] = False
        if k < n_ext and is_new[k]:
            new_pair[p] = True
    by_body = {}
    for p in pairs:
        for b in set(p):
            by_body.setdefault(b, []).append(p)
    layer = {p: 0 for p in pairs if ext_pair[p] and new_pair[p]}
    frontier = list(layer)
    lmax = 0 if layer else -1
    k = 0
    while frontier:
        nxt = []
        for p in frontier:
            for b in p:
                for q in by_body[b]:
                    if q not in layer:
                        layer[q] = k + 1
                        nxt.append(q)
        if nxt:
            lmax = k + 1
        frontier = nxt
        k += 1
    for p in pairs:
        if p not in layer:
            layer[p] = lmax + 1 if ext_pair[p] else lmax + 2
    return [layer[pair[k]] for k in range(n)]


def unmerge_condition(P, contact_states, hist_n: int, n_contacts: int) -> bool:
    """checkContactsState (:255-266) or checkMotionMetricForUnmerging (:200-205)."""
    for s in contact_states:
        if s == AM_STATE_BROKEN and P.enable_unmerge_normal:
            return True
        if s == AM_STATE_ONEDGE and P.enable_unmerge_friction:
            return True
    return hist_n >= P.step_accum_unmerging and n_contacts < 3


# ---- RigidCollection -------------------------------------------------------
def update_bb_with(C, B):
    """RigidCollection.updateBB(body) (:612-650)."""
    if B.plane:
        return
    mn = [MAXD] * 3
    mx = [-MAXD] * 3
    for b in (C, B):
        for q in range(8):
            p = mulv(b.R, b.bb[q * 3:q * 3 + 3])
            p = [p[0] + b.x[0], p[1] + b.x[1], p[2] + b.x[2]]
            for a in range(3):
                mn[a] = jmin(mn[a], p[a])
                mx[a] = jmax(mx[a], p[a])
    C.bb = bbox_corners(mn, mx)


def add_body_internal(C, B):
    """RigidCollection.addBodyInternalMethod (:477-514)."""
    C.R = list(IDENT)
    if C.pinned:
        C.v = [0.0] * 3
        C.w = [0.0] * 3
        C.I, C.I0, C.jinv, C.jinv0 = list(ZERO), list(ZERO), list(ZERO), list(ZERO)
        C.mass, C.minv = 0.0, 0.0
        return
    com = [C.x[a] * C.mass for a in range(3)]
    com = [B.mass * B.x[a] + com[a] for a in range(3)]
    inv = 1. / (B.mass + C.mass)
    com = [com[a] * inv for a in range(3)]
    v = [C.v[a] * C.mass for a in range(3)]
    v = [B.mass * B.v[a] + v[a] for a in range(3)]
    C.v = [v[a] * inv for a in range(3)]
    w = [C.w[a] * C.mass for a in range(3)]
    w = [B.mass * B.w[a] + w[a] for a in range(3)]
    C.w = [w[a] * inv for a in range(3)]
    M0 = [0.0] * 9
    for term in (C.I, get_op(C.x, C.mass, com), B.I, get_op(B.x, B.mass, com)):
        M0 = [M0[i] + term[i] for i in range(9)]
    C.I = M0
    update_bb_with(C, B)
    C.mass = C.mass + B.mass
    C.minv = inv
    C.x = com
    r = C.R
    bb = []
    for q in range(8):
        px, py, pz = C.bb[q * 3] - C.x[0], C.bb[q * 3 + 1] - C.x[1], C.bb[q * 3 + 2] - C.x[2]
        bb += [r[0] * px + r[3] * py + r[6] * pz, r[1] * px + r[4] * py + r[7] * pz,
               r[2] * px + r[5] * py + r[8] * pz]
    C.bb = bb


def finish_inertia(C):
    """updateInertiaRestAndInvert (:674-680) + updateRotationalInertiaFromTransformation."""
    if C.pinned:
        return
    C.jinv = invert3(C.I)
    C.I0 = rtmr(C.R, C.I)
    C.jinv0 = rtmr(C.R, C.jinv)
    C.I = rm0rt(C.R, C.I0)
    C.jinv = rm0rt(C.R, C.jinv0)


def body_to_collection(C, xb, Rb):
    """RigidTransform3D.multAinvB (:201-215)."""
    t = [xb[0] - C.x[0], xb[1] - C.x[1], xb[2] - C.x[2]]
    Rt = transpose(C.R)
    return mmul(Rt, Rb), mulv(Rt, t)


class MergingOracle:
    """The merging phases of advanceTime on the CPU (see the module docstring)."""

    def __init__(self, scene, params):
        self.o = Oracle(scene, params)
        self.p = params
        self.n_phys = scene.n
        self.parent = [-1] * scene.n
        self.top = list(range(scene.n))
        self.members = {}
        self.b2c = {}
        self.free = []
        self.total_steps = 0
        self.merging_event = False
        self.bpc = []          # external pairs of the last clear_body_pair_contacts
        self.ibpc = []         # internal pairs: dict(b1, b2, contacts (row dicts), hist_n, flag)
        self.n_entities = scene.n
        self.sleep_hist = {}   # entity -> RigidBody.metricHistory
        sp = scene.spring_array() if hasattr(scene, "spring_array") else []
        self.springs = [(int(r["type"]), int(r["body1"]), int(r["body2"])) for r in sp]
        kin = scene.kinematic_array() if hasattr(scene, "kinematic_array") else []
        self.kinematics = [(int(r["body"]), [float(a) for a in r["x0"]], [float(a) for a in r["dir"]],
                            float(r["amplitude"]), float(r["omega"]), float(r["phase"])) for r in kin]
        self.impulses = []     # forces held for the next external-force pass
        self.sim_time = 0.0    # RigidBodySystem.simulationTime

    # -- helpers --------------------------------------------------------
    def set_params(self, **kw):
        self.o.set_params(**kw)

    @property
    def params(self):
        return self.o.params

    def desc(self, e):
        return self.o.entity(e)

    def pinned(self, e):
        return bool(int(self.desc(e)["flags"]) & AM_BODY_PINNED)

    def sleeping(self, e):
        return bool(int(self.desc(e)["flags"]) & AM_BODY_SLEEPING)

    def kinematic(self, e):
        return bool(int(self.desc(e)["flags"]) & AM_BODY_KINEMATIC)

    # -- prescribed motion, held forces, viscous decay -----------------------
    def kinematic_update(self, t):
        """Each kinematic body at time t: x = x0 + (a sin(w t + phase)) dir,
        v = (a w cos(w t + phase)) dir, omega = 0 (fdlibm sine and cosine)."""
        if not self.kinematics:
            return
        x, R, v, w = self.o.pull_state(self.n_entities)
        x, v, w = np.array(x), np.array(v), np.array(w)
        for b, x0, d, amp, om, ph in self.kinematics:
            arg = om * t + ph
            ds = amp * self.o.lib.orc_sin(arg)
            vs = amp * om * self.o.lib.orc_cos(arg)
            for a in range(3):
                x[b][a] = x0[a] + ds * d[a]
                v[b][a] = vs * d[a]
                w[b][a] = 0.0
        self._set_state_keep_jinv(x, v, w)

    def _set_state_keep_jinv(self, x, v, w):
        """Writes x, v, omega and leaves every entity's world inverse inertia
        as it was (the oracle's set_state recomputes it from R)."""
        jinv = self.o.pull_jinv()
        self.o.set_state(x, None, v, w)
        for i in range(self.n_entities):
            self.o.set_jinv(i, jinv[i])

    def impulse_add(self, body, point_w, force):
        self.impulses.append((body, [float(a) for a in point_w], [float(a) for a in force]))

    def apply_external_forces(self):
        """RigidBodySystem.applyExternalForces (:234-257): gravity, springs,
        then the held forces once (applyImpulse, :262-267) through
        RigidBody.applyForceW (RigidBody.java:335-340) on the body's
        top-level entity."""
        self.o.apply_gravity()
        self.o.apply_springs()
        if not self.impulses:
            return
        x, _, _, _ = self.o.pull_state(self.n_entities)
        f, t = self.o.pull_forces()
        for b, p, F in self.impulses:
            e = self.top_of(b)
            r = [p[a] - float(x[e][a]) for a in range(3)]
            for a in range(3):
                f[e][a] = f[e][a] + F[a]
            tq = cross(r, F)
            for a in range(3):
                t[e][a] = t[e][a] + tq[a]
        self.o.push_forces(f, t)
        self.impulses = []

    def viscous_decay(self):
        """RigidBodySystem.applyViscousDecay (:428-435) over the top-level list."""
        a1, a2 = self.params.viscous_linear_decay, self.params.viscous_angular_decay
        if a1 == 1.0 and a2 == 1.0:
            return
        x, R, v, w = self.o.pull_state(self.n_entities)
        v, w = np.array(v), np.array(w)
        for e in self.top:
            for a in range(3):
                v[e][a] = a1 * v[e][a]
                w[e][a] = a2 * w[e][a]
        self._set_state_keep_jinv(None, v, w)

    def set_sleeping(self, e, value):
        rec = self.desc(e).copy()
        f = int(rec["flags"])
        rec["flags"] = (f | AM_BODY_SLEEPING) if value else (f & ~AM_BODY_SLEEPING)
        self.o.set_entity(e, rec)

    def top_of(self, b):
        return self.parent[b] if self.parent[b] >= 0 else b

    def state(self):
        x, R, v, w = self.o.pull_state(self.n_entities)
        return [np.array(a) for a in (x, R, v, w)]

    def load(self, e, st, jinv):
        d = self.desc(e)
        x, R, v, w = st
        pinned = bool(int(d["flags"]) & AM_BODY_PINNED)
        Rl = [float(a) for a in R[e]]
        I0 = [float(a) for a in d["inertia0"]]
        return Ent(x=[float(a) for a in x[e]], v=[float(a) for a in v[e]], w=[float(a) for a in w[e]], R=Rl,
                   jinv=[float(a) for a in jinv[e]], jinv0=[float(a) for a in d["jinv0"]], I0=I0,
                   I=list(ZERO) if pinned else rm0rt(Rl, I0), mass=0.0 if pinned else float(d["mass"]),
                   minv=0.0 if pinned else float(d["minv"]), pinned=pinned,
                   sleeping=bool(int(d["flags"]) & AM_BODY_SLEEPING),
                   plane=int(d["geom"]) == AM_GEOM_PLANE, bb=[float(a) for a in d["bbox"]])

    def store(self, e, C, st, jinv):
        x, R, v, w = st
        x[e], R[e], v[e], w[e] = C.x, C.R, C.v, C.w
        jinv[e] = C.jinv
        rec = np.zeros((), dtype=BODY_DTYPE)
        rec["geom"] = AM_GEOM_COLLECTION
        rec["flags"] = (AM_BODY_PINNED if C.pinned else 0) | (AM_BODY_SLEEPING if C.sleeping else 0)
        rec["mass"] = C.mass
        rec["minv"] = C.minv
        rec["jinv0"] = C.jinv0
        rec["inertia0"] = C.I0
        rec["bbox"] = C.bb
        rec["friction"] = 0.8
        # members share one scene (bodies of different scenes never touch)
        mem = self.members.get(e, [])
        rec["scene"] = int(self.desc(mem[0])["scene"]) if mem else 0
        self.o.set_entity(e, rec)

    def push_structure(self, st, jinv):
        n = self.n_entities
        mf = np.zeros(n, np.int32)
        mc = np.zeros(n, np.int32)
        members = []
        for e in self.top:
            if e in self.members:
                mf[e] = len(members)
                mc[e] = len(self.members[e])
                members += self.members[e]
        self.o.set_structure(np.array(self.parent, np.int32), np.array(self.top, np.int32), mf, mc,
                             np.array(members, np.int32))
        x, R, v, w = st
        self.o.set_state(x, R, v, w)
        for e in range(n):
            self.o.set_jinv(e, np.array(jinv[e], dtype=np.float64))

    def alloc(self, st, jinv):
        if not self.free:
            old = self.n_entities
            n = old + max(16, self.n_phys // 4)
            self.o.resize(n)
            self.parent += [-1] * (n - old)
            self.n_entities = n
            for i in range(4):
                pad = np.zeros((n - old, st[i].shape[1]))
                if i == 1:
                    pad[:, [0, 4, 8]] = 1.0
                st[i] = np.concatenate([st[i], pad])
            jinv.extend([list(ZERO) for _ in range(n - old)])
            self.free = list(range(n - 1, old - 1, -1))
        e = self.free.pop()
        self.members[e] = []
        self.sleep_hist[e] = []   # a new RigidCollection has an empty metric history
        return e

    def free_entity(self, e):
        self.members.pop(e, None)
        rec = np.zeros((), dtype=BODY_DTYPE)
        rec["geom"] = AM_GEOM_COLLECTION
        rec["flags"] = AM_BODY_PINNED
        self.o.set_entity(e, rec)
        self.parent[e] = -1
        self.free.append(e)

    def update_b2c(self, ce, C, st):
        x, R = st[0], st[1]
        for b in self.members[ce]:
            self.b2c[b] = body_to_collection(C, [float(a) for a in x[b]], [float(a) for a in R[b]])

    def collection_state(self, C, b, B):
        """RigidCollection.updateCollectionState (:459-470)."""
        C.pinned = C.pinned or B.pinned
        C.sleeping = C.sleeping or B.sleeping
        if B.sleeping:
            self.set_sleeping(b, False)

    def add_body(self, ce, C, b, st, jinv):
        self.parent[b] = ce
        self.members[ce].append(b)
        B = self.load(b, st, jinv)
        self.collection_state(C, b, B)
        add_body_internal(C, B)
        finish_inertia(C)
        self.update_b2c(ce, C, st)

    def add_bodies(self, ce, C, bodies, st, jinv):
        for b in bodies:
            self.parent[b] = ce
            self.members[ce].append(b)
            B = self.load(b, st, jinv)
            self.collection_state(C, b, B)
            add_body_internal(C, B)
        finish_inertia(C)
        self.update_b2c(ce, C, st)

    def new_collection(self, b1, b2, st, jinv):
        ce = self.alloc(st, jinv)
        if int(self.desc(b1)["geom"]) == AM_GEOM_PLANE:
            b1, b2 = b2, b1
        C = self.load(b1, st, jinv)
        C.plane = False
        self.parent[b1] = ce
        self.members[ce].append(b1)
        self.add_body(ce, C, b2, st, jinv)
        return ce, C

    def same(self, a, b):
        return self.parent[a] >= 0 and self.parent[a] == self.parent[b]

    # -- per-step collection physics -----------------------------------
    def push_down(self, integrated_only):
        """RigidCollection.applyVelocitiesTo for members (:925-937)."""
        if not self.members:
            return
        x, R, v, w = self.state()
        for e in self.top:
            if e not in self.members:
                continue
            if integrated_only and (self.pinned(e) or self.sleeping(e)):
                continue
            for b in self.members[e]:
                r = [x[b][a] - x[e][a] for a in range(3)]
                wxr = cross([float(a) for a in w[e]], r)
                v[b] = [v[e][a] + wxr[a] for a in range(3)]
                w[b] = w[e]
        jinv = self.o.pull_jinv()
        self.o.set_state(x, R, v, w)
        for i in range(self.n_entities):
            self.o.set_jinv(i, jinv[i])

    def member_transforms(self):
        """RigidCollection.updateBodiesPositionAndTransformations (:898-909)."""
        if not self.members:
            return
        x, R, v, w = self.state()
        jinv = self.o.pull_jinv()
        for e in self.top:
            if e not in self.members or self.pinned(e) or self.sleeping(e):
                continue
            Rc = [float(a) for a in R[e]]
            for b in self.members[e]:
                Rbc, tbc = self.b2c[b]
                t = mulv(Rc, tbc)
                x[b] = [t[a] + x[e][a] for a in range(3)]
                R[b] = mmul(Rc, Rbc)
                jinv[b] = rm0rt([float(a) for a in R[b]], [float(a) for a in self.desc(b)["jinv0"]])
        self.o.set_state(x, R, v, w)
        for i in range(self.n_entities):
            self.o.set_jinv(i, jinv[i])

    def clear_bodies(self):
        self.o.clear_bodies()
        self.push_down(False)

    def advance_velocities(self, dt):
        self.o.advance_velocities(dt)
        self.push_down(True)

    def advance_positions(self, dt):
        self.o.advance_positions(dt)
        self.member_transforms()

    def clear_top_level(self):
        f, t = self.o.pull_forces()
        dv = self.o.pull_deltav()
        for e in self.top:
            f[e] = 0.0
            t[e] = 0.0
            dv[e] = 0.0
        self.o.push_forces(f, t)
        self.o.push_deltav(dv)
        self.apply_external_forces()

    def internal_contacts(self):
        rows = [c for p in self.ibpc for c in p["contacts"]]
        return _stack(rows)

    # -- Sleeping (Sleeping.java) -------------------------------------------
    def _wake_body(self, b, woke):
        """RigidBody.wake (RigidBody.java:298-306): the body, then its collection."""
        for e in (b, self.parent[b]):
            if e >= 0 and self.sleeping(e):
                self.set_sleeping(e, False)
                self.sleep_hist[e] = []
                woke.add(e)

    def wake(self):
        """Sleeping.wake (:107-139), right after the warm start: in the order of
        the top-level list, each body still asleep wakes both ends of its pairs
        of this step between bodies that are not pinned; then each body-body
        spring with exactly one end asleep and neither end pinned wakes both."""
        if not self.params.enable_sleeping:
            return 0
        woke = set()
        c = self.o.contacts()
        pairs = {}
        for k in range(len(c["body1"])):
            a, b = int(c["body1"][k]), int(c["body2"][k])
            if self.pinned(a) or self.pinned(b):
                continue
            for e in {self.top_of(a), self.top_of(b)}:
                pairs.setdefault(e, []).append((a, b))
        for e in list(self.top):
            if not self.sleeping(e):
                continue
            for a, b in pairs.get(e, []):
                self._wake_body(a, woke)
                self._wake_body(b, woke)
        for typ, b1, b2 in self.springs:
            if typ != AM_SPRING_BODYBODY:
                continue
            t1, t2 = self.top_of(b1), self.top_of(b2)
            if self.sleeping(t1) != self.sleeping(t2) and not (self.pinned(t1) or self.pinned(t2)):
                self._wake_body(b1, woke)
                self._wake_body(b2, woke)
        return len(woke)

    def wake_all(self):
        """Sleeping.wakeAll (:144-150)."""
        woke = set()
        for e in list(self.top):
            self._wake_body(e, woke)

    def sleep(self):
        """Sleeping.sleep (:49-98) at the end of the step; returns the number of
        top-level bodies asleep afterwards."""
        P = self.params
        if P.enable_sleeping:
            c = self.o.contacts()
            ext = set()
            for k in range(len(c["body1"])):
                if not abs(c["lambda"][k][0]) > 1e-14:
                    continue
                a, b = int(c["body1"][k]), int(c["body2"][k])
                if self.pinned(a) or self.pinned(b) or self.same(a, b):
                    continue
                ext.add(self.top_of(a))
                ext.add(self.top_of(b))
            for e in list(self.top):
                if self.sleeping(e) or e in ext or self.kinematic(e):
                    continue
                h = self.sleep_hist.setdefault(e, [])
                h.append(self.o.motion_metric(e))
                if len(h) > P.sleep_step_accum:
                    h.pop(0)
                sleep = len(h) >= P.sleep_step_accum
                prev = MAXD
                for m in h:
                    if not sleep:
                        break
                    if m > prev + 5e-5 or m > P.sleep_threshold:
                        sleep = False
                    prev = m
                if sleep:
                    self.set_sleeping(e, True)
        return sum(1 for e in self.top if self.sleeping(e))

    # -- CollisionProcessor.updateInCollections (:232-303) -----------------
    def update_in_collections(self, dt, order_of_sweep):
        """order_of_sweep(): the sweep order of the combined list (external
        then internal contacts) to reproduce."""
        if not self.params.update_contacts_in_collections or not self.members:
            return 0
        ext = self.o.contacts()
        n_ext = len(ext["body1"])
        # internal contacts of the collections that are awake (:265-270, :400-408)
        awake = [p for p in self.ibpc if not self.sleeping(self.parent[p["b1"]])]
        comb = _concat(ext, _stack([c for p in awake for c in p["contacts"]]))
        # the single sweep's list and its breadth-first layers (getOrganizedContacts), kept for checks
        self.last_sweep = dict(body1=np.array(comb["body1"]), body2=np.array(comb["body2"]), n_ext=n_ext,
                               layers=organized_layers(comb["body1"], comb["body2"], comb["new_this_step"], n_ext)
                               if self.params.organize_contacts else None)
        self.o.push_contacts_raw(comb)
        used, _, _ = self.o.solve(order_of_sweep(), dt, main=False)
        res = self.o.contacts()
        k = n_ext
        for p in awake:
            for j in range(len(p["contacts"])):
                p["contacts"][j] = _row(res, k)
                k += 1
        self.o.push_contacts_raw(_slice(res, 0, n_ext))
        # members advance their velocities; top-level deltaV cleared
        x, R, v, w = self.state()
        f, t = self.o.pull_forces()
        dv = self.o.pull_deltav()
        jinv = self.o.pull_jinv()
        for e in self.top:
            if e in self.members and self.sleeping(e):
                continue
            for b in self.members.get(e, []):
                if self.pinned(b):
                    continue
                s = dt * float(self.desc(b)["minv"])
                vb = [s * f[b][a] + v[b][a] for a in range(3)]
                v[b] = [vb[a] + dv[b][a] for a in range(3)]
                dom = mulv([float(a) for a in jinv[b]], [float(a) for a in t[b]])
                dom = [dom[a] * dt for a in range(3)]
                wb = [w[b][a] + dom[a] for a in range(3)]
                w[b] = [wb[a] + dv[b][3 + a] for a in range(3)]
        for e in self.top:
            dv[e] = 0.0
        self.o.set_state(x, R, v, w)
        for i in range(self.n_entities):
            self.o.set_jinv(i, jinv[i])
        self.o.push_deltav(dv)
        return used

    def _metric(self, a, b, dt):
        """BodyPairContact.java:88-95 / :130-137."""
        if self.pinned(a):
            return self.o.motion_metric(b)
        if self.pinned(b):
            return self.o.motion_metric(a)
        if self.params.metric_position_level:
            return self.o.motion_metric_pos(a, b, dt)
        return self.o.motion_metric(a, b)

    # -- BodyPairContact histories -----------------------------------------
    def accumulate_for_unmerging(self, dt):
        """RigidBodySystem.accumulateForUnmerging + the unmerge test."""
        P = self.params
        for p in self.ibpc:
            if self.sleeping(self.parent[p["b1"]]):
                continue   # RigidBodySystem.accumulateForUnmerging skips sleeping collections (:326-329)
            m = self._metric(p["b1"], p["b2"], dt)
            p["hist_n"] = p["hist_n"] + 1 if m > P.threshold_unmerge else 0
            p["flag"] = unmerge_condition(P, [c["state"] for c in p["contacts"]], p["hist_n"], len(p["contacts"]))

    def clear_body_pair_contacts(self):
        """clearBodyPairContacts + updateBodyPairContacts bookkeeping."""
        prev = {(min(p["b1"], p["b2"]), max(p["b1"], p["b2"])): p for p in self.bpc if p["b1"] >= 0}
        c = self.o.contacts()
        n = len(c["body1"])
        out = []
        k = 0
        thr = self.params.threshold_breath
        while k < n:
            a, b = int(c["body1"][k]), int(c["body2"][k])
            e = k + 1
            while e < n and {int(c["body1"][e]), int(c["body2"][e])} == {a, b}:
                e += 1
            act = [j for j in range(k, e) if abs(c["lambda"][j][0]) > 1e-14]
            if act:
                q = prev.get((min(a, b), max(a, b)))
                out.append(dict(
                    b1=a, b2=b, first=k, end=e, active=act,
                    not_on_edge=sum(1 for j in act if c["state"][j] != AM_STATE_ONEDGE),
                    breath=any(abs(c["prev_violation"][j] - c["violation"][j]) > thr for j in act),
                    hist=list(q["hist"]) if q else [], states=list(q["states"]) if q else [], flag=False))
            k = e
        self.bpc = out
        self._contacts = c

    def accumulate_for_merging(self, dt):
        """RigidBodySystem.accumulateForMerging + static part of checkMergeCondition."""
        P = self.params
        accum = P.step_accum_merging
        for r in self.bpc:
            p1 = self.parent[r["b1"]] if self.parent[r["b1"]] >= 0 else r["b1"]
            p2 = self.parent[r["b2"]] if self.parent[r["b2"]] >= 0 else r["b2"]
            r["hist"].append(self._metric(p1, p2, dt))
            if len(r["hist"]) > accum:
                r["hist"].pop(0)
            r["states"].append(pair_state(r["not_on_edge"], len(r["active"])))
            if len(r["states"]) > accum:
                r["states"].pop(0)
            r["flag"] = merge_condition(P, r["hist"], r["states"], r["breath"], self.pinned(r["b1"]),
                                        self.pinned(r["b2"]), self.kinematic(r["b1"]) or self.kinematic(r["b2"]))

    # -- Merging.merge (:73-163) ---------------------------------------------
    def merge(self):
        def asleep(r):
            # checkMergeCondition (BodyPairContact.java:151-154): both bodies asleep
            return r["b1"] >= 0 and self.sleeping(r["b1"]) and self.sleeping(r["b2"])

        if not self.params.enable_merging or not self.bpc or not any(r["flag"] or asleep(r) for r in self.bpc):
            return 0
        st = self.state()
        jinv = [list(map(float, j)) for j in self.o.pull_jinv()]
        c = self._contacts
        adj = {}
        for i, r in enumerate(self.bpc):
            adj.setdefault(r["b1"], []).append(i)
            if r["b2"] != r["b1"]:
                adj.setdefault(r["b2"], []).append(i)
        internal = [False] * len(self.bpc)
        newly = []

        def internalize(i):
            internal[i] = True
            newly.append(i)

        def add_incomplete(body):
            for i in adj.get(body, []):
                if not internal[i] and self.same(self.bpc[i]["b1"], self.bpc[i]["b2"]):
                    internalize(i)

        colls = {}

        def coll(e):
            if e not in colls:
                colls[e] = self.load(e, st, jinv)
            return colls[e]

        merged = 0
        for i, r in enumerate(self.bpc):
            if internal[i] or not (r["flag"] or asleep(r)) or self.same(r["b1"], r["b2"]):
                continue
            self.merging_event = True
            merged += 1
            internalize(i)
            b1, b2 = r["b1"], r["b2"]
            p1, p2 = self.parent[b1], self.parent[b2]
            if p1 < 0 and p2 < 0:
                self.top.remove(b1)
                self.top.remove(b2)
                ce, C = self.new_collection(b1, b2, st, jinv)
                colls[ce] = C
                self.store(ce, C, st, jinv)
                self.top.append(ce)
            elif p1 >= 0 and p2 >= 0:
                keep, gone = (p1, p2) if len(self.members[p1]) > len(self.members[p2]) else (p2, p1)
                self.top.remove(gone)
                gone_bodies = list(self.members[gone])
                K = coll(keep)
                O = coll(gone)
                for b in gone_bodies:
                    self.parent[b] = keep
                    self.members[keep].append(b)
                K.pinned = K.pinned or O.pinned
                K.sleeping = K.sleeping or O.sleeping
                add_body_internal(K, O)
                finish_inertia(K)
                self.update_b2c(keep, K, st)
                self.store(keep, K, st, jinv)
                for b in gone_bodies:
                    add_incomplete(b)
                self.free_entity(gone)
                colls.pop(gone, None)
            else:
                host, body = (p1, b2) if p1 >= 0 else (p2, b1)
                self.top.remove(body)
                K = coll(host)
                self.add_body(host, K, body, st, jinv)
                self.store(host, K, st, jinv)
                add_incomplete(body)
        for i in newly:
            r = self.bpc[i]
            rows = []
            for j in r["active"]:
                row = _row(c, j)
                row["lambda_warm"] = np.zeros(3)
                row["b"] = np.zeros(3)
                row["diag"] = np.zeros(3)
                row["new_this_step"] = 0
                rows.append(row)
            self.ibpc.append(dict(b1=r["b1"], b2=r["b2"], contacts=rows, hist_n=0, flag=False))
            r["b1"] = -1
        self.push_structure(st, jinv)
        return merged

    # -- Merging.unmergeAll (:168-206) ------------------------------------------
    def unmerge_all(self):
        """Every awake collection returns its bodies (with its velocity at their
        position) to the end of the top-level list; its pairs restart their
        motion history and its internal contacts are dropped."""
        colls = [e for e in self.top if e in self.members and not self.sleeping(e)]
        if not colls:
            return 0
        st = self.state()
        jinv = [list(map(float, j)) for j in self.o.pull_jinv()]
        dv = self.o.pull_deltav()
        gone = set(colls)
        for r in self.bpc:
            if r["b1"] >= 0 and (self.parent[r["b1"]] in gone or self.parent[r["b2"]] in gone):
                r["hist"] = []
        self.ibpc = [p for p in self.ibpc if self.parent[p["b1"]] not in gone]
        x, R, v, w = st
        additions = []
        for ce in colls:
            C = self.load(ce, st, jinv)
            for b in self.members[ce]:
                r_ = [x[b][a] - C.x[a] for a in range(3)]
                wxr = cross(C.w, r_)
                v[b] = [C.v[a] + wxr[a] for a in range(3)]
                w[b] = C.w
                dv[b] = 0.0
                self.parent[b] = -1
                additions.append(b)
        for ce in colls:
            self.top.remove(ce)
            self.free_entity(ce)
        self.top += additions
        self.push_structure(st, jinv)
        self.o.push_deltav(dv)
        return len(colls)

    # -- Merging.unmerge (:215-273) and unmergeSelectedBpcs (:286-374) --------
    def unmerge(self, dt):
        P = self.params
        if not P.enable_unmerging or not self.ibpc:
            return 0
        if not (P.enable_unmerge_relative_motion or P.enable_unmerge_normal or P.enable_unmerge_friction):
            return 0
        if not any(p["flag"] for p in self.ibpc):
            return 0
        st = self.state()
        jinv = [list(map(float, j)) for j in self.o.pull_jinv()]
        dv = self.o.pull_deltav()
        incoll = [True] * len(self.ibpc)
        adj = {}
        for i, p in enumerate(self.ibpc):
            adj.setdefault(p["b1"], []).append(i)
            adj.setdefault(p["b2"], []).append(i)
        reexpose, additions, removals = [], [], []
        zero_dv = []
        events = 0
        for ce in list(self.top):
            if ce not in self.members or self.sleeping(ce):
                continue
            cut = [i for i, p in enumerate(self.ibpc) if p["flag"] and self.parent[p["b1"]] == ce]
            if not cut:
                continue
            for i in cut:
                incoll[i] = False
            C = self.load(ce, st, jinv)
            bodies = list(self.members[ce])
            handled, remained = set(), set()
            new_bodies = []
            remove_coll = True
            for body in bodies:
                if body in handled:
                    continue
                sub = [body]
                seen = {body}

                def dfs(xb):
                    for i in adj.get(xb, []):
                        if not incoll[i]:
                            continue
                        p = self.ibpc[i]
                        o = p["b2"] if p["b1"] == xb else p["b1"]
                        if o not in seen and o not in handled:
                            seen.add(o)
                            sub.append(o)
                            dfs(o)

                dfs(body)
                handled.update(sub)
                if len(bodies) != len(sub):
                    if len(sub) < len(bodies) // 2 + 1:
                        x, R, v, w = st
                        for b in sub:
                            r_ = [x[b][a] - C.x[a] for a in range(3)]
                            wxr = cross(C.w, r_)
                            v[b] = [C.v[a] + wxr[a] for a in range(3)]
                            w[b] = C.w
                            zero_dv.append(b)
                            self.parent[b] = -1
                        if len(sub) > 1:
                            ne, N = self.new_collection(sub[0], sub[1], st, jinv)
                            if len(sub) > 2:
                                self.add_bodies(ne, N, sub[2:], st, jinv)
                            self.store(ne, N, st, jinv)
                            new_bodies.append(ne)
                        else:
                            new_bodies.append(sub[0])
                    else:
                        remove_coll = False
                        remained.update(sub)
                else:
                    remove_coll = False
                    remained.update(sub)
                    break
            for i in cut:
                if self.same(self.ibpc[i]["b1"], self.ibpc[i]["b2"]):
                    incoll[i] = True
                else:
                    reexpose.append(i)
            if not remove_coll:
                gone = [b for b in bodies if b in handled and b not in remained]
                if gone:
                    self.remove_bodies(ce, C, gone, st, jinv)
                    self.store(ce, C, st, jinv)
            if new_bodies:
                self.merging_event = True
                events += 1
                additions += new_bodies
                if remove_coll:
                    removals.append(ce)
        self.top += additions
        for e in removals:
            self.top.remove(e)
            self.free_entity(e)
        # re-exposed contacts go back to the current list with lambda_warm = lambda
        ext = self.o.contacts()
        rows = []
        for i in reexpose:
            for c in self.ibpc[i]["contacts"]:
                row = dict(c)
                row["lambda_warm"] = np.array(c["lambda"])
                rows.append(row)
        self.ibpc = [p for i, p in enumerate(self.ibpc) if incoll[i]]
        self.push_structure(st, jinv)
        for b in zero_dv:
            dv[b] = 0.0
        if dv.shape[0] < self.n_entities:
            dv = np.concatenate([dv, np.zeros((self.n_entities - dv.shape[0], 6))])
        self.o.push_deltav(dv)
        if rows:
            self.o.push_contacts_raw(_concat(ext, _stack(rows)))
        return events

    def remove_bodies(self, ce, C, gone, st, jinv):
        """RigidCollection.removeBodies (:686-753)."""
        self.members[ce] = [b for b in self.members[ce] if b not in gone]
        mem = self.members[ce]
        was_pinned = C.pinned
        C.pinned = any(self.pinned(b) for b in mem)
        C.R = list(IDENT)
        if C.pinned:
            C.v, C.w = [0.0] * 3, [0.0] * 3
            C.I, C.I0, C.jinv, C.jinv0 = list(ZERO), list(ZERO), list(ZERO), list(ZERO)
            C.mass, C.minv = 0.0, 0.0
        elif was_pinned:
            C.x = [0.0] * 3
            C.mass = 0.0
            for b in mem:
                B = self.load(b, st, jinv)
                C.x = [B.mass * B.x[a] + C.x[a] for a in range(3)]
                C.mass = C.mass + B.mass
            C.minv = 1. / C.mass
            C.x = [C.x[a] * C.minv for a in range(3)]
            I = list(ZERO)
            for b in mem:
                B = self.load(b, st, jinv)
                I = [I[i] + B.I[i] for i in range(9)]
                op = get_op(B.x, B.mass, C.x)
                I = [I[i] + op[i] for i in range(9)]
            C.I = I
            C.jinv = invert3(C.I)
            C.I0 = rtmr(C.R, C.I)
            C.jinv0 = rtmr(C.R, C.jinv)
        else:
            for b in gone:
                B = self.load(b, st, jinv)
                com = [C.x[a] * C.mass for a in range(3)]
                C.x = [-B.mass * B.x[a] + com[a] for a in range(3)]
                s = 1. / C.mass
                com = [com[a] * s for a in range(3)]
                C.mass = C.mass - B.mass
                C.minv = 1. / C.mass
                C.x = [C.x[a] * C.minv for a in range(3)]
                for term in (B.I, get_op(C.x, C.mass, com), get_op(B.x, B.mass, com)):
                    C.I = [C.I[i] - term[i] for i in range(9)]
            C.jinv = invert3(C.I)
            C.I0 = rtmr(C.R, C.I)
            C.jinv0 = rtmr(C.R, C.jinv)
        if not C.pinned:
            C.I = rm0rt(C.R, C.I0)
            C.jinv = rm0rt(C.R, C.jinv0)
        self.update_b2c(ce, C, st)
        # updateBB() (:561-606) from the members, in collection coordinates
        mn, mx = [MAXD] * 3, [-MAXD] * 3
        for b in mem:
            d = self.desc(b)
            if int(d["geom"]) == AM_GEOM_PLANE:
                continue
            Rbc, tbc = self.b2c[b]
            for q in range(8):
                p = mulv(Rbc, [float(a) for a in d["bbox"][q * 3:q * 3 + 3]])
                p = [p[a] + tbc[a] for a in range(3)]
                for a in range(3):
                    mn[a] = jmin(mn[a], p[a])
                    mx[a] = jmax(mx[a], p[a])
        C.bb = bbox_corners(mn, mx)


# ---- contact dict helpers --------------------------------------------------
def _row(c, k):
    return {key: (np.array(val[k]) if np.ndim(val) > 1 else val[k].item()) for key, val in c.items()}


def _stack(rows):
    from adaptivemerging_amd.context import _empty_contacts
    out = _empty_contacts(len(rows))
    for k, r in enumerate(rows):
        for key in out:
            out[key][k] = r[key]
    return out


def _concat(a, b):
    return {k: np.concatenate([a[k], b[k]]) for k in a}


def _slice(c, i, j):
    return {k: v[i:j] for k, v in c.items()}
