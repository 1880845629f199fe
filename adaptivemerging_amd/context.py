"""Python handle over an am_ctx (one GPU context, one HIP stream)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MLP_PARAMS, AMError, ContactArrays, Params, StepStats, dptr, iptr, ptr


class Context:
    """Owns an am_ctx; every method raises AMError on failure."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.am_ctx_create(device, C.byref(h))
        if rc != 0:
            raise AMError(rc, "am_ctx_create failed (no usable HIP device?)")
        self.h = h
        self.params = _lib.default_params()
        self.n = 0
        self.merging_event = False  # Merging.mergingEvent as seen through merge()/unmerge()

    def close(self):
        if self.h:
            self.lib.am_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.am_last_error(self.h)
            raise AMError(rc, msg.decode() if msg else "")

    # -- configuration --------------------------------------------------
    def set_params(self, params: Params | None = None, **kw):
        if params is not None:
            self.params = params
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self.lib.am_params_set(self.h, C.byref(self.params)))

    def upload(self, scene):
        b, nd, x, R, v, w = scene.arrays()
        self._bodies, self._nodes = b, nd
        self.n = len(b)
        subs, b2c = scene.sub_arrays() if hasattr(scene, "sub_arrays") else (None, None)
        if subs is not None and len(subs):
            self._subs, self._b2c = subs, b2c
            self._check(self.lib.am_scene_upload_ex(self.h, ptr(b), len(b), ptr(nd), len(nd), ptr(subs), ptr(b2c),
                                                    len(subs)))
        else:
            self._check(self.lib.am_scene_upload(self.h, ptr(b), len(b), ptr(nd), len(nd)))
        self.push_state(x, R, v, w)
        springs = scene.spring_array() if hasattr(scene, "spring_array") else None
        if springs is not None and len(springs):
            self.set_springs(springs)
        kin = scene.kinematic_array() if hasattr(scene, "kinematic_array") else None
        if kin is not None and len(kin):
            self.set_kinematic(kin)

    def set_kinematic(self, table):
        """Prescribed motions (a KINEMATIC_DTYPE array, physical body indices)."""
        self._kin = np.ascontiguousarray(table, dtype=_lib.KINEMATIC_DTYPE)
        self._check(self.lib.am_kinematic_set(self.h, ptr(self._kin), len(self._kin)))

    def kinematic_update(self, t: float):
        self._check(self.lib.am_kinematic_update(self.h, t))

    @property
    def sim_time(self) -> float:
        t = C.c_double()
        self._check(self.lib.am_sim_time_get(self.h, C.byref(t)))
        return t.value

    @sim_time.setter
    def sim_time(self, t: float):
        self._check(self.lib.am_sim_time_set(self.h, t))

    def impulse_add(self, body: int, point_w, force):
        """A force at a world point, applied (to the body's top-level entity)
        by the next external-force pass only (RigidBodySystem.applyImpulse)."""
        p = np.ascontiguousarray(point_w, dtype=np.float64)
        f = np.ascontiguousarray(force, dtype=np.float64)
        self._check(self.lib.am_impulse_add(self.h, body, ptr(p), ptr(f)))

    def launch_shape(self):
        """(workgroups, threads per workgroup) of the last main solve's PGS kernel."""
        g, b = C.c_int32(), C.c_int32()
        self._check(self.lib.am_launch_shape_pull(self.h, C.byref(g), C.byref(b)))
        return g.value, b.value

    def scene_iterations(self):
        """(sweeps used, max |delta lambda| of the last sweep) per scene of
        the last main solve that stopped scenes one by one."""
        n = C.c_int32()
        self._check(self.lib.am_scene_iterations_pull(self.h, None, None, 0, C.byref(n)))
        used = np.zeros(n.value, np.int32)
        last = np.zeros(n.value)
        self._check(self.lib.am_scene_iterations_pull(self.h, ptr(used), ptr(last), n.value, C.byref(n)))
        return used, last

    def apply_impulses(self):
        self._check(self.lib.am_apply_impulses(self.h))

    def apply_viscous_decay(self):
        self._check(self.lib.am_apply_viscous_decay(self.h))

    def set_springs(self, springs):
        """RigidBodySystem.springs: a SPRING_DTYPE array (physical body indices)."""
        self._springs = np.ascontiguousarray(springs, dtype=_lib.SPRING_DTYPE)
        self._check(self.lib.am_springs_set(self.h, ptr(self._springs), len(self._springs)))

    def apply_spring_forces(self):
        self._check(self.lib.am_apply_spring_forces(self.h))

    # -- sleeping (Sleeping.java) ------------------------------------------------
    def wake(self) -> int:
        n = C.c_int32()
        self._check(self.lib.am_wake(self.h, C.byref(n)))
        return n.value

    def sleep(self) -> int:
        n = C.c_int32()
        self._check(self.lib.am_sleep(self.h, C.byref(n)))
        return n.value

    def wake_all(self):
        self._check(self.lib.am_wake_all(self.h))

    def unmerge_all(self) -> int:
        n = C.c_int32()
        self._check(self.lib.am_unmerge_all(self.h, C.byref(n)))
        return n.value

    def sleep_state(self, n: int | None = None):
        """(sleeping flags, history lengths, histories (n, AM_SLEEP_HIST_STRIDE))
        of the first n entities (default: all)."""
        n = self.n_entities() if n is None else n
        flags = np.zeros(n, np.uint8)
        hn = np.zeros(n, np.int32)
        hist = np.zeros((n, _lib.AM_SLEEP_HIST_STRIDE))
        self._check(self.lib.am_sleep_state_pull(self.h, ptr(flags), ptr(hn), ptr(hist), n))
        return flags, hn, hist

    def structure(self):
        """(parent per entity, top-level list) of the merged-body structure."""
        ne, nt = C.c_int32(), C.c_int32()
        self._check(self.lib.am_structure_pull(self.h, None, 0, None, 0, C.byref(ne), C.byref(nt)))
        parent = np.zeros(ne.value, np.int32)
        top = np.zeros(nt.value, np.int32)
        self._check(self.lib.am_structure_pull(self.h, ptr(parent), len(parent), ptr(top), len(top), C.byref(ne),
                                               C.byref(nt)))
        return parent, top

    def entities(self) -> np.ndarray:
        """am_body_desc of every entity (collections included)."""
        parent, _ = self.structure()
        out = np.zeros(len(parent), dtype=_lib.BODY_DTYPE)
        self._check(self.lib.am_entities_pull(self.h, ptr(out), len(out)))
        return out

    # -- state ----------------------------------------------------------
    def push_state(self, x=None, R=None, v=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, R, v, w)]
        self._check(self.lib.am_state_push(self.h, *[ptr(a) for a in arrs], self.n))

    def pull_state(self, n: int | None = None):
        """Pose and velocity of the first n entities (default: the physical bodies)."""
        n = self.n if n is None else n
        x = np.zeros((n, 3)); R = np.zeros((n, 9)); v = np.zeros((n, 3)); w = np.zeros((n, 3))
        self._check(self.lib.am_state_pull(self.h, ptr(x), ptr(R), ptr(v), ptr(w), n))
        return x, R, v, w

    def n_entities(self) -> int:
        ne, nt = C.c_int32(), C.c_int32()
        self._check(self.lib.am_structure_pull(self.h, None, 0, None, 0, C.byref(ne), C.byref(nt)))
        return ne.value

    def pull_jinv(self, n: int | None = None):
        """World inverse inertia of the first n entities (default: all)."""
        n = self.n_entities() if n is None else n
        j = np.zeros((n, 9))
        self._check(self.lib.am_jinv_pull(self.h, ptr(j), n))
        return j

    def push_forces(self, force=None, torque=None):
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
        t = None if torque is None else np.ascontiguousarray(torque, dtype=np.float64)
        n = len(f) if f is not None else (len(t) if t is not None else 0)
        self._check(self.lib.am_forces_push(self.h, ptr(f), ptr(t), n))

    def pull_forces(self, n: int | None = None):
        n = self.n_entities() if n is None else n
        f = np.zeros((n, 3)); t = np.zeros((n, 3))
        self._check(self.lib.am_forces_pull(self.h, ptr(f), ptr(t), n))
        return f, t

    def pull_deltav(self, n: int | None = None):
        n = self.n_entities() if n is None else n
        d = np.zeros((n, 6))
        self._check(self.lib.am_deltav_pull(self.h, ptr(d), n))
        return d

    def clear_bodies(self):
        self._check(self.lib.am_clear_bodies(self.h))

    def apply_gravity(self):
        self._check(self.lib.am_apply_gravity(self.h))

    # -- contacts -------------------------------------------------------
    def update_contacts_map(self):
        self._check(self.lib.am_update_contacts_map(self.h))

    def detect(self, dt: float = 0.01) -> int:
        n = C.c_int32()
        self._check(self.lib.am_detect(self.h, dt, C.byref(n)))
        return n.value

    def push_contacts(self, body1, body2, point_w, normal_w, violation, info=None, bv1=None, bv2=None):
        b1 = np.ascontiguousarray(body1, dtype=np.int32)
        b2 = np.ascontiguousarray(body2, dtype=np.int32)
        p = np.ascontiguousarray(point_w, dtype=np.float64)
        nw = np.ascontiguousarray(normal_w, dtype=np.float64)
        vi = np.ascontiguousarray(violation, dtype=np.float64)
        inf = None if info is None else np.ascontiguousarray(info, dtype=np.int32)
        v1 = None if bv1 is None else np.ascontiguousarray(bv1, dtype=np.int32)
        v2 = None if bv2 is None else np.ascontiguousarray(bv2, dtype=np.int32)
        self._check(self.lib.am_contacts_push(self.h, len(b1), ptr(b1), ptr(b2), ptr(p), ptr(nw), ptr(vi), ptr(inf),
                                              ptr(v1), ptr(v2)))

    def n_contacts(self) -> int:
        n = C.c_int32()
        self._check(self.lib.am_contacts_pull(self.h, None, 0, C.byref(n)))
        return n.value

    def contacts(self) -> dict:
        n = self.n_contacts()
        out = _empty_contacts(n)
        ca = contact_arrays(out)
        m = C.c_int32()
        self._check(self.lib.am_contacts_pull(self.h, C.byref(ca), n, C.byref(m)))
        return out

    def push_lambdas(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._check(self.lib.am_lambdas_push(self.h, ptr(lam), len(lam)))

    # -- warm start / solve ---------------------------------------------
    def warm_start(self, post: bool = False):
        bad, rep = C.c_int32(), C.c_int32()
        self._check(self.lib.am_warm_start(self.h, int(post), C.byref(bad), C.byref(rep)))
        return bad.value, rep.value

    def redo_warm_start(self):
        self._check(self.lib.am_redo_warm_start(self.h))

    def no_warm_start(self):
        self._check(self.lib.am_no_warm_start(self.h))

    def solve(self, dt: float, post: bool = False):
        it, last = C.c_int32(), C.c_double()
        self._check(self.lib.am_solve(self.h, dt, int(post), C.byref(it), C.byref(last)))
        return it.value, last.value

    def update_in_collections(self, dt: float) -> int:
        it = C.c_int32()
        self._check(self.lib.am_update_in_collections(self.h, dt, C.byref(it)))
        return it.value

    # -- merging --------------------------------------------------------
    def accumulate_for_unmerging(self, dt: float):
        self._check(self.lib.am_accumulate_for_unmerging(self.h, dt))

    def unmerge(self, dt: float) -> int:
        n = C.c_int32()
        self._check(self.lib.am_unmerge(self.h, dt, C.byref(n)))
        self.merging_event = self.merging_event or n.value > 0
        return n.value

    def clear_top_level(self):
        self._check(self.lib.am_clear_top_level(self.h))

    def clear_body_pair_contacts(self):
        self._check(self.lib.am_clear_body_pair_contacts(self.h))

    def accumulate_for_merging(self, dt: float):
        self._check(self.lib.am_accumulate_for_merging(self.h, dt))

    def merge(self) -> int:
        n = C.c_int32()
        self._check(self.lib.am_merge(self.h, C.byref(n)))
        self.merging_event = self.merging_event or n.value > 0
        return n.value

    def body_pairs(self):
        """(external, internal) body pairs as lists of dicts."""
        ne, ni = C.c_int32(), C.c_int32()
        self._check(self.lib.am_body_pairs_pull(self.h, None, 0, C.byref(ne), C.byref(ni)))
        arr = (_lib.BodyPair * max(ne.value + ni.value, 1))()
        self._check(self.lib.am_body_pairs_pull(self.h, C.cast(arr, C.c_void_p), ne.value + ni.value, C.byref(ne),
                                                C.byref(ni)))
        out = []
        for k in range(ne.value + ni.value):
            b = arr[k]
            out.append(dict(body1=b.body1, body2=b.body2, in_collection=b.in_collection, n_contacts=b.n_contacts,
                            hist_n=b.hist_n, state_n=b.state_n, flag=b.flag, collection=b.collection,
                            hist=list(b.hist)[:max(b.hist_n, 0)] if not b.in_collection else [],
                            states=list(b.states)[:max(b.state_n, 0)]))
        return out[:ne.value], out[ne.value:]

    def internal_contacts(self) -> dict:
        n = C.c_int32()
        self._check(self.lib.am_internal_contacts_pull(self.h, None, 0, C.byref(n)))
        out = _empty_contacts(n.value)
        ca = contact_arrays(out)
        self._check(self.lib.am_internal_contacts_pull(self.h, C.byref(ca), n.value, C.byref(n)))
        return out

    def solve_order(self) -> np.ndarray:
        n = C.c_int32()
        self._check(self.lib.am_solve_order_pull(self.h, None, 0, C.byref(n)))
        order = np.zeros(n.value, np.int32)
        self._check(self.lib.am_solve_order_pull(self.h, ptr(order), n.value, C.byref(n)))
        return order

    def colours(self):
        nc = self.n_contacts()
        col = np.zeros(nc, np.int32)
        k = C.c_int32()
        self._check(self.lib.am_colours_pull(self.h, ptr(col) if nc else None, nc, C.byref(k)))
        return col, k.value

    def changes(self) -> np.ndarray:
        n = C.c_int32()
        self._check(self.lib.am_changes_pull(self.h, None, 0, C.byref(n)))
        ch = np.zeros(n.value)
        self._check(self.lib.am_changes_pull(self.h, ptr(ch), n.value, C.byref(n)))
        return ch

    # -- integration ----------------------------------------------------
    def advance_velocities(self, dt: float):
        self._check(self.lib.am_advance_velocities(self.h, dt))

    def advance_positions(self, dt: float):
        self._check(self.lib.am_advance_positions(self.h, dt))

    def advance_positions_post(self, dt: float):
        self._check(self.lib.am_advance_positions_post(self.h, dt))

    def step(self, dt: float) -> dict:
        st = StepStats()
        self._check(self.lib.am_step(self.h, dt, C.byref(st)))
        return st.as_dict()

    def motion_metric(self, pairs) -> np.ndarray:
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        out = np.zeros(len(p))
        self._check(self.lib.am_motion_metric(self.h, ptr(p), len(p), ptr(out)))
        return out

    def autotune(self, dt: float, sweeps: int = 20, reps: int = 3, apply: bool = True):
        """Times the PGS kernel's launch shapes on the contacts currently held
        (am_autotune); every shape must give the first shape's bits.  Returns
        (trials as dicts, index of the fastest); with apply the fastest shape
        becomes this context's pgs_schedule / pgs_grid / pgs_block."""
        cap = 64
        trials = (_lib.TuneTrial * cap)()
        n, best = C.c_int32(), C.c_int32()
        self._check(self.lib.am_autotune(self.h, dt, sweeps, reps, int(apply), trials, cap, C.byref(n),
                                         C.byref(best)))
        out = [trials[k].as_dict() for k in range(min(n.value, cap))]
        if apply:
            b = out[best.value]
            for k in ("pgs_schedule", "pgs_grid", "pgs_block"):
                setattr(self.params, k, b[k])
        return out, best.value

    # -- learned warm start -------------------------------------------------
    def mlp_init(self, seed: int = 0):
        self._check(self.lib.am_mlp_init(self.h, C.c_uint64(seed)))

    def mlp_params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self.lib.am_mlp_get_params(self.h, ptr(p), len(p)))
        return p

    def set_mlp_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.lib.am_mlp_set_params(self.h, ptr(p), len(p)))

    def train_size(self) -> int:
        n = C.c_int32(0)
        self._check(self.lib.am_train_size(self.h, C.byref(n)))
        return n.value

    def train_clear(self):
        self._check(self.lib.am_train_clear(self.h))

    def train_epochs(self, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> float:
        loss = C.c_float(0)
        self._check(self.lib.am_train_epochs(self.h, epochs, batch, C.c_float(lr), C.c_uint64(seed),
                                             C.byref(loss)))
        return loss.value


def _empty_contacts(n: int) -> dict:
    return {
        "body1": np.zeros(n, np.int32), "body2": np.zeros(n, np.int32), "bv1": np.zeros(n, np.int32),
        "bv2": np.zeros(n, np.int32), "info": np.zeros(n, np.int32),
        "point_b1": np.zeros((n, 3)), "normal_b1": np.zeros((n, 3)), "tangent1_b1": np.zeros((n, 3)),
        "tangent2_b1": np.zeros((n, 3)), "violation": np.zeros(n), "prev_violation": np.zeros(n),
        "jacobian": np.zeros((n, 36)), "lambda": np.zeros((n, 3)), "lambda_warm": np.zeros((n, 3)),
        "b": np.zeros((n, 3)), "diag": np.zeros((n, 3)), "state": np.zeros(n, np.int32),
        "new_this_step": np.zeros(n, np.int32),
    }


def contact_arrays(out: dict) -> ContactArrays:
    """ContactArrays pointing into the numpy arrays of `out` (keys as in Context.contacts)."""
    ca = ContactArrays()
    for key, field in [("body1", "body1"), ("body2", "body2"), ("bv1", "bv1"), ("bv2", "bv2"), ("info", "info"),
                       ("state", "state"), ("new_this_step", "new_this_step")]:
        setattr(ca, field, iptr(out[key]))
    for key, field in [("point_b1", "point_b1"), ("normal_b1", "normal_b1"), ("tangent1_b1", "tangent1_b1"),
                       ("tangent2_b1", "tangent2_b1"), ("violation", "violation"),
                       ("prev_violation", "prev_violation"), ("jacobian", "jacobian"), ("lambda", "lambda_"),
                       ("lambda_warm", "lambda_warm"), ("b", "b"), ("diag", "diag")]:
        setattr(ca, field, dptr(out[key]))
    return ca
