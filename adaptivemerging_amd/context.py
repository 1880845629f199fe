"""Python handle over an am_ctx (one GPU context, one HIP stream)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import AMError, ContactArrays, Params, StepStats, dptr, iptr, ptr


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
        self._check(self.lib.am_scene_upload(self.h, ptr(b), len(b), ptr(nd), len(nd)))
        self.push_state(x, R, v, w)

    def set_collections(self, parent, top, member_first, member_count, members):
        parent = np.ascontiguousarray(parent, dtype=np.int32)
        top = np.ascontiguousarray(top, dtype=np.int32)
        mf = np.ascontiguousarray(member_first, dtype=np.int32)
        mc = np.ascontiguousarray(member_count, dtype=np.int32)
        mem = np.ascontiguousarray(members, dtype=np.int32)
        self._check(self.lib.am_collections_set(self.h, ptr(parent), ptr(top), len(top), ptr(mf), ptr(mc),
                                                ptr(mem) if len(mem) else None, len(mem)))

    # -- state ----------------------------------------------------------
    def push_state(self, x=None, R=None, v=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, R, v, w)]
        self._check(self.lib.am_state_push(self.h, *[ptr(a) for a in arrs], self.n))

    def pull_state(self):
        x = np.zeros((self.n, 3)); R = np.zeros((self.n, 9)); v = np.zeros((self.n, 3)); w = np.zeros((self.n, 3))
        self._check(self.lib.am_state_pull(self.h, ptr(x), ptr(R), ptr(v), ptr(w), self.n))
        return x, R, v, w

    def pull_jinv(self):
        j = np.zeros((self.n, 9))
        self._check(self.lib.am_jinv_pull(self.h, ptr(j), self.n))
        return j

    def push_forces(self, force=None, torque=None):
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
        t = None if torque is None else np.ascontiguousarray(torque, dtype=np.float64)
        self._check(self.lib.am_forces_push(self.h, ptr(f), ptr(t), self.n))

    def pull_forces(self):
        f = np.zeros((self.n, 3)); t = np.zeros((self.n, 3))
        self._check(self.lib.am_forces_pull(self.h, ptr(f), ptr(t), self.n))
        return f, t

    def pull_deltav(self):
        d = np.zeros((self.n, 6))
        self._check(self.lib.am_deltav_pull(self.h, ptr(d), self.n))
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
        out = {
            "body1": np.zeros(n, np.int32), "body2": np.zeros(n, np.int32), "bv1": np.zeros(n, np.int32),
            "bv2": np.zeros(n, np.int32), "info": np.zeros(n, np.int32),
            "point_b1": np.zeros((n, 3)), "normal_b1": np.zeros((n, 3)), "tangent1_b1": np.zeros((n, 3)),
            "tangent2_b1": np.zeros((n, 3)), "violation": np.zeros(n), "prev_violation": np.zeros(n),
            "jacobian": np.zeros((n, 36)), "lambda": np.zeros((n, 3)), "lambda_warm": np.zeros((n, 3)),
            "b": np.zeros((n, 3)), "diag": np.zeros((n, 3)), "state": np.zeros(n, np.int32),
            "new_this_step": np.zeros(n, np.int32),
        }
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

    def single_sweep(self, dt: float) -> int:
        it = C.c_int32()
        self._check(self.lib.am_single_sweep(self.h, dt, C.byref(it)))
        return it.value

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

    def step(self, dt: float) -> dict:
        st = StepStats()
        self._check(self.lib.am_step(self.h, dt, C.byref(st)))
        return st.as_dict()

    def motion_metric(self, pairs) -> np.ndarray:
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        out = np.zeros(len(p))
        self._check(self.lib.am_motion_metric(self.h, ptr(p), len(p), ptr(out)))
        return out


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
