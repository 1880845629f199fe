"""ctypes binding of the CPU oracle (oracle/_build/liboracle_am.so).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from adaptivemerging_amd._lib import BODY_DTYPE, SPRING_DTYPE, ContactArrays, Params, default_params_host, ptr  # noqa: F401
from adaptivemerging_amd.context import contact_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle_am.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_libs: dict = {}


def load(path: str | None = None) -> C.CDLL:
    """The oracle library at `path` (default: the in-tree build).  Another
    path lets a test load a copy of the oracle built with a planted mistake."""
    path = path or LIB
    if path in _libs:
        return _libs[path]
    if path == LIB and not os.path.exists(LIB):
        build()
    lib = C.CDLL(path)
    vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
    sig = {
        "orc_create": (vp, [vp, i32, vp, i32]),
        "orc_create_ex": (vp, [vp, i32, vp, i32, vp, vp, i32]),
        "orc_destroy": (None, [vp]),
        "orc_set_params": (None, [vp, C.POINTER(Params)]),
        "orc_set_structure": (None, [vp, vp, vp, i32, vp, vp, vp]),
        "orc_set_state": (None, [vp, vp, vp, vp, vp]),
        "orc_get_state": (None, [vp, vp, vp, vp, vp]),
        "orc_set_jinv": (None, [vp, i32, vp]),
        "orc_get_jinv": (None, [vp, vp]),
        "orc_set_forces": (None, [vp, vp, vp]),
        "orc_get_forces": (None, [vp, vp, vp]),
        "orc_get_deltav": (None, [vp, vp]),
        "orc_set_deltav": (None, [vp, vp]),
        "orc_resize": (None, [vp, i32]),
        "orc_entities": (i32, [vp]),
        "orc_set_entity": (None, [vp, i32, vp]),
        "orc_get_entity": (None, [vp, i32, vp]),
        "orc_contacts_set_raw": (None, [vp, i32, C.POINTER(ContactArrays)]),
        "orc_clear_bodies": (None, [vp]),
        "orc_apply_gravity": (None, [vp]),
        "orc_set_springs": (None, [vp, vp, C.c_int32]),
        "orc_apply_springs": (None, [vp]),
        "orc_update_contacts_map": (None, [vp]),
        "orc_detect": (i32, [vp]),
        "orc_contacts_set": (None, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
        "orc_contacts_get": (i32, [vp, C.POINTER(ContactArrays), i32]),
        "orc_lambdas_set": (None, [vp, vp, i32]),
        "orc_box_box": (i32, [vp, i32, i32]),
        "orc_box_plane": (i32, [vp, i32, i32]),
        "orc_box_sphere": (i32, [vp, i32, i32, vp, dbl]),
        "orc_box_sphere_test": (i32, [vp, i32, vp, dbl]),
        "orc_warm_start": (None, [vp, i32, C.POINTER(i32), C.POINTER(i32)]),
        "orc_redo_warm_start": (None, [vp]),
        "orc_no_warm_start": (None, [vp]),
        "orc_pgs_solve": (i32, [vp, vp, i32, dbl, i32, i32, vp, C.POINTER(dbl)]),
        "orc_advance_velocities": (None, [vp, dbl]),
        "orc_advance_positions": (None, [vp, dbl]),
        "orc_advance_positions_post": (None, [vp, dbl]),
        "orc_motion_metric": (dbl, [vp, i32, i32]),
        "orc_motion_metric_pos": (dbl, [vp, i32, i32, dbl]),
        "orc_step": (i32, [vp, dbl, C.POINTER(i32), C.POINTER(dbl)]),
        "orc_scene_iterations": (i32, [vp, vp, vp, i32]),
        "orc_sin": (dbl, [dbl]),
        "orc_cos": (dbl, [dbl]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs[path] = lib
    return lib


class Oracle:
    """CPU restatement of the reference path over the same Scene arrays."""

    def __init__(self, scene, params: Params | None = None, lib_path: str | None = None):
        self.lib = load(lib_path)
        b, nd, x, R, v, w = scene.arrays()
        self._b, self._nd = b, nd
        self.n = len(b)
        subs, b2c = scene.sub_arrays() if hasattr(scene, "sub_arrays") else (None, None)
        if subs is not None and len(subs):
            self._subs, self._b2c = subs, b2c
            self.h = C.c_void_p(self.lib.orc_create_ex(ptr(b), len(b), ptr(nd), len(nd), ptr(subs), ptr(b2c),
                                                       len(subs)))
        else:
            self.h = C.c_void_p(self.lib.orc_create(ptr(b), len(b), ptr(nd), len(nd)))
        self.params = params if params is not None else default_params_host()
        self.lib.orc_set_params(self.h, C.byref(self.params))
        self.set_state(x, R, v, w)
        springs = scene.spring_array() if hasattr(scene, "spring_array") else None
        if springs is not None and len(springs):
            self.set_springs(springs)

    def __del__(self):
        try:
            if self.h:
                self.lib.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_params(self, params: Params | None = None, **kw):
        if params is not None:
            self.params = params
        for k, v in kw.items():
            setattr(self.params, k, v)
        self.lib.orc_set_params(self.h, C.byref(self.params))

    def set_structure(self, parent, top, member_first, member_count, members):
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (parent, top, member_first, member_count, members)]
        self._struct = arrs
        self.lib.orc_set_structure(self.h, ptr(arrs[0]), ptr(arrs[1]), len(arrs[1]), ptr(arrs[2]), ptr(arrs[3]),
                                   ptr(arrs[4]) if len(arrs[4]) else None)

    def set_state(self, x=None, R=None, v=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, R, v, w)]
        self.lib.orc_set_state(self.h, *[ptr(a) for a in arrs])

    def set_jinv(self, entity, jinv):
        j = np.ascontiguousarray(jinv, dtype=np.float64)
        self.lib.orc_set_jinv(self.h, entity, ptr(j))

    @property
    def n_entities(self) -> int:
        return self.lib.orc_entities(self.h)

    def pull_state(self, n: int | None = None):
        """Pose and velocity of the first n entities (default: physical bodies)."""
        m = self.n_entities
        x = np.zeros((m, 3)); R = np.zeros((m, 9)); v = np.zeros((m, 3)); w = np.zeros((m, 3))
        self.lib.orc_get_state(self.h, ptr(x), ptr(R), ptr(v), ptr(w))
        n = self.n if n is None else n
        return x[:n], R[:n], v[:n], w[:n]

    def pull_jinv(self):
        j = np.zeros((self.n_entities, 9))
        self.lib.orc_get_jinv(self.h, ptr(j))
        return j

    def resize(self, n: int):
        self.lib.orc_resize(self.h, n)

    def set_entity(self, e: int, rec):
        r = np.ascontiguousarray(np.asarray(rec, dtype=BODY_DTYPE).reshape(()))
        self.lib.orc_set_entity(self.h, e, ptr(r))

    def entity(self, e: int):
        r = np.zeros((), dtype=BODY_DTYPE)
        self.lib.orc_get_entity(self.h, e, ptr(r))
        return r

    def push_deltav(self, dv):
        d = np.ascontiguousarray(dv, dtype=np.float64)
        self.lib.orc_set_deltav(self.h, ptr(d))

    def push_contacts_raw(self, c: dict):
        c = {k: np.ascontiguousarray(v) for k, v in c.items()}
        self._raw = c
        ca = contact_arrays(c)
        self.lib.orc_contacts_set_raw(self.h, len(c["body1"]), C.byref(ca))

    def push_forces(self, force=None, torque=None):
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
        t = None if torque is None else np.ascontiguousarray(torque, dtype=np.float64)
        self.lib.orc_set_forces(self.h, ptr(f), ptr(t))

    def pull_forces(self):
        f = np.zeros((self.n_entities, 3)); t = np.zeros((self.n_entities, 3))
        self.lib.orc_get_forces(self.h, ptr(f), ptr(t))
        return f, t

    def pull_deltav(self):
        d = np.zeros((self.n_entities, 6))
        self.lib.orc_get_deltav(self.h, ptr(d))
        return d

    def clear_bodies(self):
        self.lib.orc_clear_bodies(self.h)

    def apply_gravity(self):
        self.lib.orc_apply_gravity(self.h)

    def set_springs(self, springs):
        self._springs = np.ascontiguousarray(springs, dtype=SPRING_DTYPE)
        self.lib.orc_set_springs(self.h, ptr(self._springs), len(self._springs))

    def apply_springs(self):
        self.lib.orc_apply_springs(self.h)

    def update_contacts_map(self):
        self.lib.orc_update_contacts_map(self.h)

    def detect(self) -> int:
        return self.lib.orc_detect(self.h)

    def push_contacts(self, body1, body2, point_w, normal_w, violation, info=None, bv1=None, bv2=None):
        b1 = np.ascontiguousarray(body1, dtype=np.int32)
        b2 = np.ascontiguousarray(body2, dtype=np.int32)
        p = np.ascontiguousarray(point_w, dtype=np.float64)
        nw = np.ascontiguousarray(normal_w, dtype=np.float64)
        vi = np.ascontiguousarray(violation, dtype=np.float64)
        opt = [None if a is None else np.ascontiguousarray(a, dtype=np.int32) for a in (info, bv1, bv2)]
        self.lib.orc_contacts_set(self.h, len(b1), ptr(b1), ptr(b2), ptr(p), ptr(nw), ptr(vi), *[ptr(a) for a in opt])

    def contacts(self) -> dict:
        n = self.lib.orc_contacts_get(self.h, C.byref(ContactArrays()), 0)
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
        self.lib.orc_contacts_get(self.h, C.byref(ca), n)
        return out

    def push_lambdas(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self.lib.orc_lambdas_set(self.h, ptr(lam), len(lam))

    def warm_start(self, post=False):
        bad, rep = C.c_int32(), C.c_int32()
        self.lib.orc_warm_start(self.h, int(post), C.byref(bad), C.byref(rep))
        return bad.value, rep.value

    def redo_warm_start(self):
        self.lib.orc_redo_warm_start(self.h)

    def no_warm_start(self):
        self.lib.orc_no_warm_start(self.h)

    def solve(self, order, dt, main=True, post=False):
        order = np.ascontiguousarray(order, dtype=np.int32)
        iters = max(self.params.iterations, self.params.iterations_in_collection, 1)
        ch = np.zeros(iters)
        last = C.c_double()
        used = self.lib.orc_pgs_solve(self.h, ptr(order), len(order), dt, int(main), int(post), ptr(ch),
                                      C.byref(last))
        return used, last.value, ch[:used] if main else ch[:0]

    def scene_iterations(self):
        """(sweeps used, max |dl| of the last sweep) per scene of the last
        main solve that stopped scenes one by one."""
        n = self.lib.orc_scene_iterations(self.h, None, None, 0)
        used = np.zeros(n, np.int32)
        last = np.zeros(n)
        self.lib.orc_scene_iterations(self.h, ptr(used), ptr(last), n)
        return used, last

    def box_box(self, b1, b2):
        return self.lib.orc_box_box(self.h, b1, b2)

    def box_plane(self, box, plane):
        return self.lib.orc_box_plane(self.h, box, plane)

    def box_sphere(self, box, other, c, r):
        c = np.ascontiguousarray(c, dtype=np.float64)
        return self.lib.orc_box_sphere(self.h, box, other, ptr(c), r)

    def box_sphere_test(self, box, c, r):
        c = np.ascontiguousarray(c, dtype=np.float64)
        return self.lib.orc_box_sphere_test(self.h, box, ptr(c), r)

    def advance_velocities(self, dt):
        self.lib.orc_advance_velocities(self.h, dt)

    def advance_positions(self, dt):
        self.lib.orc_advance_positions(self.h, dt)

    def advance_positions_post(self, dt):
        self.lib.orc_advance_positions_post(self.h, dt)

    def motion_metric(self, a, b=-1) -> float:
        return self.lib.orc_motion_metric(self.h, a, b)

    def motion_metric_pos(self, a, b, dt) -> float:
        return self.lib.orc_motion_metric_pos(self.h, a, b, dt)

    def step(self, dt):
        it, last = C.c_int32(), C.c_double()
        n = self.lib.orc_step(self.h, dt, C.byref(it), C.byref(last))
        return n, it.value, last.value
