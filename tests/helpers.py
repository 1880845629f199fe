"""Shared helpers for the parity tests (compare libam with the CPU oracle)."""
import numpy as np

CONTACT_INT_KEYS = ["body1", "body2", "bv1", "bv2", "info", "state", "new_this_step"]
CONTACT_FLOAT_KEYS = ["point_b1", "normal_b1", "tangent1_b1", "tangent2_b1", "violation", "prev_violation",
                      "jacobian", "lambda", "lambda_warm", "b", "diag"]


def assert_bits_equal(name, a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{name}: shape {a.shape} != {b.shape}"
    if a.dtype.kind == "f":
        same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
    else:
        same = a == b
    if not np.all(same):
        idx = np.argwhere(~same)[:5]
        detail = "; ".join(f"{tuple(i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}" for i in idx)
        raise AssertionError(f"{name}: {int((~same).sum())} of {same.size} differ ({detail})")


def assert_contacts_equal(g, o, keys=None):
    keys = keys or CONTACT_INT_KEYS + CONTACT_FLOAT_KEYS
    assert len(g["body1"]) == len(o["body1"]), f"contact count {len(g['body1'])} != {len(o['body1'])}"
    for k in keys:
        assert_bits_equal(k, g[k], o[k])


def twin(scene, **params):
    """A GPU context and an oracle on the same scene and parameters."""
    from adaptivemerging_amd.context import Context
    from oracle.pyoracle import Oracle

    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(**params)
    orc = Oracle(scene)
    for k, _ in type(ctx.params)._fields_:
        setattr(orc.params, k, getattr(ctx.params, k))
    orc.set_params()
    return ctx, orc


def oracle_step_like_gpu(ctx, orc, dt):
    """One advanceTime (no merging) on both, the oracle sweeping in the GPU's
    sweep order; external forces are gravity then the scene's springs.
    Returns (iterations gpu, iterations oracle)."""
    ctx.clear_bodies(); orc.clear_bodies()
    ctx.apply_gravity(); orc.apply_gravity()
    ctx.apply_spring_forces(); orc.apply_springs()
    ctx.update_contacts_map(); orc.update_contacts_map()
    ctx.detect(dt); orc.detect()
    ctx.warm_start(); orc.warm_start()
    ctx.redo_warm_start(); orc.redo_warm_start()
    it_g, _ = ctx.solve(dt)
    order = ctx.solve_order()
    it_o, _, _ = orc.solve(order, dt)
    ctx.advance_velocities(dt); orc.advance_velocities(dt)
    ctx.advance_positions(dt); orc.advance_positions(dt)
    return it_g, it_o


def merging_oracle_step(mo, dt, step, order=None, kick=None):
    """One RigidBodySystem.advanceTime (RigidBodySystem.java:102-185) on the
    merging oracle alone: every phase including sleeping, contacts swept in
    list order unless order(n) gives the sweep order of the single sweep.
    kick = (body, fy) adds an upward force after the external forces.
    Returns (merged, cut, asleep)."""
    import numpy as np

    mo.total_steps += 1
    mo.clear_bodies()
    mo.o.apply_gravity()
    mo.o.apply_springs()
    if kick is not None:
        f, t = mo.o.pull_forces()
        f[kick[0]][1] += kick[1]
        mo.o.push_forces(f, t)
    mo.o.update_contacts_map()
    n = mo.o.detect()
    mo.o.warm_start()
    mo.wake()
    n_int = sum(len(q["contacts"]) for q in mo.ibpc if not mo.sleeping(mo.parent[q["b1"]]))
    mo.update_in_collections(dt, (lambda: order(n + n_int)) if order else
                             (lambda: np.arange(n + n_int, dtype=np.int32)))
    mo.accumulate_for_unmerging(dt)
    cut = mo.unmerge(dt)
    if mo.merging_event:
        mo.clear_top_level()
    mo.o.redo_warm_start()
    mo.o.solve(np.arange(len(mo.o.contacts()["body1"]), dtype=np.int32), dt)
    mo.clear_body_pair_contacts()
    mo.advance_velocities(dt)
    mo.accumulate_for_merging(dt)
    mo.advance_positions(dt)
    merged = 0
    steps = mo.params.steps_between_merge_events
    if steps > 0 and step % steps == 0:
        merged = mo.merge()
    asleep = mo.sleep()
    return merged, cut, asleep


def load_scene_fixture(name):
    """A scene written by tests/golden/make_scene_fixtures.py, as a Scene, and
    its <collision> settings."""
    import json
    import os

    import numpy as np

    from adaptivemerging_amd.scene import Scene

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", f"scene_{name}.npz"))
    s = Scene()
    s.bodies = list(z["bodies"])
    s.nodes = list(z["nodes"])
    s.x = [np.array(a) for a in z["x"]]
    s.R = [np.array(a).reshape(3, 3) for a in z["R"]]
    s.v = [np.array(a) for a in z["v"]]
    s.w = [np.array(a) for a in z["w"]]
    s.names = [str(a) for a in z["names"]]
    s.subs = list(z["subs"])
    s.sub_b2c = [list(a) for a in z["sub_b2c"]]
    s.springs = list(z["springs"])
    return s, json.loads(str(z["collision"]))
