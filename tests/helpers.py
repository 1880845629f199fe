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
