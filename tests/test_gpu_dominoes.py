"""Dominoes on kinematic platforms with merging on (the domino configuration:
rows of dominoes on oscillating platforms, a 5 m/s push on the first domino of
each row), on the GPU against the merging oracle phase by phase and bit for
bit at a reduced size, am_step against the phases, and the full size (2,000
dominoes on 10 platforms) through the push and the start of the cascade."""
import math

import numpy as np
import pytest

from adaptivemerging_amd import workloads
from adaptivemerging_amd.context import Context
from oracle.merging import MergingOracle
from oracle.pyoracle import load as load_oracle
from tests.helpers import assert_bits_equal, assert_contacts_equal
from tests.test_gpu_merging import check_organized, compare_world

pytestmark = pytest.mark.gpu

DT = 0.01
PARAMS = dict(gravity_amount=9.81, iterations=50, enable_sleeping=0)


def make(scene, **extra):
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(**{**PARAMS, **extra})
    mo = MergingOracle(scene, type(ctx.params)())
    for k, _ in type(ctx.params)._fields_:
        setattr(mo.params, k, getattr(ctx.params, k))
    mo.set_params()
    return ctx, mo


def prescribed(scene, t):
    """(body, position, velocity) of every kinematic body at time t, from
    fdlibm's sine and cosine as the kernel uses them."""
    lib = load_oracle()
    out = []
    for q in scene.kinematic_array():
        arg = float(q["omega"]) * t + float(q["phase"])
        ds = float(q["amplitude"]) * lib.orc_sin(arg)
        vs = float(q["amplitude"]) * float(q["omega"]) * lib.orc_cos(arg)
        x = np.array([float(q["x0"][a]) + ds * float(q["dir"][a]) for a in range(3)])
        v = np.array([vs * float(q["dir"][a]) for a in range(3)])
        out.append((int(q["body"]), x, v))
    return out


def step_both(ctx, mo, t, step, push=None):
    """One advanceTime on both, phase by phase (sleeping off)."""
    mo.total_steps += 1
    ctx.clear_bodies(); mo.clear_bodies()
    ctx.kinematic_update(t); mo.kinematic_update(t)
    if push:
        for body, point, force in push:
            ctx.impulse_add(body, point, force)
            mo.impulse_add(body, point, force)
    ctx.apply_gravity(); ctx.apply_spring_forces(); ctx.apply_impulses()
    mo.apply_external_forces()
    ctx.update_contacts_map(); mo.o.update_contacts_map()
    ctx.detect(DT); mo.o.detect()
    ctx.warm_start(); mo.o.warm_start()
    ctx.update_in_collections(DT)
    mo.update_in_collections(DT, ctx.solve_order)
    check_organized(ctx, mo)
    assert_contacts_equal(ctx.internal_contacts(), mo.internal_contacts())
    ctx.accumulate_for_unmerging(DT); mo.accumulate_for_unmerging(DT)
    cut_g = ctx.unmerge(DT)
    cut_o = mo.unmerge(DT)
    assert cut_g == cut_o, f"step {step}: unmerge {cut_g} vs {cut_o}"
    if mo.merging_event:
        ctx.clear_top_level(); mo.clear_top_level()
    ctx.redo_warm_start(); mo.o.redo_warm_start()
    it_g, _ = ctx.solve(DT)
    it_o, _, _ = mo.o.solve(ctx.solve_order(), DT)
    assert it_g == it_o
    ctx.clear_body_pair_contacts(); mo.clear_body_pair_contacts()
    ctx.advance_velocities(DT); mo.advance_velocities(DT)
    ctx.accumulate_for_merging(DT); mo.accumulate_for_merging(DT)
    ctx.advance_positions(DT); mo.advance_positions(DT)
    ext, _ = ctx.body_pairs()
    assert [(p["body1"], p["body2"], p["flag"], p["hist_n"]) for p in ext] == \
        [(r["b1"], r["b2"], int(r["flag"]), len(r["hist"])) for r in mo.bpc]
    merged = 0
    if step % 10 == 0:
        merged = ctx.merge()
        assert merged == mo.merge()
    ctx.apply_viscous_decay(); mo.viscous_decay()
    return merged, cut_g


def test_dominoes_phase_by_phase_match_oracle(gpu_lib):
    """Two platforms with two rows of six dominoes; the rows are pushed at
    step 15, topple onto each other and merge; every phase of 200 steps is
    compared bit for bit, and the platforms follow their prescribed motion."""
    scene = workloads.domino_platforms(seed=1, platforms=2, rows_per_platform=2, per_row=6)
    ctx, mo = make(scene)
    merges = 0
    for step in range(1, 201):
        t = (step - 1) * DT
        push = None
        if step == 15:
            push = workloads.domino_push(scene, ctx.pull_state()[0], DT)
        m, cut = step_both(ctx, mo, t, step, push)
        merges += m
        if m or cut:
            print(f"step {step}: merged {m} cut {cut} collections {len(mo.members)}", flush=True)
        compare_world(ctx, mo, step)
        x, _, v, _ = ctx.pull_state()
        # the kinematic update set these before the step; pinned bodies do not integrate
        for b, xp, vp in prescribed(scene, t):
            assert_bits_equal(f"step {step} platform {b} x", x[b], xp)
            assert_bits_equal(f"step {step} platform {b} v", v[b], vp)
    heads = scene.row_heads
    x, R, _, _ = ctx.pull_state()
    # every head was knocked over: its vertical axis tilted by more than 30 degrees
    for b in heads:
        assert R[b].reshape(3, 3)[1, 1] < math.cos(math.radians(30)), f"row head {b} still standing"
    assert merges >= 1


def test_dominoes_step_equals_phases(gpu_lib):
    """am_step (kinematic update, held forces, viscous decay, simulation time)
    runs exactly the phases above."""
    scene = workloads.domino_platforms(seed=4, platforms=2, rows_per_platform=2, per_row=5)
    extra = dict(viscous_linear_decay=0.999, viscous_angular_decay=0.998)
    a = Context(); a.upload(scene); a.set_params(**PARAMS, **extra)
    b = Context(); b.upload(scene); b.set_params(**PARAMS, **extra)
    for step in range(1, 36):
        t = b.sim_time
        assert a.sim_time == t
        if step == 12:
            push = workloads.domino_push(scene, a.pull_state()[0], DT)
            for body, point, force in push:
                a.impulse_add(body, point, force)
                b.impulse_add(body, point, force)
        a.step(DT)
        b.clear_bodies(); b.kinematic_update(t); b.apply_gravity(); b.apply_spring_forces(); b.apply_impulses()
        b.update_contacts_map(); b.detect(DT); b.warm_start(); b.wake()
        b.update_in_collections(DT); b.accumulate_for_unmerging(DT); b.unmerge(DT)
        if b.merging_event:
            b.clear_top_level()
        b.redo_warm_start(); b.solve(DT); b.clear_body_pair_contacts(); b.advance_velocities(DT)
        b.accumulate_for_merging(DT); b.advance_positions(DT)
        if step % 10 == 0:
            b.merge()
        b.sleep()
        b.apply_viscous_decay()
        b.sim_time = t + DT
        pa, ta = a.structure()
        pb, tb = b.structure()
        assert list(ta) == list(tb) and list(pa) == list(pb)
        n = len(pa)
        for name, x, y in zip("xRvw", a.pull_state(n), b.pull_state(n)):
            assert_bits_equal(f"step {step} {name}", x, y)
    assert a.sim_time == b.sim_time


def test_dominoes_full_size(gpu_lib):
    """2,000 dominoes in 20 rows on 10 platforms (50 sweeps, merging on at the
    reference defaults): the push at step 300 starts every row; 200 steps
    later the platforms have followed their motion exactly, every head has
    toppled, the contact count has stayed in the configuration's range and
    merges have happened."""
    scene = workloads.domino_platforms(seed=0)
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(**PARAMS)
    counts, colls = [], []
    for step in range(1, 501):
        if step == 300:
            for body, point, force in workloads.domino_push(scene, ctx.pull_state()[0], DT):
                ctx.impulse_add(body, point, force)
        st = ctx.step(DT)
        counts.append(st["n_contacts"] + st["n_internal_contacts"])
        colls.append(st["n_collections"])
        if step % 50 == 0:
            print(f"step {step}: contacts {st['n_contacts']} internal {st['n_internal_contacts']} "
                  f"collections {st['n_collections']} sweeps {st['iterations_used']} {st['ms_total']:.2f} ms",
                  flush=True)
    x, R, v, w = ctx.pull_state()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    t = ctx.sim_time
    for b, xp, vp in prescribed(scene, t - DT):
        assert_bits_equal(f"platform {b} x", x[b], xp)
    # standing dominoes touch their platform at 4 points: about 8,000 contacts before the push
    assert 7000 <= counts[298] <= 25000, counts[298]
    assert max(counts) <= 25000
    for b in scene.row_heads:
        assert R[b].reshape(3, 3)[1, 1] < math.cos(math.radians(30)), f"row head {b} still standing"
    assert max(colls) >= 1
