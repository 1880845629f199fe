"""Sleeping on the GPU (am_wake / am_sleep, after Sleeping.java:49-160 and its
call sites) against the merging oracle, phase by phase and bit for bit: the
sleeping flags and metric histories of every entity, the state, the
structure and the contacts, without merging and with it; am_step with
sleeping against the same phases."""
import numpy as np
import pytest

from adaptivemerging_amd.context import Context
from oracle.merging import MergingOracle
from tests.helpers import assert_bits_equal, assert_contacts_equal
from tests.test_gpu_merging import compare_world
from tests.test_sleeping_oracle import resting_scene

pytestmark = pytest.mark.gpu

BASE = dict(gravity_amount=9.81, iterations=100, tolerance=1e-10, sleep_threshold=1e-3)


def make(scene, **kw):
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(**{**BASE, **kw})
    mo = MergingOracle(scene, type(ctx.params)())
    for k, _ in type(ctx.params)._fields_:
        setattr(mo.params, k, getattr(ctx.params, k))
    mo.set_params()
    return ctx, mo


def compare_sleep(ctx, mo, step):
    n = mo.n_entities
    flags, hn, hist = ctx.sleep_state(n)
    want = np.array([mo.sleeping(e) for e in range(n)], dtype=np.uint8)
    assert np.array_equal(flags, want), f"step {step}: sleeping flags {flags} vs {want}"
    for e in range(n):
        h = mo.sleep_hist.get(e, [])
        assert hn[e] == len(h), f"step {step}: entity {e} history length {hn[e]} vs {len(h)}"
        if h:
            assert_bits_equal(f"step {step} entity {e} history", hist[e, :hn[e]], np.array(h))


def step_phases(ctx, mo, dt, step):
    """One advanceTime on both, phase by phase, with the oracle sweeping in the
    GPU's orders.  Returns (merged, cut, woken, asleep) of the GPU."""
    mo.total_steps += 1
    ctx.clear_bodies(); mo.clear_bodies()
    ctx.apply_gravity(); mo.o.apply_gravity()
    ctx.apply_spring_forces(); mo.o.apply_springs()
    ctx.update_contacts_map(); mo.o.update_contacts_map()
    ctx.detect(dt); mo.o.detect()
    assert_contacts_equal(ctx.contacts(), mo.o.contacts())
    ctx.warm_start(); mo.o.warm_start()
    woken = ctx.wake()
    assert woken == mo.wake(), f"step {step}: woken"
    compare_sleep(ctx, mo, step)
    ctx.update_in_collections(dt)
    mo.update_in_collections(dt, ctx.solve_order)
    assert_contacts_equal(ctx.internal_contacts(), mo.internal_contacts())
    ctx.accumulate_for_unmerging(dt); mo.accumulate_for_unmerging(dt)
    cut = ctx.unmerge(dt)
    assert cut == mo.unmerge(dt), f"step {step}: unmerge"
    if mo.merging_event:
        ctx.clear_top_level(); mo.clear_top_level()
    ctx.redo_warm_start(); mo.o.redo_warm_start()
    it_g, _ = ctx.solve(dt)
    it_o, _, _ = mo.o.solve(ctx.solve_order(), dt)
    assert it_g == it_o
    ctx.clear_body_pair_contacts(); mo.clear_body_pair_contacts()
    ctx.advance_velocities(dt); mo.advance_velocities(dt)
    ctx.accumulate_for_merging(dt); mo.accumulate_for_merging(dt)
    ctx.advance_positions(dt); mo.advance_positions(dt)
    merged = 0
    if step % ctx.params.steps_between_merge_events == 0:
        merged = ctx.merge()
        assert merged == mo.merge(), f"step {step}: merge"
    asleep = ctx.sleep()
    assert asleep == mo.sleep(), f"step {step}: asleep"
    compare_sleep(ctx, mo, step)
    return merged, cut, woken, asleep


def run(ctx, mo, steps, dt=0.01):
    log = []
    for step in range(1, steps + 1):
        print("sleeping step", step, flush=True)
        out = step_phases(ctx, mo, dt, step)
        compare_world(ctx, mo, step)
        log.append(out)
    return log


def test_falling_box_wakes_sleeping_box(gpu_lib):
    ctx, mo = make(resting_scene(), enable_merging=0, enable_unmerging=0)
    log = run(ctx, mo, 50)
    assert max(a for _, _, _, a in log) == 4 and any(w for _, _, w, _ in log)


def test_spring_wakes_sleeping_end(gpu_lib):
    ctx, mo = make(resting_scene(spring=True), enable_merging=0, enable_unmerging=0)
    log = run(ctx, mo, 30)
    assert sum(w for _, _, w, _ in log) >= 2


def test_sleeping_with_merging(gpu_lib):
    """Resting boxes merge with the plane, the collection falls asleep, a
    falling box wakes it, a member is cut loose and merges again."""
    ctx, mo = make(resting_scene())
    log = run(ctx, mo, 70)
    assert sum(m for m, _, _, _ in log) >= 3 and sum(c for _, c, _, _ in log) >= 1
    assert any(a for _, _, _, a in log) and any(w for _, _, w, _ in log)


def test_collection_born_asleep(gpu_lib):
    ctx, mo = make(resting_scene(falling=False), steps_between_merge_events=20, sleep_threshold=1e-9)
    log = run(ctx, mo, 25)
    assert log[19][0] == 3
    parent, top = ctx.structure()
    (ce,) = [e for e in top if e >= mo.n_phys]
    flags, _, _ = ctx.sleep_state()
    assert flags[ce] == 1


def test_am_step_runs_the_sleeping_phases(gpu_lib):
    scene = resting_scene()
    a = Context(); a.upload(scene); a.set_params(**BASE)
    b = Context(); b.upload(scene); b.set_params(**BASE)
    dt = 0.01
    for step in range(1, 61):
        print("am_step with sleeping, step", step, flush=True)
        st = a.step(dt)
        b.clear_bodies(); b.apply_gravity(); b.apply_spring_forces()
        b.update_contacts_map(); b.detect(dt); b.warm_start(); b.wake()
        b.update_in_collections(dt); b.accumulate_for_unmerging(dt); b.unmerge(dt)
        if b.merging_event:
            b.clear_top_level()
        b.redo_warm_start(); b.solve(dt); b.clear_body_pair_contacts(); b.advance_velocities(dt)
        b.accumulate_for_merging(dt); b.advance_positions(dt)
        if step % 10 == 0:
            b.merge()
        asleep = b.sleep()
        assert st["n_sleeping"] == asleep
        pa, ta = a.structure()
        pb, tb = b.structure()
        assert list(ta) == list(tb) and list(pa) == list(pb)
        n = len(pa)
        for name, x, y in zip("xRvw", a.pull_state(n), b.pull_state(n)):
            assert_bits_equal(f"step {step} {name}", x, y)
        fa, na, ha = a.sleep_state()
        fb, nb, hb = b.sleep_state()
        assert_bits_equal(f"step {step} flags", fa, fb)
        assert_bits_equal(f"step {step} history lengths", na, nb)
        for e in range(len(na)):
            assert_bits_equal(f"step {step} entity {e} history", ha[e, :na[e]], hb[e, :nb[e]])


def test_wake_all(gpu_lib):
    ctx, mo = make(resting_scene(falling=False), enable_merging=0, enable_unmerging=0)
    run(ctx, mo, 12)
    flags, hn, _ = ctx.sleep_state()
    assert flags.any()
    ctx.wake_all(); mo.wake_all()
    compare_sleep(ctx, mo, 13)
    flags, hn, _ = ctx.sleep_state()
    assert not flags.any()


def test_facade_advance_time_equals_am_step(gpu_lib):
    """RigidBodySystem.advanceTime (the Python mirror of the Java surface, with
    its Sleeping settings) gives the same steps as am_step."""
    from adaptivemerging_amd.system import RigidBodySystem

    scene = resting_scene()
    system = RigidBodySystem(scene)
    system.gravityAmount = 9.81
    system.collision.iterations = 100
    system.collision.tolerance = 1e-10
    system.sleeping.threshold = 1e-3
    a = Context(); a.upload(scene); a.set_params(**BASE)
    for step in range(1, 61):
        print("facade with sleeping, step", step, flush=True)
        system.advanceTime(0.01)
        a.step(0.01)
        pa, _ = a.structure()
        n = len(pa)
        for name, x, y in zip("xRvw", a.pull_state(n), system.ctx.pull_state(n)):
            assert_bits_equal(f"step {step} {name}", x, y)
        assert_bits_equal(f"step {step} flags", a.sleep_state()[0], system.ctx.sleep_state()[0])
