"""Parity at the full sizes of the measured configurations: the 20,000-box
pile, the 33-layer jenga tower, 512 sphere-tree bodies of 512 leaves (dropped
without overlap, and placed interpenetrating for about a million contacts)
and the batched scenes.

The GPU runs the scene until it is in contact; its state then goes to the
oracle, both detect from it (the contacts must agree bit for bit), and both
step on from there with the oracle sweeping in the GPU's order, comparing
the contacts, lambdas and body states bit for bit.  Those steps use few
sweeps so that the CPU side stays within seconds.  On the pile the launch-shape
tuner then checks that every launch shape of the PGS kernel (both schedules)
gives the same bits on the full contact set."""
import numpy as np
import pytest

from adaptivemerging_amd import workloads
from adaptivemerging_amd.context import Context
from oracle.pyoracle import Oracle
from tests.helpers import assert_bits_equal, assert_contacts_equal, oracle_step_like_gpu

pytestmark = pytest.mark.gpu

GEOM_KEYS = ["body1", "body2", "bv1", "bv2", "info", "point_b1", "normal_b1", "tangent1_b1", "tangent2_b1",
             "violation", "jacobian"]


def settle(scene, steps, dt, **params):
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(**params)
    st = None
    for k in range(steps):
        st = ctx.step(dt)
        if k % 25 == 0 or k == steps - 1:
            print(f"settle step {k}: contacts {st['n_contacts']} sweeps {st['iterations_used']}", flush=True)
    return ctx, st


def sync_and_step(ctx, scene, dt, steps, **params):
    """Hands the GPU's state to a fresh oracle, detects on both from it, then
    steps both; returns the contact count of the last step."""
    ctx.set_params(**params)
    orc = Oracle(scene)
    for k, _ in type(ctx.params)._fields_:
        setattr(orc.params, k, getattr(ctx.params, k))
    orc.set_params()
    orc.set_state(*ctx.pull_state())
    # both maps start from the same contacts: detect once from the synced state
    ctx.update_contacts_map()
    n_g = ctx.detect(dt)
    n_o = orc.detect()
    print("synced detection:", n_g, "contacts on the GPU,", n_o, "in the oracle", flush=True)
    assert n_g == n_o > 0
    assert_contacts_equal(ctx.contacts(), orc.contacts(), keys=GEOM_KEYS)
    n = 0
    for step in range(steps):
        it_g, it_o = oracle_step_like_gpu(ctx, orc, dt)
        c = ctx.contacts()
        n = len(c["body1"])
        print(f"step {step}: {n} contacts, sweeps {it_g} / {it_o}", flush=True)
        assert it_g == it_o
        assert_contacts_equal(c, orc.contacts())
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
    return n


def test_pile_20k(gpu_lib):
    scene = workloads.box_pile(seed=0, n=20000)
    dt = 0.01
    ctx, st = settle(scene, 100, dt, iterations=100, tolerance=1e-6, gravity_amount=9.81, enable_merging=0,
                     enable_unmerging=0, enable_sleeping=0)
    assert st["n_contacts"] > 20000
    n = sync_and_step(ctx, scene, dt, 2, iterations=5, tolerance=0.0)
    assert n > 20000
    trials, best = ctx.autotune(dt, sweeps=5, reps=1, apply=False)
    print("launch shapes with identical bits:", len(trials), flush=True)
    assert {t["pgs_schedule"] for t in trials} == {0, 1} and len(trials) >= 10


def test_jenga_33_layers(gpu_lib):
    """The small plain configuration at its size: 100 boxes, mu 0.6, 100
    sweeps, every step bit for bit from the start."""
    scene = workloads.jenga_tower(seed=0)
    assert scene.n == 101
    from tests.helpers import twin

    ctx, orc = twin(scene, iterations=100, tolerance=1e-10, friction_override=1, friction=0.6, gravity_amount=9.81,
                    enable_merging=0, enable_unmerging=0, enable_sleeping=0)
    dt = 0.01
    for step in range(100):
        it_g, it_o = oracle_step_like_gpu(ctx, orc, dt)
        assert it_g == it_o
        if step % 20 == 19:
            print(f"jenga step {step}: {len(ctx.contacts()['body1'])} contacts, sweeps {it_g}", flush=True)
            assert_contacts_equal(ctx.contacts(), orc.contacts())
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
    assert len(ctx.contacts()["body1"]) > 300


def test_sphere_trees_512_interpenetrating(gpu_lib):
    """512 bodies of 512 leaves placed at random in the open box with no
    overlap check: their trees interpenetrate, about a million leaf contacts
    (the widest tree-tree detection load)."""
    scene = workloads.sphere_tree_drop(seed=0, n=512)
    dt = 0.005
    ctx, st = settle(scene, 4, dt, iterations=20, tolerance=1e-6, gravity_amount=9.81, enable_merging=0,
                     enable_unmerging=0, enable_sleeping=0)
    assert st["n_contacts"] > 100000
    sync_and_step(ctx, scene, dt, 1, iterations=5, tolerance=0.0)


def test_sphere_tree_pile_512(gpu_lib):
    """The sphere-tree configuration: 512 trees of 512 leaves spawned with no
    leaf overlap and dropped into the open box (dt 0.005, 100 sweeps); once
    the pile has formed (about 4,000 leaf contacts: the trees rest on a few
    leaves each), the GPU's state goes to the oracle and both step on bit for
    bit."""
    scene = workloads.sphere_tree_pile(seed=0, n=512)
    dt = 0.005
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(iterations=100, tolerance=1e-6, gravity_amount=9.81, friction_override=1, friction=0.4,
                   enable_merging=0, enable_unmerging=0, enable_sleeping=0)
    first = ctx.step(dt)
    assert first["n_contacts"] == 0, "the spawn must be contact free"
    st = first
    for k in range(1, 500):
        st = ctx.step(dt)
        if k % 100 == 0:
            print(f"pile step {k}: contacts {st['n_contacts']} sweeps {st['iterations_used']} "
                  f"{st['ms_total']:.2f} ms", flush=True)
    assert st["n_contacts"] > 2000
    n = sync_and_step(ctx, scene, dt, 1, iterations=5, tolerance=0.0)
    assert n > 2000


def test_batched_256_scenes(gpu_lib):
    """256 scenes of 50 to 2,000 bodies in one context (156k bodies).  After
    the GPU has piled them up, a sample of scenes (every 16th and the
    largest) goes to one oracle each; both detect from the GPU's state and
    take one step with the oracle sweeping in the GPU's order restricted to
    its scene.  Merging is off here (the merging oracle is too slow at this
    size); a fixed sweep count makes the batched solve comparable with one
    scene solved alone."""
    scenes = workloads.batched_pile_scenes(seed=0)
    scene = workloads.batch(scenes)
    offs = np.cumsum([0] + [s.n for s in scenes])
    params = dict(iterations=50, tolerance=0.0, gravity_amount=9.81, enable_merging=0, enable_unmerging=0,
                  enable_sleeping=0)
    dt = 0.01
    ctx, st = settle(scene, 60, dt, **params)
    assert st["n_contacts"] > 100000
    sample = sorted(set(range(0, len(scenes), 16)) | {int(np.argmax([s.n for s in scenes]))})
    x, R, v, w = ctx.pull_state()
    orcs = {}
    for k in sample:
        o = Oracle(scenes[k])
        o.set_params(**params)
        sl = slice(offs[k], offs[k + 1])
        o.set_state(x[sl], R[sl], v[sl], w[sl])
        orcs[k] = o

    def split(g):
        sc = np.searchsorted(offs, g["body1"], side="right") - 1
        return sc

    def scene_contacts(g, sc, k):
        mine = np.nonzero(sc == k)[0]
        gk = {key: val[mine] for key, val in g.items()}
        for key in ("body1", "body2"):
            gk[key] = gk[key] - offs[k]
        return gk

    ctx.update_contacts_map()
    ctx.detect(dt)
    g = ctx.contacts()
    sc = split(g)
    for k, o in orcs.items():
        o.detect()
        assert_contacts_equal(scene_contacts(g, sc, k), o.contacts(), keys=GEOM_KEYS)
    print("synced detection:", len(g["body1"]), "contacts; sampled scenes agree", flush=True)
    ctx.clear_bodies(); ctx.apply_gravity(); ctx.update_contacts_map(); ctx.detect(dt)
    ctx.warm_start(); ctx.redo_warm_start()
    it, _ = ctx.solve(dt)
    assert it == 50
    order = ctx.solve_order()
    g = ctx.contacts()
    sc = split(g)
    ctx.advance_velocities(dt); ctx.advance_positions(dt)
    x, R, v, w = ctx.pull_state()
    for k, o in orcs.items():
        o.clear_bodies(); o.apply_gravity(); o.update_contacts_map(); o.detect()
        o.warm_start(); o.redo_warm_start()
        mine = np.nonzero(sc == k)[0]
        local = -np.ones(len(sc), dtype=np.int64)
        local[mine] = np.arange(len(mine))
        o.solve(local[order[sc[order] == k]].astype(np.int32), dt)
        assert_contacts_equal(scene_contacts(g, sc, k), o.contacts())
        o.advance_velocities(dt); o.advance_positions(dt)
        sl = slice(offs[k], offs[k + 1])
        for name, a, b in zip("xRvw", (x[sl], R[sl], v[sl], w[sl]), o.pull_state()):
            assert_bits_equal(f"scene {k} {name}", a, b)
        print(f"scene {k}: {scenes[k].n} bodies, {len(mine)} contacts bit-exact", flush=True)
