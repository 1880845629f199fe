"""Post-stabilization (RigidBodySystem.postStabilization, :354-378): after
the positions move, a second detection, a warm start with the post flag, a
solve with the post flag, and a position update driven by deltaV
(am_advance_positions_post).  Phase by phase against the oracle, bit for bit;
am_step with enable_post_stabilization runs the same phases."""
import pytest

from adaptivemerging_amd import _lib, workloads
from adaptivemerging_amd.context import Context
from tests.helpers import assert_bits_equal, assert_contacts_equal, oracle_step_like_gpu, twin

pytestmark = pytest.mark.gpu

PARAMS = dict(iterations=50, tolerance=1e-10, gravity_amount=9.81, enable_post_stabilization=1, enable_merging=0,
              enable_unmerging=0, enable_sleeping=0)


def post_pass(ctx, orc, dt):
    ctx.clear_bodies(); orc.clear_bodies()
    ctx.update_contacts_map(); orc.update_contacts_map()
    ctx.detect(dt); orc.detect()
    ctx.warm_start(True); orc.warm_start(True)
    it_g, _ = ctx.solve(dt, True)
    it_o, _, _ = orc.solve(ctx.solve_order(), dt, post=True)
    ctx.advance_positions_post(dt); orc.advance_positions_post(dt)
    return it_g, it_o


def test_post_stabilization_matches_oracle(gpu_lib):
    scene = workloads.jenga_tower(seed=3, layers=8, overlap=0.004)
    ctx, orc = twin(scene, **PARAMS)
    dt = 0.01
    for step in range(20):
        print("post-stabilization step", step, flush=True)
        oracle_step_like_gpu(ctx, orc, dt)
        it_g, it_o = post_pass(ctx, orc, dt)
        assert it_g == it_o
        assert_contacts_equal(ctx.contacts(), orc.contacts())
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
    assert len(ctx.contacts()["body1"]) > 0


def test_am_step_runs_post_stabilization(gpu_lib):
    scene = workloads.box_cluster(seed=31, n=125, spheres=12)
    a = Context(); a.upload(scene); a.set_params(**PARAMS)
    b = Context(); b.upload(scene); b.set_params(**PARAMS)
    dt = 0.01
    for step in range(10):
        print("am_step with post-stabilization, step", step, flush=True)
        a.step(dt)
        b.clear_bodies(); b.apply_gravity(); b.apply_spring_forces(); b.update_contacts_map(); b.detect(dt)
        b.warm_start(); b.redo_warm_start(); b.solve(dt); b.advance_velocities(dt); b.advance_positions(dt)
        b.clear_bodies(); b.update_contacts_map(); b.detect(dt); b.warm_start(True); b.solve(dt, True)
        b.advance_positions_post(dt)
        for name, x, y in zip("xRvw", a.pull_state(), b.pull_state()):
            assert_bits_equal(f"step {step} {name}", x, y)


def test_am_step_refuses_post_stabilization_with_merging(gpu_lib):
    ctx = Context(); ctx.upload(workloads.merge_stack(seed=1))
    ctx.set_params(enable_post_stabilization=1)
    with pytest.raises(_lib.AMError) as e:
        ctx.step(0.01)
    assert e.value.code == _lib.AM_ERR_INVALID


def test_facade_post_stabilization_equals_am_step(gpu_lib):
    """RigidBodySystem.advanceTime with collision.enablePostStabilization
    gives the same steps as am_step with enable_post_stabilization."""
    from adaptivemerging_amd.system import RigidBodySystem

    scene = workloads.box_cluster(seed=32, n=64, spheres=6)
    system = RigidBodySystem(scene)
    system.gravityAmount = 9.81
    system.collision.iterations = 50
    system.collision.tolerance = 1e-10
    system.collision.enablePostStabilization = True
    system.merging.enableMerging = False
    system.merging.enableUnmerging = False
    system.sleeping.enableSleeping = False
    a = Context(); a.upload(scene); a.set_params(**PARAMS)
    for step in range(8):
        print("facade post-stabilization step", step, flush=True)
        system.advanceTime(0.01)
        a.step(0.01)
        for name, x, y in zip("xRvw", a.pull_state(), system.ctx.pull_state()):
            assert_bits_equal(f"step {step} {name}", x, y)
