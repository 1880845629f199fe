"""The Python mirror of the Java surface (RigidBodySystem / CollisionProcessor
/ PGS) drives libam and reproduces the oracle's step."""
import numpy as np
import pytest

from adaptivemerging_amd import workloads
from adaptivemerging_amd.system import RigidBodySystem
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def test_system_steps_match_oracle_and_tower_stands(gpu_lib):
    scene = workloads.jenga_tower(seed=0, layers=6)
    system = RigidBodySystem(scene)
    system.gravityAmount = 9.81
    system.collision.iterations = 60
    system.collision.frictionOverride = True
    system.collision.friction = 0.6
    system.merging.enableMerging = False     # the oracle steps below do not merge
    system.merging.enableUnmerging = False
    orc = Oracle(scene)
    orc.set_params(gravity_amount=9.81, iterations=60, friction_override=1, friction=0.6)
    for step in range(20):
        system.advanceTime(0.01)
        orc.clear_bodies(); orc.apply_gravity(); orc.update_contacts_map(); orc.detect()
        orc.warm_start(); orc.redo_warm_start()
        order = system.ctx.solve_order()
        orc.solve(order, 0.01)
        orc.advance_velocities(0.01); orc.advance_positions(0.01)
    st = system.state()
    xo, _, _, _ = orc.pull_state()
    np.testing.assert_allclose(st["x"], xo, atol=1e-9)
    assert len(system.collision.contacts) > 0
    # the tower is standing: the top block settles into its support by a few
    # centimetres (penetration under Baumgarte feedback) but does not topple
    assert abs(st["x"][-1][1] - scene.x[-1][1]) < 0.05
    assert np.hypot(st["x"][-1][0] - scene.x[-1][0], st["x"][-1][2] - scene.x[-1][2]) < 0.01


def test_reset_restores_the_initial_scene(gpu_lib):
    """RigidBodySystem.reset (RigidBodySystem.java:389-425): collections
    dissolve, bodies return to their initial state, springs reset, and the
    run after a reset repeats the first one bit for bit."""
    scene = workloads.merge_stack(seed=6)
    scene.add_spring(scene.n - 2, [0.3, 0.1, 0.0], k=30.0, d=2.0)
    system = RigidBodySystem(scene)
    system.gravityAmount = 9.81
    system.collision.iterations = 60

    def run():
        out = []
        for _ in range(25):
            system.advanceTime(0.01)
            out.append(system.ctx.pull_state())
        return out, system.ctx.structure()

    first, (parent1, top1) = run()
    assert any(p >= 0 for p in parent1[:scene.n]), "the boards should have merged"
    system.reset()
    parent, top = system.ctx.structure()
    assert list(top) == list(range(scene.n)) and all(p < 0 for p in parent[:scene.n])
    x, R, v, w = system.ctx.pull_state()
    np.testing.assert_array_equal(x, np.array(scene.x))
    second, (parent2, top2) = run()
    assert list(parent1) == list(parent2) and list(top1) == list(top2)
    for k, (a, b) in enumerate(zip(first, second)):
        for name, p, q in zip("xRvw", a, b):
            np.testing.assert_array_equal(p, q, err_msg=f"step {k} {name}")


def test_csv_export_of_the_step_metrics(gpu_lib, tmp_path):
    """saveCSV on: one header, then a line per step after the first, with the
    top-level body count and the contact count of each step."""
    from adaptivemerging_amd.csv_export import COLUMNS, read_csv

    scene = workloads.merge_stack(seed=2)
    system = RigidBodySystem(scene)
    system.gravityAmount = 9.81
    system.sceneName = str(tmp_path / "stack")
    system.saveCSV = True
    counts = []
    for _ in range(12):
        system.advanceTime(0.01)
        counts.append((len(system.ctx.structure()[1]), system.ctx.n_contacts()))
    system.saveCSV = False
    system.advanceTime(0.01)
    rows = read_csv(str(tmp_path / "stack_merged.csv"))
    assert len(rows) == 11 and all(list(r) == list(COLUMNS) for r in rows)
    assert [(r["#bodies"], r["#contacts"]) for r in rows] == counts[1:]
    assert all(r["computeTime"] > 0 for r in rows)
