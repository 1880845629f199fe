"""Springs on the GPU (k_springs after RigidBodySystem.applySpringForces,
RigidBodySystem.java:309-322, and Spring.java:139-209) against the oracle,
bit for bit: the accumulated forces and torques, whole steps of a scene whose
contacts and springs interact, springs on collection members (force to the
member and to its collection), and am_step's external-force phase."""
import numpy as np
import pytest

from adaptivemerging_amd import workloads, xml_scene
from adaptivemerging_amd.context import Context
from adaptivemerging_amd.scene import random_rotation
from oracle.merging import MergingOracle
from tests.helpers import assert_bits_equal, assert_contacts_equal, oracle_step_like_gpu, twin

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mods", [(1.0, 1.0), (2.5, 0.3), (0.0, 1.0)])
def test_spring_forces_match_oracle(gpu_lib, mods):
    scene = workloads.hung_platform(seed=1)
    ctx, orc = twin(scene, gravity_amount=9.81, spring_stiffness_mod=mods[0], spring_damping_mod=mods[1])
    rng = np.random.default_rng(7)
    n = scene.n
    x = np.array(scene.x) + rng.normal(scale=0.1, size=(n, 3))
    R = np.stack([random_rotation(rng).reshape(9) for _ in range(n)])
    v = rng.normal(size=(n, 3))
    w = rng.normal(size=(n, 3))
    ctx.push_state(x, R, v, w)
    orc.set_state(x, R, v, w)
    print("spring forces, mods", mods, flush=True)
    ctx.clear_bodies(); orc.clear_bodies()
    ctx.apply_gravity(); orc.apply_gravity()
    ctx.apply_spring_forces(); orc.apply_springs()
    fg, tg = ctx.pull_forces()
    fo, to = orc.pull_forces()
    assert_bits_equal("force", fg, fo[:len(fg)])
    assert_bits_equal("torque", tg, to[:len(tg)])
    assert np.any(tg != 0.0)


def test_hung_platform_steps_match_oracle(gpu_lib):
    scene = workloads.hung_platform(seed=0)
    ctx, orc = twin(scene, gravity_amount=9.81, iterations=100)
    dt = 0.01
    for step in range(40):
        print("hung platform step", step, flush=True)
        it_g, it_o = oracle_step_like_gpu(ctx, orc, dt)
        assert it_g == it_o
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
    c = ctx.contacts()
    assert_contacts_equal(c, orc.contacts())
    assert len(c["body1"]) > 0


def test_am_step_applies_springs(gpu_lib):
    """am_step's external forces are gravity and then the springs: it equals
    the phase-by-phase step above."""
    scene = workloads.hung_platform(seed=2)
    a = Context(); a.upload(scene); a.set_params(gravity_amount=9.81, iterations=80)
    ctx, orc = twin(scene, gravity_amount=9.81, iterations=80)
    dt = 0.01
    for step in range(30):
        print("am_step with springs, step", step, flush=True)
        a.step(dt)
        oracle_step_like_gpu(ctx, orc, dt)
        for name, p, q in zip("xRvw", a.pull_state(), ctx.pull_state()):
            assert_bits_equal(f"step {step} {name}", p, q)
    assert len(a.contacts()["body1"]) > 0


def test_springs_on_collection_members(gpu_lib):
    """Boards with springs settle and merge; a spring on a member adds its
    force to the member and to the collection (Spring.applyOneBody and
    applyTwoBodies), and the merged run matches the oracle step by step."""
    from tests.test_gpu_merging import PARAMS, compare_world, step_both

    scene = workloads.merge_stack(seed=5)
    top = scene.n - 2   # the highest board
    scene.add_spring(top, [0.4, 0.1, 0.0], k=50.0, d=5.0)
    scene.add_spring(top, [-0.4, 0.1, 0.2], pW=[-0.4, 1.5, 0.2], k=20.0, d=1.0)
    scene.add_spring(1, [0.5, 0.0, 0.0], body2=scene.n - 1, pB2=[-0.15, 0.0, 0.0], k=10.0, d=1.0)
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(**PARAMS)
    mo = MergingOracle(scene, type(ctx.params)())
    for k, _ in type(ctx.params)._fields_:
        setattr(mo.params, k, getattr(ctx.params, k))
    mo.set_params()
    dt = 0.01
    merges = 0
    for step in range(1, 31):
        print("merging with springs, step", step, flush=True)
        m, _ = step_both(ctx, mo, dt, step)
        merges += m
        compare_world(ctx, mo, step)
    assert merges >= 1
    parent, _ = ctx.structure()
    assert parent[top] >= 0


def test_spring_scene_from_xml_steps_match_oracle(gpu_lib):
    """A scene file in the reference's format with world, zero and body-body
    springs loads with its springs and steps like the oracle."""
    text = """<root>
      <collision iterations="60"/>
      <body type="plane" name="ground" p="0 0 0" n="0 1 0"/>
      <body type="box" name="board" dim="2 0.2 1" density="200">
        <x>0 1.0 0</x>
        <spring pB="1 0.1 0.5" pW="1 3 0.5" k="3000" d="100"/>
        <spring pB="-1 0.1 -0.5" k="3000" d="100"/>
      </body>
      <body type="box" name="cube" dim="0.3 0.3 0.3" density="500"><x>0.2 1.26 0</x></body>
      <body type="sphere" name="ball" r="0.2"><x>1.8 1.0 0</x>
        <spring pB="0 0.2 0" pB2="1 0 0" body2="board" k="400" d="4" ls="0.8"/>
      </body>
    </root>"""
    scene, coll = xml_scene.parse_scene(text)
    assert len(scene.springs) == 3
    ctx, orc = twin(scene, gravity_amount=9.81, iterations=coll["iterations"])
    dt = 0.01
    for step in range(30):
        print("xml spring scene step", step, flush=True)
        oracle_step_like_gpu(ctx, orc, dt)
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
