"""Composite bodies on the GPU (sub-body shapes in the narrow phase, their
sphere nodes' world centres, contacts on the composite with the sub-body in
the bv id) against the oracle, bit for bit over whole steps; also loaded
from scene XML, and merged into collections."""
import numpy as np
import pytest

from adaptivemerging_amd import xml_scene
from adaptivemerging_amd.context import Context
from adaptivemerging_amd.scene import Scene, random_rotation, rot_axis_angle
from oracle.merging import MergingOracle
from tests.helpers import assert_bits_equal, assert_contacts_equal, oracle_step_like_gpu, twin

pytestmark = pytest.mark.gpu


def table(s, x, R=None):
    legs = [dict(kind="box", size=[0.06, 0.4, 0.06], x=[sx * 0.3, 0.2, sz * 0.2], density=600)
            for sx in (-1, 1) for sz in (-1, 1)]
    return s.add_composite([dict(kind="box", size=[0.8, 0.05, 0.55], x=[0, 0.425, 0], density=600)] + legs, x=x, R=R)


def dumbbell(s, x, R=None):
    return s.add_composite([dict(kind="sphere", r=0.12, x=[-0.25, 0, 0], density=2000),
                            dict(kind="box", size=[0.4, 0.05, 0.05], x=[0, 0, 0], density=2000),
                            dict(kind="sphere", r=0.12, x=[0.25, 0, 0], density=2000)], x=x, R=R)


def composite_scene(seed=0):
    rng = np.random.default_rng(seed)
    s = Scene()
    s.add_plane([0, 0, 0], [0, 1, 0])
    table(s, [0, 0.001, 0])
    table(s, [1.2, 0.3, 0.1], R=rot_axis_angle([0.2, 1, 0.1], 0.5))
    for k in range(4):
        dumbbell(s, [rng.uniform(-0.3, 0.3), 0.7 + 0.3 * k, rng.uniform(-0.2, 0.2)], R=random_rotation(rng))
    s.add_box([0.2, 0.2, 0.2], [1.2, 1.2, 0.1], R=random_rotation(rng))
    s.add_sphere(0.15, [-0.1, 2.2, 0.0])
    return s


def test_composites_step_like_the_oracle(gpu_lib):
    ctx, orc = twin(composite_scene(), gravity_amount=9.81, iterations=60, enable_sleeping=0)
    total = 0
    for step in range(50):
        print("composite step", step, flush=True)
        it_g, it_o = oracle_step_like_gpu(ctx, orc, 0.01)
        assert it_g == it_o
        c = ctx.contacts()
        assert_contacts_equal(c, orc.contacts())
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
        total += int(np.sum(c["bv1"] <= -3) + np.sum(c["bv2"] <= -3))
    assert total > 0, "no contact on a composite's sub-body"


def test_composite_xml_scene_steps_like_the_oracle(gpu_lib):
    text = """<root>
      <body type="plane" name="ground" p="0 0 0" n="0 1 0"/>
      <body type="composite" name="chair"><x>0 0.05 0</x>
        <body type="box" name="seat" dim="0.5 0.06 0.5" density="500"><x>0 0.45 0</x></body>
        <body type="box" name="back" dim="0.5 0.5 0.06" density="500"><x>0 0.7 -0.22</x><R>1 0 0 0.1</R></body>
        <body type="box" name="l1" dim="0.05 0.45 0.05" density="500"><x>0.2 0.22 0.2</x></body>
        <body type="box" name="l2" dim="0.05 0.45 0.05" density="500"><x>-0.2 0.22 0.2</x></body>
        <body type="box" name="l3" dim="0.05 0.45 0.05" density="500"><x>0.2 0.22 -0.2</x></body>
        <body type="box" name="l4" dim="0.05 0.45 0.05" density="500"><x>-0.2 0.22 -0.2</x></body>
        <body type="sphere" name="knob" r="0.05"><x>0 0.98 -0.22</x></body>
      </body>
      <body type="box" name="crate" dim="0.3 0.3 0.3" density="300"><x>0.05 1.0 0</x></body>
    </root>"""
    scene, coll = xml_scene.parse_scene(text)
    assert len(scene.subs) == 7
    ctx, orc = twin(scene, gravity_amount=9.81, iterations=60, enable_sleeping=0)
    for step in range(40):
        print("xml composite step", step, flush=True)
        oracle_step_like_gpu(ctx, orc, 0.01)
        assert_contacts_equal(ctx.contacts(), orc.contacts())
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)


def test_composites_merge_like_the_oracle(gpu_lib):
    """Tables resting on the ground and a box on a table merge into
    collections and the run matches the merging oracle phase by phase."""
    from tests.test_gpu_merging import PARAMS, compare_world, step_both

    s = Scene()
    s.add_plane([0, 0, 0], [0, 1, 0])
    table(s, [0, 0.0005, 0])
    s.add_box([0.2, 0.2, 0.2], [0.05, 0.5505, 0.0], density=300)
    table(s, [1.5, 0.0005, 0.0])
    ctx = Context()
    ctx.upload(s)
    ctx.set_params(**PARAMS, enable_sleeping=0)
    mo = MergingOracle(s, type(ctx.params)())
    for k, _ in type(ctx.params)._fields_:
        setattr(mo.params, k, getattr(ctx.params, k))
    mo.set_params()
    merges = 0
    for step in range(1, 26):
        print("composite merging step", step, flush=True)
        m, _ = step_both(ctx, mo, 0.01, step)
        merges += m
        compare_world(ctx, mo, step)
    assert merges >= 1
