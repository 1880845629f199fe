"""Mesh bodies (XMLParser.createMesh: sphere-tree collision geometry and a
full rest-pose inertia tensor from the OBJ mesh) on the GPU against the
oracle, bit for bit over whole steps."""
import pytest

from adaptivemerging_amd import xml_scene
from tests.helpers import assert_bits_equal, assert_contacts_equal, oracle_step_like_gpu, twin
from tests.test_mesh import write_mesh_files

pytestmark = pytest.mark.gpu


def test_mesh_bodies_step_like_the_oracle(gpu_lib, tmp_path):
    write_mesh_files(tmp_path)
    text = """<root>
      <body type="plane" name="ground" p="0 0 0" n="0 1 0"/>
      <body type="mesh" name="m1" scale="1.5" density="300" obj="block.obj" st="block.sph">
        <x>0 0.5 0</x><R>1 1 0 0.4</R><omega>0.5 1 0</omega></body>
      <body type="mesh" name="m2" scale="1.0" density="800" obj="block.obj" st="block.sph">
        <x>0.3 1.4 0.1</x><R>0 1 1 1.1</R></body>
      <body type="box" name="b" dim="0.3 0.3 0.3" density="500"><x>-0.8 0.3 0</x></body>
    </root>"""
    scene, _ = xml_scene.parse_scene(text, data_root=str(tmp_path))
    ctx, orc = twin(scene, gravity_amount=9.81, iterations=80)
    n_contacts = 0
    for step in range(40):
        print("mesh step", step, flush=True)
        it_g, it_o = oracle_step_like_gpu(ctx, orc, 0.01)
        assert it_g == it_o
        for name, a, b in zip("xRvw", ctx.pull_state(), orc.pull_state()):
            assert_bits_equal(f"step {step} {name}", a, b)
        c = ctx.contacts()
        assert_contacts_equal(c, orc.contacts())
        n_contacts += len(c["body1"])
    assert n_contacts > 0
