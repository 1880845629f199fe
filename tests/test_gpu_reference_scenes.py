"""Scenes of the reference (scenes3D: chariot, stackOnSwing, funnel, seesaw,
tower, venus; loaded by this package's XML loader into the fixtures of
tests/golden/make_scene_fixtures.py) stepped on the GPU with the reference's
default settings (gravity 1, dt 0.05, 30 PGS iterations, merging and sleeping
on) against the merging oracle, phase by phase and bit for bit: contacts,
single sweep, merge/unmerge decisions, sleeping flags and histories, state."""
import pytest

from adaptivemerging_amd.xml_scene import apply_collision_settings
from tests.helpers import load_scene_fixture
from tests.test_gpu_sleeping import make, run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["chariot", "stackOnSwing", "funnel", "seesaw", "tower", "venus"])
def test_reference_scene_matches_oracle(gpu_lib, name):
    scene, coll = load_scene_fixture(name)
    ctx, mo = make(scene)
    # the reference defaults, then the file's <collision> settings
    from adaptivemerging_amd._lib import default_params_host

    p = default_params_host()
    apply_collision_settings(coll, p)
    ctx.set_params(p)
    for k, _ in type(ctx.params)._fields_:
        setattr(mo.params, k, getattr(ctx.params, k))
    mo.set_params()
    print("reference scene", name, flush=True)
    log = run(ctx, mo, 40, dt=0.05)
    assert len(log) == 40
