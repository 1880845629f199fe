"""The dataflow schedule of the PGS sweeps (per-body update counters instead
of grid barriers) against the CPU oracle's sequential sweep in the exported
order: lambdas, deltaV, contact states and the per-sweep max |dl| bit-exact.

The schedule double-buffers lambda and deltaV by sweep parity and may run one
sweep past the one that converges, so the cases cover an early exit on odd and
even sweeps, an exit one sweep before the cap and on the cap itself, no exit,
and lanes that own several groups (small grids)."""
import numpy as np
import pytest

from adaptivemerging_amd import workloads
from tests.helpers import assert_bits_equal, assert_contacts_equal, twin

pytestmark = pytest.mark.gpu


def _solve_pair(scene, lam_seed, **params):
    ctx, orc = twin(scene, **params)
    ctx.apply_gravity(); orc.apply_gravity()
    n = ctx.detect(0.01); orc.detect()
    if lam_seed is not None:
        lam = np.random.default_rng(lam_seed).uniform(0, 0.5, size=(n, 3))
        ctx.push_lambdas(lam); orc.push_lambdas(lam)
    it_g, last_g = ctx.solve(0.01)
    order = ctx.solve_order()
    it_o, last_o, ch_o = orc.solve(order, 0.01)
    print("contacts", n, "sweeps", it_g, it_o, "params", params, flush=True)
    assert it_g == it_o
    assert_bits_equal("last_change", np.array([last_g]), np.array([last_o]))
    assert_contacts_equal(ctx.contacts(), orc.contacts())
    assert_bits_equal("deltaV", ctx.pull_deltav(), orc.pull_deltav())
    np.testing.assert_allclose(ctx.changes(), ch_o, rtol=1e-14, atol=0)
    ctx.close()
    return it_g


@pytest.mark.parametrize("grid", [0, 1, 3])
@pytest.mark.parametrize("iterations", [1, 2, 3, 8])
def test_flow_fixed_sweeps(gpu_lib, grid, iterations):
    scene = workloads.box_cluster(seed=11, n=512, spheres=40)
    _solve_pair(scene, 3, iterations=iterations, tolerance=0.0, pgs_grid=grid, pgs_schedule=0)


def test_flow_early_exit_around_the_cap(gpu_lib):
    scene = workloads.jenga_tower(seed=12, layers=8, overlap=0.002)
    kw = dict(tolerance=3e-3, friction_override=1, friction=0.6, pgs_schedule=0)
    t = _solve_pair(scene, None, iterations=1000, **kw)
    assert 4 < t < 1000, f"expected an early exit, got {t} sweeps"
    for cap in (t - 2, t - 1, t, t + 1, t + 4):
        got = _solve_pair(scene, None, iterations=cap, **kw)
        assert got == min(cap, t)


@pytest.mark.parametrize("tolerance", [0.5, 0.3, 0.2, 0.12])
def test_flow_early_exit_odd_and_even(gpu_lib, tolerance):
    scene = workloads.box_cluster(seed=13, n=343, spheres=30)
    _solve_pair(scene, 5, iterations=500, tolerance=tolerance, pgs_schedule=0, pgs_grid=2)


def test_flow_equals_coloured_schedule(gpu_lib):
    from adaptivemerging_amd.context import Context

    scene = workloads.box_cluster(seed=14, n=729, spheres=60)
    out = []
    for schedule in (0, 1):
        ctx = Context()
        ctx.upload(scene)
        ctx.set_params(iterations=30, tolerance=1e-6, gravity_amount=9.81, pgs_schedule=schedule)
        for _ in range(3):
            ctx.step(0.01)
        out.append((ctx.pull_state(), ctx.contacts()["lambda"]))
        ctx.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert_bits_equal("state", a, b)
    assert_bits_equal("lambda", out[0][1], out[1][1])


def test_flow_groups_beyond_the_lds_slots(gpu_lib):
    """Sphere trees resting on planes give body pairs with hundreds of
    contacts: the records past the kernel's LDS slots come from global memory."""
    import collections

    scene = workloads.sphere_tree_drop(seed=5, n=16, heights=(0.15, 0.9), box=(1.2, 1.0, 1.2))
    ctx, orc = twin(scene, iterations=12, tolerance=0.0, gravity_amount=9.81, pgs_schedule=0)
    ctx.apply_gravity(); orc.apply_gravity()
    n = ctx.detect(0.01); orc.detect()
    c = ctx.contacts()
    sizes = collections.Counter(zip(np.minimum(c["body1"], c["body2"]), np.maximum(c["body1"], c["body2"])))
    print("contacts", n, "largest group", max(sizes.values()), flush=True)
    assert max(sizes.values()) > 8
    it_g, _ = ctx.solve(0.01)
    order = ctx.solve_order()
    it_o, _, _ = orc.solve(order, 0.01)
    assert it_g == it_o
    assert_contacts_equal(ctx.contacts(), orc.contacts())
    assert_bits_equal("deltaV", ctx.pull_deltav(), orc.pull_deltav())


@pytest.mark.parametrize("schedule", [0, 1])
def test_flow_unclamped_contacts_take_the_general_update(gpu_lib, schedule):
    """Contacts of magnet-active bodies are not clamped (PGS.java:119-120).
    The dataflow kernel's fast update covers clamped contacts only, so every
    step holding one runs again with the general update; results stay
    bit-exact against the oracle's sequential sweep, at every sweep."""
    from adaptivemerging_amd._lib import AM_BODY_MAGNET_ACTIVE

    scene = workloads.jenga_tower(seed=3, layers=8, overlap=0.002)
    dyn = [i for i in range(scene.n) if not int(scene.bodies[i]["flags"]) & 1]
    for i in dyn[::5]:
        scene.bodies[i]["flags"] = int(scene.bodies[i]["flags"]) | AM_BODY_MAGNET_ACTIVE
    for iterations in (1, 2, 9):
        _solve_pair(scene, 7, iterations=iterations, tolerance=0.0, gravity_amount=9.81, pgs_schedule=schedule)
