"""The launch-shape autotuner (am_autotune): every candidate shape of the PGS
kernel gives the same bits (the call checks this itself and fails otherwise),
the contacts, lambdas and deltaV it found are put back, the tuned context
then solves and steps bit for bit like the oracle and like an untuned
context, and the call refuses an empty contact set."""
import numpy as np
import pytest

from adaptivemerging_amd import _lib, workloads
from adaptivemerging_amd.context import Context
from tests.helpers import assert_bits_equal, assert_contacts_equal, twin

pytestmark = pytest.mark.gpu


def _prepared(scene, **params):
    ctx, orc = twin(scene, **params)
    ctx.apply_gravity(); orc.apply_gravity()
    n = ctx.detect(0.01); orc.detect()
    ctx.warm_start(); orc.warm_start()
    return ctx, orc, n


def test_autotune_restores_and_solves_like_the_oracle(gpu_lib):
    scene = workloads.box_cluster(seed=21, n=729, spheres=60)
    ctx, orc, n = _prepared(scene, iterations=40, tolerance=1e-8, gravity_amount=9.81)
    lam = np.random.default_rng(4).uniform(0, 0.5, size=(n, 3))
    ctx.push_lambdas(lam); orc.push_lambdas(lam)
    before = ctx.contacts()
    dv_before = ctx.pull_deltav()
    print("autotune on", n, "contacts", flush=True)
    trials, best = ctx.autotune(0.01, sweeps=12, reps=2)
    for t in trials:
        print("trial", t, flush=True)
    assert len(trials) >= 4
    assert {t["pgs_schedule"] for t in trials} == {0, 1}
    assert all(t["ms_per_sweep"] > 0 for t in trials)
    fastest = min(t["ms_per_sweep"] for t in trials)
    assert trials[best]["ms_per_sweep"] == fastest or (best == 0 and trials[0]["ms_per_sweep"] <= fastest * 1.02)
    assert (trials[0]["pgs_schedule"], trials[0]["pgs_grid"], trials[0]["pgs_block"]) == (0, 0, 0)
    for k in ("pgs_schedule", "pgs_grid", "pgs_block"):
        assert getattr(ctx.params, k) == trials[best][k]
    assert_contacts_equal(ctx.contacts(), before)
    assert_bits_equal("deltaV", ctx.pull_deltav(), dv_before)
    it_g, _ = ctx.solve(0.01)
    it_o, _, _ = orc.solve(ctx.solve_order(), 0.01)
    assert it_g == it_o
    assert_contacts_equal(ctx.contacts(), orc.contacts())
    assert_bits_equal("deltaV", ctx.pull_deltav(), orc.pull_deltav())


def test_tuned_steps_equal_untuned_steps(gpu_lib):
    scene = workloads.box_cluster(seed=22, n=512, spheres=40)
    kw = dict(iterations=60, tolerance=1e-7, gravity_amount=9.81, enable_sleeping=0)
    a = Context(); a.upload(scene); a.set_params(**kw)
    b = Context(); b.upload(scene); b.set_params(**kw)
    for step in range(12):
        print("step", step, flush=True)
        a.step(0.01); b.step(0.01)
        if step == 4:
            trials, best = a.autotune(0.01, sweeps=10, reps=1)
            print("chosen", trials[best], flush=True)
            # the last step's statistics survive the trials
            assert_bits_equal("changes", a.changes(), b.changes())
            assert_contacts_equal(a.contacts(), b.contacts())
        for name, x, y in zip("xRvw", a.pull_state(), b.pull_state()):
            assert_bits_equal(f"step {step} {name}", x, y)


def test_autotune_without_contacts_is_refused(gpu_lib):
    scene = workloads.box_cluster(seed=23, n=8, spheres=0)
    ctx = Context(); ctx.upload(scene)
    with pytest.raises(_lib.AMError) as e:
        ctx.autotune(0.01)
    assert e.value.code == _lib.AM_ERR_STATE
    with pytest.raises(_lib.AMError) as e:
        ctx.autotune(0.01, sweeps=0)
    assert e.value.code == _lib.AM_ERR_INVALID
