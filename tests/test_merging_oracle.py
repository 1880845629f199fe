"""Pins the CPU restatement of adaptive merging (oracle/merging.py) with
checks that do not retype its formulas: rule tables for the merge and
unmerge conditions (BodyPairContact.java:83-266), mass properties of a
collection against the parallel-axis theorem and momentum conservation,
rigid push-down, and a 3D analogue of the reference's merge/unmerge
scenario assertions (src/tests/LCPAppTests.java:151-178, 216-247)."""
import math

import numpy as np
import pytest

from adaptivemerging_amd import workloads
from adaptivemerging_amd._lib import AM_STATE_BROKEN, AM_STATE_CLEAR, AM_STATE_ONEDGE, default_params_host
from adaptivemerging_amd.scene import random_rotation
from oracle import merging as M

B, O, C = AM_STATE_BROKEN, AM_STATE_ONEDGE, AM_STATE_CLEAR


def params(**kw):
    p = default_params_host()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("not_on_edge,n_active,expected", [
    (0, 1, O), (0, 4, O),          # no contact off the friction-cone edge
    (2, 2, C), (3, 5, C),          # two or more stable contacts
    (1, 1, C),                     # a single contact decides
    (1, 2, O), (1, 4, O),          # one stable contact among several
])
def test_pair_state_table(not_on_edge, n_active, expected):
    assert M.pair_state(not_on_edge, n_active) == expected


def test_merge_condition_table():
    P = params()
    good = dict(hist=[1e-3, 2e-3, 5e-3], states=[C, C, C], breath=False, pinned1=False, pinned2=False)
    assert M.merge_condition(P, **good)
    assert not M.merge_condition(P, **{**good, "hist": [1e-3, 2e-3]})               # history too short
    assert not M.merge_condition(P, **{**good, "hist": [1e-3, 2e-2, 1e-3]})         # one metric above 1e-2
    assert M.merge_condition(P, **{**good, "hist": [1e-2, 1e-2, 1e-2]})             # threshold is inclusive
    assert not M.merge_condition(P, **{**good, "states": [C, O, C]})                # unstable contacts
    assert not M.merge_condition(P, **{**good, "breath": True})                     # violation still changing
    assert M.merge_condition(P, **{**good, "pinned1": True})                        # pinned merges by default
    assert not M.merge_condition(params(enable_merge_pinned=0), **{**good, "pinned2": True})
    assert M.merge_condition(params(enable_merge_stable_contact=0), **{**good, "states": [O, O, O]})
    assert M.merge_condition(params(enable_merge_let_it_breathe=0), **{**good, "breath": True})
    assert M.merge_condition(params(step_accum_merging=2), **{**good, "hist": [0, 0], "states": [C, C]})


def test_unmerge_condition_table():
    P = params()
    assert M.unmerge_condition(P, [C, C, B], 0, 3)            # a broken contact
    assert M.unmerge_condition(P, [O], 0, 1)                  # a contact on the cone's edge
    assert not M.unmerge_condition(P, [C, C, C], 5, 3)        # moving apart but three contacts hold
    assert M.unmerge_condition(P, [C, C], 3, 2)               # moving apart for three steps, two contacts
    assert not M.unmerge_condition(P, [C, C], 2, 2)
    assert not M.unmerge_condition(params(enable_unmerge_normal=0), [B, C, C], 0, 3)
    assert not M.unmerge_condition(params(enable_unmerge_friction=0), [O, C, C], 0, 3)


def _ent(rng, mass, pinned=False):
    R = random_rotation(rng).reshape(-1).tolist()
    I0 = [0.0] * 9
    d = rng.uniform(0.5, 2.0, 3)
    I0[0], I0[4], I0[8] = d
    return M.Ent(x=rng.normal(size=3).tolist(), v=rng.normal(size=3).tolist(), w=rng.normal(size=3).tolist(),
                 R=R, I0=I0, I=M.rm0rt(R, I0), jinv=M.rm0rt(R, M.invert3(I0)), jinv0=M.invert3(I0),
                 mass=mass, minv=1.0 / mass, pinned=pinned, plane=False,
                 bb=[s * 0.5 for _ in range(1) for s in [1, 1, 1] * 8])


def test_collection_mass_properties_closed_form():
    """Two bodies merged: mass, centre of mass, linear momentum and the
    world inertia about the new centre (parallel-axis theorem) computed
    independently with numpy."""
    rng = np.random.default_rng(0)
    a, b = _ent(rng, 2.0), _ent(rng, 5.0)
    Cc = a.copy()
    M.add_body_internal(Cc, b)
    M.finish_inertia(Cc)
    m = a.mass + b.mass
    com = (a.mass * np.array(a.x) + b.mass * np.array(b.x)) / m
    assert Cc.mass == m
    np.testing.assert_allclose(Cc.x, com, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(m * np.array(Cc.v), a.mass * np.array(a.v) + b.mass * np.array(b.v), rtol=1e-13)
    I = np.zeros((3, 3))
    for e in (a, b):
        R = np.array(e.R).reshape(3, 3)
        d = np.array(e.x) - com
        I += R @ np.diag([e.I0[0], e.I0[4], e.I0[8]]) @ R.T + e.mass * (d @ d * np.eye(3) - np.outer(d, d))
    np.testing.assert_allclose(np.array(Cc.I).reshape(3, 3), I, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.array(Cc.jinv).reshape(3, 3) @ I, np.eye(3), atol=1e-12)
    # the collection frame is world-aligned (updateTheta)
    assert Cc.R == M.IDENT


def test_pinned_collection_is_massless():
    rng = np.random.default_rng(1)
    a, b = _ent(rng, 2.0), _ent(rng, 3.0, pinned=True)
    Cc = a.copy()
    Cc.pinned = Cc.pinned or b.pinned
    M.add_body_internal(Cc, b)
    assert Cc.mass == 0 and Cc.minv == 0 and Cc.v == [0.0] * 3 and Cc.jinv == [0.0] * 9
    assert Cc.x == a.x


def test_remove_inverts_add():
    """Adding then removing a body restores the mass and centre of mass."""
    rng = np.random.default_rng(2)
    a, b, c = _ent(rng, 2.0), _ent(rng, 5.0), _ent(rng, 1.5)
    Cc = a.copy()
    M.add_body_internal(Cc, b)
    M.finish_inertia(Cc)
    m2, x2 = Cc.mass, list(Cc.x)
    M.add_body_internal(Cc, c)
    M.finish_inertia(Cc)
    # RigidCollection.removeBodies' mass/com update, written out for one body
    com = [Cc.x[k] * Cc.mass for k in range(3)]
    x = [-c.mass * c.x[k] + com[k] for k in range(3)]
    mass = Cc.mass - c.mass
    x = [x[k] * (1.0 / mass) for k in range(3)]
    assert math.isclose(mass, m2, rel_tol=1e-15)
    np.testing.assert_allclose(x, x2, rtol=1e-12, atol=1e-13)


def test_body_to_collection_round_trip():
    rng = np.random.default_rng(3)
    Cc = _ent(rng, 1.0)
    xb = rng.normal(size=3).tolist()
    Rb = random_rotation(rng).reshape(-1).tolist()
    Rbc, tbc = M.body_to_collection(Cc, xb, Rb)
    # B2W = C2W * B2C must give the body's transform back
    t = M.mulv(Cc.R, tbc)
    np.testing.assert_allclose([t[k] + Cc.x[k] for k in range(3)], xb, atol=1e-14)
    np.testing.assert_allclose(M.mmul(Cc.R, Rbc), Rb, atol=1e-15)


def test_scenario_settles_merges_and_unmerges():
    """3D analogue of the reference's merge/unmerge scenarios: resting bodies
    merge after the accumulation window; a force that lifts a member breaks its
    contacts inside the collection and it leaves."""
    scene = workloads.merge_stack(seed=3)
    p = params(gravity_amount=9.81, iterations=100, tolerance=1e-10)
    mo = M.MergingOracle(scene, p)
    dt = 0.01
    merged_at, cut_at = [], []
    for step in range(1, 36):
        mo.total_steps += 1
        mo.clear_bodies()
        mo.o.apply_gravity()
        if step == 31:
            f, t = mo.o.pull_forces()
            f[4][1] += 2000.0
            mo.o.push_forces(f, t)
        mo.o.update_contacts_map()
        n = mo.o.detect()
        mo.o.warm_start()
        n_int = sum(len(q["contacts"]) for q in mo.ibpc)
        mo.update_in_collections(dt, lambda: np.arange(n + n_int, dtype=np.int32))
        mo.accumulate_for_unmerging(dt)
        if mo.unmerge(dt):
            cut_at.append(step)
        if mo.merging_event:
            mo.clear_top_level()
        mo.o.redo_warm_start()
        mo.o.solve(np.arange(len(mo.o.contacts()["body1"]), dtype=np.int32), dt)
        mo.clear_body_pair_contacts()
        mo.advance_velocities(dt)
        # right after the push-down, members move rigidly with their collection
        for e, mem in mo.members.items():
            if e in mo.top and len(mem) >= 2 and not mo.pinned(e):
                assert mo._metric(mem[0], mem[1], dt) < 1e-12
        mo.accumulate_for_merging(dt)
        mo.advance_positions(dt)
        if step % 10 == 0 and mo.merge():
            merged_at.append(step)
    assert merged_at and merged_at[0] == 10
    assert cut_at == [31]
    assert 4 in mo.top   # the lifted board is a free body again


def test_unmerge_all_returns_every_body():
    """Merging.unmergeAll (Merging.java:168-206): the collections leave the
    top-level list, their bodies follow the bodies that stayed, in collection
    order, each moving with its collection's velocity at its position; the
    internal pairs are gone and merging starts over."""
    from tests.helpers import merging_oracle_step

    scene = workloads.merge_stack(seed=3)
    p = params(gravity_amount=9.81, iterations=100, tolerance=1e-10, enable_sleeping=0)
    mo = M.MergingOracle(scene, p)
    for step in range(1, 13):
        merging_oracle_step(mo, 0.01, step)
    colls = [e for e in mo.top if e in mo.members]
    assert colls
    stay = [e for e in mo.top if e not in mo.members]
    members = [b for e in colls for b in mo.members[e]]
    x, R, v, w = mo.state()
    want_v = {}
    for e in colls:
        for b in mo.members[e]:
            want_v[b] = np.array(v[e]) + np.cross(w[e], np.array(x[b]) - np.array(x[e]))
    assert mo.unmerge_all() == len(colls)
    assert mo.top == stay + members and not mo.ibpc and not mo.members
    assert all(mo.parent[b] == -1 for b in members)
    x2, _, v2, w2 = mo.state()
    for b, vb in want_v.items():
        np.testing.assert_allclose(v2[b], vb, rtol=1e-12, atol=1e-14)
    merged = 0
    for step in range(13, 31):
        m, _, _ = merging_oracle_step(mo, 0.01, step)
        merged += m
    assert merged >= 1


def test_position_level_metric_dispatch():
    """Merging.metricPositionLevel (BodyPairContact.java:88-95, :130-137): the
    relative metric of two unpinned entities is the position-level one, a
    pair with a pinned body keeps the other body's own velocity metric; the
    stack still settles and merges with it."""
    from tests.helpers import merging_oracle_step

    scene = workloads.merge_stack(seed=3)
    p = params(gravity_amount=9.81, iterations=100, tolerance=1e-10, metric_position_level=1)
    mo = M.MergingOracle(scene, p)
    dt = 0.01
    merged = free_pairs = 0
    for step in range(1, 13):
        m, _, _ = merging_oracle_step(mo, dt, step)
        merged += m
        if step == 3:
            free_pairs = sum(1 for r in mo.bpc if not (mo.pinned(r["b1"]) or mo.pinned(r["b2"])))
            for r in mo.bpc:
                a, b = r["b1"], r["b2"]
                if mo.pinned(a) or mo.pinned(b):
                    assert mo._metric(a, b, dt) == mo.o.motion_metric(b if mo.pinned(a) else a)
                else:
                    assert mo._metric(a, b, dt) == mo.o.motion_metric_pos(a, b, dt)
                    assert mo._metric(a, b, dt) != mo.o.motion_metric(a, b)
    assert merged >= 1 and free_pairs >= 2


def test_organized_layers_stack_closed_form():
    """getOrganizedContacts (CollisionProcessor.java:346-441) on a column of
    boxes 1..5 on the plane 0 with a new contact only between 4 and 5: the
    layers run down the column, reach box 6 on the plane through the plane,
    then the external pair never reached, then the internal pair."""
    from oracle.merging import organized_layers

    pairs = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 6), (9, 10), (7, 8)]
    b1, b2, new = [], [], []
    for i, (a, b) in enumerate(pairs):
        for j in range(2):          # two contacts per pair
            b1.append(b if i % 2 else a)
            b2.append(a if i % 2 else b)
            new.append(1 if (a, b) == (4, 5) and j == 1 else 0)
    n_ext = 14                      # the last pair (7, 8) is internal
    layers = organized_layers(b1, b2, new, n_ext)
    want = {(4, 5): 0, (3, 4): 1, (2, 3): 2, (1, 2): 3, (0, 1): 4, (0, 6): 5, (9, 10): 6, (7, 8): 7}
    for k, (a, b) in enumerate(zip(b1, b2)):
        assert layers[k] == want[(min(a, b), max(a, b))]
    # no new contact: every external pair comes before every internal pair
    layers = organized_layers(b1, b2, [0] * len(b1), n_ext)
    assert set(layers[:n_ext]) == {0} and set(layers[n_ext:]) == {1}
    # a new contact on an internal pair does not seed a layer
    new_int = [0] * n_ext + [1, 1]
    assert organized_layers(b1, b2, new_int, n_ext) == layers
