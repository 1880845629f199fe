"""The C ABI of libam (include/am.h): exported symbols, struct layouts shared
with the ctypes binding, reference defaults and error behaviour without a GPU
(CPU only; no kernel runs)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from adaptivemerging_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "am.h")


@pytest.fixture(scope="module")
def libam():
    if not os.path.exists(_lib.LIB_PATH):
        from adaptivemerging_amd import build

        build.build()
    return _lib.load()


def header_functions():
    text = open(HEADER).read()
    return set(re.findall(r"^(?:const char\*|int)\s+(am_\w+)\s*\(", text, flags=re.M))


def test_header_matches_binding_list():
    assert header_functions() == set(_lib.SYMBOLS)


def test_library_exports_every_header_symbol(libam):
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = header_functions() - exported
    assert not missing, f"not exported: {sorted(missing)}"
    for name in header_functions():
        assert hasattr(libam, name)


LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "am.h"
#define F(T, f) printf(#T " " #f " %zu %zu\n", offsetof(T, f), sizeof(((T*)0)->f));
int main(void) {
  printf("sizeof am_body_desc %zu\n", sizeof(am_body_desc));
  printf("sizeof am_sphere_node %zu\n", sizeof(am_sphere_node));
  printf("sizeof am_params %zu\n", sizeof(am_params));
  printf("sizeof am_contact_arrays %zu\n", sizeof(am_contact_arrays));
  printf("sizeof am_step_stats %zu\n", sizeof(am_step_stats));
  printf("sizeof am_body_pair %zu\n", sizeof(am_body_pair));
  printf("sizeof am_spring %zu\n", sizeof(am_spring));
  printf("sizeof am_tune_trial %zu\n", sizeof(am_tune_trial));
  F(am_body_desc, size) F(am_body_desc, radius) F(am_body_desc, jinv0) F(am_body_desc, plane_d) F(am_body_desc, bbox)
  F(am_body_desc, inertia0)
  F(am_body_pair, flag) F(am_body_pair, collection) F(am_body_pair, hist) F(am_body_pair, states)
  F(am_sphere_node, r) F(am_sphere_node, first_child) F(am_sphere_node, n_children)
  F(am_params, tolerance) F(am_params, compliance) F(am_params, sliding_threshold) F(am_params, gravity_amount)
  F(am_params, threshold_merge) F(am_params, threshold_breath) F(am_params, pgs_block) F(am_params, pgs_grid)
  F(am_step_stats, last_change) F(am_step_stats, ms_total) F(am_step_stats, ms_setup_kernel)
  F(am_step_stats, used_learned) F(am_step_stats, n_internal_contacts) F(am_step_stats, ms_merging)
  F(am_spring, type) F(am_spring, body2) F(am_spring, pb1) F(am_spring, pb2) F(am_spring, pw) F(am_spring, l0)
  F(am_spring, k) F(am_spring, d) F(am_spring, ls)
  F(am_params, spring_stiffness_mod) F(am_params, spring_damping_mod) F(am_params, pgs_schedule) F(am_params, organize_contacts)
  F(am_params, host_threads)
  F(am_tune_trial, grid_used) F(am_tune_trial, block_used) F(am_tune_trial, ms_per_sweep)
  return 0;
}
"""


def test_struct_layouts_match_ctypes(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    sizes = {ln.split()[1]: int(ln.split()[2]) for ln in lines if ln.startswith("sizeof")}
    assert sizes["am_body_desc"] == _lib.BODY_DTYPE.itemsize
    assert sizes["am_sphere_node"] == _lib.NODE_DTYPE.itemsize
    assert sizes["am_params"] == C.sizeof(_lib.Params)
    assert sizes["am_contact_arrays"] == C.sizeof(_lib.ContactArrays)
    assert sizes["am_step_stats"] == C.sizeof(_lib.StepStats)
    assert sizes["am_body_pair"] == C.sizeof(_lib.BodyPair)
    assert sizes["am_spring"] == _lib.SPRING_DTYPE.itemsize
    assert sizes["am_tune_trial"] == C.sizeof(_lib.TuneTrial)
    for ln in lines:
        if ln.startswith("sizeof"):
            continue
        struct, field, off, size = ln.split()
        if struct == "am_body_desc":
            assert _lib.BODY_DTYPE.fields[field][1] == int(off), field
        elif struct == "am_sphere_node":
            assert _lib.NODE_DTYPE.fields[field][1] == int(off), field
        elif struct == "am_params":
            assert getattr(_lib.Params, field).offset == int(off), field
        elif struct == "am_step_stats":
            assert getattr(_lib.StepStats, field).offset == int(off), field
        elif struct == "am_body_pair":
            assert getattr(_lib.BodyPair, field).offset == int(off), field
        elif struct == "am_spring":
            assert _lib.SPRING_DTYPE.fields[field][1] == int(off), field
        elif struct == "am_tune_trial":
            assert getattr(_lib.TuneTrial, field).offset == int(off), field


def test_params_defaults_are_reference_defaults(libam):
    p = _lib.Params()
    assert libam.am_params_default(C.byref(p)) == 0
    h = _lib.default_params_host()
    for name, _ in _lib.Params._fields_:
        assert getattr(p, name) == getattr(h, name), name
    assert p.iterations == 30 and p.tolerance == 1e-5 and p.compliance == 1e-3 and p.feedback_stiffness == 0.5
    assert p.gravity_amount == 1.0 and p.gravity_angle_deg == 90.0 and p.sliding_threshold == 0.01
    assert p.step_accum_merging == 3 and p.steps_between_merge_events == 10 and p.threshold_unmerge == 2e-2
    # RigidBodySystem.springStiffnessMod / springDampingMod (RigidBodySystem.java:553-554)
    assert p.spring_stiffness_mod == 1.0 and p.spring_damping_mod == 1.0


def test_error_codes_without_gpu(libam):
    assert libam.am_last_error(None) == b"null context"
    assert libam.am_ctx_destroy(None) == _lib.AM_ERR_INVALID
    assert libam.am_params_default(None) == _lib.AM_ERR_INVALID
    assert libam.am_detect(None, 0.01, None) == _lib.AM_ERR_INVALID
    assert libam.am_step(None, 0.01, None) == _lib.AM_ERR_INVALID
    assert libam.am_springs_set(None, None, 0) == _lib.AM_ERR_INVALID
    assert libam.am_apply_spring_forces(None) == _lib.AM_ERR_INVALID
    assert libam.am_autotune(None, 0.01, 20, 3, 1, None, 0, None, None) == _lib.AM_ERR_INVALID
    h = C.c_void_p()
    rc = libam.am_ctx_create(-1, C.byref(h))
    assert rc != 0 and not h.value
    assert libam.am_version().startswith(b"libam")


def test_scene_builder_matches_reference_box():
    # XMLParser.createBox: mass = x*y*z*density, I = m/12 (y^2+z^2, ...),
    # boundingBoxB order, radius = half diagonal
    from adaptivemerging_amd import Scene

    s = Scene()
    s.add_box([1.0, 0.15, 0.25], [0, 0, 0], density=1000.0)
    b, _, _, _, _, _ = s.arrays()
    m = 1.0 * 0.15 * 0.25 * 1000.0
    assert b["mass"][0] == m
    assert b["jinv0"][0][0] == 1.0 / (1.0 / 12 * m * (0.15 * 0.15 + 0.25 * 0.25))
    bb = b["bbox"][0].reshape(8, 3)
    assert np.array_equal(bb[0], [-0.5, -0.075, -0.125]) and np.array_equal(bb[1], [-0.5, -0.075, 0.125])
    assert np.array_equal(bb[4], [0.5, -0.075, -0.125]) and np.array_equal(bb[7], [0.5, 0.075, 0.125])
    assert b["radius"][0] == np.sqrt(0.5 ** 2 + 0.075 ** 2 + 0.125 ** 2)
