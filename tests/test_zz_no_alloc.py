"""am_motion_metric_noalloc (the allocation-free variant of am_motion_metric,
staging through a fixed device array) against am_motion_metric, bit for bit,
called through the C ABI: two calls in a row, the second with fewer pairs, and
a call past the array's capacity refused."""
import ctypes as C

import numpy as np
import pytest

from adaptivemerging_amd import _lib, workloads
from adaptivemerging_amd.context import Context

pytestmark = pytest.mark.gpu


def test_noalloc_metric_equals_metric(gpu_lib):
    scene = workloads.box_cluster(seed=21, n=64, spheres=8)
    ctx = Context()
    ctx.upload(scene)
    rng = np.random.default_rng(5)
    v = rng.normal(size=(scene.n, 3))
    w = rng.normal(size=(scene.n, 3))
    ctx.push_state(None, None, v, w)
    lib = ctx.lib
    for n in (3000, 700):
        a = rng.integers(0, scene.n, size=n)
        b = rng.integers(-1, scene.n, size=n)
        pairs = np.ascontiguousarray(np.stack([a, b], axis=1), dtype=np.int32)
        ref = np.zeros(n)
        got = np.full(n, np.nan)
        assert lib.am_motion_metric(ctx.h, _lib.ptr(pairs), n, _lib.ptr(ref)) == 0
        assert lib.am_motion_metric_noalloc(ctx.h, _lib.ptr(pairs), n, _lib.ptr(got)) == 0
        assert np.array_equal(ref.view(np.int64), got.view(np.int64)), f"{n} pairs differ"
        assert np.all(ref >= 0) and np.any(ref > 0)
    big = np.zeros((_lib.AM_METRIC_NOALLOC_MAX + 1, 2), dtype=np.int32)
    out = np.zeros(len(big))
    rc = lib.am_motion_metric_noalloc(ctx.h, _lib.ptr(big), len(big), _lib.ptr(out))
    assert rc == _lib.AM_ERR_CAPACITY
    assert b"at most" in lib.am_last_error(ctx.h)
    ctx.close()
