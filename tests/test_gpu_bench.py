"""bench.py end to end at a small size: every phase (settle, warm-up, tuning,
training on held-out piles, plain and learned timed steps, the dominoes leg
with merging on, CPU baseline) runs
and the one JSON line carries the fields the measurement needs.  It runs in
this process: a second process on the GPU would share it with the test
runner's own contexts."""
import contextlib
import io
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_small_end_to_end(gpu_lib):
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench

    argv = ["--bodies", "400", "--settle", "40", "--warmup", "1", "--steps", "3", "--iterations", "200",
            "--train-scenes", "1", "--eval-scenes", "1", "--train-bodies", "40", "--train-settle", "40",
            "--train-steps", "8", "--epochs", "2", "--cpu-sweeps", "5", "--merging-settle", "20",
            "--merging-steps", "5", "--batched-settle", "30", "--batched-steps", "3"]
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        rc = bench.main(argv)
    assert rc == 0
    line = [ln for ln in buf.getvalue().splitlines() if ln.startswith("{")][-1]
    rec = json.loads(line)
    for key in ("metric", "value", "unit", "ms_per_step", "roofline", "cpu_baseline", "config"):
        assert key in rec
    assert rec["value"] > 0 and rec["roofline"]["frac"] > 0
    assert rec["cpu_baseline"]["value"] > 0 and rec["cpu_baseline"]["cores"] == 1
    lrn = rec["config"]["learned"]
    assert lrn["train_samples"] > 0
    assert lrn["replay_identical"] is True
    assert lrn["held_out"]["sweeps_plain"] > 0 and lrn["held_out"]["sweeps_learned"] > 0
    assert rec["config"]["plain"]["contacts_mean"] > 0
    mrg = rec["config"]["merging"]
    assert mrg["ms_per_step"] > 0 and mrg["contacts_mean"] > 1000 and mrg["sweeps_mean"] > 0
    bat = rec["config"]["batched"]
    assert bat["scenes"] == 256 and bat["ms_per_step"] > 0 and bat["contacts_mean"] > 1000
    assert bat["sweeps_mean"] > 0 and bat["contact_updates_per_s"] > 0 and bat["launch_shape"] is not None
