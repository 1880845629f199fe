#!/usr/bin/env python3
"""Benchmark: one advanceTime-equivalent step of the 3D contact path on a
settled 20,000-box pile (edges U(0.1, 0.5) m, densities log-uniform
100-10,000 kg/m^3, 4 x 4 m bin, about 290k contacts at rest), PGS run towards
residual 1e-6 (max |delta lambda| per sweep) with an iteration cap.

A "step" = clear, gravity, contacts map, broad + narrow phase, warm start,
FP64 PGS to the residual or the cap, velocity and position update, all on the
GPU with body state resident in HBM.

Phases (only the last is timed):
  1. settle: the pile is built shelf-packed and falls into the bin for
     --settle steps at --settle-iterations sweeps, so every later phase runs
     on the resting pile whatever --warmup is;
  2. warm-up: --warmup steps at the full sweep cap;
  3. launch-shape tuning of the PGS kernel on the live contacts;
  4. learned start: the warm-start network is trained on held-out small
     piles solved to the tolerance (converged targets), then a twin context
     replays phases 1-2 (the GPU path is deterministic, checked) and runs the
     timed steps with the learned start and its plain-start fallback;
  5. timed: --steps steps with the plain warm start (the headline), then the
     same steps from the same state with the learned start;
  6. merging leg: the dominoes-on-platforms configuration with merging on,
     timed on its own after its push (reported under config.merging);
  7. batched leg: 256 independent piles of boxes and spheres in one context
     with merging on, 50 sweeps, timed on its own after they settle
     (reported under config.batched).

Prints one JSON line: contact-updates/s (N_contacts x sweeps / solve time)
and ms/step of the plain run as the headline, both runs' figures (sweeps,
residual reached, residual at the cap) under config, the PGS kernel's
fraction of the HBM roofline, and the CPU oracle timed on this host on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/step + contact-updates/s at PGS residual 1e-6, learned vs plain warm-start"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_sweep(n_contacts: int, n_dyn: int) -> int:
    """Compulsory FP64 traffic of one sweep: per contact J 288 + b 24 + D 24 +
    lambda read+write 48 + mu 8 + two body indices 8 = 400 B; per dynamic body
    deltaV read+write 96 + minv 8 + jinv 72 = 176 B."""
    return 400 * n_contacts + 176 * n_dyn


TRAFFIC_PROFILES = {
    "dataflow": os.path.join(ROOT, "profiles", "k_pgs_flow_pile20k_traffic.md"),
    "coloured": os.path.join(ROOT, "profiles", "k_pgs_coloured_pile20k_traffic.md"),
}


def workload_key(args) -> dict:
    return dict(bodies=args.bodies, settle=args.settle, settle_iterations=args.settle_iterations,
                warmup=args.warmup, steps=args.steps, iterations=args.iterations, tolerance=args.tolerance,
                dt=args.dt, seed=args.seed, schedule=args.pgs_schedule)


def measured_traffic(args) -> dict | None:
    """HBM bytes per PGS launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this same workload (counters scaled by a plain copy
    of known size); None when the workload arguments differ from the
    profiled ones.  The bytes moved per sweep do not depend on the wave
    count the tuner settles on (it varies between runs)."""
    try:
        with open(TRAFFIC_PROFILES[args.pgs_schedule]) as f:
            line = next(ln for ln in f if ln.startswith('{"workload"'))
    except (OSError, StopIteration):
        return None
    rec = json.loads(line)
    if rec["workload"] != workload_key(args):
        return None
    return rec


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_cpu_baseline(scene, state, params_kw, dt, sweeps: int, budget_s: float):
    """The CPU oracle (single thread) on the GPU's current pile state: one
    step builds the contacts and the warm-start map, then a step whose solve
    runs `sweeps` sequential sweeps (no early exit) gives contact-updates/s."""
    from oracle.pyoracle import Oracle

    orc = Oracle(scene)
    orc.set_params(**{**params_kw, "iterations": sweeps, "tolerance": 0.0})
    x, R, v, w = state
    orc.set_state(x, R, v, w)
    t_start = time.perf_counter()
    samples = []
    for _ in range(2):
        t0 = time.perf_counter()
        orc.clear_bodies()
        orc.apply_gravity()
        orc.update_contacts_map()
        n = orc.detect()
        orc.warm_start()
        orc.redo_warm_start()
        t1 = time.perf_counter()
        it, last, _ = orc.solve(np.arange(n, dtype=np.int32), dt)
        t2 = time.perf_counter()
        orc.advance_velocities(dt)
        orc.advance_positions(dt)
        t3 = time.perf_counter()
        samples.append(dict(n=n, sweeps=it, solve_s=t2 - t1, step_s=t3 - t0))
        if time.perf_counter() - t_start > budget_s:
            break
    s = samples[-1]
    return {
        "value": s["n"] * s["sweeps"] / s["solve_s"] if s["solve_s"] > 0 else 0.0,
        "unit": "contact-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"C oracle (oracle/am_oracle.c, -O2 -ffp-contract=off), 1 thread on '{cpu_model()}' "
                   f"(nproc {os.cpu_count()}): step {len(samples)} from the GPU's settled pile state, "
                   f"{s['n']} contacts, {s['sweeps']} sequential sweeps (a bounded sample of the capped solve); "
                   f"solve {s['solve_s']:.3f} s, whole step {s['step_s']:.3f} s"),
        "ms_per_step": s["step_s"] * 1e3,
        "sweeps": s["sweeps"],
        "contacts": s["n"],
    }


def run_steps(ctx, steps: int, dt: float):
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.step(dt))
    return stats, time.perf_counter() - t0


def summarize(stats, wall_s: float, n_dyn: int, tolerance: float) -> dict:
    n_c = np.array([s["n_contacts"] for s in stats], dtype=np.float64)
    sweeps = np.array([s["iterations_used"] for s in stats], dtype=np.float64)
    solve_ms = np.array([s["ms_solve"] for s in stats])
    pgs_ms = np.array([s["ms_pgs_kernel"] for s in stats])
    updates = float(np.sum(n_c * sweeps))
    bytes_pgs = float(sum(algorithmic_bytes_per_sweep(int(s["n_contacts"]), n_dyn) * s["iterations_used"]
                          for s in stats))
    return {
        "contact_updates_per_s": updates / (float(np.sum(solve_ms)) * 1e-3) if np.sum(solve_ms) > 0 else 0.0,
        "ms_per_step": wall_s * 1e3 / max(len(stats), 1),
        "contacts_mean": float(np.mean(n_c)),
        "sweeps_mean": float(np.mean(sweeps)),
        "residual_final_mean": float(np.mean([s["last_change"] for s in stats])),
        "residual_reached_frac": float(np.mean([s["last_change"] < tolerance for s in stats])),
        "colours_mean": float(np.mean([s["n_colours"] for s in stats])),
        "ms_detect": float(np.mean([s["ms_detect"] for s in stats])),
        "ms_warm_start": float(np.mean([s["ms_warm_start"] for s in stats])),
        "ms_learned_start": float(np.mean([s["ms_learned"] for s in stats])),
        "ms_solve": float(np.mean(solve_ms)),
        "ms_pgs_kernel": float(np.mean(pgs_ms)),
        "ms_integrate": float(np.mean([s["ms_integrate"] for s in stats])),
        "ms_step_events": float(np.mean([s["ms_total"] for s in stats])),
        "learned_used_frac": float(np.mean([s["used_learned"] for s in stats])),
        "start_residual_plain": float(np.mean([s["start_residual_plain"] for s in stats])),
        "start_residual_learned": float(np.mean([s["start_residual_learned"] for s in stats])),
        "_pgs_bytes": bytes_pgs,
        "_pgs_ms": float(np.sum(pgs_ms)),
        "_pgs_launches": len(stats),
    }


def settle_and_warm(ctx, args, params_kw, label):
    """Phases 1-2 on one context; returns the last step's statistics."""
    ctx.set_params(**{**params_kw, "iterations": args.settle_iterations})
    st = None
    t0 = time.perf_counter()
    for k in range(args.settle):
        st = ctx.step(args.dt)
        if k % 50 == 0 or k == args.settle - 1:
            print(f"# {label} settle {k}: contacts {st['n_contacts']} sweeps {st['iterations_used']} "
                  f"residual {st['last_change']:.3g} step {st['ms_total']:.2f} ms ({time.perf_counter() - t0:.1f} s)",
                  file=sys.stderr, flush=True)
    ctx.set_params(**params_kw)
    for k in range(args.warmup):
        st = ctx.step(args.dt)
        print(f"# {label} warmup {k}: contacts {st['n_contacts']} sweeps {st['iterations_used']} "
              f"residual {st['last_change']:.3g} colours {st['n_colours']} step {st['ms_total']:.2f} ms",
              file=sys.stderr, flush=True)
    return st


def run_merging_leg(args) -> dict:
    """The dominoes-on-platforms configuration with merging on (the
    reference's adaptive-merging path): 2,000 dominoes in 20 rows on 10
    kinematic platforms, 50 sweeps at the default tolerance, merging and
    unmerging at the reference defaults, sleeping off.  --merging-settle
    untimed steps, the push on each row's first domino (5 m/s), then
    --merging-steps timed steps through the start of the cascade."""
    from adaptivemerging_amd import workloads
    from adaptivemerging_amd.context import Context

    scene = workloads.domino_platforms(seed=args.seed)
    n_dominoes = int(np.sum((scene.arrays()[0]["flags"] & 1) == 0))
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(gravity_amount=9.81, iterations=50, enable_sleeping=0,
                   pgs_plain_launch=int(args.pgs_launch == "plain"))
    t0 = time.perf_counter()
    for _ in range(args.merging_settle):
        ctx.step(args.dt)
    for body, point, force in workloads.domino_push(scene, ctx.pull_state()[0], args.dt):
        ctx.impulse_add(body, point, force)
    stats, wall = run_steps(ctx, args.merging_steps, args.dt)
    ctx.close()
    out = {
        "workload": f"{n_dominoes} dominoes (0.05 x 0.5 x 0.25 m) in 20 rows on 10 kinematic platforms, dt "
                    f"{args.dt}, 50 sweeps, merging and unmerging on at the reference defaults; "
                    f"{args.merging_settle} untimed steps, a 5 m/s push on each row's first domino, then "
                    f"{args.merging_steps} timed steps",
        "ms_per_step": wall * 1e3 / max(len(stats), 1),
        "ms_merging": float(np.mean([s["ms_merging"] for s in stats])),
        "ms_detect": float(np.mean([s["ms_detect"] for s in stats])),
        "ms_solve": float(np.mean([s["ms_solve"] for s in stats])),
        "contacts_mean": float(np.mean([s["n_contacts"] for s in stats])),
        "internal_contacts_mean": float(np.mean([s["n_internal_contacts"] for s in stats])),
        "collections_max": int(max(s["n_collections"] for s in stats)),
        "merged": int(sum(s["n_merged"] for s in stats)),
        "unmerged": int(sum(s["n_unmerged"] for s in stats)),
        "sweeps_mean": float(np.mean([s["iterations_used"] for s in stats])),
        "untimed_s": time.perf_counter() - t0 - wall,
    }
    print(f"# merging: {out['ms_per_step']:.2f} ms/step (merging {out['ms_merging']:.2f} ms), "
          f"{out['contacts_mean']:.0f} contacts, {out['internal_contacts_mean']:.0f} internal, "
          f"{out['merged']} merged, {out['unmerged']} unmerged", file=sys.stderr, flush=True)
    return out


def run_batched_leg(args) -> dict:
    """The batched configuration: 256 independent scenes held by one context
    (one detection, one colouring and one persistent PGS launch per step over
    all of them), body counts log-uniform 50-2,000, each a pile of boxes
    (edges 0.1-0.5 m) and spheres (radius 0.05-0.3 m) dropped on its own
    plane; 50 sweeps at the default tolerance with per-scene exits, merging
    and unmerging on at the reference defaults, sleeping off.
    --batched-settle untimed steps, the launch-shape tuner (with --autotune),
    then --batched-steps timed steps."""
    from adaptivemerging_amd import workloads
    from adaptivemerging_amd.context import Context

    scene = workloads.batched_piles(seed=args.seed)
    n_scenes = int(np.max(scene.arrays()[0]["scene"])) + 1
    ctx = Context()
    ctx.upload(scene)
    ctx.set_params(gravity_amount=9.81, iterations=50, enable_sleeping=0,
                   pgs_plain_launch=int(args.pgs_launch == "plain"))
    t0 = time.perf_counter()
    for _ in range(args.batched_settle):
        ctx.step(args.dt)
    shape = None
    if args.autotune == "on":
        trials, best = ctx.autotune(args.dt, sweeps=20, reps=3)
        b = trials[best]
        shape = {"schedule": "dataflow" if b["pgs_schedule"] == 0 else "coloured", "grid": b["grid_used"],
                 "block": b["block_used"], "ms_per_sweep": b["ms_per_sweep"]}
    stats, wall = run_steps(ctx, args.batched_steps, args.dt)
    used, _ = ctx.scene_iterations()
    ctx.close()
    n_c = np.array([s["n_contacts"] for s in stats], dtype=np.float64)
    sweeps = np.array([s["iterations_used"] for s in stats], dtype=np.float64)
    solve_ms = np.array([s["ms_solve"] for s in stats])
    out = {
        "workload": f"{n_scenes} independent scenes ({scene.n} bodies), body counts log-uniform 50-2000, piles of "
                    f"boxes and spheres on a plane each, dt {args.dt}, 50 sweeps, merging and unmerging on at the "
                    f"reference defaults; {args.batched_settle} untimed steps, then {args.batched_steps} timed steps",
        "scenes": n_scenes,
        "bodies": scene.n,
        "ms_per_step": wall * 1e3 / max(len(stats), 1),
        "contact_updates_per_s": float(np.sum(n_c * sweeps)) / (float(np.sum(solve_ms)) * 1e-3)
        if np.sum(solve_ms) > 0 else 0.0,
        "ms_merging": float(np.mean([s["ms_merging"] for s in stats])),
        "ms_detect": float(np.mean([s["ms_detect"] for s in stats])),
        "ms_solve": float(np.mean(solve_ms)),
        "ms_pgs_kernel": float(np.mean([s["ms_pgs_kernel"] for s in stats])),
        "contacts_mean": float(np.mean(n_c)),
        "internal_contacts_mean": float(np.mean([s["n_internal_contacts"] for s in stats])),
        "collections_max": int(max(s["n_collections"] for s in stats)),
        "merged": int(sum(s["n_merged"] for s in stats)),
        "unmerged": int(sum(s["n_unmerged"] for s in stats)),
        "colours_mean": float(np.mean([s["n_colours"] for s in stats])),
        "sweeps_mean": float(np.mean(sweeps)),
        "scene_sweeps_last_step": {"min": int(np.min(used)), "mean": float(np.mean(used)), "max": int(np.max(used))},
        "launch_shape": shape,
        "untimed_s": time.perf_counter() - t0 - wall,
    }
    print(f"# batched: {out['ms_per_step']:.2f} ms/step (solve {out['ms_solve']:.2f}, merging "
          f"{out['ms_merging']:.2f}, detect {out['ms_detect']:.2f}), {out['contacts_mean']:.0f} contacts, "
          f"{out['internal_contacts_mean']:.0f} internal, {out['merged']} merged, {out['unmerged']} unmerged",
          file=sys.stderr, flush=True)
    return out


def small_pile(args, k: int):
    """Scene k of the small piles the learned start is trained (k >= 0) and
    evaluated (k < 0) on: the configuration's boxes and densities, few
    enough in a wide bin to form a shallow pile whose solves reach the
    tolerance."""
    from adaptivemerging_amd import workloads

    return workloads.box_pile(seed=1000 + args.seed * 97 + k, n=args.train_bodies, bin_size=args.train_bin)


def train_learned_start(args, params_kw):
    """Trains the warm-start network on small piles (other seeds than the
    measured pile) whose solves reach the tolerance: each converged main
    solve adds (features, converged lambda) pairs.  Returns (flat parameters,
    info dict)."""
    from adaptivemerging_amd.context import Context

    t0 = time.perf_counter()
    tr = Context()
    n_train = 0
    collected = 0
    for k in range(args.train_scenes):
        scene = small_pile(args, k)
        tr.upload(scene)
        tr.set_params(**{**params_kw, "iterations": args.train_iterations, "collect_training": 0,
                         "learned_warm_start": 0})
        for _ in range(args.train_settle):
            tr.step(args.dt)
        tr.set_params(collect_training=1)
        for _ in range(args.train_steps):
            st = tr.step(args.dt)
            collected += int(st["last_change"] < args.tolerance)
        tr.set_params(collect_training=0)
    n_train = tr.train_size()
    info = dict(train_scenes=args.train_scenes, train_bodies=args.train_bodies, converged_solves=collected,
                train_samples=n_train)
    if n_train < 256:
        tr.close()
        info["error"] = "too few converged training solves"
        return None, info
    tr.mlp_init(args.seed)
    losses = [tr.train_epochs(1, 256, 1e-3, args.seed + e) for e in range(args.epochs)]
    params = tr.mlp_params()
    tr.train_clear()
    tr.close()
    info.update(epochs=args.epochs, loss_first=losses[0], loss_last=losses[-1], train_s=time.perf_counter() - t0)
    print(f"# learned start: {n_train} pairs from {collected} converged solves, loss {losses[0]:.4g} -> "
          f"{losses[-1]:.4g} in {info['train_s']:.1f} s", file=sys.stderr, flush=True)
    return params, info


def evaluate_learned_start(args, params_kw, mlp):
    """Sweeps to the tolerance with the plain and with the learned start on
    held-out small piles (seeds the network was not trained on), where the
    tolerance is reached, while the boxes land (steps --eval-from onwards of
    the drop: contacts keep appearing, which is where a start for new
    contacts can matter): two contexts per pile replay the same steps (the
    GPU path is deterministic), then run the same steps, one of them with the
    learned start and its plain-start fallback."""
    from adaptivemerging_amd.context import Context

    kw = {**params_kw, "iterations": args.train_iterations, "learned_warm_start": 0, "collect_training": 0}
    out = {"plain": [], "learned": [], "ms_plain": [], "ms_learned": [], "used": [], "converged_plain": [],
           "converged_learned": []}
    for k in range(1, args.eval_scenes + 1):
        scene = small_pile(args, -k)
        ctxs = []
        for learned in (False, True):
            c = Context()
            c.upload(scene)
            c.set_params(**kw)
            for _ in range(args.eval_from):
                c.step(args.dt)
            if learned:
                c.set_mlp_params(mlp)
                c.set_params(learned_warm_start=1)
            ctxs.append(c)
        for name, c in zip(("plain", "learned"), ctxs):
            for _ in range(args.train_steps):
                st = c.step(args.dt)
                out[name].append(st["iterations_used"])
                out["ms_" + name].append(st["ms_total"])
                out["converged_" + name].append(st["last_change"] < args.tolerance)
                if name == "learned":
                    out["used"].append(st["used_learned"])
            c.close()
    res = dict(eval_scenes=args.eval_scenes, steps_per_scene=args.train_steps,
               sweeps_plain=float(np.mean(out["plain"])), sweeps_learned=float(np.mean(out["learned"])),
               ms_per_step_plain=float(np.mean(out["ms_plain"])), ms_per_step_learned=float(np.mean(out["ms_learned"])),
               converged_frac_plain=float(np.mean(out["converged_plain"])),
               converged_frac_learned=float(np.mean(out["converged_learned"])),
               learned_used_frac=float(np.mean(out["used"])))
    res["sweep_ratio"] = res["sweeps_plain"] / res["sweeps_learned"] if res["sweeps_learned"] > 0 else None
    res["ms_ratio"] = res["ms_per_step_plain"] / res["ms_per_step_learned"] if res["ms_per_step_learned"] > 0 else None
    print(f"# held-out piles: sweeps to {args.tolerance}: plain {res['sweeps_plain']:.1f}, learned "
          f"{res['sweeps_learned']:.1f} (learned start used {res['learned_used_frac']:.2f})", file=sys.stderr,
          flush=True)
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bodies", type=int, default=20000)
    ap.add_argument("--settle", type=int, default=350, help="steps the pile falls and settles before the warm-up")
    ap.add_argument("--settle-iterations", type=int, default=100, help="PGS sweep cap while settling")
    ap.add_argument("--iterations", type=int, default=1000, help="PGS iteration cap")
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--learned", choices=["on", "off"], default="on",
                    help="also run the learned warm start from the same state")
    ap.add_argument("--train-scenes", type=int, default=4, help="small piles the network is trained on")
    ap.add_argument("--eval-scenes", type=int, default=2, help="held-out small piles the gain is measured on")
    ap.add_argument("--eval-from", type=int, default=10, help="step of the drop the held-out measurement starts at")
    ap.add_argument("--train-bodies", type=int, default=100)
    ap.add_argument("--train-bin", type=float, default=4.0)
    ap.add_argument("--train-settle", type=int, default=80)
    ap.add_argument("--train-steps", type=int, default=30)
    ap.add_argument("--train-iterations", type=int, default=3000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--merging-leg", choices=["on", "off"], default="on",
                    help="also time the dominoes configuration with merging on (reported under config.merging)")
    ap.add_argument("--merging-settle", type=int, default=300, help="untimed domino steps before the push")
    ap.add_argument("--merging-steps", type=int, default=100, help="timed domino steps from the push on")
    ap.add_argument("--batched-leg", choices=["on", "off"], default="on",
                    help="also time 256 batched scenes with merging on (reported under config.batched)")
    ap.add_argument("--batched-settle", type=int, default=200, help="untimed steps of the batched scenes")
    ap.add_argument("--batched-steps", type=int, default=50, help="timed steps of the batched scenes")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-sweeps", type=int, default=40, help="sequential sweeps of the CPU baseline sample")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--pgs-grid", type=int, default=0)
    ap.add_argument("--pgs-block", type=int, default=0)
    ap.add_argument("--pgs-launch", choices=["cooperative", "plain"], default="cooperative",
                    help="plain: ordinary launch of the persistent kernel (for profilers without cooperative-launch "
                         "support)")
    ap.add_argument("--pgs-schedule", choices=["dataflow", "coloured"], default=None,
                    help="dataflow: each body's deltaV handed from update to update; coloured: a grid barrier between colours "
                         "(bit-identical results)")
    ap.add_argument("--autotune", choices=["auto", "on", "off"], default="auto",
                    help="after the warm-up, time the PGS kernel's launch shapes on the live contacts (both "
                         "schedules; bit-identical results checked) and run the timed steps with the fastest. "
                         "auto: on unless --pgs-schedule, --pgs-grid or --pgs-block is given")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (calibrates the HBM byte counters) and exit")
    args = ap.parse_args(argv)
    explicit_shape = args.pgs_schedule is not None or args.pgs_grid != 0 or args.pgs_block != 0
    if args.pgs_schedule is None:
        args.pgs_schedule = "dataflow"
    if args.autotune == "auto":
        args.autotune = "off" if explicit_shape else "on"

    if args.copy_probe > 0:
        import torch

        src = torch.empty(args.copy_probe // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4, "launches": 3}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent scenes would shard across GPUs; one GPU is measured here)", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0 or args.settle < 0:
        print("bench.py: --steps must be >= 1, --warmup and --settle >= 0", file=sys.stderr)
        return 2

    from adaptivemerging_amd import workloads
    from adaptivemerging_amd.context import Context

    t0 = time.perf_counter()
    scene = workloads.box_pile(seed=args.seed, n=args.bodies)
    b, _, _, _, _, _ = scene.arrays()
    n_dyn = int(np.sum((b["flags"] & 1) == 0))
    # the pile runs without merging and, like every measured configuration,
    # without sleeping
    params_kw = dict(iterations=args.iterations, tolerance=args.tolerance, gravity_amount=9.81,
                     pgs_grid=args.pgs_grid, pgs_block=args.pgs_block, enable_merging=0, enable_unmerging=0,
                     enable_sleeping=0, pgs_plain_launch=int(args.pgs_launch == "plain"),
                     pgs_schedule=0 if args.pgs_schedule == "dataflow" else 1)
    print(f"# scene built: {scene.n} bodies in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    learned = args.learned == "on"
    mlp, train_info = (None, None)
    if learned:
        mlp, train_info = train_learned_start(args, params_kw)
        if mlp is not None and args.eval_scenes > 0:
            train_info["held_out"] = evaluate_learned_start(args, params_kw, mlp)

    plain = Context()
    plain.upload(scene)
    settle_and_warm(plain, args, params_kw, "plain")

    tuned = None
    if args.autotune == "on":
        t1 = time.perf_counter()
        trials, best = plain.autotune(args.dt, sweeps=20, reps=3)
        chosen = trials[best]
        auto = next(t for t in trials if t["pgs_schedule"] == 0 and t["pgs_grid"] == 0)
        tuned = dict(chosen=chosen, trials=len(trials), default_ms_per_sweep=auto["ms_per_sweep"],
                     tune_s=time.perf_counter() - t1)
        args.pgs_schedule = "dataflow" if chosen["pgs_schedule"] == 0 else "coloured"
        args.pgs_grid, args.pgs_block = chosen["pgs_grid"], chosen["pgs_block"]
        shape = {k: chosen[k] for k in ("pgs_schedule", "pgs_grid", "pgs_block")}
        params_kw.update(shape)
        for t in trials:
            print(f"# tune: schedule {t['pgs_schedule']} grid {t['grid_used']} block {t['block_used']}: "
                  f"{t['ms_per_sweep'] * 1e3:.2f} us/sweep", file=sys.stderr, flush=True)
        print(f"# tuned in {tuned['tune_s']:.1f} s: schedule {args.pgs_schedule}, grid {chosen['grid_used']}, "
              f"block {chosen['block_used']}", file=sys.stderr, flush=True)

    state0 = plain.pull_state()
    stats_p, wall_p = run_steps(plain, args.steps, args.dt)
    res_p = summarize(stats_p, wall_p, n_dyn, args.tolerance)
    print(f"# plain: {res_p['ms_per_step']:.2f} ms/step, {res_p['contacts_mean']:.0f} contacts, sweeps "
          f"{res_p['sweeps_mean']:.0f}, residual {res_p['residual_final_mean']:.3g}", file=sys.stderr, flush=True)

    merging = run_merging_leg(args) if args.merging_leg == "on" else None
    batched = run_batched_leg(args) if args.batched_leg == "on" else None

    cpu = None
    if args.cpu_baseline == "on":
        cpu = run_cpu_baseline(scene, plain.pull_state(), params_kw, args.dt, args.cpu_sweeps, args.cpu_budget)
        print(f"# cpu baseline: {cpu['value'] / 1e6:.1f} M contact-updates/s ({cpu['contacts']} contacts)",
              file=sys.stderr, flush=True)

    res_l = None
    if learned:
        res_l = dict(train_info)
        if mlp is not None:
            twin = Context()
            twin.upload(scene)
            settle_and_warm(twin, args, params_kw, "twin")
            # the twin ran the plain context's phases 1-2: the GPU path is
            # deterministic, so its state must equal the plain one's before its timed steps
            res_l["replay_identical"] = all(np.array_equal(a, c) for a, c in zip(state0, twin.pull_state()))
            twin.set_mlp_params(mlp)
            twin.set_params(learned_warm_start=1)
            stats_l, wall_l = run_steps(twin, args.steps, args.dt)
            res_l.update(summarize(stats_l, wall_l, n_dyn, args.tolerance))
            twin.close()
            print(f"# learned: {res_l['ms_per_step']:.2f} ms/step, sweeps {res_l['sweeps_mean']:.0f}, residual "
                  f"{res_l['residual_final_mean']:.3g}, used {res_l['learned_used_frac']:.2f}", file=sys.stderr,
                  flush=True)

    grid_used, block_used = plain.launch_shape()
    traffic = measured_traffic(args)
    achieved = res_p["_pgs_bytes"] / (res_p["_pgs_ms"] * 1e-3) / 1e9 if res_p["_pgs_ms"] > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": ("am::k_pgs_flow<false> (persistent dataflow PGS)" if args.pgs_schedule == "dataflow"
                   else "am::k_pgs<false> (persistent coloured PGS)"),
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic["traffic_bytes_per_launch"] if traffic else None,
        "traffic_source": os.path.relpath(TRAFFIC_PROFILES[args.pgs_schedule], ROOT) if traffic else None,
        "grid_used": grid_used,
        "block_used": block_used,
        "bytes_per_launch": res_p["_pgs_bytes"] / max(res_p["_pgs_launches"], 1),
        "ms_per_launch": res_p["_pgs_ms"] / max(res_p["_pgs_launches"], 1),
    }
    for r in (res_p, res_l):
        if r is not None:
            for k in [k for k in r if k.startswith("_")]:
                del r[k]
    out = {
        "metric": METRIC,
        "value": res_p["contact_updates_per_s"],
        "unit": "contact-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res_p["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.bodies}-box pile at rest ({res_p['contacts_mean']:.0f} contacts), edges U(0.1,0.5) m, "
                        f"density log-uniform 100-10000 kg/m^3, 4x4 m bin (floor + 4 walls), dt {args.dt}, g 9.81, "
                        f"settled {args.settle} steps at {args.settle_iterations} sweeps, then PGS to residual "
                        f"{args.tolerance} capped at {args.iterations} sweeps (the residual is not reached on this "
                        f"pile: see plain.residual_final_mean); value and ms_per_step: plain warm start",
            "bodies": scene.n,
            "plain": res_p,
            "learned": res_l,
            "parallelism": "none (1 GPU)",
            "launch_shape": tuned,
            "merging": merging,
            "batched": batched,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plain.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
