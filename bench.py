"""Benchmark of the resolve2d step on one MI355X.

Workload (default): 1024 independent worlds, body counts uniform in [50, 500]
(mixed discs / rectangles), each with 0-20 distance joints and 0-10 anchored
springs, stepped together (dt = 1/60, 4 sub-steps, 4 solver iterations).
A step is one Solver.process over every world of the batch.  Inputs are
resident on the device before timing starts.

Printed JSON line: body-steps/s of the batched step (value), p50/p99 step
latency, the same scenes stepped one world per launch ("plain"), a roofline
entry for the dominant kernel (timed with HIP events on the library's stream),
and a CPU baseline (the oracle restatement on a bounded sample, one core).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def checked(rc, what):
    from resolve2d_amd import _native
    _native.check(rc, what)


def build_worlds(ctx, n_worlds, seed0):
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    worlds, specs = [], []
    for i in range(n_worlds):
        sc = scenes.batched_world(seed0 + i)
        w = GpuWorld(ctx)
        apply_scene(sc, w)
        worlds.append(w)
        specs.append(sc)
    return worlds, specs


def count_bodies(sc):
    return sum(1 for op in sc if op[0] in ("disc", "rect"))


def cpu_baseline(specs, steps, budget_s):
    """Oracle (f32 restatement of the reference step), one core, on worlds spread
    over the batch, each over the full horizon, until the time budget is spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    body_steps, t_total, used = 0, 0.0, 0
    order = list(range(0, len(specs), 37)) + [i for i in range(len(specs)) if i % 37]
    for idx in order:
        sc = specs[idx]
        o = pyoracle.OracleWorld("f32")
        apply_scene(sc, o)
        n = count_bodies(sc)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        t_total += time.perf_counter() - t0
        body_steps += n * steps
        used += 1
        o.close()
        if t_total > budget_s:
            break
    return body_steps / t_total, used, t_total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--plain-worlds", type=int, default=32)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--single-steps", type=int, default=300)
    ap.add_argument("--tuned-hash", action="store_true")
    ap.add_argument("--variant", type=int, default=-1, help="solve-kernel variant (-1: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: worlds are independent and would be sharded whole across GPUs "
              f"with no exchange; --gpus {args.gpus} is not run here.", file=sys.stderr)
        sys.exit(2)

    from resolve2d_amd import scenes
    from resolve2d_amd.solver import Context
    ctx = Context(0)
    ctx.set_solve_variant(args.variant)
    t0 = time.perf_counter()
    worlds, specs = build_worlds(ctx, args.worlds, args.seed)
    handles = [w.handle for w in worlds]
    n_bodies = sum(count_bodies(sc) for sc in specs)
    setup_s = time.perf_counter() - t0
    dt, S, K = scenes.DT, scenes.SUB_STEPS, scenes.ITERS

    log(f"built {args.worlds} worlds, {n_bodies} bodies in {setup_s:.1f} s")
    for _ in range(args.warmup):
        ctx.process_batch_handles(handles, dt, S, K, 1)
    ctx.sync()

    # timed region: exactly `steps` batched steps; per-step HIP events on the library stream
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    lat = []
    ctx.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        ctx.process_batch_handles(handles, dt, S, K, 1)
        lat.append(time.perf_counter() - a)
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    phase_ms, phase_steps = ctx.phase_times()
    counts = ctx.last_counts()
    ctx.set_profiling(False)

    value = n_bodies * args.steps / elapsed
    lat_ms = np.array(lat) * 1e3
    log(f"batched: {args.steps} steps in {elapsed:.3f} s")

    # plain: the same scenes, one world per launch
    from resolve2d_amd.engine import GpuWorld, apply_scene
    plain_body_steps, plain_t = 0, 0.0
    for sc in specs[:args.plain_worlds]:
        w = GpuWorld(ctx)
        apply_scene(sc, w)
        checked(w.process(dt, S, K), "plain step")
        ctx.sync()
        a = time.perf_counter()
        for _ in range(args.plain_steps):
            checked(w.process(dt, S, K), "plain step")
        ctx.sync()
        plain_t += time.perf_counter() - a
        plain_body_steps += count_bodies(sc) * args.plain_steps
        w.close()
    plain_value = plain_body_steps / plain_t
    log("plain done")

    # roofline of the dominant kernel (solve): algorithmic bytes per launch
    # = 112 B per body (pose, momentum, force, torque, mass, inertia, flags, shape read;
    #   pose, momentum, force, torque, AABB written) + 120 B per manifold
    #   (geometry read once, impulses and effective masses written) + 16 B per solver item
    names = ["broadphase", "narrowphase", "schedule", "solve"]
    per_step = {n: phase_ms[i] / max(phase_steps, 1) for i, n in enumerate(names)}
    dom = max(per_step, key=per_step.get)
    solve_bytes = 112 * n_bodies + 120 * counts["manifolds"] + 16 * counts["items"]
    solve_s = per_step["solve"] / 1e3
    achieved = solve_bytes / solve_s / 1e9 if solve_s > 0 else 0.0

    # configuration 2: one 4,096-body world, full 300-step horizon, one world per launch,
    # reference order vs the coloured schedule
    single = {}
    if not args.no_single:
        for mode, label in ((0, "reference_order"), (1, "colored")):
            w = GpuWorld(ctx)
            apply_scene(scenes.mixed(4096, seed=1, hash_cell=2.0 if args.tuned_hash else None), w)
            if mode:
                w.set_schedule(1)
            ctx.sync()
            lat1 = []
            a0 = time.perf_counter()
            for i in range(args.single_steps):
                a = time.perf_counter()
                checked(w.process(dt, S, K), f"single-world step ({label})")
                lat1.append(time.perf_counter() - a)
                if i % 100 == 99:
                    log(f"single world {label}: {i + 1} steps")
            ctx.sync()
            t1 = time.perf_counter() - a0
            single[label] = {"value": 4097 * args.single_steps / t1, "unit": "body-steps/s",
                             "ms_per_step": t1 / args.single_steps * 1e3,
                             "p50_step_ms": float(np.percentile(np.array(lat1) * 1e3, 50)),
                             "p99_step_ms": float(np.percentile(np.array(lat1) * 1e3, 99))}
            w.close()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        rate, used, tcpu = cpu_baseline(specs, args.steps, args.cpu_budget)
        cpu = {"value": rate, "unit": "body-steps/s", "cores": 1, "kind": "port",
               "sample": f"oracle f32 restatement, {used} worlds spread over the batch x {args.steps} steps "
                         f"({tcpu:.1f} s)"}

    out = {
        "metric": "body-steps/sec batched vs plain on 1 MI355X; p50/p99 step latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes)",
        "config": {"workload": "batched worlds: 1024 worlds x U[50,500] mixed discs/rects, 0-20 distance joints, "
                               "0-10 springs per world; dt=1/60, 4 sub-steps, 4 iterations",
                   "worlds": args.worlds, "bodies": n_bodies, "manifolds_last_step": counts["manifolds"],
                   "schedule": "reference order (bit-exact)", "parallelism": "1 GPU"},
        "p50_step_ms": float(np.percentile(lat_ms, 50)),
        "p99_step_ms": float(np.percentile(lat_ms, 99)),
        "plain": {"value": plain_value, "unit": "body-steps/s",
                  "sample": f"first {args.plain_worlds} worlds, {args.plain_steps} steps each, one world per launch"},
        "batched_vs_plain": value / plain_value,
        "single_world_4096": single,
        "phase_ms_per_step": per_step,
        "roofline": {"kernel": "k_solve", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "dominant_phase": dom},
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
