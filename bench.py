"""Benchmark of the resolve2d step on one MI355X.

Workload (the headline value): 1024 independent worlds, body counts uniform
in [50, 500] (mixed discs / rectangles), each with 0-20 distance joints and
0-10 anchored springs, stepped together (dt = 1/60, 4 sub-steps, 4 solver
iterations).  A step is one Solver.process over every world of the batch;
inputs are resident on the device before timing starts.

The timed region is one call of `steps` steps.  Worlds of at most 1,024
bodies take whole steps per workgroup: one launch of the fused kernel runs
every step of every world; `value` is bodies x steps / the call's wall time,
`p50/p99_step_ms` are the world-steps' latencies (each workgroup's wall
clock).  The launch shape (solve variant, fused or phase kernels) is picked
by the autotuner on copies of the batch.

The printed JSON line also carries: the same scenes stepped one world per
launch ("plain"); the 4,096-body single world in the reference order and in
the coloured schedule; the 65,536-body constrained world; a served Poisson
stream of scenes of 16-8192 bodies (continuous batching, p50/p99 step and
request latency); accuracy (bit equality with the f32 oracle, error against
an f64 run of the same method, maximum penetration, energy drift of the
resting pyramid); a roofline entry for the dominant kernel (its launch timed
with HIP events on the library's stream; HBM traffic from the committed
rocprofv3 counter measurement; an LDS figure); a CPU baseline (the
oracle restatement on a bounded sample, worlds in parallel over host cores);
and the opt-in warm starting (the batch warm started: its throughput against
the cold run; the resting pyramid warm against cold).
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# LDS: 256 CUs x 128 B per clock x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4
# autotuner: samples of 30 steps from step 60 of a copy of the batch; the 3
# fastest timed again over 120 steps from the same state
TUNE_AT, TUNE_STEPS, TUNE_REFINE, TUNE_REFINE_STEPS = 60, 30, 3, 120
G = 9.81
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


# fused kernel shape of each solve variant: <slots, workgroup size, waves per SIMD>
FUSED_SHAPES = {1: "4,256,3", 2: "2,512,3", 3: "3,256,3", 4: "4,256,2", 5: "8,128,2", 7: "4,256,3",
                9: "0,256,3", 10: "0,256,4", 11: "0,192,3", 12: "0,128,2"}


def traffic_for(kernel, variant, args):
    """HBM bytes per launch of the kernel on this workload as counted by
    rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size),
    from the committed measurement; None when there is none for it."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    shape = FUSED_SHAPES.get(variant) if kernel == "k_fused" else None
    name = f"{kernel}<{shape}>" if shape else kernel
    return t.get(f"{name}/worlds{args.worlds}/steps{args.steps}/seed{args.seed}")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def checked(rc, what):
    from resolve2d_amd import _native
    _native.check(rc, what)


def count_bodies(sc):
    return sum(1 for op in sc if op[0] in ("disc", "rect"))


# ----------------------------------------------------------------- CPU baseline
def _cpu_worker(args):
    """One host core: the oracle (f32) on its share of the sampled worlds, each
    over the full horizon, until its time budget is spent."""
    specs, steps, budget = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    body_steps, t_total, used = 0, 0.0, 0
    for sc in specs:
        o = pyoracle.OracleWorld("f32")
        apply_scene(sc, o)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        t_total += time.perf_counter() - t0
        body_steps += count_bodies(sc) * steps
        used += 1
        o.close()
        if t_total > budget:
            break
    return body_steps, t_total, used


def cpu_baseline(specs, steps, budget_s, cores):
    """Worlds in parallel over `cores` host processes (started before the GPU is
    touched); each process steps whole worlds spread over the batch."""
    order = list(range(0, len(specs), 7)) + [i for i in range(len(specs)) if i % 7]
    shares = [[specs[order[j]] for j in range(k, len(order), cores)] for k in range(cores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(sh, steps, budget_s) for sh in shares])
    wall = time.perf_counter() - t0
    body_steps = sum(r[0] for r in res)
    busy = max(r[1] for r in res)
    return body_steps / busy, sum(r[2] for r in res), busy, wall


# ------------------------------------------------------------------- accuracy
def energy(b):
    """Kinetic + gravitational potential energy of the dynamic bodies of a
    body array (engine layout: static col 2, pos 3-4, momentum 5-6, mass 9,
    angular momentum 11, inertia 13)."""
    dyn = b[:, 2] == 0
    m, I = b[dyn, 9], b[dyn, 13]
    p2 = b[dyn, 5] ** 2 + b[dyn, 6] ** 2
    return float(np.sum(0.5 * p2 / m + 0.5 * b[dyn, 11] ** 2 / I + m * G * b[dyn, 4]))


def state_error(gpu, ref):
    """Max position / velocity error of gpu against ref (engine body arrays),
    absolute and relative to the largest magnitude in ref."""
    dyn = ref[:, 2] == 0
    dp = np.abs(gpu[dyn, 3:5] - ref[dyn, 3:5]).max(initial=0.0)
    v_g = gpu[dyn, 5:7] / gpu[dyn, 9:10]
    v_r = ref[dyn, 5:7] / ref[dyn, 9:10]
    dv = np.abs(v_g - v_r).max(initial=0.0)
    return dict(max_abs_pos_m=float(dp), max_abs_vel_ms=float(dv),
                rel_pos=float(dp / max(1.0, np.abs(ref[dyn, 3:5]).max(initial=0.0))),
                rel_vel=float(dv / max(1e-6, np.abs(v_r).max(initial=0.0))))


def max_penetration(man):
    """Deepest contact (m) among the valid points of a manifold array."""
    best = 0.0
    for p in range(2):
        base = 7 + 13 * p
        valid = man[:, base] != 0
        if valid.any():
            best = max(best, float(-man[valid, base + 9].min()))
    return best


def accuracy(ctx, specs, sample, steps, snaps):
    """The sampled worlds stepped on the GPU (batched together) against the
    oracle in f32 (must be bit-equal) and in f64 (error of the method in f32)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    gw = [GpuWorld(ctx) for _ in sample]
    o32 = [pyoracle.OracleWorld("f32") for _ in sample]
    o64 = [pyoracle.OracleWorld("f64") for _ in sample]
    for i, idx in enumerate(sample):
        for e in (gw[i], o32[i], o64[i]):
            apply_scene(specs[idx], e)
    res = {}
    bit_equal = True
    for s in range(1, steps + 1):
        ctx.process_batch_handles([w.handle for w in gw], scenes.DT, scenes.SUB_STEPS, scenes.ITERS, 1)
        for i in range(len(sample)):
            o32[i].process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
            o64[i].process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        if s in snaps:
            errs = []
            for i in range(len(sample)):
                g = gw[i].bodies()
                a32 = o32[i].bodies()
                bit_equal &= bool((g.astype(np.float32).view(np.uint32) ==
                                   a32.astype(np.float32).view(np.uint32)).all())
                errs.append(state_error(g, o64[i].bodies()))
            res[f"step_{s}"] = {k: max(e[k] for e in errs) for k in errs[0]}
    pen = max(max_penetration(w.manifolds()) for w in gw)
    for e in gw + o32 + o64:
        e.close()
    return res, bit_equal, pen


def cpu_single(ops, steps, budget_s):
    """The f32 oracle (the reference's algorithm restated in C, single
    threaded like the reference) stepping one world on one host thread, from
    the start, for `steps` steps or until budget_s has passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    o = pyoracle.OracleWorld("f32")
    apply_scene(ops, o)
    n = o.bodies().shape[0]
    done, t0 = 0, time.perf_counter()
    while done < steps and time.perf_counter() - t0 < budget_s:
        o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        done += 1
    t = time.perf_counter() - t0
    o.close()
    return {"value": n * done / t, "unit": "body-steps/s", "ms_per_step": t / done * 1e3, "cores": 1,
            "kind": "port", "sample": f"oracle f32 restatement, 1 thread, the first {done} of the {steps} steps"}


def pyramid_energy(ctx, steps):
    """Energy drift of the resting 20-row pyramid (configuration 1) over the
    horizon: GPU f32 against the f64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    g = GpuWorld(ctx)
    o = pyoracle.OracleWorld("f64")
    apply_scene(scenes.pyramid(20), g)
    apply_scene(scenes.pyramid(20), o)
    e0 = energy(g.bodies())
    worst, gpu_s = 0.0, 0.0
    for s in range(1, steps + 1):
        a = time.perf_counter()
        checked(g.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), "pyramid step")
        gpu_s += time.perf_counter() - a
        o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        if s % 20 == 0 or s == steps:
            worst = max(worst, abs(energy(g.bodies()) - e0) / abs(e0))
    b, r = g.bodies(), o.bodies()
    out = dict(steps=steps, drift=(energy(b) - e0) / abs(e0), max_abs_drift=worst,
               drift_f64=(energy(r) - e0) / abs(e0), max_penetration_m=max_penetration(g.manifolds()),
               vs_f64=state_error(b, r), gpu_ms_per_step=gpu_s / steps * 1e3,
               cpu_single=cpu_single(scenes.pyramid(20), steps, 5.0))
    g.close()
    o.close()
    return out


def warm_start_runs(ctx, specs, args, cold_value, factor=0.8):
    """Opt-in warm starting (r2d_set_warm_start): the batched workload with
    every world warm started (same warm-up and timed steps as the cold run
    above, so the ratio is the feature's cost), and the resting 20-row
    pyramid over 600 steps warm against cold: deepest contact, energy drift,
    contacts kept, GPU time per step and bit equality with the f32 oracle's
    warm starting."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    dt, S, K = scenes.DT, scenes.SUB_STEPS, scenes.ITERS
    worlds = []
    for sc in specs:
        w = GpuWorld(ctx)
        apply_scene(sc, w)
        w.set_warm_start(factor)
        worlds.append(w)
    n_bodies = sum(count_bodies(sc) for sc in specs)
    batch = ctx.prepare_batch([w.handle for w in worlds])
    for _ in range(args.warmup):
        batch.process(dt, S, K, 1)
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    ctx.sync()
    t0 = time.perf_counter()
    batch.process(dt, S, K, args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kernel_ms, launches, _ = ctx.fused_times()
    ctx.set_profiling(False)
    value = n_bodies * args.steps / elapsed
    for w in worlds:
        w.close()
    out = {"factor": factor, "batched": {"value": value, "unit": "body-steps/s", "vs_cold": value / cold_value,
                                         "fused_launches": launches, "kernel_ms": kernel_ms,
                                         "note": "every world warm started; same warm-up and timed steps as the "
                                                 "cold headline run"}}
    pyr = {}
    for f in (0.0, factor):
        g = GpuWorld(ctx)
        o = pyoracle.OracleWorld("f32")
        for e in (g, o):
            apply_scene(scenes.pyramid(20), e)
            if f:
                e.set_warm_start(f)
        e0 = energy(g.bodies())
        pen, worst, gpu_s = 0.0, 0.0, 0.0
        for s in range(1, 601):
            a = time.perf_counter()
            checked(g.process(dt, S, K), "pyramid step")
            gpu_s += time.perf_counter() - a
            checked(o.process(dt, S, K), "oracle pyramid step")
            if s % 20 == 0:
                pen = max(pen, max_penetration(g.manifolds()))
                worst = max(worst, abs(energy(g.bodies()) - e0) / abs(e0))
        b = g.bodies()
        pyr["warm" if f else "cold"] = {
            "max_penetration_m": pen, "drift": (energy(b) - e0) / abs(e0), "max_abs_drift": worst,
            "manifolds_at_600": int(g.manifolds().shape[0]), "gpu_ms_per_step": gpu_s / 600 * 1e3,
            "bit_equal_to_f32_oracle": bool(np.array_equal(b.astype(np.float32).view(np.uint32),
                                                           o.bodies().astype(np.float32).view(np.uint32)))}
        g.close()
        o.close()
    out["pyramid_600"] = pyr
    out["pyramid_600"]["note"] = ("20-row pyramid (210 boxes, ground), 600 steps; penetration and drift sampled "
                                  "every 20 steps; 590 contacts at rest")
    return out


def large_world(ctx, steps, snaps, with_f64):
    """The large constrained world (65,536 bodies, ropes, motors, walls): GPU
    step time over the horizon and the error against an f64 run of the same
    method at the snapshot steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    ops = scenes.large_constrained()
    g = GpuWorld(ctx)
    apply_scene(ops, g)
    n = g.bodies().shape[0]
    ctx.sync()
    shots, lat, t = {}, [], 0.0
    for s in range(1, steps + 1):
        a = time.perf_counter()
        checked(g.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), "large-world step")
        lat.append(time.perf_counter() - a)
        if s in snaps:
            shots[s] = g.bodies()
    man = g.manifolds()
    g.close()
    out = {"bodies": n, "steps": steps, "ms_per_step": float(np.mean(lat) * 1e3),
           "value": n * steps / float(np.sum(lat)), "unit": "body-steps/s",
           "p50_step_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
           "p99_step_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
           "max_penetration_m": max_penetration(man)}
    if with_f64:
        o = pyoracle.OracleWorld("f64")
        apply_scene(ops, o)
        errs = {}
        for s in range(1, steps + 1):
            o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
            if s in snaps:
                errs[f"step_{s}"] = state_error(shots[s], o.bodies())
        o.close()
        out["vs_f64_oracle"] = errs
    out["cpu_single"] = cpu_single(ops, steps, 8.0)
    return out


# --------------------------------------------------------------------- serving
def served_arrivals(capacity, seconds, load, seed, steps):
    """A Poisson stream of scenes (16-8192 bodies, log-uniform) offering `load`
    x `capacity` body-steps/s for `seconds`."""
    from resolve2d_amd import scenes
    mean_n = (8192 - 16) / math.log(8192 / 16)
    rate = load * capacity / (mean_n * steps)
    arr = scenes.poisson_stream(seed, rate, max(1, int(rate * seconds)))
    return rate, [(t, (n, s)) for t, n, s in arr]


def serve_stream(ctx, rate, arrivals, steps, load, per_tick):
    """Serve a stream with continuous batching; returns its figures."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld
    from resolve2d_amd.serve import Server, run_stream

    def build(payload):
        n, s = payload
        return scenes.build_served(GpuWorld(ctx), s, n)

    def done(payload, w, st):
        w.close()

    sv = Server(ctx, hbm_budget_bytes=64 << 30, max_worlds=4096, steps_per_tick=per_tick)
    try:
        r = run_stream(sv, arrivals, build, steps, on_done=done)
    finally:
        sv.close()
    st = r["stats"]
    return {"value": st["body_steps"] / r["wall_s"], "unit": "body-steps/s", "offered_load": load,
            "rate_scenes_per_s": rate, "scenes": st["completed"], "steps_per_scene": steps, "wall_s": r["wall_s"],
            "p50_step_ms": st["p50_step_ms"], "p99_step_ms": st["p99_step_ms"],
            "p50_batch_ms": st["p50_batch_ms"], "p99_batch_ms": st["p99_batch_ms"],
            "p50_request_ms": r.get("request_p50_ms"), "p99_request_ms": r.get("request_p99_ms"),
            "mean_scenes_per_batch": st["scene_steps"] / max(1, st["batch_steps"]), "lanes": st["lanes"],
            "steps_per_launch": per_tick,
            "peak_resident_bytes": st["peak_resident_bytes"]}


def plain_served(ctx, arrivals, steps, count):
    """The same kind of scenes stepped one world per launch over the full
    horizon: `count` scenes spread over the stream's sizes."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld
    pick = sorted(arrivals, key=lambda a: a[1][0])
    pick = [pick[int(i * (len(pick) - 1) / max(1, count - 1))] for i in range(count)]
    bs, t = 0, 0.0
    for _, (n, s) in pick:
        w = scenes.build_served(GpuWorld(ctx), s, n)
        ctx.sync()
        a = time.perf_counter()
        for _ in range(steps):
            checked(w.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), "plain served step")
        ctx.sync()
        t += time.perf_counter() - a
        bs += (n + 1) * steps
        w.close()
    return bs / t, [a[1][0] for a in pick]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--plain-worlds", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=0, help="plain steps per world (0: the full horizon)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--cpu-cores", type=int, default=0, help="host processes for the CPU baseline (0: up to 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--single-steps", type=int, default=300)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--serve-sat-seconds", type=float, default=4.0,
                    help="also serve an overloaded stream (3x capacity) for this long (0: skip)")
    ap.add_argument("--serve-plain-scenes", type=int, default=6)
    ap.add_argument("--serve-steps-per-tick", type=int, default=8,
                    help="steps a serving lane's launch takes (fewer when a running scene has fewer left)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-starting runs")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--call-steps", type=int, default=40, help="one-step calls timed after the timed region")
    ap.add_argument("--tuned-hash", action="store_true")
    ap.add_argument("--variant", type=int, default=-1, help="solve-kernel variant (-1: library default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB between two device buffers (a kernel of known HBM traffic, for "
                         "scaling rocprofv3's FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        assert bool(torch.equal(dst, src))
        print(json.dumps({"copy_mib": args.calibrate_copy, "launches": 3, "bytes_read_per_launch": n * 4,
                          "bytes_written_per_launch": n * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: worlds are independent and would be sharded whole across GPUs "
              f"with no exchange; --gpus {args.gpus} is not run here.", file=sys.stderr)
        sys.exit(2)

    from resolve2d_amd import scenes
    t0 = time.perf_counter()
    specs = [scenes.batched_world(args.seed + i) for i in range(args.worlds)]
    n_bodies = sum(count_bodies(sc) for sc in specs)
    log(f"generated {args.worlds} worlds, {n_bodies} bodies in {time.perf_counter() - t0:.1f} s")

    # CPU baseline first: its worker processes start before anything touches the GPU
    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or max(1, min(16, os.cpu_count() or 1))
        log(f"cpu baseline on {cores} processes")
        rate, used, busy, wall = cpu_baseline(specs, args.steps, args.cpu_budget, cores)
        cpu = {"value": rate, "unit": "body-steps/s", "cores": cores, "kind": "port",
               "sample": f"oracle f32 restatement (the reference's algorithm in C), {used} worlds spread over "
                         f"the batch, each x {args.steps} steps, worlds in parallel over {cores} processes "
                         f"({busy:.1f} s per process)"}

    from resolve2d_amd.engine import GpuWorld, apply_scene
    from resolve2d_amd.solver import Context
    ctx = Context(0)
    ctx.set_solve_variant(args.variant)
    t0 = time.perf_counter()
    worlds = []
    for sc in specs:
        w = GpuWorld(ctx)
        apply_scene(sc, w)
        worlds.append(w)
    handles = [w.handle for w in worlds]
    setup_s = time.perf_counter() - t0
    dt, S, K = scenes.DT, scenes.SUB_STEPS, scenes.ITERS
    log(f"built {args.worlds} device worlds in {setup_s:.1f} s")
    batch = ctx.prepare_batch(handles)
    for _ in range(args.warmup):
        batch.process(dt, S, K, 1)
    tuned = None
    if args.variant < 0 and not args.no_tune:
        # launch shape picked on copies of the batch advanced to step 60 (every
        # shape gives the same bits, so the copies change nothing measured):
        # each candidate runs one 30-step call from that same state
        from resolve2d_amd import autotune

        def fresh_copy():
            copy = []
            for sc in specs:
                w = GpuWorld(ctx)
                apply_scene(sc, w)
                copy.append(w)
            ctx.process_batch_handles([w.handle for w in copy], dt, S, K, TUNE_AT)
            return copy

        cands = autotune.candidates(lds_limits=(0,))
        tuned = autotune.tune(ctx, None, dt, S, K, rounds=1, cands=cands, cache_path=None,
                              steps_per_sample=TUNE_STEPS, make_batch=fresh_copy, refine=TUNE_REFINE,
                              refine_steps=TUNE_REFINE_STEPS)
        log(f"autotune: variant {tuned['variant']} fused {tuned['fused']} ({tuned['step_ms']:.3f} ms per step)")
    ctx.sync()

    # timed region: exactly `steps` batched steps in one call.  Worlds of at
    # most 1,024 bodies take whole steps per workgroup: one fused kernel runs
    # every step of every world, with HIP events around the launch and each
    # workgroup's wall clock per step (the world-step latency).  Otherwise the
    # hipGraph of a step is replayed back to back with HIP events between the
    # phases of every step on the library's stream.
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    ctx.sync()
    t_start = time.perf_counter()
    batch.process(dt, S, K, args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    phase_ms, phase_steps = ctx.phase_times()
    fused_ms, fused_launches, world_steps = ctx.fused_times()
    fused_phases = ctx.fused_phase_times()
    work = ctx.fused_counts()
    lat_ms = np.array(ctx.step_times())
    counts = ctx.last_counts()
    ctx.set_profiling(False)
    value = n_bodies * args.steps / elapsed
    log(f"batched: {args.steps} steps in {elapsed:.3f} s ({value / 1e6:.1f} M body-steps/s)")
    # interactive use: one step per call, host work and synchronisation included
    call_lat = []
    for _ in range(args.call_steps):
        a = time.perf_counter()
        batch.process(dt, S, K, 1)
        call_lat.append((time.perf_counter() - a) * 1e3)
    batch_pen = max(max_penetration(w.manifolds()) for w in worlds[:: max(1, len(worlds) // 64)])
    for w in worlds:
        w.close()

    # plain: the same scenes, one world per launch, over the same horizon
    plain_steps = args.plain_steps or args.steps
    plain_body_steps, plain_t = 0, 0.0
    for sc in specs[:: max(1, len(specs) // args.plain_worlds)][:args.plain_worlds]:
        w = GpuWorld(ctx)
        apply_scene(sc, w)
        ctx.sync()
        a = time.perf_counter()
        for _ in range(plain_steps):
            checked(w.process(dt, S, K), "plain step")
        ctx.sync()
        plain_t += time.perf_counter() - a
        plain_body_steps += count_bodies(sc) * plain_steps
        w.close()
    plain_value = plain_body_steps / plain_t
    log("plain done")

    # roofline of the dominant kernel.  Fused: k_fused, one launch for the
    # whole timed region; its algorithmic HBM bytes are the step's
    # 72 B per body (pose, momentum, angle read; mass, inertia, mu, shape,
    # flags read; pose, momentum written) + 16 B per cell entry (written and
    # read once) + 8 B per raw candidate pair + 176 B per manifold (written by
    # the narrowphase, read by the solve), summed over the launch's
    # world-steps; its LDS bytes are S x K x (64 B per manifold: two bodies'
    # momentum and inverse masses read, momentum written; 48 B per joint).
    # Phase kernels: the solve phase, 112 B per body + 120 B per manifold +
    # 16 B per solver item.
    joint_items = max(0, work["items"] - work["manifolds"])
    if fused_launches:
        kernel = "k_fused"
        launch_s = fused_ms / fused_launches / 1e3
        launch_bytes = (72 * work["bodies"] + 16 * work["entries"] + 8 * work["raw_candidates"] +
                        176 * work["manifolds"]) / fused_launches
        lds_bytes = S * K * (64 * work["manifolds"] + 48 * joint_items) / fused_launches
        per_step = {k: v / max(world_steps, 1) for k, v in fused_phases.items()}
        per_step_note = "per world-step, workgroup wall clock"
    else:
        kernel = "k_solve"
        names = ["broadphase", "narrowphase", "schedule", "solve"]
        per_step = {n: phase_ms[i] / max(phase_steps, 1) for i, n in enumerate(names)}
        launch_s = per_step["solve"] / 1e3
        launch_bytes = 112 * n_bodies + 120 * counts["manifolds"] + 16 * counts["items"]
        lds_bytes = S * K * (64 * counts["manifolds"] + 48 * max(0, counts["items"] - counts["manifolds"]))
        per_step_note = "per step, HIP events between phase launches"
    dom = max(per_step, key=per_step.get)
    used_variant = tuned["variant"] if tuned else (args.variant if args.variant >= 0 else 1)
    achieved = launch_bytes / launch_s / 1e9 if launch_s > 0 else 0.0
    lds_achieved = lds_bytes / launch_s / 1e9 if launch_s > 0 else 0.0

    # configuration 2: one 4,096-body world, full horizon, one world per launch,
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
        single["cpu_single"] = cpu_single(scenes.mixed(4096, seed=1, hash_cell=2.0 if args.tuned_hash else None),
                                          args.single_steps, 6.0)

    large = None
    warm = None
    if not args.no_warm:
        log("warm starting")
        warm = warm_start_runs(ctx, specs, args, value)
        log(f"warm starting: batched {warm['batched']['value'] / 1e6:.1f} M body-steps/s "
            f"({warm['batched']['vs_cold']:.2f}x cold)")
    if not args.no_large:
        log("large constrained world")
        large = large_world(ctx, 120, {1, 10, 120}, not args.no_accuracy)

    acc = None
    if not args.no_accuracy:
        log("accuracy")
        sample = list(range(0, args.worlds, max(1, args.worlds // 4)))[:4]
        horizon = args.steps
        errs, bit_eq, pen = accuracy(ctx, specs, sample, horizon, {1, 60, horizon})
        acc = {"worlds": sample, "bit_equal_to_f32_oracle": bit_eq, "tolerance_vs_reference_f32": 1e-5,
               "vs_f64_oracle": errs, "max_penetration_m": pen, "batch_max_penetration_m": batch_pen,
               "pyramid_600": pyramid_energy(ctx, 600)}

    serve = None
    if not args.no_serve:
        log("serving stream")
        rate, arrivals = served_arrivals(value, args.serve_seconds, args.serve_load, args.seed, args.steps)
        serve = serve_stream(ctx, rate, arrivals, args.steps, args.serve_load, args.serve_steps_per_tick)
        log(f"served {serve['scenes']} scenes, {serve['value'] / 1e6:.1f} M body-steps/s")
        if args.serve_sat_seconds > 0:
            rate2, arr2 = served_arrivals(value, args.serve_sat_seconds, 3.0, args.seed + 1, args.steps)
            sat = serve_stream(ctx, rate2, arr2, args.steps, 3.0, args.serve_steps_per_tick)
            serve["saturated"] = {k: sat[k] for k in ("value", "offered_load", "scenes", "wall_s", "p50_step_ms",
                                                      "p99_step_ms", "p50_request_ms", "p99_request_ms")}
            log(f"saturated: {sat['value'] / 1e6:.1f} M body-steps/s")
        pv, sizes = plain_served(ctx, arrivals, args.steps, args.serve_plain_scenes)
        serve["plain"] = {"value": pv, "unit": "body-steps/s",
                          "sample": f"{len(sizes)} served scenes of {sizes} bodies, {args.steps} steps each, one "
                                    f"world per launch"}
        serve["served_vs_plain"] = serve["value"] / pv
        if "saturated" in serve:
            serve["saturated_vs_plain"] = serve["saturated"]["value"] / pv

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
                   "schedule": "reference order (bit-exact)", "parallelism": "1 GPU",
                   "solve_launch": ({"autotuned": True, "variant": tuned["variant"], "fused": tuned["fused"],
                                     "tuned_at_step": TUNE_AT, "sample_steps": TUNE_STEPS,
                                     "candidates_step_ms (variant/lds_limit/fused)": tuned["all_ms"],
                                     "refine_steps": TUNE_REFINE_STEPS,
                                     "refined_step_ms": tuned.get("refined_ms")}
                                    if tuned else {"autotuned": False, "variant": args.variant})},
        "p50_step_ms": float(np.percentile(lat_ms, 50)) if len(lat_ms) else None,
        "p99_step_ms": float(np.percentile(lat_ms, 99)) if len(lat_ms) else None,
        "p999_step_ms": float(np.percentile(lat_ms, 99.9)) if len(lat_ms) else None,
        "max_step_ms": float(lat_ms.max()) if len(lat_ms) else None,
        "step_latency": ("each world's step, wall clock of its workgroup in the fused kernel (all world-steps of the "
                         "timed region)" if fused_launches else
                         "device time per step from HIP events (the timed region's steps)"),
        "per_call": {"p50_ms": float(np.percentile(call_lat, 50)) if call_lat else None,
                     "p99_ms": float(np.percentile(call_lat, 99)) if call_lat else None,
                     "calls": len(call_lat), "note": "one step per r2d_process_batch call after the timed "
                                                     "region, host preparation and synchronisation included"},
        "plain": {"value": plain_value, "unit": "body-steps/s",
                  "sample": f"{args.plain_worlds} worlds spread over the batch, {plain_steps} steps each, "
                            f"one world per launch"},
        "batched_vs_plain": value / plain_value,
        "served": serve,
        "single_world_4096": single,
        "large_world_65536": large,
        "accuracy": acc,
        "warm_start": warm,
        "mode": ("fused: one kernel runs every step of every world, one workgroup per world" if fused_launches else
                 "phase kernels: one launch sequence (hipGraph) per step"),
        "phase_ms": per_step,
        "phase_ms_note": per_step_note,
        "roofline": {"kernel": (f"{kernel}<{FUSED_SHAPES[used_variant]}>" if kernel == "k_fused" and
                                used_variant in FUSED_SHAPES else kernel), "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_for(kernel, used_variant, args),
                     "launch_ms": launch_s * 1e3, "algorithmic_bytes_per_launch": launch_bytes,
                     "lds": {"achieved": lds_achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                             "frac": lds_achieved / LDS_PEAK_GBS, "bytes_per_launch": lds_bytes},
                     "dominant_phase": dom},
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
