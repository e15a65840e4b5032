"""Benchmark of the resolve2d step on one MI355X; prints one JSON line.

Headline (configuration 3): 1024 independent worlds, body counts uniform in
[50, 500] (mixed discs / rectangles), each with 0-20 distance joints and 0-10
anchored springs, stepped together over a 240-step horizon (dt = 1/60,
4 sub-steps, 4 solver iterations).  A step is one Solver.process over every
world of the batch; inputs are resident on the device before timing starts.

The timed region is one call of `--steps` K steps taken from the settled end
of the horizon: the batch is first advanced max(W, 240 - K) steps untimed
(W = --warmup), so with K = 20 the window is steps 221-240, where the piles
hold the most contacts.  Worlds of at most 1,024 bodies take whole steps per
workgroup: one launch of the fused kernel runs every step of every world of
the call; `value` is bodies x K / the call's wall time; `p50/p99_step_ms`
are the world-steps' latencies (each workgroup's wall clock).  `horizon`
reports the whole 240-step job (a fresh batch, all 240 steps in one call).
The launch shape is picked by the autotuner on copies of the batch at the
window's start.

Also in the line: the same scenes one world per launch over the same window
and over the horizon ("plain", the reference's one-world process call); the
4,096-body world (configuration 2, tuned hash of cell 2x the largest body
extent) in the reference order and the coloured schedule with its FP64
error, penetration and energy change; the 65,540-body constrained world
(configuration 4) with its FP64 error at steps 1, 10, 120; a served Poisson
stream of 240-step scenes of 16-8192 bodies (configuration 5: capacity from
an overloaded stream, then a stream at 70 % of it; p50/p99 step and request
latency; accuracy of sampled scenes); accuracy of the batch (bit equality
with the f32 oracle, FP64 error, penetration) and of the resting 20-row
pyramid (configuration 1: energy drift); the roofline of the dominant kernel
(its launch timed with HIP events on the library's stream; HBM traffic from
the committed rocprofv3 counters; the HBM copy rate measured in the same
run); a CPU baseline (the oracle restatement, worlds in parallel over the
host cores, median of 5 passes, for configurations 3 and 5); and the opt-in
warm starting.
"""
import argparse
import ctypes
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
# LDS: 256 CUs x 256 B per clock x 2.4 GHz.  gfx950 serves 128-bit LDS reads
# at twice the 128 B per clock of gfx942: the library's own LDS read
# microbenchmark (r2d_debug_peaks) measures ~200 B per clock per CU
LDS_PEAK_GBS = 256 * 256 * 2.4
HORIZON = 240           # steps of a configuration-3 world and of a served scene
G = 9.81
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
# host processes at most: the host CPU share that comes with one GPU of an MI355X node
CPU_SHARE = 16

# fused kernel shape of each solve variant: <slots, workgroup size, waves per SIMD>
FUSED_SHAPES = {1: "4,256,3", 2: "2,512,3", 3: "3,256,3", 4: "4,256,2", 5: "8,128,2", 7: "4,256,3",
                9: "0,256,3", 10: "0,256,4", 11: "0,192,3", 12: "0,128,2", 13: "0,512,2",
                14: "0,768,3"}


def traffic_for(name, window, args):
    """HBM bytes per launch of the kernel on this workload and window as
    counted by rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size), from the committed measurement; None when there is none."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get(f"{name}/worlds{args.worlds}/steps{window[0]}-{window[1]}/seed{args.seed}")


LDS_FILE = os.path.join(ROOT, "profiles", "lds.json")


def lds_counted_for(name, window, args):
    """LDS instructions of one launch of the kernel on this workload and
    window as counted by rocprofv3 (SQ_INSTS_LDS) and the bytes they move at
    least, from the committed measurement; None when there is none."""
    try:
        with open(LDS_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get(f"{name}/worlds{args.worlds}/steps{window[0]}-{window[1]}/seed{args.seed}")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def checked(rc, what):
    from resolve2d_amd import _native
    _native.check(rc, what)


def count_bodies(sc):
    return sum(1 for op in sc if op[0] in ("disc", "rect"))


def host_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# ----------------------------------------------------------------- CPU baseline
def _cpu_worker(job):
    """One host process: the f32 oracle stepping its share of the sample, every
    world from its start over `steps` steps, `repeats` times; returns
    (body-steps, seconds) per pass (world construction untimed)."""
    specs, steps, repeats = job
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    out = []
    for _ in range(repeats):
        bs, t = 0, 0.0
        for sc in specs:
            o = pyoracle.OracleWorld("f32")
            apply_scene(sc, o)
            a = time.perf_counter()
            for _ in range(steps):
                o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
            t += time.perf_counter() - a
            bs += count_bodies(sc) * steps
            o.close()
        out.append((bs, t))
    return out


def cpu_baseline(specs, steps, cores, repeats):
    """The sample's worlds in parallel over `cores` host processes (started
    before the GPU is touched; shares balanced by body count), `repeats`
    passes; a pass's rate is its body-steps over its slowest process's time.
    Returns the median pass rate and every pass's rate."""
    order = sorted(range(len(specs)), key=lambda i: -count_bodies(specs[i]))
    shares, load = [[] for _ in range(cores)], [0] * cores
    for i in order:
        k = int(np.argmin(load))
        shares[k].append(specs[i])
        load[k] += count_bodies(specs[i])
    shares = [s for s in shares if s]
    with mp.get_context("fork").Pool(len(shares)) as pool:
        res = pool.map(_cpu_worker, [(sh, steps, repeats) for sh in shares])
    rates = []
    for p in range(repeats):
        bs = sum(r[p][0] for r in res)
        t = max(r[p][1] for r in res)
        rates.append(bs / t)
    return float(np.median(rates)), rates, len(shares)


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


def bits_equal(a, b):
    return bool(np.array_equal(a.astype(np.float32).view(np.uint32), b.astype(np.float32).view(np.uint32)))


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
                bit_equal &= bits_equal(g, o32[i].bodies())
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
    worst, gpu_s, lat = 0.0, 0.0, []
    for s in range(1, steps + 1):
        a = time.perf_counter()
        checked(g.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), "pyramid step")
        lat.append(time.perf_counter() - a)
        o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        if s % 20 == 0 or s == steps:
            worst = max(worst, abs(energy(g.bodies()) - e0) / abs(e0))
    b, r = g.bodies(), o.bodies()
    lat = np.array(lat) * 1e3
    out = dict(steps=steps, drift=(energy(b) - e0) / abs(e0), max_abs_drift=worst,
               drift_f64=(energy(r) - e0) / abs(e0), max_penetration_m=max_penetration(g.manifolds()),
               vs_f64=state_error(b, r), gpu_ms_per_step=float(lat.mean()),
               gpu_p50_step_ms=float(np.percentile(lat, 50)), gpu_p99_step_ms=float(np.percentile(lat, 99)),
               cpu_single=cpu_single(scenes.pyramid(20), steps, 5.0))
    g.close()
    o.close()
    return out


def single_world(ctx, steps, tuned_hash):
    """Configuration 2: the 4,096-body world over its horizon, one world per
    call, in the reference order and in the coloured schedule; the tuned
    spatial hash (cell = twice the largest body extent, 2N buckets) unless
    the reference hash is asked for.  Each mode: throughput, per-call p50/p99,
    its error against the f64 oracle run of the same mode at steps 1, 60 and
    the last, the deepest contact and the change of mechanical energy."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    ops = scenes.mixed(4096, seed=1, hash_cell="auto" if tuned_hash else None)
    out = {"hash": ({"mode": "tuned", "cell_width_m": ops[0][1], "table": "2N"} if tuned_hash else
                    {"mode": "reference", "cell_width_m": 4.0, "table": "2N"})}
    snaps = {1, 60, steps}
    for mode, label in ((0, "reference_order"), (1, "colored")):
        w = GpuWorld(ctx)
        o = pyoracle.OracleWorld("f64")
        for e in (w, o):
            apply_scene(ops, e)
            if mode:
                e.set_schedule(1)
        e0 = energy(w.bodies())
        n = w.bodies().shape[0]
        ctx.sync()
        lat, errs, pen, o_steps = [], {}, 0.0, 0
        for i in range(1, steps + 1):
            a = time.perf_counter()
            checked(w.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), f"single-world step ({label})")
            lat.append(time.perf_counter() - a)
            if i in snaps:
                g = w.bodies()
                while o_steps < i:
                    o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
                    o_steps += 1
                errs[f"step_{i}"] = state_error(g, o.bodies())
                pen = max(pen, max_penetration(w.manifolds()))
        t1 = float(np.sum(lat))
        lat = np.array(lat) * 1e3
        out[label] = {"value": n * steps / t1, "unit": "body-steps/s", "ms_per_step": t1 / steps * 1e3,
                      "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
                      "vs_f64_oracle": errs, "max_penetration_m": pen,
                      "energy_change": (energy(w.bodies()) - e0) / abs(e0)}
        log(f"single world {label}: {t1 / steps * 1e3:.3f} ms per step")
        w.close()
        o.close()
    out["cpu_single"] = cpu_single(ops, steps, 6.0)
    out["note"] = ("one r2d_process call per step, host preparation and synchronisation included; energy_change "
                   "is the relative change of kinetic + potential energy of a pile falling onto the ground "
                   "(collisions dissipate it)")
    return out


def warm_start_runs(ctx, specs, window, cold_value, factor=0.8):
    """Opt-in warm starting (r2d_set_warm_start): the batched workload with
    every world warm started, timed over the same window as the cold run, and
    the resting 20-row pyramid over 600 steps warm against cold: deepest
    contact, energy drift, contacts kept, GPU time per step and bit equality
    with the f32 oracle's warm starting."""
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
    first, last = window
    batch.process(dt, S, K, first - 1)
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    ctx.sync()
    t0 = time.perf_counter()
    batch.process(dt, S, K, last - first + 1)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kernel_ms, launches, _ = ctx.fused_times()
    ctx.set_profiling(False)
    value = n_bodies * (last - first + 1) / elapsed
    for w in worlds:
        w.close()
    out = {"factor": factor, "batched": {"value": value, "unit": "body-steps/s", "vs_cold": value / cold_value,
                                         "fused_launches": launches, "kernel_ms": kernel_ms,
                                         "note": f"every world warm started; steps {first}-{last} as the cold "
                                                 f"headline window"}}
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
            "bit_equal_to_f32_oracle": bits_equal(b, o.bodies())}
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
    shots, lat = {}, []
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
           "max_penetration_m": max_penetration(man),
           "note": "one r2d_process call per step from the start, host synchronisation per step included"}
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


# ------------------------------------------------------- critical path of a step
def chain_depth(ops, states, manifolds, iters):
    """Longest dependency chain of one sub-step's solver in the reference order
    (lib.zig:226-236: collision_iters passes of the constraints in order, then
    the manifolds in discovery order).  An item waits for every earlier item
    (over the passes) that shares one of its ordering bodies: both bodies of a
    distance / offset joint, the body of a fixed-position / motor joint, and
    those bodies of a manifold that are dynamic or held by a distance / offset
    joint (a contact's zero change of any other static body orders nothing).
    Bodies as scene ids; states: the world's body rows (id col 0, static col 2);
    manifolds: (ref id, inc id) rows.  Returns (items, depth)."""
    static = {int(r[0]) for r in states if r[2]}
    items, jointed = [], set()
    for op in ops:
        if op[0] != "joint":
            continue
        t, a, b = op[1], op[5], op[6]
        if t <= 1:
            items.append((a, b))
            jointed |= {a, b}
        else:
            items.append((a,))
    for ref, inc in manifolds:
        items.append(tuple(x for x in (int(ref), int(inc)) if x not in static or x in jointed))
    last, depth = {}, 0
    for _ in range(iters):
        for bodies in items:
            lv = 1 + max((last.get(x, 0) for x in bodies), default=0)
            for x in bodies:
                last[x] = lv
            depth = max(depth, lv)
    return len(items), depth


def critical_path(ctx, label, ops, warm_steps, timed_steps, cpu_ms):
    """One world stepped one call per step: after warm_steps, timed_steps
    with HIP-event / workgroup-clock profiling, then the chain depth of its
    last step from the world's own manifold readback; the solve's device time
    per step over (sub-steps x depth) is the GPU's time per chain link."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    g = GpuWorld(ctx)
    apply_scene(ops, g)
    for s in range(warm_steps):
        if s == warm_steps - 1:
            ctx.set_profiling(True)   # the timing buffers exist before the timed steps
        checked(g.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), f"{label} step")
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    lat = []
    for _ in range(timed_steps):
        a = time.perf_counter()
        checked(g.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), f"{label} step")
        lat.append(time.perf_counter() - a)
    fused_ms, launches, wsteps = ctx.fused_times()
    if launches:
        solve_ms = ctx.fused_phase_times()["solve"] / max(wsteps, 1)
        kernel_ms = fused_ms / launches
    else:
        ph, n = ctx.phase_times()
        solve_ms = ph[3] / max(n, 1)
        kernel_ms = sum(ph) / max(n, 1)
    ctx.set_profiling(False)
    states, man = g.bodies(), g.manifolds()
    items, depth = chain_depth(ops, states, man[:, :2], scenes.ITERS)
    g.close()
    links = scenes.SUB_STEPS * depth
    return {"bodies": int(states.shape[0]), "items": items, "manifolds": int(man.shape[0]),
            "depth_per_substep": depth, "links_per_step": links,
            "solve_ms_per_step": solve_ms, "device_ms_per_step": kernel_ms,
            "call_ms_per_step": float(np.mean(lat) * 1e3), "ns_per_link": solve_ms * 1e6 / max(links, 1),
            "cpu_ms_per_step": cpu_ms, "at_step": warm_steps + timed_steps,
            "mode": "fused kernel (one workgroup)" if launches else "phase kernels"}


# --------------------------------------------------------------------- serving
MEAN_SERVED_N = (8192 - 16) / math.log(8192 / 16)   # mean of log-uniform [16, 8192]


def served_arrivals(capacity, seconds, load, seed, steps):
    """A Poisson stream of scenes (16-8192 bodies, log-uniform) offering
    `load` x `capacity` body-steps/s for `seconds`."""
    from resolve2d_amd import scenes
    rate = load * capacity / (MEAN_SERVED_N * steps)
    arr = scenes.poisson_stream(seed, rate, max(1, int(rate * seconds)))
    return rate, [(t, (n, s)) for t, n, s in arr]


def saturated_rate(trace, t1, t2):
    """Body-steps per second collected between t1 and t2 seconds of a served
    stream (the slope of its (time, body-steps) trace): an overloaded
    stream's rate while every lane stays busy, without its ramp-up and the
    drain of its last requests."""
    pts = [(t, b) for t, b in trace if t1 <= t <= t2]
    if len(pts) < 2 or pts[-1][0] <= pts[0][0]:
        return None
    return (pts[-1][1] - pts[0][1]) / (pts[-1][0] - pts[0][0])


def serve_stream(ctx, rate, arrivals, steps, load, per_tick, keep=(), trace=None):
    """Serve a stream with continuous batching; returns its figures (and the
    final body / manifold arrays of the scenes whose payloads are in keep)."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld
    from resolve2d_amd.serve import Server, run_stream
    kept = {}

    def build(payload):
        n, s = payload
        return scenes.build_served(GpuWorld(ctx), s, n)

    def done(payload, w, st):
        if payload in keep:
            kept[payload] = w.bodies()
        w.close()

    sv = Server(ctx, hbm_budget_bytes=64 << 30, max_worlds=4096, steps_per_tick=per_tick)
    try:
        r = run_stream(sv, arrivals, build, steps, on_done=done, trace=trace)
    finally:
        sv.close()
    st = r["stats"]
    out = {"value": st["body_steps"] / r["wall_s"], "unit": "body-steps/s", "offered_load": load,
           "rate_scenes_per_s": rate, "scenes": st["completed"], "steps_per_scene": steps, "wall_s": r["wall_s"],
           "admitted_scenes_per_s": st["completed"] / r["wall_s"],
           "p50_step_ms": st["p50_step_ms"], "p99_step_ms": st["p99_step_ms"],
           "p50_batch_ms": st["p50_batch_ms"], "p99_batch_ms": st["p99_batch_ms"],
           "p50_request_ms": r.get("request_p50_ms"), "p99_request_ms": r.get("request_p99_ms"),
           "mean_scenes_per_batch": st["scene_steps"] / max(1, st["batch_steps"]), "lanes": st["lanes"],
           "steps_per_launch": per_tick, "peak_resident_bytes": st["peak_resident_bytes"]}
    return out, kept


def served_accuracy(kept, steps):
    """Served scenes against the oracle over their whole request: bit
    equality with the f32 restatement, error against f64, and the deepest
    contact of the last step (from the f32 restatement's manifolds, whose
    bodies the served world equals bit for bit; a served world's manifolds
    live in its serving lane's scratch, which later steps reuse)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import apply_scene
    rows, bit_eq = [], True
    for (n, s), gb in sorted(kept.items()):
        ops = scenes.served_ops(s, n)
        o32, o64 = pyoracle.OracleWorld("f32"), pyoracle.OracleWorld("f64")
        for o in (o32, o64):
            apply_scene(ops, o)
            for _ in range(steps):
                o.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS)
        eq = bits_equal(gb, o32.bodies())
        bit_eq &= eq
        rows.append({"bodies": n, "bit_equal_to_f32_oracle": eq, "vs_f64_oracle": state_error(gb, o64.bodies()),
                     "max_penetration_m": max_penetration(o32.manifolds())})
        o32.close()
        o64.close()
    return {"scenes": rows, "bit_equal_to_f32_oracle": bit_eq, "steps": steps}


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


# ------------------------------------------------------------- HBM copy rate
def hbm_copy_rate(mib=1024, reps=8):
    """Device-to-device copy of `mib` MiB (hipMemcpy, timed with HIP events):
    bytes read + written per second, the measured HBM rate the roofline is
    also reported against."""
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    P = ctypes.c_void_p
    n = mib << 20
    a, b = P(), P()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        return None
    e0, e1 = P(), P()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    try:
        hip.hipMemset(a, 1, ctypes.c_size_t(n))
        for _ in range(2):
            hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)
        hip.hipDeviceSynchronize()
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    return {"value": 2.0 * n * reps / (ms.value / 1e3) / 1e9, "unit": "GB/s",
            "sample": f"hipMemcpy device to device, {mib} MiB x {reps}, read + write bytes"}


def one_world(args):
    """One configuration's single world over its horizon, one r2d_process
    per step (the plain mode of configurations 1, 2 and 4), for profiling."""
    from resolve2d_amd import scenes
    from resolve2d_amd.engine import GpuWorld, apply_scene
    from resolve2d_amd.solver import Context
    ops, steps = {"config1": (scenes.pyramid(20), 600),
                  "config2": (scenes.mixed(4096, seed=1, hash_cell=None if args.reference_hash else "auto"), 300),
                  "config4": (scenes.large_constrained(), 120)}[args.only_world]
    ctx = Context(0)
    ctx.set_solve_variant(args.variant)
    g = GpuWorld(ctx)
    apply_scene(ops, g)
    ctx.sync()
    lat = []
    for i in range(steps):
        a = time.perf_counter()
        checked(g.process(scenes.DT, scenes.SUB_STEPS, scenes.ITERS), f"{args.only_world} step")
        lat.append(time.perf_counter() - a)
        if i % 100 == 99:
            log(f"{args.only_world}: {i + 1} steps")
    n = g.bodies().shape[0]
    g.close()
    lat = np.array(lat) * 1e3
    print(json.dumps({"world": args.only_world, "bodies": n, "steps": steps, "ms_per_step": float(lat.mean()),
                      "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99))}),
          flush=True)
    ctx.close()


# ---------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps: the last K of the 240-step horizon")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps before the window, at least")
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--plain-worlds", type=int, default=16)
    ap.add_argument("--cpu-cores", type=int, default=0, help="host processes for the CPU baseline (0: all, <= 16)")
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--cpu-worlds", type=int, default=128, help="configuration-3 worlds in the CPU sample")
    ap.add_argument("--cpu-scenes", type=int, default=32, help="configuration-5 scenes in the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--single-steps", type=int, default=300)
    ap.add_argument("--reference-hash", action="store_true",
                    help="configuration 2 with the reference's hash (cell 4.0) instead of the tuned one")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=16.0,
                    help="seconds of arrivals of the served stream at the chosen load (the drain of its last "
                         "requests counts in its wall time, so a longer stream is closer to steady state)")
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--serve-sat-seconds", type=float, default=4.0,
                    help="seconds of arrivals of the overloaded stream that measures the served capacity")
    ap.add_argument("--serve-plain-scenes", type=int, default=6)
    ap.add_argument("--serve-steps-per-tick", type=int, default=8,
                    help="steps a serving lane's launch takes (fewer when a running scene has fewer left)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-critical", action="store_true", help="skip the one-world critical-path measurements")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-starting runs")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--tune-fused-only", action="store_true",
                    help="the autotuner tries the fused shapes only (for counter runs of a profiler: the "
                         "phase-kernel candidates add thousands of dispatches)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline window and the horizon call (for profiling runs)")
    ap.add_argument("--only-world", choices=["config1", "config2", "config4"], default=None,
                    help="only step one configuration's single world over its horizon, one call per step, and "
                         "print its time per step (for profiling runs)")
    ap.add_argument("--call-steps", type=int, default=40, help="one-step calls timed after the horizon")
    ap.add_argument("--variant", type=int, default=-1, help="solve-kernel variant (-1: library default)")
    ap.add_argument("--peaks-only", action="store_true",
                    help="only run the library's peak microbenchmarks (HBM read, copy, LDS) and print them")
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
    if args.peaks_only:
        from resolve2d_amd.solver import Context
        pctx = Context(0)
        print(json.dumps(dict(pctx.measured_peaks(), unit="GB/s")), flush=True)
        pctx.close()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: worlds are independent and would be sharded whole across GPUs "
              f"with no exchange; --gpus {args.gpus} is not run here.", file=sys.stderr)
        sys.exit(2)
    if args.only_world:
        one_world(args)
        return
    if args.headline_only:
        args.no_cpu = args.no_single = args.no_serve = args.no_accuracy = args.no_large = args.no_warm = True
        args.no_critical = True
        args.plain_worlds = 0
    K = max(1, args.steps)
    first = max(args.warmup, HORIZON - K) + 1 if K < HORIZON else args.warmup + 1
    window = (first, first + K - 1)

    from resolve2d_amd import scenes
    t0 = time.perf_counter()
    specs = [scenes.batched_world(args.seed + i) for i in range(args.worlds)]
    n_bodies = sum(count_bodies(sc) for sc in specs)
    log(f"generated {args.worlds} worlds, {n_bodies} bodies in {time.perf_counter() - t0:.1f} s; "
        f"timed window: steps {window[0]}-{window[1]}")

    # CPU baseline first: its worker processes start before anything touches the GPU
    cpu = None
    if not args.no_cpu:
        ncpu = host_cpus()
        cores = args.cpu_cores or min(CPU_SHARE, ncpu)
        stride = max(1, len(specs) // args.cpu_worlds)
        sample = specs[::stride][:args.cpu_worlds]
        log(f"cpu baseline: {len(sample)} worlds x {HORIZON} steps on {cores} processes, {args.cpu_repeats} passes")
        rate, rates, used = cpu_baseline(sample, HORIZON, cores, args.cpu_repeats)
        s_bodies = sum(count_bodies(sc) for sc in sample)
        cpu = {"value": rate, "unit": "body-steps/s", "cores": used, "host_cpus": ncpu, "kind": "port",
               "repeats": args.cpu_repeats, "pass_values": rates,
               "sample": f"oracle f32 restatement (the reference's algorithm in C), {len(sample)} of the "
                         f"{args.worlds} configuration-3 worlds ({s_bodies} bodies, every {stride}th) from their "
                         f"start over the {HORIZON}-step horizon, worlds in parallel over {used} processes (the "
                         f"host's CPU share per GPU is {CPU_SHARE}; the host shows {ncpu}); median of "
                         f"{args.cpu_repeats} passes"}
        # configuration 5: served scenes sampled from the same size law
        rng = np.random.default_rng(args.seed + 7)
        sizes = np.exp(np.linspace(math.log(16), math.log(8192), args.cpu_scenes)).round().astype(int)
        sc5 = [scenes.served_ops(int(rng.integers(0, 2 ** 31)), int(n)) for n in sizes]
        rate5, rates5, used5 = cpu_baseline(sc5, HORIZON, cores, args.cpu_repeats)
        cpu["config5"] = {"value": rate5, "unit": "body-steps/s", "cores": used5, "pass_values": rates5,
                          "sample": f"{len(sc5)} served scenes of 16-8192 bodies (log-spaced), each from its start "
                                    f"over {HORIZON} steps, worlds in parallel; median of {args.cpu_repeats} passes"}
        log(f"cpu baseline: {rate / 1e6:.1f} M (config 3), {rate5 / 1e6:.1f} M (config 5) body-steps/s")

    from resolve2d_amd.engine import GpuWorld, apply_scene
    from resolve2d_amd.solver import Context
    ctx = Context(0)
    ctx.set_solve_variant(args.variant)
    dt, S, K_IT = scenes.DT, scenes.SUB_STEPS, scenes.ITERS

    def build_batch():
        ws = []
        for sc in specs:
            w = GpuWorld(ctx)
            apply_scene(sc, w)
            ws.append(w)
        return ws

    t0 = time.perf_counter()
    worlds = build_batch()
    setup_s = time.perf_counter() - t0
    log(f"built {args.worlds} device worlds in {setup_s:.1f} s")
    batch = ctx.prepare_batch([w.handle for w in worlds])

    tuned = None
    if args.variant < 0 and not args.no_tune:
        # launch shape picked on copies of the batch advanced to the window's
        # start (every shape gives the same bits, so the copies change nothing
        # measured): each candidate times one call over the window's steps,
        # and the three fastest are timed again from the same state
        from resolve2d_amd import autotune

        def fresh_copy():
            copy = build_batch()
            ctx.process_batch_handles([w.handle for w in copy], dt, S, K_IT, window[0] - 1)
            return copy

        cands = [c for c in autotune.candidates(lds_limits=(0,))
                 if c[2] or (c[0] in (1, 11) and not args.tune_fused_only)]
        tuned = autotune.tune(ctx, None, dt, S, K_IT, rounds=1, cands=cands, cache_path=None,
                              steps_per_sample=K, make_batch=fresh_copy, refine=3, refine_steps=K, log=log)
        log(f"autotune: variant {tuned['variant']} fused {tuned['fused']} ({tuned['step_ms']:.3f} ms per step)")

    # timed region: the window, one call of K steps after the untimed advance
    # (profiled too, so the timing buffers exist before the timed launch)
    ctx.set_profiling(True)
    batch.process(dt, S, K_IT, window[0] - 1)
    ctx.sync()
    ctx.reset_phase_times()
    ctx.sync()
    t_start = time.perf_counter()
    batch.process(dt, S, K_IT, K)
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    phase_ms, phase_steps = ctx.phase_times()
    fused_ms, fused_launches, world_steps = ctx.fused_times()
    fused_phases = ctx.fused_phase_times()
    work = ctx.fused_counts()
    lat_ms = np.array(ctx.step_times())
    counts = ctx.last_counts()
    ctx.set_profiling(False)
    value = n_bodies * K / elapsed
    log(f"batched: steps {window[0]}-{window[1]} in {elapsed:.4f} s ({value / 1e6:.1f} M body-steps/s)")
    batch_pen = max(max_penetration(w.manifolds()) for w in worlds[:: max(1, len(worlds) // 64)])
    for w in worlds:
        w.close()

    # the whole job: a fresh batch, all 240 steps in one call.  A call
    # without sub-steps first makes the batch resident on the device: it runs
    # only the broadphase and narrowphase (lib.zig:194), which change no body
    # state, so the 240 timed steps start from the built state
    worlds = build_batch()
    batch = ctx.prepare_batch([w.handle for w in worlds])
    batch.process(dt, 0, 0, 1)
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    ctx.sync()
    t_h = time.perf_counter()
    batch.process(dt, S, K_IT, HORIZON)
    ctx.sync()
    el_h = time.perf_counter() - t_h
    h_fused_ms, h_launches, _ = ctx.fused_times()
    h_lat = np.array(ctx.step_times())
    ctx.set_profiling(False)
    horizon = {"steps": f"1-{HORIZON}", "value": n_bodies * HORIZON / el_h, "unit": "body-steps/s",
               "ms_per_step": el_h / HORIZON * 1e3,
               "launch_ms": h_fused_ms / h_launches if h_launches else None,
               "p50_step_ms": float(np.percentile(h_lat, 50)) if len(h_lat) else None,
               "p99_step_ms": float(np.percentile(h_lat, 99)) if len(h_lat) else None,
               "note": "the whole configuration-3 job: a fresh batch, all 240 steps in one call"}
    log(f"horizon: {HORIZON} steps in {el_h:.3f} s ({horizon['value'] / 1e6:.1f} M body-steps/s)")
    # interactive use: one step per call, host work and synchronisation included
    call_lat = []
    for _ in range(args.call_steps):
        a = time.perf_counter()
        batch.process(dt, S, K_IT, 1)
        call_lat.append((time.perf_counter() - a) * 1e3)
    for w in worlds:
        w.close()

    # one world per call (plain mode, the single-world configurations, the
    # accuracy runs): a context of its own, whose launch shape is the
    # library's default for a few worlds (the batch's tuned shape is for
    # 1,024 worlds sharing the CUs)
    octx = Context(0)
    octx.set_solve_variant(args.variant)

    # plain: the same scenes, one world per launch (the reference's
    # one-world process call), over the same window and over the horizon
    plain = None
    if args.plain_worlds:
        pick = specs[:: max(1, len(specs) // args.plain_worlds)][:args.plain_worlds]
        pw_bs, pw_t, ph_bs, ph_t = 0, 0.0, 0, 0.0
        for sc in pick:
            w = GpuWorld(octx)
            apply_scene(sc, w)
            octx.sync()
            a = time.perf_counter()
            for s in range(1, window[1] + 1):
                if s == window[0]:
                    octx.sync()
                    b0 = time.perf_counter()
                checked(w.process(dt, S, K_IT), "plain step")
            octx.sync()
            b1 = time.perf_counter()
            pw_t += b1 - b0
            pw_bs += count_bodies(sc) * K
            if window[1] >= HORIZON:
                ph_t += b1 - a
                ph_bs += count_bodies(sc) * window[1]
            w.close()
        plain = {"value": pw_bs / pw_t, "unit": "body-steps/s",
                 "sample": f"{len(pick)} worlds spread over the batch, one world per launch (r2d_process per "
                           f"step); steps {window[0]}-{window[1]} timed",
                 "launch": "the library's default shape for one world (k_fused<0,768,3>: a lane per solver item "
                           "up to 768 items)" if args.variant < 0 else f"variant {args.variant}"}
        if ph_t:
            plain["horizon"] = {"value": ph_bs / ph_t, "unit": "body-steps/s",
                                "sample": f"the same worlds over steps 1-{window[1]}"}
        log(f"plain: {plain['value'] / 1e6:.2f} M body-steps/s")

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
    used_variant = tuned["variant"] if tuned else (args.variant if args.variant >= 0 else 1)
    if fused_launches:
        kernel = f"k_fused<{FUSED_SHAPES.get(used_variant, '?')}>"
        launch_s = fused_ms / fused_launches / 1e3
        launch_bytes = (72 * work["bodies"] + 16 * work["entries"] + 8 * work["raw_candidates"] +
                        176 * work["manifolds"]) / fused_launches
        lds_bytes = S * K_IT * (64 * work["manifolds"] + 48 * joint_items) / fused_launches
        per_step = {k: v / max(world_steps, 1) for k, v in fused_phases.items()}
        per_step_note = "per world-step, workgroup wall clock"
    else:
        kernel = "k_solve"
        names = ["broadphase", "narrowphase", "schedule", "solve"]
        per_step = {n: phase_ms[i] / max(phase_steps, 1) for i, n in enumerate(names)}
        launch_s = per_step["solve"] / 1e3
        launch_bytes = 112 * n_bodies + 120 * counts["manifolds"] + 16 * counts["items"]
        lds_bytes = S * K_IT * (64 * counts["manifolds"] + 48 * max(0, counts["items"] - counts["manifolds"]))
        per_step_note = "per step, HIP events between phase launches"
    dom = max(per_step, key=per_step.get)
    achieved = launch_bytes / launch_s / 1e9 if launch_s > 0 else 0.0
    lds_achieved = lds_bytes / launch_s / 1e9 if launch_s > 0 else 0.0
    lds_counted = lds_counted_for(kernel, window, args)
    if lds_counted and launch_s > 0:
        gbs = lds_counted["bytes_lower_bound"] / launch_s / 1e9
        lds_counted = dict(lds_counted, achieved_lower_bound=gbs, frac_lower_bound=gbs / LDS_PEAK_GBS,
                           note="SQ_INSTS_LDS of the committed rocprofv3 run of this window, x 256 B (64 lanes x "
                                "4 B, the least a wave's LDS instruction moves), over this run's launch time")
    copy = hbm_copy_rate()
    try:
        peaks = ctx.measured_peaks()
        log(f"measured peaks: HBM read {peaks['hbm_read']:.0f}, copy {peaks['hbm_copy']:.0f}, "
            f"LDS {peaks['lds_read']:.0f} GB/s")
    except RuntimeError as e:   # the microbenchmarks need 2 GiB of free device memory
        log(f"measured peaks unavailable: {e}")
        peaks = None

    single = None
    if not args.no_single:
        single = single_world(octx, args.single_steps, not args.reference_hash)

    warm = None
    if not args.no_warm:
        log("warm starting")
        warm = warm_start_runs(ctx, specs, window, value)
        log(f"warm starting: batched {warm['batched']['value'] / 1e6:.1f} M body-steps/s "
            f"({warm['batched']['vs_cold']:.2f}x cold)")
    large = None
    if not args.no_large:
        log("large constrained world")
        large = large_world(octx, 120, {1, 10, 120}, not args.no_accuracy)

    acc = None
    if not args.no_accuracy:
        log("accuracy")
        sample = list(range(0, args.worlds, max(1, args.worlds // 4)))[:4]
        errs, bit_eq, pen = accuracy(octx, specs, sample, HORIZON, {1, 60, HORIZON})
        acc = {"worlds": sample, "bit_equal_to_f32_oracle": bit_eq, "tolerance_vs_reference_f32": 1e-5,
               "vs_f64_oracle": errs, "max_penetration_m": pen, "batch_max_penetration_m": batch_pen,
               "pyramid_600": pyramid_energy(octx, 600)}

    crit = None
    if not args.no_critical:
        # the reference-order critical path of one-world steps, with the
        # GPU's solve time per chain link (one world per call)
        log("critical paths")
        crit = {}
        cpu1 = acc["pyramid_600"]["cpu_single"]["ms_per_step"] if acc else None
        crit["pyramid_20_rows"] = critical_path(octx, "pyramid", scenes.pyramid(20), 590, 10, cpu1)
        if single:
            cpu2 = single["cpu_single"]["ms_per_step"]
            ops2 = scenes.mixed(4096, seed=1, hash_cell=None if args.reference_hash else "auto")
            crit["world_4096"] = critical_path(octx, "4,096-body world", ops2, args.single_steps - 10, 10, cpu2)
        if large:
            crit["world_65540"] = critical_path(octx, "65,540-body world", scenes.large_constrained(), 110, 10,
                                                large["cpu_single"]["ms_per_step"])
        crit["note"] = ("depth: longest dependency chain of a sub-step's solver items (4 passes of joints then "
                        "manifolds in discovery order, items ordered by their shared dynamic or jointed bodies), from "
                        "the world's manifolds at the last timed step; links_per_step = 4 sub-steps x depth; "
                        "ns_per_link = the solve's device time per step / links_per_step")

    serve = None
    if not args.no_serve:
        # capacity: an overloaded stream (2x the batched horizon rate) drained
        # to the end; then a stream offering serve_load x that capacity
        log("serving: overloaded stream")
        rate2, arr2 = served_arrivals(horizon["value"], args.serve_sat_seconds, 2.0, args.seed + 1, HORIZON)
        sat_trace = []
        sat, _ = serve_stream(ctx, rate2, arr2, HORIZON, 2.0, args.serve_steps_per_tick, trace=sat_trace)
        # capacity: the saturated part of the overloaded stream, from 1 s (the
        # lanes have filled) to its last arrival (then it only drains)
        t_sat = (1.0, arr2[-1][0] if arr2 else 0.0)
        capacity = saturated_rate(sat_trace, *t_sat) or sat["value"]
        sat["whole_run_value"] = sat["value"]
        sat["value"] = capacity
        sat["saturated_window_s"] = list(t_sat)
        log(f"served capacity: {capacity / 1e6:.1f} M body-steps/s (whole overloaded run "
            f"{sat['whole_run_value'] / 1e6:.1f} M)")
        rate, arrivals = served_arrivals(capacity, args.serve_seconds, args.serve_load, args.seed, HORIZON)
        by_size = sorted(arrivals, key=lambda a: a[1][0])
        keep = {by_size[0][1], by_size[len(by_size) // 2][1],
                min(by_size, key=lambda a: abs(a[1][0] - 2000))[1]}
        serve, kept = serve_stream(ctx, rate, arrivals, HORIZON, args.serve_load, args.serve_steps_per_tick,
                                   keep=keep)
        log(f"served {serve['scenes']} scenes, {serve['value'] / 1e6:.1f} M body-steps/s")
        serve["capacity"] = {k: sat[k] for k in ("value", "whole_run_value", "saturated_window_s", "offered_load",
                                                 "rate_scenes_per_s", "scenes", "wall_s", "admitted_scenes_per_s",
                                                 "p50_step_ms", "p99_step_ms", "p50_request_ms", "p99_request_ms")}
        serve["capacity"]["note"] = ("value: body-steps/s collected while the overloaded stream kept every lane busy "
                                     "(from 1 s to its last arrival); whole_run_value: all its body-steps over its "
                                     "whole wall time, ramp-up and drain included")
        serve["offered_body_steps_per_s"] = args.serve_load * capacity
        pv, sizes = plain_served(octx, arrivals, HORIZON, args.serve_plain_scenes)
        serve["plain"] = {"value": pv, "unit": "body-steps/s",
                          "sample": f"{len(sizes)} served scenes of {sizes} bodies, {HORIZON} steps each, one "
                                    f"world per launch"}
        serve["served_vs_plain"] = serve["value"] / pv
        serve["capacity_vs_plain"] = capacity / pv
        serve["accuracy"] = served_accuracy(kept, HORIZON)

    out = {
        "metric": "body-steps/sec batched vs plain on 1 MI355X; p50/p99 step latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes)",
        "config": {"workload": f"batched worlds (configuration 3): 1024 worlds x U[50,500] mixed discs/rects, 0-20 "
                               f"distance joints, 0-10 springs per world; dt=1/60, 4 sub-steps, 4 iterations; "
                               f"{HORIZON}-step horizon, timed window steps {window[0]}-{window[1]} (one call of "
                               f"{K} steps after {window[0] - 1} untimed)",
                   "worlds": args.worlds, "bodies": n_bodies, "window": list(window),
                   "manifolds_last_step": counts["manifolds"],
                   "schedule": "reference order (bit-exact)", "parallelism": "1 GPU",
                   "solve_launch": ({"autotuned": True, "variant": tuned["variant"], "fused": tuned["fused"],
                                     "tuned_at_step": window[0], "sample_steps": K,
                                     "candidates_step_ms (variant/lds_limit/fused)": tuned["all_ms"],
                                     "refined_step_ms": tuned.get("refined_ms")}
                                    if tuned else {"autotuned": False, "variant": args.variant})},
        "p50_step_ms": float(np.percentile(lat_ms, 50)) if len(lat_ms) else None,
        "p99_step_ms": float(np.percentile(lat_ms, 99)) if len(lat_ms) else None,
        "p999_step_ms": float(np.percentile(lat_ms, 99.9)) if len(lat_ms) else None,
        "max_step_ms": float(lat_ms.max()) if len(lat_ms) else None,
        "step_latency": ("each world's step, wall clock of its workgroup in the fused kernel (all world-steps of the "
                         "timed region)" if fused_launches else
                         "device time per step from HIP events (the timed region's steps)"),
        "horizon": horizon,
        "per_call": {"p50_ms": float(np.percentile(call_lat, 50)) if call_lat else None,
                     "p99_ms": float(np.percentile(call_lat, 99)) if call_lat else None,
                     "calls": len(call_lat), "note": "one step per r2d_process_batch call after the horizon, "
                                                     "host preparation and synchronisation included"},
        "plain": plain,
        "batched_vs_plain": value / plain["value"] if plain else None,
        "horizon_vs_plain": (horizon["value"] / plain["horizon"]["value"]
                             if plain and "horizon" in plain else None),
        "served": serve,
        "single_world_4096": single,
        "critical_path": crit,
        "large_world_65536": large,
        "accuracy": acc,
        "warm_start": warm,
        "mode": ("fused: one kernel runs every step of every world, one workgroup per world" if fused_launches else
                 "phase kernels: one launch sequence (hipGraph) per step"),
        "phase_ms": per_step,
        "phase_ms_note": per_step_note,
        "roofline": {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_for(kernel, window, args),
                     "launch_ms": launch_s * 1e3, "algorithmic_bytes_per_launch": launch_bytes,
                     "measured_copy_rate": copy,
                     "frac_of_measured_copy": achieved / copy["value"] if copy else None,
                     "measured_peaks": dict(peaks, unit="GB/s", sample=(
                         "the library's own microbenchmarks (r2d_debug_peaks): a 1 GiB read stream and a 1 GiB "
                         "copy (read + written bytes) by 16-byte loads, and conflict-free 16-byte LDS reads on "
                         "every CU")) if peaks else None,
                     "frac_of_measured_read": achieved / peaks["hbm_read"] if peaks else None,
                     "lds": {"achieved": lds_achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                             "frac": lds_achieved / LDS_PEAK_GBS, "bytes_per_launch": lds_bytes,
                             "measured_peak": peaks["lds_read"] if peaks else None,
                             "frac_of_measured": lds_achieved / peaks["lds_read"] if peaks else None,
                             "counted": (dict(lds_counted, frac_lower_bound_of_measured=lds_counted[
                                 "achieved_lower_bound"] / peaks["lds_read"])
                                         if peaks and lds_counted and "achieved_lower_bound" in lds_counted
                                         else lds_counted)},
                     "dominant_phase": dom},
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    print(json.dumps(out), flush=True)
    octx.close()
    ctx.close()


if __name__ == "__main__":
    main()
