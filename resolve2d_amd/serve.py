"""Continuous-batching server (Python mirror of r2d_serve_* in include/r2d.h).

The reference steps one world per `Solver.process` call (src/core/lib.zig:189);
serving has no counterpart there.  The server is native (r2d_api.hip): it
admits requests while their footprints fit an HBM budget, keeps admitted
worlds resident on the device, and steps them in lanes by world size -- each
lane one batched launch per step, at its own pace, on its own stream.
`run_stream` drives it with timed arrivals the way a request loop would.
"""
import ctypes
import time

import numpy as np

from . import _native as N
from .scenes import DT, ITERS, SUB_STEPS


class Server:
    def __init__(self, ctx, hbm_budget_bytes: int, max_worlds: int, dt: float = DT, sub_steps: int = SUB_STEPS,
                 collision_iters: int = ITERS, split_bodies=(1024, 4096), steps_per_tick: int = 1):
        """split_bodies: lanes by world size (n <= split[0], n <= split[1],
        larger); (0, 0) steps every world in one lane.  steps_per_tick: steps
        a lane's launch takes (fewer when a running world needs fewer)."""
        self.ctx = ctx
        self.lib = ctx.lib
        o = N.ServeOptions()
        o.hbm_budget_bytes, o.max_worlds, o.dt = hbm_budget_bytes, max_worlds, dt
        o.sub_steps, o.collision_iters = sub_steps, collision_iters
        o.split_bodies[0], o.split_bodies[1] = split_bodies
        o.steps_per_tick = steps_per_tick
        h = ctypes.c_void_p()
        N.check(self.lib.r2d_serve_create(ctx.handle, ctypes.byref(o), ctypes.byref(h)), "r2d_serve_create")
        self.handle = h
        ctx.adopt(self, server=True)

    def close(self):
        if self.handle:
            N.check(self.lib.r2d_serve_destroy(self.handle), "r2d_serve_destroy")
            self.handle = None
            self.ctx.release(self)

    def submit(self, world, steps: int) -> int:
        t = ctypes.c_uint64()
        N.check(self.lib.r2d_serve_submit(self.handle, world.handle, steps, ctypes.byref(t)), "r2d_serve_submit")
        return t.value

    def tick(self) -> dict:
        info = N.ServeTickInfo()
        N.check(self.lib.r2d_serve_tick(self.handle, ctypes.byref(info)), "r2d_serve_tick")
        return {f: getattr(info, f) for f, _ in N.ServeTickInfo._fields_}

    def drain(self):
        N.check(self.lib.r2d_serve_drain(self.handle), "r2d_serve_drain")

    def poll(self, ticket: int) -> dict:
        st = N.RequestStatus()
        N.check(self.lib.r2d_serve_poll(self.handle, ticket, ctypes.byref(st)), "r2d_serve_poll")
        return {f: getattr(st, f) for f, _ in N.RequestStatus._fields_}

    def completed(self) -> list:
        """Tickets of the requests finished since the last call, oldest first."""
        out = []
        buf = (ctypes.c_uint64 * 256)()
        n = ctypes.c_uint32()
        while True:
            N.check(self.lib.r2d_serve_completed(self.handle, buf, 256, ctypes.byref(n)), "r2d_serve_completed")
            out.extend(buf[:n.value])
            if n.value < 256:
                return out

    def stats(self) -> dict:
        s = N.ServeSummary()
        N.check(self.lib.r2d_serve_stats(self.handle, ctypes.byref(s)), "r2d_serve_stats")
        return {f: getattr(s, f) for f, _ in N.ServeSummary._fields_}


def run_stream(server: Server, arrivals, build, steps: int, time_limit: float = None, on_done=None, trace=None):
    """Serve a timed stream.  `arrivals` is a list of (t_seconds, payload);
    `build(payload)` returns a new world for it (built on the host while the
    device steps the running batch); `on_done(payload, world, status)` is
    called as requests finish (the caller may read and free the world).
    Arrivals after `time_limit` seconds are not submitted.  trace, if a list,
    gets (seconds since the start, body-steps collected so far) after every
    tick.  Returns a dict of wall time, request latencies and the server's
    statistics."""
    t0 = time.perf_counter()
    pending = list(arrivals)
    pending.reverse()
    live = {}
    lat = []
    submitted = 0
    while pending or live:
        now = time.perf_counter() - t0
        while pending and pending[-1][0] <= now:
            t_arr, payload = pending.pop()
            if time_limit is not None and t_arr > time_limit:
                pending.clear()
                break
            w = build(payload)
            live[server.submit(w, steps)] = (payload, w)
            submitted += 1
        info = server.tick()
        if trace is not None:
            trace.append((time.perf_counter() - t0, server.stats()["body_steps"]))
        if info["completed"]:
            for t in server.completed():
                if t not in live:
                    continue
                st = server.poll(t)
                payload, w = live.pop(t)
                lat.append(st["total_ms"])
                if on_done:
                    on_done(payload, w, st)
        if not live and pending:
            # idle until the next arrival
            time.sleep(max(0.0, min(pending[-1][0] - (time.perf_counter() - t0), 0.05)))
    wall = time.perf_counter() - t0
    out = dict(wall_s=wall, submitted=submitted, stats=server.stats())
    if lat:
        out["request_p50_ms"] = float(np.percentile(lat, 50))
        out["request_p99_ms"] = float(np.percentile(lat, 99))
    return out
