"""The f32 oracle run in host processes started with `spawn` (they import
no GPU code and touch no device), for checks over many worlds: test
infrastructure only."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_worlds(job):
    """(scene specs, steps, dt, sub_steps, iters) -> each world's body array
    after `steps` oracle steps from its start."""
    specs, steps, dt, sub_steps, iters = job
    for p in (ROOT, os.path.join(ROOT, "oracle")):
        if p not in sys.path:
            sys.path.insert(0, p)
    import pyoracle
    from resolve2d_amd.engine import apply_scene
    out = []
    for sc in specs:
        o = pyoracle.OracleWorld("f32")
        apply_scene(sc, o)
        for _ in range(steps):
            if o.process(dt, sub_steps, iters) != 0:
                raise RuntimeError("oracle step failed")
        out.append(o.bodies())
        o.close()
    return out


def oracle_bodies(specs, steps, dt, sub_steps, iters, procs):
    """Body arrays of every spec after `steps` oracle steps, computed over
    `procs` spawned processes (shares balanced by the number of operations)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    order = sorted(range(len(specs)), key=lambda i: -len(specs[i]))
    shares, load = [[] for _ in range(procs)], [0] * procs
    for i in order:
        k = load.index(min(load))
        shares[k].append(i)
        load[k] += len(specs[i])
    shares = [s for s in shares if s]
    # an executor, not a Pool: a worker that dies raises here instead of
    # being restarted forever
    with ProcessPoolExecutor(max_workers=len(shares), mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(run_worlds, [([specs[i] for i in sh], steps, dt, sub_steps, iters) for sh in shares]))
    out = [None] * len(specs)
    for sh, arrays in zip(shares, res):
        for i, a in zip(sh, arrays):
            out[i] = a
    return out
