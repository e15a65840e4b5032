"""Autotuner of the step's launch shape for a batch of worlds.

Knobs (all give bit-identical results, so tuning may run on live steps):
  - solve variant: the register-slot kernel's slots x workgroup size, and the
    record kernels' workgroup sizes (r2d_ctx_set_solve_variant);
  - LDS limit: the largest world whose body state the solve kernel keeps in
    LDS (r2d_ctx_set_lds_limit).
Each candidate is timed on real steps of the batch (the solve phase's HIP
events), candidates interleaved round by round so the slowly changing state
does not favour any of them; the median decides.  The winner is kept per
batch shape -- worlds, largest world and total bodies, each rounded up to a
power of two -- in a JSON file, so later runs of the same shape reuse it.
"""
import json
import os
from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

SOLVE_VARIANTS = 9
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "resolve2d_amd", "autotune.json")
SOLVE_PHASE = 3


def _pow2(x: int) -> int:
    p = 1
    while p < x:
        p *= 2
    return p


def batch_shape(ctx, handles: Sequence) -> Tuple[int, int, int]:
    """(worlds, largest world, total bodies), each rounded up to a power of two."""
    import ctypes
    sizes = []
    for h in handles:
        n = ctypes.c_size_t()
        N.check(ctx.lib.r2d_num_bodies(h, ctypes.byref(n)), "r2d_num_bodies")
        sizes.append(n.value)
    return _pow2(len(sizes)), _pow2(max(sizes, default=1)), _pow2(sum(sizes))


def shape_key(shape: Tuple[int, int, int]) -> str:
    return "w{}_n{}_b{}".format(*shape)


def load_cache(path: str = DEFAULT_CACHE) -> Dict[str, dict]:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(cache: Dict[str, dict], path: str = DEFAULT_CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def candidates(variants: Optional[Iterable[int]] = None, lds_limits: Iterable[int] = (0, 1024)):
    vs = list(range(SOLVE_VARIANTS)) if variants is None else list(variants)
    return [(v, lim) for v in vs for lim in lds_limits]


def apply(ctx, choice: dict):
    ctx.set_solve_variant(int(choice["variant"]))
    ctx.set_lds_limit(int(choice["lds_limit"]))


def pick(times: Dict[Tuple[int, int], list]) -> Tuple[int, int]:
    """The candidate with the smallest median time (ties: the earlier one)."""
    best, best_t = None, None
    for c, ts in times.items():
        m = float(np.median(ts))
        if best_t is None or m < best_t:
            best, best_t = c, m
    return best


def tune(ctx, handles: Sequence, dt: float, sub_steps: int, iters: int, rounds: int = 2,
         cands=None, cache_path: Optional[str] = DEFAULT_CACHE, use_cache: bool = True) -> dict:
    """Pick the launch shape for stepping `handles` together and apply it to
    the context.  Spends len(cands) * rounds real steps of the batch (untimed
    by the caller's measure; the results are the same as with any other
    shape).  Returns {"variant", "lds_limit", "solve_ms", "key", "cached"}."""
    key = shape_key(batch_shape(ctx, handles))
    cache = load_cache(cache_path) if (cache_path and use_cache) else {}
    if key in cache:
        apply(ctx, cache[key])
        return dict(cache[key], key=key, cached=True)
    cands = cands or candidates()
    times = {c: [] for c in cands}
    was = ctx.profiling
    ctx.set_profiling(True)
    try:
        for _ in range(rounds):
            for c in cands:
                ctx.set_solve_variant(c[0])
                ctx.set_lds_limit(c[1])
                ctx.reset_phase_times()
                ctx.process_batch_handles(handles, dt, sub_steps, iters, 1)
                ms, n = ctx.phase_times()
                times[c].append(ms[SOLVE_PHASE] / max(n, 1))
    finally:
        ctx.set_profiling(was)
    v, lim = pick(times)
    choice = {"variant": v, "lds_limit": lim, "solve_ms": float(np.median(times[(v, lim)])),
              "all_ms": {f"{c[0]}/{c[1]}": float(np.median(t)) for c, t in times.items()}}
    apply(ctx, choice)
    if cache_path:
        cache = load_cache(cache_path)
        cache[key] = choice
        save_cache(cache, cache_path)
    return dict(choice, key=key, cached=False)
