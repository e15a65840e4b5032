"""Autotuner of the step's launch shape for a batch of worlds.

Knobs (all give bit-identical results, so tuning may run on live steps):
  - solve variant: the register-slot kernel's slots x workgroup size, and the
    record kernels' workgroup sizes (r2d_ctx_set_solve_variant);
  - LDS limit: the largest world whose body state the solve kernel keeps in
    LDS (r2d_ctx_set_lds_limit);
  - whole steps per workgroup (r2d_ctx_set_fused) for variants that have a
    fused kernel.  A fused launch ends with its slowest world, so it is timed
    on calls of several steps, as the caller's calls would run.
Each candidate is timed on real steps of the batch -- the device time of a
whole step: the fused kernel's launch (HIP events around it) when the batch
takes whole steps per workgroup, else the sum of the four phases' HIP
events -- candidates interleaved round by round so the slowly changing state
does not favour any of them; the median decides.  The winner is kept per
batch shape -- worlds, largest world and total bodies, each rounded up to a
power of two -- in a JSON file, so later runs of the same shape reuse it.
"""
import json
import os
from typing import Callable, Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

SOLVE_VARIANTS = 15
# variants that differ on the phase kernels (9 and 10 are variant 6 there, 12
# is variant 7; 11 is the record kernel with 192 threads, 13 with 512)
PHASE_VARIANTS = tuple(range(9)) + (11, 13)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "resolve2d_amd", "autotune.json")


def step_ms(ctx, steps: int) -> float:
    """Device time per step of the profiled calls since the last reset."""
    fused_ms, launches, _ = ctx.fused_times()
    if launches:
        return fused_ms / max(steps, 1)
    ms, n = ctx.phase_times()
    return sum(ms) / max(n, 1)


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


# variants with a fused kernel (the register-slot shapes, lockstep walk aside,
# and the records-only shapes)
FUSED_VARIANTS = (1, 2, 3, 4, 5, 7, 9, 10, 11, 12, 13, 14)


def candidates(variants: Optional[Iterable[int]] = None, lds_limits: Iterable[int] = (0, 1024),
               fused: Iterable[bool] = (True, False)):
    """(variant, lds_limit, fused) triples; fused only where the variant has a
    fused kernel, and the LDS limit only where the phase kernels run."""
    vs = list(range(SOLVE_VARIANTS)) if variants is None else list(variants)
    out = []
    for f in fused:
        for v in vs:
            if (f and v not in FUSED_VARIANTS) or (not f and v not in PHASE_VARIANTS):
                continue
            for lim in ((0,) if f else lds_limits):
                out.append((v, lim, bool(f)))
    return out


def apply(ctx, choice: dict):
    ctx.set_solve_variant(int(choice["variant"]))
    ctx.set_lds_limit(int(choice["lds_limit"]))
    ctx.set_fused(bool(choice.get("fused", True)))


def pick(times: Dict[tuple, list]) -> tuple:
    """The candidate with the smallest median time (ties: the earlier one)."""
    best, best_t = None, None
    for c, ts in times.items():
        m = float(np.median(ts))
        if best_t is None or m < best_t:
            best, best_t = c, m
    return best


def tune(ctx, handles: Sequence, dt: float, sub_steps: int, iters: int, rounds: int = 2,
         cands=None, cache_path: Optional[str] = DEFAULT_CACHE, use_cache: bool = True,
         steps_per_sample: int = 30, make_batch: Optional[Callable[[], list]] = None,
         refine: int = 0, refine_steps: int = 120, log: Optional[Callable[[str], None]] = None) -> dict:
    """Pick the launch shape for stepping `handles` together and apply it to
    the context.  Spends len(cands) * rounds * steps_per_sample real steps of
    the batch (untimed by the caller's measure; the results are the same as
    with any other shape), each sample one call of steps_per_sample steps.
    With make_batch (returns fresh worlds in the same state each time, with
    handle and close()), every sample starts from that same state instead of
    where the previous sample left the batch.  refine > 1: the `refine`
    fastest candidates are timed again on one call of refine_steps steps
    (a short sample can rank shapes differently from a long run, whose later
    steps hold more contacts), and that time decides.  Returns {"variant",
    "lds_limit", "fused", "step_ms", "key", "cached", "all_ms"} (plus
    "refined_ms" after a refinement).  log, if given, gets a line per sample."""
    probe = None
    if handles is None:
        probe = make_batch()
        handles = [w.handle for w in probe]
    key = shape_key(batch_shape(ctx, handles))
    if probe is not None:
        for w in probe:
            w.close()
    cache = load_cache(cache_path) if (cache_path and use_cache) else {}
    if key in cache:
        apply(ctx, cache[key])
        return dict(cache[key], key=key, cached=True)
    cands = cands or candidates()
    times = {c: [] for c in cands}
    was = ctx.profiling

    def sample(c, steps):
        ctx.set_solve_variant(c[0])
        ctx.set_lds_limit(c[1])
        ctx.set_fused(c[2] if len(c) > 2 else True)
        fresh = make_batch() if make_batch else None
        hs = [w.handle for w in fresh] if fresh else handles
        ctx.reset_phase_times()
        ctx.process_batch_handles(hs, dt, sub_steps, iters, steps)
        t = step_ms(ctx, steps)
        for w in fresh or ():
            w.close()
        return t

    ctx.set_profiling(True)
    refined = {}
    try:
        for _ in range(rounds):
            for c in cands:
                times[c].append(sample(c, steps_per_sample))
                if log:
                    log(f"autotune sample {c}: {times[c][-1]:.3f} ms per step")
        if refine > 1:
            ranked = sorted(times, key=lambda c: float(np.median(times[c])))[:refine]
            refined = {}
            for c in ranked:
                refined[c] = [sample(c, refine_steps)]
                if log:
                    log(f"autotune refine {c}: {refined[c][0]:.3f} ms per step")
    finally:
        ctx.set_profiling(was)
    best = pick(refined) if refined else pick(times)
    v, lim = best[0], best[1]
    fused = bool(best[2]) if len(best) > 2 else True
    choice = {"variant": v, "lds_limit": lim, "fused": fused,
              "step_ms": float(np.median((refined or times)[best])),
              "all_ms": {"/".join(str(int(x)) for x in c): float(np.median(t)) for c, t in times.items()}}
    if refined:
        choice["refined_ms"] = {"/".join(str(int(x)) for x in c): float(t[0]) for c, t in refined.items()}
    apply(ctx, choice)
    if cache_path:
        cache = load_cache(cache_path)
        cache[key] = choice
        save_cache(cache, cache_path)
    return dict(choice, key=key, cached=False)
