"""Autotuner for the persistent traversal kernel.

Times rrCmdIntersect on the caller's own BVH and rays for every combination
of the kernel's run-time knobs (rrxSetTuning): ray indices taken per queue
atomic (32/64/128), LDS stack entries per lane (8/16/32) and the refill
threshold (lanes still traversing below which a wave fetches new rays),
checks that every combination returns the same hits as the current setting,
and leaves the fastest one set.

Not swept, with the reason: the radix digit width (8 bits; 4-bit digits
double the passes over memory for a sort that is bandwidth-bound) and the
treelet size (7 leaves; the treelet kernels' subset dynamic programme is
written for 2^7 masks and 9 leaves would take 4x the LDS per treelet).
"""
from __future__ import annotations

import contextlib
import itertools
from ctypes import byref, c_uint32

import numpy as np

from . import rr

FETCH = (32, 64, 128)
DEPTH = (8, 16, 32)
MIN_BUSY = (16, 24, 32, 40, 48)


def get(lib, key) -> int:
    v = c_uint32()
    rr.check(lib.rrxGetTuning(key, byref(v)), "rrxGetTuning")
    return v.value


def set_(lib, key, value) -> None:
    rr.check(lib.rrxSetTuning(key, value), "rrxSetTuning")


def current(lib) -> dict:
    return {"rays_per_fetch": get(lib, rr.RRX_TUNING_TRACE_RAYS_PER_FETCH),
            "lds_stack": get(lib, rr.RRX_TUNING_TRACE_LDS_STACK),
            "min_busy": get(lib, rr.RRX_TUNING_TRACE_MIN_BUSY)}


def apply(lib, cfg: dict) -> None:
    set_(lib, rr.RRX_TUNING_TRACE_RAYS_PER_FETCH, cfg["rays_per_fetch"])
    set_(lib, rr.RRX_TUNING_TRACE_LDS_STACK, cfg["lds_stack"])
    set_(lib, rr.RRX_TUNING_TRACE_MIN_BUSY, cfg["min_busy"])


@contextlib.contextmanager
def plain_traversal(lib, plain: bool = True):
    """Run the enclosed queries on the plain kernels (one lane per ray for its
    whole traversal) instead of the persistent ones, then restore the mode."""
    before = get(lib, rr.RRX_TUNING_TRACE_PLAIN)
    set_(lib, rr.RRX_TUNING_TRACE_PLAIN, int(plain))
    try:
        yield
    finally:
        set_(lib, rr.RRX_TUNING_TRACE_PLAIN, before)


def tune_traversal(ctx, accel, rays, queries=("closest", "any"), reps: int = 5, fetch=FETCH, depth=DEPTH,
                   min_busy=MIN_BUSY) -> dict:
    """Sweep the knobs on (accel, rays); returns {"best", "default", "table"}
    and leaves the best configuration set.  Raises if any configuration's
    hits differ from the starting configuration's."""
    import torch

    lib = ctx.lib
    start = current(lib)
    n = int(rays.shape[0])
    hits = ctx.empty(n * 16)
    scratch = ctx.trace_scratch(n)

    def run():
        for q in queries:
            ctx.intersect(accel, rays, q, "full", hits=hits, scratch=scratch)

    def sample():
        outs = []
        for q in queries:
            ctx.intersect(accel, rays, q, "full", hits=hits, scratch=scratch)
            torch.cuda.synchronize()
            outs.append(hits.cpu().numpy().copy())
        return outs

    reference = sample()
    table = []
    for f, d, m in itertools.product(fetch, depth, min_busy):
        cfg = {"rays_per_fetch": f, "lds_stack": d, "min_busy": m}
        apply(lib, cfg)
        got = sample()
        if not all(np.array_equal(a, b) for a, b in zip(got, reference)):
            apply(lib, start)
            raise RuntimeError(f"hits differ under tuning {cfg}")
        run()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record()
            run()
            b.record()
        torch.cuda.synchronize()
        ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
        table.append({**cfg, "ms": round(ms, 4), "mrays_s": round(len(queries) * n / ms / 1e3, 1)})
    best = min(table, key=lambda r: r["ms"])
    default = next((r for r in table if all(r[k] == start[k] for k in start)), None)
    apply(lib, {k: best[k] for k in start})
    return {"best": best, "default": default, "table": table}
