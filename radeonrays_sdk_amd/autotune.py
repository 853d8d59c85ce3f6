"""Autotuner for the persistent traversal kernel.

Times rrCmdIntersect on the caller's own BVH and rays for every combination
of the kernel's run-time knobs (rrxSetTuning): ray indices taken per queue
atomic (32/64/128), LDS stack entries per lane (8/16/32) and the refill
threshold (lanes still traversing below which a wave fetches new rays),
checks that every combination returns the same hits as the current setting,
and leaves the fastest one set.

The sort's variants are swept by tune_sort on a caller-given number of
keys: the LSD passes with 4-, 8- or 10-bit digits (eight, four or three
passes over 30-bit Morton codes) and, on 2^22..2^24 keys, the hybrid (one
most-significant-digit pass, then every bucket sorted in LDS); all give the
same stable order.

The treelet size of the restructuring (5, 7 or 9 leaves) is swept by
tune_treelet: it trades build time (the subset dynamic programme grows as
3^leaves) against tree quality (traversal time); closest hits are the same
for every size.
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


DIGIT_BITS = (4, 8, 10)


def tune_sort(ctx, n: int = 16 << 20, reps: int = 5, digit_bits=DIGIT_BITS, seed: int = 0x52520005) -> dict:
    """Sweep the sort's variants on n random 30-bit keys with index values
    (rrxCmdSortPairs): the LSD passes with each digit width, and the hybrid
    (one most-significant-digit pass, then each bucket sorted in LDS; only
    on 2^22..2^24 keys).  Returns {"best", "default", "table"} and leaves the
    fastest variant set.  Raises if any variant's output differs."""
    import torch

    lib = ctx.lib
    start = (get(lib, rr.RRX_TUNING_SORT_DIGIT_BITS), get(lib, rr.RRX_TUNING_SORT_HYBRID))
    dev = ctx.torch_device
    g = torch.Generator(device="cpu").manual_seed(seed)
    keys = torch.randint(0, 1 << 30, (n,), dtype=torch.int32, generator=g).to(dev)
    vals = torch.arange(n, dtype=torch.int32, device=dev)
    k2, v2 = keys.clone(), vals.clone()
    tmp = (torch.empty_like(keys), torch.empty_like(vals))

    def run():
        k2.copy_(keys)
        v2.copy_(vals)
        ctx.sort_pairs(k2, v2, 30, tmp=tmp)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in ev]))

    def restore(setting):
        set_(lib, rr.RRX_TUNING_SORT_DIGIT_BITS, setting[0])
        set_(lib, rr.RRX_TUNING_SORT_HYBRID, setting[1])

    copy_ms = timed(lambda: (k2.copy_(keys), v2.copy_(vals)))
    variants = [(bits, 0) for bits in digit_bits] + ([(8, 1)] if (1 << 22) <= n <= (1 << 24) else [])
    reference = None
    table = []
    for bits, hybrid in variants:
        restore((bits, hybrid))
        run()
        torch.cuda.synchronize()
        got = (k2.cpu().numpy(), v2.cpu().numpy())
        if reference is None:
            reference = got
        elif not (np.array_equal(got[0], reference[0]) and np.array_equal(got[1], reference[1])):
            restore(start)
            raise RuntimeError(f"sort output differs with {bits}-bit digits, hybrid={hybrid}")
        ms = timed(run) - copy_ms
        table.append({"digit_bits": bits, "hybrid": hybrid, "ms": round(ms, 4)})
    best = min(table, key=lambda r: r["ms"])
    default = next((r for r in table if (r["digit_bits"], r["hybrid"]) == start), None)
    restore((best["digit_bits"], best["hybrid"]))
    return {"best": best, "default": default, "table": table}


TREELET_LEAVES = (5, 7, 9)


def tune_treelet(ctx, vertices, indices, rays, reps: int = 3, leaves=TREELET_LEAVES) -> dict:
    """Sweep the treelet size on one mesh: for each size, the restructured
    build time and the closest + any-hit traversal time of `rays`; the score
    is their sum (one build per pass over the rays).  Closest hits must be
    identical for every size (raises otherwise).  Leaves the best size set;
    returns {"best", "default", "table"}."""
    import torch

    lib = ctx.lib
    start = get(lib, rr.RRX_TUNING_TREELET_LEAVES)
    n = int(rays.shape[0])
    hits = ctx.empty(n * 16)
    scratch = ctx.trace_scratch(n)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in ev]))

    reference = None
    table = []
    for size in leaves:
        set_(lib, rr.RRX_TUNING_TREELET_LEAVES, size)
        accel = ctx.build_geometry(vertices, indices, restructure=True)
        build_ms = timed(lambda: ctx.build_geometry(vertices, indices, restructure=True, result=accel.buffer))
        ctx.intersect(accel, rays, "closest", "full", hits=hits, scratch=scratch)
        torch.cuda.synchronize()
        got = hits.cpu().numpy().copy()
        if reference is None:
            reference = got
        elif not np.array_equal(got, reference):
            set_(lib, rr.RRX_TUNING_TREELET_LEAVES, start)
            raise RuntimeError(f"closest hits differ with {size}-leaf treelets")
        trace_ms = timed(lambda: [ctx.intersect(accel, rays, q, "full", hits=hits, scratch=scratch)
                                  for q in ("closest", "any")])
        table.append({"treelet_leaves": size, "build_ms": round(build_ms, 4), "trace_ms": round(trace_ms, 4),
                      "score_ms": round(build_ms + trace_ms, 4)})
    best = min(table, key=lambda r: r["score_ms"])
    default = next((r for r in table if r["treelet_leaves"] == start), None)
    set_(lib, rr.RRX_TUNING_TREELET_LEAVES, best["treelet_leaves"])
    return {"best": best, "default": default, "table": table}
