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

The two-level kernel has its own refill threshold and rays per fetch
(tune_two_level, on the caller's scene and rays), and the batch kernel's
geometry requests their own refill threshold (tune_batch, on the caller's
batch of requests): the best values depend on how long the rays are.
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


def _timed_median(run, reps: int) -> float:
    import torch

    run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record()
        run()
        b.record()
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


TWO_LEVEL_FETCH = (32, 64, 128)
TWO_LEVEL_MIN_BUSY = (16, 24, 32, 40)


def tune_two_level(ctx, scene, rays, queries=("closest",), reps: int = 5, fetch=TWO_LEVEL_FETCH,
                   min_busy=TWO_LEVEL_MIN_BUSY) -> dict:
    """Sweep the two-level kernel's rays per fetch and refill threshold
    (RRX_TUNING_TRACE_2L_RAYS_PER_FETCH, RRX_TUNING_TRACE_2L_MIN_BUSY) on
    rrCmdIntersect of (scene, rays); returns {"best", "default", "table"} and
    leaves the best set.  Raises if any setting's hits differ from the
    starting setting's."""
    import torch

    lib = ctx.lib
    keys = {"rays_per_fetch": rr.RRX_TUNING_TRACE_2L_RAYS_PER_FETCH, "min_busy": rr.RRX_TUNING_TRACE_2L_MIN_BUSY}
    start = {k: get(lib, v) for k, v in keys.items()}
    n = int(rays.shape[0])
    hits = ctx.empty(n * 16)
    scratch = ctx.trace_scratch(n)

    def run():
        for q in queries:
            ctx.intersect(scene, rays, q, "full", hits=hits, scratch=scratch)

    def sample():
        outs = []
        for q in queries:
            ctx.intersect(scene, rays, q, "full", hits=hits, scratch=scratch)
            torch.cuda.synchronize()
            outs.append(hits.cpu().numpy().copy())
        return outs

    reference = sample()
    table = []
    try:
        for f, m in itertools.product(fetch, min_busy):
            cfg = {"rays_per_fetch": f, "min_busy": m}
            for k, v in cfg.items():
                set_(lib, keys[k], v)
            if not all(np.array_equal(a, b) for a, b in zip(sample(), reference)):
                raise RuntimeError(f"hits differ under two-level tuning {cfg}")
            ms = _timed_median(run, reps)
            table.append({**cfg, "ms": round(ms, 4), "mrays_s": round(len(queries) * n / ms / 1e3, 1)})
    except BaseException:
        for k, v in start.items():
            set_(lib, keys[k], v)
        raise
    best = min(table, key=lambda r: r["ms"])
    default = next((r for r in table if all(r[k] == start[k] for k in start)), None)
    for k in start:
        set_(lib, keys[k], best[k])
    return {"best": best, "default": default, "table": table}


BATCH_MIN_BUSY = (8, 12, 16, 24, 32)


def _graph_of(ctx, fn):
    """One HIP graph of fn's work on a side stream (the bench's capture recipe)."""
    import torch

    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s, capture_error_mode="relaxed"):
            with ctx.on_stream(s):
                fn()
    torch.cuda.synchronize()
    return g


def tune_batch(ctx, requests, reps: int = 5, min_busy=BATCH_MIN_BUSY, scratch=None) -> dict:
    """Sweep the batch kernel's refill threshold for geometry requests
    (RRX_TUNING_TRACE_BATCH_MIN_BUSY) on rrxCmdIntersectBatch of the caller's
    requests ([(accel, rays, query, output[, hits])], as
    api.Context.intersect_batch takes them); returns {"best", "default",
    "table"} and leaves the best set.  Each setting is captured in a HIP
    graph (the threshold is a launch argument) and timed by replays, so the
    host's per-call table set-up of large batches does not enter the times.
    Raises if any setting's hits differ from the starting setting's."""
    import torch

    lib = ctx.lib
    key = rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY
    start = get(lib, key)
    if scratch is None:
        counts = [int(r[1].shape[0]) for r in requests]
        scratch = ctx.batch_scratch(len(requests), sum(counts), max(counts + [1]))
    outs = ctx.intersect_batch([r[:4] for r in requests], scratch=scratch)
    reqs = [tuple(r[:4]) + (o,) for r, o in zip(requests, outs)]
    torch.cuda.synchronize()
    reference = [o.cpu().numpy().copy() for o in outs]
    table = []
    try:
        for m in min_busy:
            set_(lib, key, m)
            g = _graph_of(ctx, lambda: ctx.intersect_batch(reqs, scratch=scratch))
            for o in outs:
                o.zero_()
            g.replay()
            torch.cuda.synchronize()
            if not all(np.array_equal(o.cpu().numpy(), r) for o, r in zip(outs, reference)):
                raise RuntimeError(f"batch hits differ at refill threshold {m}")
            table.append({"min_busy": m, "ms": round(_timed_median(g.replay, reps), 4)})
    except BaseException:
        set_(lib, key, start)
        raise
    best = min(table, key=lambda r: r["ms"])
    default = next((r for r in table if r["min_busy"] == start), None)
    set_(lib, key, best["min_busy"])
    return {"best": best, "default": default, "table": table}
