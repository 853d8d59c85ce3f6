"""GPU tests of failure reporting: device-side failures come back as errors
instead of wrong results given silently, and every kernel still exits.

* A BLAS whose node chain is deeper than the 80-entry traversal stack
  (crafted on the host): the overflowing rays end early, the stack-overflow
  word is set, rrWaitEvent fails with RR_ERROR_INTERNAL and
  rrxGetDeviceErrors names it; rays that do not overflow still equal the
  oracle (whose stack is unbounded).
* The sort's bounded look-back wait, with its bound lowered to 0 by the
  host: tiles give up, the look-back word is set; with the default bound the
  same sort is right again.
* A query on a buffer without a valid header, and queries on copies of
  valid buffers the context never built (the level is read from the header
  on the device, so recording never waits).
"""
import ctypes
from ctypes import byref, c_uint32, c_void_p

import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, autotune, rr, scenes
from tests.gpu_util import RR_INVALID, hits_equal, say, to_dev

pytestmark = pytest.mark.gpu


def _device_errors(ctx) -> int:
    v = c_uint32()
    rr.check(ctx.lib.rrxGetDeviceErrors(ctx.ctx, byref(v)), "rrxGetDeviceErrors")
    return v.value


def deep_chain_blas(n: int) -> tuple[np.ndarray, np.ndarray]:
    """A BLAS of n triangles whose internal node i has leaf i as child0 and
    internal node i + 1 as child1 (the last has two leaves): a chain n - 1
    deep.  Triangle k lies in the plane x = n - k around the x axis, so a
    ray from the origin along +x enters the internal child first at every
    level and defers the leaf: it needs n - 2 stack entries.  Returns the
    buffer (header + nodes) and the triangles' vertices [n, 3, 3]."""
    tri = np.zeros((n, 3, 3), np.float32)
    for k in range(n):
        x = np.float32(n - k)
        tri[k] = [[x, -1, -1], [x, 1, -1], [x, 0, 2]]
    nodes = np.zeros((2 * n - 1, 16), np.uint32)
    f = nodes.view(np.float32)
    lo, hi = tri.min(axis=1), tri.max(axis=1)
    sub_lo, sub_hi = lo.copy(), hi.copy()  # box of triangles k..n-1
    for k in range(n - 2, -1, -1):
        sub_lo[k], sub_hi[k] = np.minimum(lo[k], sub_lo[k + 1]), np.maximum(hi[k], sub_hi[k + 1])
    for k in range(n):
        leaf = n - 1 + k
        f[leaf, 0:3], f[leaf, 4:7], f[leaf, 8:11] = tri[k]
        nodes[leaf, 3], nodes[leaf, 7] = RR_INVALID, k
        nodes[leaf, 11] = k if k < n - 1 else n - 2
    for i in range(n - 1):
        c1 = i + 1 if i < n - 2 else n - 1 + (n - 1)
        b1_lo, b1_hi = (sub_lo[i + 1], sub_hi[i + 1])
        f[i, 0:3], f[i, 4:7], f[i, 8:11], f[i, 12:15] = lo[i], hi[i], b1_lo, b1_hi
        nodes[i, 3], nodes[i, 7] = n - 1 + i, c1
        nodes[i, 11] = RR_INVALID if i == 0 else i - 1
        nodes[i, 15] = 0
    hdr = np.zeros(32, np.uint32)
    hdr[0], hdr[1], hdr[2], hdr[3] = 0x53415252, 1, n, 2 * n - 1
    hdr[4:6] = np.array([128], np.uint64).view(np.uint32)
    hdr[11] = 1  # build serial
    buf = np.concatenate([hdr.view(np.uint8), nodes.view(np.uint8).reshape(-1)])
    return buf, tri


def _rays(origins, dirs):
    return scenes.make_rays(np.asarray(origins, np.float32), np.asarray(dirs, np.float32))


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_stack_overflow_is_reported(ctx, dev, mode):
    import torch

    say(f"traversal stack overflow on a 200-deep chain, {mode} kernels")
    n = 200
    buf, tri = deep_chain_blas(n)
    assert oracle.is_valid(buf) and oracle.check_consistency(buf)
    assert oracle.height(buf) >= n - 1
    # 64 rays down the chain (each needs 198 entries) and 64 that miss it.
    deep = _rays([[0, 0, 0]] * 64, [[1, 0, 0]] * 64)
    miss = _rays([[0, 5, 5]] * 64, [[1, 0, 0]] * 64)
    rays = np.concatenate([deep, miss])
    accel = api.Accel(to_dev(buf, dev), "blas", n)
    assert _device_errors(ctx) == 0
    saved = autotune.current(ctx.lib)
    try:
        if mode == "plain":
            autotune.set_(ctx.lib, rr.RRX_TUNING_TRACE_PLAIN, 1)
        for query in ("closest", "any"):
            hits = ctx.intersect(accel, to_dev(rays, dev), query, "full_t")
            torch.cuda.synchronize()
            assert _device_errors(ctx) == rr.RRX_DEVICE_ERROR_STACK_OVERFLOW, query
            g = api.decode_hits(hits, "full_t")
            ref = oracle.trace(buf, rays, any_hit=(query == "any"))
            # The oracle's stack is unbounded: the deep rays hit; the GPU's
            # ended early, and said so.  Rays that never overflowed agree.
            assert (ref["prim_id"][:64] != RR_INVALID).all()
            ok, msg = hits_equal({k: a[64:] for k, a in g.items()}, ref[64:])
            assert ok, msg
        # Errors are cleared once read: a normal query reports nothing.
        hits = ctx.intersect(accel, to_dev(miss, dev), "closest", "full")
        torch.cuda.synchronize()
        assert _device_errors(ctx) == 0
    finally:
        autotune.apply(ctx.lib, saved)


def test_stack_overflow_fails_rr_wait_event(ctx, dev):
    """Through the reference's own submission path: record into an internal
    command stream, submit, wait: the wait returns RR_ERROR_INTERNAL."""
    import torch

    say("stack overflow reported by rrWaitEvent, one- and two-level")
    n = 200
    buf, tri = deep_chain_blas(n)
    rays = to_dev(_rays([[0, 0, 0]] * 256, [[1, 0, 0]] * 256), dev)
    blas = api.Accel(to_dev(buf, dev), "blas", n)
    tlas = ctx.build_scene([blas], np.eye(3, 4, dtype=np.float32)[None])
    torch.cuda.synchronize()
    lib = ctx.lib
    for accel in (blas, tlas):
        hits = ctx.empty(256 * 16)
        scratch = ctx.trace_scratch(256)
        cs, ev = c_void_p(), c_void_p()
        rr.check(lib.rrAllocateCommandStream(ctx.ctx, byref(cs)), "rrAllocateCommandStream")
        try:
            with ctx.ptrs(accel.address, rays.data_ptr(), hits.data_ptr(), scratch.data_ptr()) as (a, r, h, s):
                rr.check(lib.rrCmdIntersect(ctx.ctx, a, rr.RR_INTERSECT_QUERY_CLOSEST, r, 256, None,
                                            rr.RR_INTERSECT_QUERY_OUTPUT_FULL_HIT, h, s, cs), "rrCmdIntersect")
                rr.check(lib.rrSumbitCommandStream(ctx.ctx, cs, None, byref(ev)), "rrSumbitCommandStream")
                assert lib.rrWaitEvent(ctx.ctx, ev) == rr.RR_ERROR_INTERNAL
                rr.check(lib.rrReleaseEvent(ctx.ctx, ev), "rrReleaseEvent")
        finally:
            lib.rrReleaseCommandStream(ctx.ctx, cs)
        assert _device_errors(ctx) == 0  # the wait consumed the report


def test_sort_lookback_timeout_is_reported(ctx, dev):
    import torch

    say("sort look-back bound 0: give-ups reported, default bound sorts right")
    n = 1 << 24
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    keys0 = torch.randint(0, 1 << 30, (n,), device=dev, dtype=torch.int32, generator=g)
    vals0 = torch.arange(n, device=dev, dtype=torch.int32)
    lib = ctx.lib
    assert _device_errors(ctx) == 0
    seen = 0
    try:
        rr.check(lib.rrxSetTuning(rr.RRX_TUNING_SORT_SPIN_LIMIT, 0), "rrxSetTuning")
        for _ in range(3):  # a give-up needs a tile to find a predecessor unpublished
            k, v = keys0.clone(), vals0.clone()
            ctx.sort_pairs(k, v, key_bits=30)
            torch.cuda.synchronize()
            seen |= _device_errors(ctx)
            if seen:
                break
    finally:
        rr.check(lib.rrxSetTuning(rr.RRX_TUNING_SORT_SPIN_LIMIT, 1 << 22), "rrxSetTuning")
    assert seen == rr.RRX_DEVICE_ERROR_SORT_LOOKBACK
    k, v = keys0.clone(), vals0.clone()
    ctx.sort_pairs(k, v, key_bits=30)
    torch.cuda.synchronize()
    assert _device_errors(ctx) == 0
    hk = keys0.cpu().numpy().view(np.uint32)
    order = np.argsort(hk, kind="stable")
    assert np.array_equal(k.cpu().numpy().view(np.uint32), hk[order])
    assert np.array_equal(v.cpu().numpy(), order.astype(np.int32))


def test_buffer_without_header_is_reported(ctx, dev):
    import torch

    say("queries on a buffer without a valid header")
    junk = api.Accel(ctx.empty(4096).zero_(), "blas", 0)
    rays = to_dev(_rays([[0, 0, 0]] * 100, [[1, 0, 0]] * 100), dev)
    ctx.intersect(junk, rays, "closest", "full")
    torch.cuda.synchronize()
    assert _device_errors(ctx) == rr.RRX_DEVICE_ERROR_INVALID_BUFFER
    # In a batch: reported too, and the other requests are answered.
    v, i = scenes.triangle_soup(2000, scenes.rng_for(60))
    good = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev))
    r2 = scenes.camera_rays(32, 32, (0.5, 0.5, -1.5))
    d2 = to_dev(r2, dev)
    outs = ctx.intersect_batch([(junk, rays, "closest", "full"), (good, d2, "closest", "full_t")])
    torch.cuda.synchronize()
    assert _device_errors(ctx) == rr.RRX_DEVICE_ERROR_INVALID_BUFFER
    ok, msg = hits_equal(api.decode_hits(outs[1], "full_t"), oracle.trace(good.buffer.cpu().numpy(), r2))
    assert ok, msg


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_buffers_the_context_did_not_build(ctx, dev, mode):
    """Copies of a BLAS and a TLAS (addresses the context never recorded a
    build into) are traced at the level their header names, decided on the
    device: the hits equal those of the originals."""
    import torch

    say(f"queries on copied BLAS/TLAS buffers, {mode} kernels")
    rng = scenes.rng_for(61)
    v, i = scenes.triangle_soup(3000, rng, sigma=0.08)
    blas = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True)
    xf = scenes.random_transforms(rng, 20, trans=3.0)
    tlas = ctx.build_scene([blas] * 20, xf)
    torch.cuda.synchronize()
    o = rng.uniform(-4, 4, size=(30000, 3)).astype(np.float32)
    rays = to_dev(scenes.make_rays(o, scenes.unit_sphere(30000, rng)), dev)
    saved = autotune.current(ctx.lib)
    try:
        if mode == "plain":
            autotune.set_(ctx.lib, rr.RRX_TUNING_TRACE_PLAIN, 1)
        for orig in (blas, tlas):
            copy = api.Accel(orig.buffer.clone(), orig.kind, orig.prim_count, orig.children)
            for query in ("closest", "any"):
                for out in ("full", "instance", "full_t"):
                    a = ctx.intersect(orig, rays, query, out)
                    b = ctx.intersect(copy, rays, query, out)
                    torch.cuda.synchronize()
                    assert torch.equal(a, b), f"{orig.kind} {query}/{out}"
        assert _device_errors(ctx) == 0
    finally:
        autotune.apply(ctx.lib, saved)


def test_unknown_primitive_type_is_not_implemented(ctx, dev):
    say("unknown primitive type on a live context")
    lib = ctx.lib
    v = to_dev(np.zeros((3, 3), np.float32), dev)
    i = to_dev(np.zeros((1, 3), np.int32), dev)
    out = ctx.empty(4096)
    with ctx.ptrs(v.data_ptr(), i.data_ptr(), out.data_ptr()) as (pv, pi, po):
        prim = rr.RRTriangleMeshPrimitive(pv, 3, 12, pi, 1, rr.RR_INDEX_TYPE_UINT32)
        inp = rr.RRGeometryBuildInput(7, 1, ctypes.cast(ctypes.pointer(prim), c_void_p))
        assert (lib.rrCmdBuildGeometry(ctx.ctx, rr.RR_BUILD_OPERATION_BUILD, byref(inp), None, po, po, ctx.cs) ==
                rr.RR_ERROR_NOT_IMPLEMENTED)


def test_released_handles_are_rejected(ctx, dev):
    say("released device pointer and event handles")
    lib = ctx.lib
    t = ctx.empty(256)
    p = c_void_p()
    rr.check(lib.rrGetDevicePtrFromHipPtr(ctx.ctx, c_void_p(t.data_ptr()), 0, byref(p)), "rrGetDevicePtrFromHipPtr")
    assert lib.rrGetHipAddress(p) == t.data_ptr()
    rr.check(lib.rrReleaseDevicePtr(ctx.ctx, p), "rrReleaseDevicePtr")
    assert lib.rrReleaseDevicePtr(ctx.ctx, p) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrGetHipAddress(p) is None
    cs, ev = c_void_p(), c_void_p()
    rr.check(lib.rrAllocateCommandStream(ctx.ctx, byref(cs)), "rrAllocateCommandStream")
    rr.check(lib.rrSumbitCommandStream(ctx.ctx, cs, None, byref(ev)), "rrSumbitCommandStream")
    rr.check(lib.rrWaitEvent(ctx.ctx, ev), "rrWaitEvent")
    rr.check(lib.rrReleaseEvent(ctx.ctx, ev), "rrReleaseEvent")
    assert lib.rrReleaseEvent(ctx.ctx, ev) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrWaitEvent(ctx.ctx, ev) == rr.RR_ERROR_INVALID_PARAMETER
    rr.check(lib.rrReleaseCommandStream(ctx.ctx, cs), "rrReleaseCommandStream")
    assert lib.rrReleaseCommandStream(ctx.ctx, cs) == rr.RR_ERROR_INVALID_PARAMETER
