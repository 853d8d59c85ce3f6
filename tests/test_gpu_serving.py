"""GPU tests of the batched query path and the serving loop.

rrxCmdIntersectBatch and both serving modes must write, for every request,
exactly the bytes an rrCmdIntersect call with the same arguments writes on
its own (and those match the oracle, checked here on a subset).  The server
tests also check the bookkeeping: every request completes once, times are
ordered, batched mode coalesces requests, plain mode never does, and the
memory budget bounds what is admitted at once.
"""
import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, scenes, serving
from tests.gpu_util import hits_equal, say, to_dev

pytestmark = pytest.mark.gpu

HIT_BYTES = {"full": 16, "instance": 4, "full_t": 32}


@pytest.fixture(scope="module")
def small_world(ctx, dev):
    """Eight BLASes (soups and heightfields, one empty), one TLAS over two of
    them, and the four ray pools at 64K rays each."""
    import torch

    blas, bufs = [], []
    specs = [("soup", 5000), ("terrain", 9), ("soup", 1), ("soup", 70000), ("terrain", 120), ("soup", 0),
             ("soup", 300), ("terrain", 40)]
    for k, (kind, n) in enumerate(specs):
        if kind == "soup":
            v, i = scenes.triangle_soup(n, scenes.rng_for(40, k))
            if n == 0:
                v = np.zeros((3, 3), np.float32)
        else:
            v, i = scenes.heightfield_mesh(n, seed=scenes.SEED_BASE + 40 + k)
        a = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev).reshape(-1, 3), restructure=(k % 2 == 0))
        blas.append(a)
    tlas = ctx.build_scene([blas[0], blas[3]], scenes.random_transforms(scenes.rng_for(40, 99), 2, trans=0.5))
    torch.cuda.synchronize()
    bufs = [a.buffer.cpu().numpy() for a in blas]
    pools = scenes.serving_ray_pools(1 << 16)
    kinds = [k for k, _ in specs]
    return blas, bufs, tlas, pools, [to_dev(p, dev) for p in pools], kinds


def _alone(ctx, accel, rays, query, output):
    import torch

    h = ctx.intersect(accel, rays, query, output)
    torch.cuda.synchronize()
    return h.cpu().numpy()


def test_batch_matches_individual_calls(ctx, dev, small_world):
    import torch

    blas, bufs, tlas, pools, dpools, kinds = small_world
    rng = np.random.default_rng(5)
    counts = [0, 1, 63, 64, 65, 127, 4096, 20000, 65536, 3, 777, 1000]
    reqs, meta = [], []
    for k in range(36):
        target = tlas if k % 9 == 8 else blas[k % len(blas)]
        kind = "soup" if target is tlas else kinds[k % len(blas)]
        pool = (0 if kind == "soup" else 2) + int(rng.integers(0, 2))
        n = counts[k % len(counts)]
        off = int(rng.integers(0, (1 << 16) - n + 1))
        query = "closest" if rng.uniform() < 0.7 else "any"
        output = ("full", "full_t", "instance")[k % 3]
        rays = dpools[pool][off:off + n]
        reqs.append((target, rays, query, output))
        meta.append((target, pool, off, n, query, output))
    say(f"rrxCmdIntersectBatch over {len(reqs)} mixed requests")
    outs = ctx.intersect_batch(reqs)
    torch.cuda.synchronize()
    got = [o.cpu().numpy() for o in outs]
    for k, (target, pool, off, n, query, output) in enumerate(meta):
        if n == 0:
            continue
        ref = _alone(ctx, target, dpools[pool][off:off + n], query, output)
        assert np.array_equal(got[k][: n * HIT_BYTES[output]], ref[: n * HIT_BYTES[output]]), f"request {k}"
    # A few against the oracle directly.
    for k in (5, 6, 7, 21):
        target, pool, off, n, query, output = meta[k]
        if target is tlas or output == "instance":
            continue
        idx = next(j for j, b in enumerate(blas) if b is target)
        ref = oracle.trace(bufs[idx], pools[pool][off:off + n], any_hit=(query == "any"))
        ok, msg = hits_equal(api.decode_hits(outs[k][: n * HIT_BYTES[output]], output), ref,
                             ("prim_id", "inst_id", "u", "v") + (("t",) if output == "full_t" else ()))
        assert ok, f"request {k}: {msg}"


def _stream(ctx, dev, small_world, n_requests, seed):
    blas, bufs, tlas, pools, dpools, kinds = small_world
    fake_scenes = [scenes.ServingScene(k, None, None) for k in kinds]
    reqs = scenes.serving_requests(n_requests, fake_scenes, 1 << 16, lo=10, hi=1 << 16, stream=seed)
    hits = [ctx.empty(q.count * 16) for q in reqs]
    served = [serving.Request(blas[q.scene], dpools[q.pool], q.offset * 32, q.count, q.query, "full", h)
              for q, h in zip(reqs, hits)]
    return reqs, served, hits


def _check_against_alone(ctx, small_world, reqs, hits):
    blas, bufs, tlas, pools, dpools, kinds = small_world
    for k, q in enumerate(reqs):
        ref = _alone(ctx, blas[q.scene], dpools[q.pool][q.offset:q.offset + q.count], q.query, "full")
        assert np.array_equal(hits[k].cpu().numpy()[: q.count * 16], ref[: q.count * 16]), f"request {k}"


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_server_results_same_alone_and_in_the_mix(ctx, dev, small_world, mode):
    import torch

    say(f"server mode={mode}: 300 requests, Poisson arrivals at 50k QPS")
    reqs, served, hits = _stream(ctx, dev, small_world, 300, seed=1)
    torch.cuda.synchronize()
    with serving.Server(ctx, mode=mode, max_inflight=3, max_batch_requests=64, max_batch_rays=1 << 18) as srv:
        t0 = srv.now() + 2_000_000
        arrivals = scenes.poisson_arrivals(len(served), 50_000.0, stream=1)
        first = srv.submit(served, [t0 + int(a) for a in arrivals])
        srv.drain()
        t = srv.timings(first, len(served))
    assert (t["arrival_ns"] <= t["dispatch_ns"]).all() and (t["dispatch_ns"] <= t["complete_ns"]).all()
    assert (t["arrival_ns"] >= t0).all()
    if mode == "plain":
        assert (t["batch_size"] == 1).all()
        assert len(np.unique(t["batch_id"])) == len(served)
    else:
        assert (t["batch_size"] >= 1).all() and (t["batch_size"] <= 64).all()
        # Requests of one dispatch share its id and size.
        for b in np.unique(t["batch_id"]):
            sel = t["batch_id"] == b
            assert (t["batch_size"][sel] == sel.sum()).all()
    _check_against_alone(ctx, small_world, reqs, hits)


def test_server_memory_budget_bounds_admission(ctx, dev, small_world):
    import torch

    say("server with a 4 MiB admission budget, all requests released at once")
    reqs, served, hits = _stream(ctx, dev, small_world, 120, seed=2)
    torch.cuda.synchronize()
    budget = 4 << 20
    with serving.Server(ctx, mode="batched", max_inflight=2, max_batch_requests=32, max_batch_rays=1 << 16,
                        memory_budget=budget) as srv:
        first = srv.submit(served)
        srv.drain()
        t = srv.timings(first, len(served))
        peak = srv.peak_admitted()
    total = sum(q.count * 48 for q in reqs)
    assert total > 2 * budget
    assert 0 < peak <= budget
    assert t["batch_size"].max() > 1  # released together: coalesced
    _check_against_alone(ctx, small_world, reqs, hits)


def test_server_rejects_oversized_requests(ctx, dev, small_world):
    blas, bufs, tlas, pools, dpools, kinds = small_world
    h = ctx.empty(1000 * 16)
    big = serving.Request(blas[0], dpools[0], 0, 1000, "closest", "full", h)
    with serving.Server(ctx, mode="batched", max_batch_rays=999) as srv:
        with pytest.raises(Exception):
            srv.submit([big])
    with serving.Server(ctx, mode="batched", max_batch_rays=1 << 16, memory_budget=1000 * 48 - 1) as srv:
        with pytest.raises(Exception):
            srv.submit([big])


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_host_requests_staged_through_a_bounded_arena(ctx, dev, small_world, mode):
    """Host-memory requests (rrxServerSubmitHost) under a staging arena far
    smaller than the requests in all: every request's hits equal its own
    rrCmdIntersect call, admission never holds more than the arena, and the
    server coalesces requests that were released together."""
    import torch

    blas, bufs, tlas, pools, dpools, kinds = small_world
    say(f"host requests, {mode} server, 2 MiB staging arena")
    fake_scenes = [scenes.ServingScene(k, None, None) for k in kinds]
    reqs = scenes.serving_requests(150, fake_scenes, 1 << 16, lo=10, hi=1 << 15, stream=7)
    outputs = [("full", "full_t", "instance")[k % 3] for k in range(len(reqs))]
    host_rays = [torch.from_numpy(pools[q.pool][q.offset:q.offset + q.count].copy()).pin_memory() for q in reqs]
    host_hits = [torch.zeros(q.count * HIT_BYTES[o], dtype=torch.uint8).pin_memory() for q, o in zip(reqs, outputs)]
    served = [serving.HostRequest(blas[q.scene], r, q.query, o, h)
              for q, r, o, h in zip(reqs, host_rays, outputs, host_hits)]
    arena = 2 << 20
    total = sum(q.count * (32 + HIT_BYTES[o]) for q, o in zip(reqs, outputs))
    assert total > 4 * arena
    with serving.Server(ctx, mode=mode, max_inflight=3, max_batch_requests=64, max_batch_rays=1 << 15,
                        memory_budget=arena, staging_bytes=arena) as srv:
        first = srv.submit_host(served)
        srv.drain()
        t = srv.timings(first, len(served))
        peak = srv.peak_admitted()
    assert 0 < peak <= arena
    if mode == "batched":
        assert t["batch_size"].max() > 1
    else:
        assert (t["batch_size"] == 1).all()
    for k, (q, o) in enumerate(zip(reqs, outputs)):
        ref = _alone(ctx, blas[q.scene], dpools[q.pool][q.offset:q.offset + q.count], q.query, o)
        assert np.array_equal(host_hits[k].numpy(), ref[: q.count * HIT_BYTES[o]]), f"request {k}"


def test_host_requests_need_an_arena(ctx, dev, small_world):
    import torch

    blas, bufs, tlas, pools, dpools, kinds = small_world
    rays = torch.from_numpy(pools[0][:100].copy())
    hits = torch.zeros(100 * 16, dtype=torch.uint8)
    req = serving.HostRequest(blas[0], rays, "closest", "full", hits)
    with serving.Server(ctx, mode="batched") as srv:  # no staging arena
        with pytest.raises(Exception):
            srv.submit_host([req])
    with serving.Server(ctx, mode="batched", staging_bytes=4096) as srv:  # arena smaller than the request
        with pytest.raises(Exception):
            srv.submit_host([req])


def _owned_copy(ctx, accel):
    """A byte copy of `accel` in memory from rrAllocateDeviceBuffer, written
    through rrMapDevicePtr / rrUnmapDevicePtr: an address this context
    recorded no level for.  Returns (handle, accel-like view)."""
    import ctypes
    from ctypes import byref, c_void_p
    from types import SimpleNamespace

    from radeonrays_sdk_amd import rr

    data = accel.buffer.cpu().numpy()
    h, m = c_void_p(), c_void_p()
    rr.check(ctx.lib.rrAllocateDeviceBuffer(ctx.ctx, data.nbytes, byref(h)), "rrAllocateDeviceBuffer")
    rr.check(ctx.lib.rrMapDevicePtr(ctx.ctx, h, byref(m)), "rrMapDevicePtr")
    ctypes.memmove(m.value, data.ctypes.data, data.nbytes)
    rr.check(ctx.lib.rrUnmapDevicePtr(ctx.ctx, h, byref(m)), "rrUnmapDevicePtr")
    view = SimpleNamespace(address=ctx.lib.rrGetHipAddress(h), kind=accel.kind, prim_count=accel.prim_count,
                           children=accel.children)
    return h, view


def _device_errors(ctx) -> int:
    from ctypes import byref, c_uint32

    from radeonrays_sdk_amd import rr

    e = c_uint32()
    rr.check(ctx.lib.rrxGetDeviceErrors(ctx.ctx, byref(e)), "rrxGetDeviceErrors")
    return e.value


def test_scene_requests_share_the_batch_launch(ctx, dev, small_world):
    """Scene (TLAS) requests go through the batched two-level kernel, mixed
    with geometry requests in one table: every request equals its own
    rrCmdIntersect call byte for byte, for both queries and all outputs, on
    a scene of 300 overlapping instances that also holds instances of an
    empty and of a one-triangle BLAS, and in batches of scenes only and of
    geometries only.  Buffers at addresses the context recorded no level for
    (copies in rrAllocateDeviceBuffer memory) are sent to the right kernel
    by their headers; one at an address recorded as the other level is
    reported (RRX_DEVICE_ERROR_INVALID_BUFFER) and answered with misses, and the
    rest of its batch is unaffected; releasing an owned buffer forgets the
    levels recorded in it."""
    import torch

    from radeonrays_sdk_amd import rr

    blas, bufs, tlas_small, pools, dpools, kinds = small_world
    rng = scenes.rng_for(40, 500)
    empty = ctx.build_geometry(to_dev(np.zeros((3, 3), np.float32), dev),
                               to_dev(np.zeros((0, 3), np.int32), dev).reshape(0, 3))
    v1, i1 = scenes.triangle_soup(1, rng, sigma=0.5)
    one = ctx.build_geometry(to_dev(v1, dev), to_dev(i1.reshape(-1, 3), dev))
    members = [blas[k] for k in rng.integers(0, len(blas), size=300)] + [empty, one] * 10
    big = ctx.build_scene(members, scenes.random_transforms(rng, len(members), trans=0.6))
    torch.cuda.synchronize()
    scenes_ = [big, tlas_small]
    h_geo, geo_copy = _owned_copy(ctx, blas[6])
    h_scene, scene_copy = _owned_copy(ctx, tlas_small)
    originals = {id(geo_copy): blas[6], id(scene_copy): tlas_small}

    def batch(targets, seed, stale=()):
        r = np.random.default_rng(seed)
        reqs = []
        for k, target in enumerate(targets):
            n = int((0, 1, 64, 65, 1000, 4097, 20000)[k % 7])
            off = int(r.integers(0, (1 << 16) - n + 1))
            pool = int(r.integers(0, 4))
            reqs.append((target, dpools[pool][off:off + n], ("closest", "any")[k % 2], ("full", "full_t", "instance")[k % 3]))
        assert _device_errors(ctx) == 0
        outs = ctx.intersect_batch(reqs)
        torch.cuda.synchronize()
        assert _device_errors(ctx) == (rr.RRX_DEVICE_ERROR_INVALID_BUFFER if stale else 0)
        for k, (target, rays, query, output) in enumerate(reqs):
            n = int(rays.shape[0])
            if n == 0:
                continue
            got = api.decode_hits(outs[k][: n * HIT_BYTES[output]], output)
            if k in stale:
                assert (got["prim_id" if output != "instance" else "id"] == 0xFFFFFFFF).all(), f"request {k}"
                continue
            ref = _alone(ctx, originals.get(id(target), target), rays, query, output)
            assert np.array_equal(outs[k].cpu().numpy()[: n * HIT_BYTES[output]], ref[: n * HIT_BYTES[output]]), \
                f"request {k} ({target.kind}, {query}, {output}, {n} rays)"

    say("batched scene + geometry requests, scene-only and geometry-only batches, unrecorded and stale addresses")
    batch([scenes_[k % 2] if k % 3 else blas[k % len(blas)] for k in range(30)], 1)
    batch([scenes_[k % 2] for k in range(14)], 2)
    batch([blas[k % len(blas)] for k in range(14)], 3)
    batch([blas[1], big, geo_copy, scene_copy, geo_copy, scene_copy, blas[4]], 4)
    batch([blas[1], blas[4], scene_copy, scene_copy, scene_copy], 5)  # no recorded scene: the two-level kernel
    batch([big, tlas_small, geo_copy, geo_copy, geo_copy], 6)         # still runs, for the unrecorded ones
    # A geometry at an address recorded as a scene: reported, answered with misses.
    recorded = ctx.build_scene([blas[0]] * 200, scenes.random_transforms(rng, 200, trans=0.5))
    assert blas[6].buffer.numel() <= recorded.buffer.numel()
    recorded.buffer[: blas[6].buffer.numel()].copy_(blas[6].buffer)
    stale_geo = api.Accel(recorded.buffer, "blas", blas[6].prim_count)
    torch.cuda.synchronize()
    batch([blas[1], blas[4], stale_geo, stale_geo, big], 7, stale=(2, 3))
    # Releasing owned memory forgets what was built into it: a geometry built
    # into rrAllocateDeviceBuffer memory (recorded as a geometry), the memory
    # released, a scene copied into the next allocation (often at the same
    # address): traced as the scene.
    from ctypes import byref, c_void_p

    class _Owned:  # what build_geometry needs of a result buffer
        def __init__(self, addr, nbytes):
            self.addr, self.n = addr, nbytes
        def numel(self):
            return self.n
        def element_size(self):
            return 1
        def data_ptr(self):
            return self.addr

    vv, ii = scenes.triangle_soup(10, rng)
    nbytes = max(int(tlas_small.buffer.numel()), int(ctx.geometry_requirements(10, 30, False).result_buffer_size))
    h = c_void_p()
    rr.check(ctx.lib.rrAllocateDeviceBuffer(ctx.ctx, nbytes, byref(h)), "rrAllocateDeviceBuffer")
    out = ctx.lib.rrGetHipAddress(h)
    ctx.build_geometry(to_dev(vv, dev), to_dev(ii, dev), result=_Owned(out, nbytes))
    torch.cuda.synchronize()
    rr.check(ctx.lib.rrReleaseDevicePtr(ctx.ctx, h), "rrReleaseDevicePtr")
    h_again, again = _owned_copy(ctx, tlas_small)
    originals[id(again)] = tlas_small
    say(f"owned address reused after release: {again.address == out}")
    batch([blas[1], again, again, again], 8)
    for hh in (h_geo, h_scene, h_again):
        rr.check(ctx.lib.rrReleaseDevicePtr(ctx.ctx, hh), "rrReleaseDevicePtr")
    # The big scene's hits against the oracle on a sample.
    rays = pools[0][:3000]
    hits = ctx.intersect_batch([(big, to_dev(rays, dev), "closest", "full")])[0]
    torch.cuda.synchronize()
    host = {b.address: b.buffer.cpu().numpy() for b in members}
    ref = oracle.trace(big.buffer.cpu().numpy(), rays, blas_map=host)
    ok, msg = hits_equal(api.decode_hits(hits, "full"), ref, ("prim_id", "inst_id", "u", "v"))
    assert ok, msg
    assert (ref["inst_id"] != 0xFFFFFFFF).sum() > 300
