"""GPU parity of the TLAS build and two-level traversal."""
import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, autotune, scenes
from tests.gpu_util import RR_INVALID, hits_equal, nodes_equal, say, tlas_tables_equal, to_dev

pytestmark = pytest.mark.gpu


def _scene(ctx, dev, n_meshes, n_inst, tris_lo, tris_hi, seed):
    import torch

    rng = scenes.rng_for(2, seed)
    counts = scenes.log_uniform_ints(rng, n_meshes, tris_lo, tris_hi)
    blases, meshes = [], []
    for c in counts:
        v, i = scenes.triangle_soup(int(c), rng, sigma=0.08)
        meshes.append((v, i))
        blases.append(ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True))
    pick = rng.integers(0, n_meshes, size=n_inst)
    xf = scenes.random_transforms(rng, n_inst, trans=5.0)
    tlas = ctx.build_scene([blases[k] for k in pick], xf)
    torch.cuda.synchronize()
    host = {b.address: b.buffer.cpu().numpy() for b in blases}
    return tlas, blases, pick, xf, host


@pytest.mark.parametrize("n_inst", [1, 2, 37, 300, 4096])
def test_tlas_build_and_trace(ctx, dev, n_inst):
    say(f"tlas build + two-level traversal, {n_inst} instances")
    import torch

    tlas, blases, pick, xf, host = _scene(ctx, dev, 5, n_inst, 50, 2000, n_inst)
    buf = tlas.buffer.cpu().numpy()
    ref = oracle.build_tlas(xf, [blases[k].address for k in pick], [host[blases[k].address] for k in pick])
    ok, msg = nodes_equal(buf, ref)
    assert ok, msg
    # The GPU's inverse and forward transforms and its BLAS address table
    # equal the oracle's bit for bit (lbvh_fit_aabb_scene.comp:118-119).
    ok, msg = tlas_tables_equal(buf, ref)
    assert ok, msg
    assert oracle.is_valid(buf) and oracle.check_consistency(buf)
    rng = scenes.rng_for(2, 1000 + n_inst)
    o = rng.uniform(-8, 8, size=(20000, 3)).astype(np.float32)
    rays = scenes.make_rays(o, scenes.unit_sphere(20000, rng))
    drays = to_dev(rays, dev)
    for query in ("closest", "any"):
        hits = ctx.intersect(tlas, drays, query, "full_t")
        torch.cuda.synchronize()
        gpu = api.decode_hits(hits, "full_t")
        # The oracle traverses its own TLAS (its own transforms), not the GPU's.
        r = oracle.trace(ref, rays, any_hit=(query == "any"), blas_map=host)
        ok, msg = hits_equal(gpu, r)
        assert ok, f"{query}: {msg}"
        if query == "closest":
            bf = oracle.brute_force(ref, rays[:3000], blas_map=host)
            ok, msg = hits_equal({k: a[:3000] for k, a in gpu.items()}, bf)
            assert ok, "brute force: " + msg
            assert (gpu["inst_id"] != RR_INVALID).sum() > 50
    ids = ctx.intersect(tlas, drays, "closest", "instance")
    torch.cuda.synchronize()
    r = oracle.trace(ref, rays, blas_map=host)
    assert np.array_equal(api.decode_hits(ids, "instance")["id"], r["inst_id"])


@pytest.mark.parametrize("depth", [8, 16, 32])
def test_persistent_two_level_equals_plain(ctx, dev, depth):
    """The persistent two-level kernel (rays handed out from queues, leaves
    postponed, instances entered and left mid-wave) returns exactly the plain
    kernel's hits, for every query and output and each LDS stack depth, on a
    scene whose instances overlap heavily (many instance entries per ray)."""
    say(f"persistent vs plain two-level traversal, LDS stack {depth}")
    import torch

    tlas, blases, pick, xf, host = _scene(ctx, dev, 6, 500, 1, 3000, 77)
    # Add instances of an empty BLAS and of a one-triangle BLAS (root = leaf).
    empty = ctx.build_geometry(to_dev(np.zeros((3, 3), np.float32), dev),
                               to_dev(np.zeros((0, 3), np.int32), dev).reshape(0, 3))
    v1, i1 = scenes.triangle_soup(1, scenes.rng_for(2, 3077), sigma=0.5)
    one = ctx.build_geometry(to_dev(v1, dev), to_dev(i1.reshape(-1, 3), dev))
    extra = [empty, one] * 20
    xf = np.concatenate([xf, scenes.random_transforms(scenes.rng_for(2, 4077), len(extra), trans=5.0)])
    all_blases = [blases[k] for k in pick] + extra
    tlas = ctx.build_scene(all_blases, xf)
    torch.cuda.synchronize()
    host.update({b.address: b.buffer.cpu().numpy() for b in (empty, one)})
    rng = scenes.rng_for(2, 2077)
    n = 50000
    o = rng.uniform(-8, 8, size=(n, 3)).astype(np.float32)
    rays = scenes.make_rays(o, scenes.unit_sphere(n, rng))
    rays[::7, 7] = -1.0  # some rays with an empty interval
    drays = to_dev(rays, dev)
    lib = ctx.lib
    saved = autotune.current(lib)
    try:
        autotune.set_(lib, autotune.rr.RRX_TUNING_TRACE_LDS_STACK, depth)
        for query in ("closest", "any"):
            for out in ("full", "instance", "full_t"):
                got = ctx.intersect(tlas, drays, query, out)
                with autotune.plain_traversal(lib):
                    ref = ctx.intersect(tlas, drays, query, out)
                torch.cuda.synchronize()
                assert torch.equal(got, ref), f"{query}/{out}"
        buf = tlas.buffer.cpu().numpy()
        hits = ctx.intersect(tlas, drays, "any", "full_t")
        torch.cuda.synchronize()
        ok, msg = hits_equal(api.decode_hits(hits, "full_t"), oracle.trace(buf, rays, any_hit=True, blas_map=host))
        assert ok, msg
    finally:
        autotune.apply(lib, saved)


@pytest.mark.parametrize("plain", [False, True])
def test_trace_follows_the_header_not_the_recorded_level(ctx, dev, plain):
    """The context records the level of every buffer it builds, by address,
    and uses it to size the launches; the caller may free such a buffer and
    put another at the same address (a copy, a build elsewhere, a host
    upload), so traversal must follow the buffer's own header.  A BLAS
    copied over a recorded TLAS, and a TLAS copied over a recorded BLAS,
    trace exactly as the originals, in the persistent and the plain kernels
    and with visit counters."""
    say(f"level from the header over a recycled address, {'plain' if plain else 'persistent'} kernels")
    import torch

    tlas, blases, pick, xf, host = _scene(ctx, dev, 3, 300, 50, 400, 91)
    rng = scenes.rng_for(2, 5091)
    v, i = scenes.triangle_soup(100, rng, sigma=0.3)
    blas = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev))
    assert blas.buffer.numel() < tlas.buffer.numel(), "the BLAS must fit in the TLAS's buffer"
    n = 20000
    rays = scenes.make_rays(rng.uniform(-2, 2, size=(n, 3)).astype(np.float32), scenes.unit_sphere(n, rng))
    drays = to_dev(rays, dev)
    # A BLAS at an address recorded as a TLAS.
    recorded = ctx.build_scene([blases[k] for k in pick], xf)
    over_tlas = recorded.buffer
    over_tlas[: blas.buffer.numel()].copy_(blas.buffer)
    as_blas = api.Accel(over_tlas, "blas", 100)
    # A TLAS at an address recorded as a BLAS (built into a buffer the TLAS fits in).
    big = ctx.empty(tlas.buffer.numel() + 4096)
    ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), result=big)
    big[: tlas.buffer.numel()].copy_(tlas.buffer)
    as_tlas = api.Accel(big, "tlas", len(pick), tlas.children)
    torch.cuda.synchronize()
    with autotune.plain_traversal(ctx.lib, plain):
        for copy, orig in ((as_blas, blas), (as_tlas, tlas)):
            for query in ("closest", "any"):
                for out in ("full", "instance", "full_t"):
                    got = ctx.intersect(copy, drays, query, out)
                    ref = ctx.intersect(orig, drays, query, out)
                    torch.cuda.synchronize()
                    assert torch.equal(got, ref), f"{copy.kind} copy {query}/{out}"
            gh, gs = ctx.intersect_with_stats(copy, drays, "closest")
            rh, rs = ctx.intersect_with_stats(orig, drays, "closest")
            torch.cuda.synchronize()
            assert torch.equal(gh, rh) and torch.equal(gs, rs), f"{copy.kind} copy with stats"
        for orig in (blas, tlas):
            ref = api.decode_hits(ctx.intersect(orig, drays, "closest", "full"), "full")
            assert (ref["prim_id"] != RR_INVALID).sum() > 200, orig.kind
