"""Full-size parity of the batched-build and serving workloads.

Batched build / refit (8,192 soups, log-uniform 64-16,384 triangles, about
24.4M in all): every mesh of one rrxCmdBuildGeometryBatch equals, byte for
byte, what rrCmdBuildGeometry builds for it alone (fast and restructured
builds), and sampled meshes equal the oracle's build.  Then 16 frames, each
jittering every vertex by N(0, 1e-3^2) from the base mesh, refitting all
BLASes and tracing 256 any-hit rays per mesh, captured once in a HIP graph
and replayed: per frame, the sampled meshes' BVHs equal the oracle's refit
of the same vertices, and their any-hit results equal the oracle's any-hit
traversal of that BVH.

Serving stream (32 resident scenes, log-uniform 10k-2M triangles, requests
of log-uniform 1k-1M rays, 70 % closest / 30 % any): every request served by
the batched server equals its own rrCmdIntersect call byte for byte, and
sampled requests equal the oracle.
"""
import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, scenes, serving
from tests.gpu_util import hits_equal, nodes_equal, say, to_dev

pytestmark = pytest.mark.gpu

SAMPLE = 32


@pytest.fixture(scope="module")
def batched_world(ctx, dev):
    import torch

    say("generating the 8,192-mesh batched workload")
    w = scenes.batched()
    V = torch.from_numpy(w.vertices).to(dev)
    I = torch.arange(3 * int(w.tri_counts.max()), dtype=torch.int32, device=dev).reshape(-1, 3)
    spans = [(int(o), int(n)) for o, n in zip(w.vertex_offsets, w.tri_counts)]
    meshes = [(V[o:o + 3 * n], I[:n]) for o, n in spans]
    rng = np.random.default_rng(3)
    sample = sorted(set(rng.choice(len(spans), SAMPLE, replace=False).tolist()) | {0, len(spans) - 1})
    yield w, V, I, spans, meshes, sample
    del V, I, meshes
    torch.cuda.empty_cache()


def _packed(ctx, spans, restructure):
    """One zeroed device buffer holding every mesh's result buffer (each
    256-byte aligned), and the per-mesh views into it."""
    sizes = [ctx.geometry_requirements(n, 3 * n, restructure).result_buffer_size for _, n in spans]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    big = ctx.empty(int(sum(sizes))).zero_()
    return big, [big[int(o):int(o) + s] for o, s in zip(offs, sizes)]


@pytest.mark.parametrize("restructure", [False, True])
def test_full_batched_build_equals_per_mesh_builds_and_oracle(ctx, dev, batched_world, restructure):
    import torch

    w, V, I, spans, meshes, sample = batched_world
    say(f"batched build of {len(spans)} meshes ({int(w.tri_counts.sum())} tris), restructure={restructure}")
    big_a, res_a = _packed(ctx, spans, restructure)
    big_b, res_b = _packed(ctx, spans, restructure)
    accels = ctx.build_geometry_batch(meshes, restructure=restructure, results=res_a)
    scratch = ctx.empty(ctx.geometry_requirements(int(w.tri_counts.max()), restructure=restructure)
                        .temporary_build_buffer_size)
    for (v, i), r in zip(meshes, res_b):
        ctx.build_geometry(v, i, restructure, scratch=scratch, result=r)
    torch.cuda.synchronize()
    # Both buffers started zeroed, so headers (build serial 1) and padding agree too.
    if not torch.equal(big_a, big_b):
        bad = [k for k, (a, b) in enumerate(zip(res_a, res_b)) if not torch.equal(a, b)]
        pytest.fail(f"{len(bad)} meshes differ from their per-mesh builds; first {bad[:5]}")
    for k in sample:
        o, n = spans[k]
        v = w.vertices[o:o + 3 * n]
        ok, msg = nodes_equal(accels[k].buffer.cpu().numpy(), oracle.build_blas(v, np.arange(3 * n).reshape(-1, 3),
                                                                                   restructure))
        assert ok, f"oracle, mesh {k} (n={n}): {msg}"


def test_full_batched_refit_frames_from_a_graph(ctx, dev, batched_world):
    """16 jittered frames of (batched refit of all 8,192 BLASes, 256 any-hit
    rays per mesh in one batched query) replayed from one captured graph."""
    import torch

    w, V, I, spans, meshes, sample = batched_world
    M, rpm = len(spans), w.rays_per_mesh
    say(f"16 graph-replayed refit + any-hit frames over {M} meshes")
    accels = ctx.build_geometry_batch(meshes)
    built = {k: accels[k].buffer.cpu().numpy() for k in sample}
    cur = V.clone()
    cur_meshes = [(cur[o:o + 3 * n], I[:n]) for o, n in spans]
    rays = torch.from_numpy(w.rays).to(dev)
    hits = ctx.empty(rays.shape[0] * 32)
    requests = [(accels[k], rays[k * rpm:(k + 1) * rpm], "any", "full_t", hits[k * rpm * 32:(k + 1) * rpm * 32])
                for k in range(M)]
    bscratch = ctx.empty(ctx.batch_build_requirements(cur_meshes))
    qscratch = ctx.batch_scratch(M, rays.shape[0], rpm)

    def frame():
        ctx.update_geometry_batch(accels, cur_meshes, scratch=bscratch)
        ctx.intersect_batch(requests, scratch=qscratch)

    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        with ctx.on_stream(s):
            frame()  # warm-up outside capture
        s.synchronize()
        with torch.cuda.graph(g, stream=s, capture_error_mode="relaxed"):
            with ctx.on_stream(s):
                frame()
    torch.cuda.synchronize()
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x52520003)
    for f in range(16):
        torch.add(V, torch.randn(V.shape, generator=gen, device=dev), alpha=1e-3, out=cur)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        h = api.decode_hits(hits, "full_t")
        n_hit = 0
        for k in sample:
            o, n = spans[k]
            v = cur[o:o + 3 * n].cpu().numpy()
            ref = oracle.refit(built[k], v, np.arange(3 * n).reshape(-1, 3))
            ok, msg = nodes_equal(accels[k].buffer.cpu().numpy(), ref)
            assert ok, f"frame {f} mesh {k}: {msg}"
            r = oracle.trace(ref, w.rays[k * rpm:(k + 1) * rpm], any_hit=True)
            ok, msg = hits_equal({key: a[k * rpm:(k + 1) * rpm] for key, a in h.items()}, r)
            assert ok, f"frame {f} mesh {k} any-hit: {msg}"
            n_hit += int((r["prim_id"] != 0xFFFFFFFF).sum())
        assert n_hit > SAMPLE * rpm // 10


def test_full_serving_stream_same_alone_and_in_the_mix(ctx, dev):
    import torch

    say("serving: 32 scenes (10k-2M tris), 400 requests of 1k-1M rays")
    world = scenes.serving_scenes()
    blas = [ctx.build_geometry(to_dev(s.vertices, dev), to_dev(s.indices, dev).reshape(-1, 3), restructure=True)
            for s in world]
    pool_rays = 1 << 20
    pools = scenes.serving_ray_pools(pool_rays)
    dpools = [to_dev(p, dev) for p in pools]
    reqs = scenes.serving_requests(400, world, pool_rays, stream=9)
    hits = [ctx.empty(q.count * 32) for q in reqs]
    served = [serving.Request(blas[q.scene], dpools[q.pool], q.offset * 32, q.count, q.query, "full_t", h)
              for q, h in zip(reqs, hits)]
    torch.cuda.synchronize()
    with serving.Server(ctx, mode="batched", max_inflight=3, max_batch_requests=256, max_batch_rays=4 << 20) as srv:
        t0 = srv.now() + 2_000_000
        arrivals = scenes.poisson_arrivals(len(served), 3000.0, stream=9)
        first = srv.submit(served, [t0 + int(a) for a in arrivals])
        srv.drain()
        t = srv.timings(first, len(served))
    assert t["batch_size"].max() > 1  # requests were coalesced
    assert sum(q.count for q in reqs) > 20_000_000
    for k, q in enumerate(reqs):
        alone = ctx.intersect(blas[q.scene], dpools[q.pool][q.offset:q.offset + q.count], q.query, "full_t")
        torch.cuda.synchronize()
        assert torch.equal(hits[k][: q.count * 32], alone[: q.count * 32]), f"request {k}"
    # Sampled requests (first 20,000 rays of each) against the oracle on the
    # downloaded BVH, whose nodes equal the oracle's build for two scenes.
    for si in (0, 1):
        ok, msg = nodes_equal(blas[si].buffer.cpu().numpy(), oracle.build_blas(world[si].vertices, world[si].indices,
                                                                                True))
        assert ok, f"scene {si}: {msg}"
    bufs = {}
    for k in range(0, len(reqs), 25):
        q = reqs[k]
        n = min(q.count, 20_000)
        if q.scene not in bufs:
            bufs[q.scene] = blas[q.scene].buffer.cpu().numpy()
        ref = oracle.trace(bufs[q.scene], pools[q.pool][q.offset:q.offset + n], any_hit=(q.query == "any"))
        got = api.decode_hits(hits[k][: n * 32], "full_t")
        ok, msg = hits_equal(got, ref)
        assert ok, f"request {k} (scene {q.scene}, {q.query}): {msg}"
