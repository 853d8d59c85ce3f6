"""GPU parity of AABB-list geometries (RR_PRIMITIVE_TYPE_AABB_LIST): build,
refit and traversal (single level, batched, and under a scene next to a
triangle geometry) against the oracle."""
import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, scenes
from tests.gpu_util import RR_INVALID, hits_equal, nodes_equal, say, to_dev

pytestmark = pytest.mark.gpu


def _boxes(n, rng, size=0.03):
    c = rng.uniform(0, 1, size=(n, 3))
    e = rng.uniform(0.001, size, size=(n, 3))
    return np.concatenate([c - e, c + e], axis=1).astype(np.float32)


def _rays(n, rng):
    o = rng.uniform(-0.5, 1.5, size=(n, 3)).astype(np.float32)
    return scenes.make_rays(o, scenes.unit_sphere(n, rng))


@pytest.mark.parametrize("n", [1, 2, 65, 5000, 70000])
@pytest.mark.parametrize("restructure", [False, True])
def test_aabb_build_trace_refit(ctx, dev, n, restructure):
    import torch

    say(f"aabb list n={n} restructure={restructure}")
    rng = scenes.rng_for(61, n)
    b = _boxes(n, rng)
    accel = ctx.build_aabbs(to_dev(b, dev), restructure=restructure)
    torch.cuda.synchronize()
    buf = accel.buffer.cpu().numpy()
    assert buf[:128].view(np.uint32)[10] & 2
    ok, msg = nodes_equal(buf, oracle.build_blas_aabbs(b, restructure))
    assert ok, msg
    rays = _rays(20000, rng)
    drays = to_dev(rays, dev)
    for query in ("closest", "any"):
        hits = ctx.intersect(accel, drays, query, "full_t")
        torch.cuda.synchronize()
        g = api.decode_hits(hits, "full_t")
        ok, msg = hits_equal(g, oracle.trace(buf, rays, any_hit=(query == "any")))
        assert ok, f"{query}: {msg}"
    # Refit to moved boxes.
    b2 = b + rng.normal(0, 1e-3, size=(n, 1)).astype(np.float32)
    ctx.build_aabbs(to_dev(b2, dev), update=accel)
    torch.cuda.synchronize()
    v2 = np.stack([b2[:, 0:3], b2[:, 3:6], b2[:, 0:3]], axis=1).reshape(-1, 3)
    ok, msg = nodes_equal(accel.buffer.cpu().numpy(),
                          oracle.refit(buf, v2, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)))
    assert ok, "refit: " + msg


def test_aabbs_in_batches_and_scenes(ctx, dev):
    import torch

    say("aabb geometry in a batched query and under a scene with a triangle geometry")
    rng = scenes.rng_for(62)
    boxes = ctx.build_aabbs(to_dev(_boxes(4000, rng, 0.08), dev), restructure=True)
    v, i = scenes.triangle_soup(4000, rng, sigma=0.08)
    tris = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True)
    rays = _rays(30000, rng)
    drays = to_dev(rays, dev)
    alone = [ctx.intersect(boxes, drays, "closest", "full_t"), ctx.intersect(tris, drays, "any", "full")]
    batch = ctx.intersect_batch([(boxes, drays, "closest", "full_t"), (tris, drays, "any", "full")])
    torch.cuda.synchronize()
    for a, b in zip(alone, batch):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    xf = scenes.random_transforms(rng, 12, scale=(0.5, 1.0), trans=1.0)
    pick = [boxes if k % 2 == 0 else tris for k in range(12)]
    tlas = ctx.build_scene(pick, xf)
    torch.cuda.synchronize()
    host = {boxes.address: boxes.buffer.cpu().numpy(), tris.address: tris.buffer.cpu().numpy()}
    buf = tlas.buffer.cpu().numpy()
    o = rng.uniform(-3, 3, size=(20000, 3)).astype(np.float32)
    srays = scenes.make_rays(o, scenes.unit_sphere(20000, rng))
    for query in ("closest", "any"):
        hits = ctx.intersect(tlas, to_dev(srays, dev), query, "full_t")
        torch.cuda.synchronize()
        g = api.decode_hits(hits, "full_t")
        ok, msg = hits_equal(g, oracle.trace(buf, srays, any_hit=(query == "any"), blas_map=host))
        assert ok, f"{query}: {msg}"
    g = api.decode_hits(ctx.intersect(tlas, to_dev(srays, dev), "closest", "full_t"), "full_t")
    torch.cuda.synchronize()
    ok, msg = hits_equal(g, oracle.brute_force(buf, srays, blas_map=host))
    assert ok, "brute force: " + msg
    assert (g["inst_id"] != RR_INVALID).sum() > 100
