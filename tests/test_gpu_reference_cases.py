"""The reference's own test inputs on the HIP path, bit-exact against the oracle.

* resources/cornell_box.obj (as tests/golden/cornell_box.npz): axis-aligned
  walls, so zero-extent boxes and duplicate Morton codes, and negative zeros
  among its coordinates.
* BuildObj (test/test_vk/basic_test.h:373-750): one BLAS over the mesh, fast
  and restructured builds, traced with the reference's ray fan
  (basic_test.h:527-549: origin (0, 15, 0), direction (-1, -1 + 2y/R,
  -1 + 2x/R), t in [0.001, 1e5], R = 2048).  The fan was written for
  Sponza's scale: around the box at its own size every ray misses, so the
  box is also traced scaled by 20, which puts the fan's origin inside it.
* BuildObj2Level (basic_test.h:752-1069): one BLAS per object of the file
  and one identity instance each (:858-866), the same rays.
* BuildSingleTriangle (basic_test.h:275-372).
* UpdateObj (basic_test.h:1071-1180): y -= 40 on every vertex, refit, the
  640 x 640 fan of :1196-1208; UpdateTest (test/test_vk/hlbvh_test.h:430-448):
  y += 10, refit.
* Rays whose direction has exactly-zero components and whose origins lie on
  the planes of the boxes' faces (the slab test's 0 * inf = NaN path).
* The oracle's CPU edge cases run on the GPU: a shared edge is watertight,
  vertical rays through a heightfield's grid vertices and edges never leak,
  equal distances resolve to the lowest primitive id, and the AABB-leaf
  known answers.
Every GPU hit is compared with the oracle's traversal of its own build;
closest hits also with the oracle's FP32 brute force over all triangles.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, scenes
from tests.gpu_util import RR_INVALID, hits_equal, nodes_equal, say, tlas_tables_equal, to_dev

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def cornell(scale: float = 1.0):
    d = np.load(GOLDEN / "cornell_box.npz")
    v = (d["vertices"] * np.float32(scale)).astype(np.float32)
    return v, d["triangles"].astype(np.int32), d["object"], list(d["names"])


def fan_rays(res: int, origin=(0.0, 15.0, 0.0), tmin=0.001, tmax=100000.0) -> np.ndarray:
    """The reference's ray fan: ray i = res * y + x has direction
    (-1, -1 + (2 / res) * y, -1 + (2 / res) * x), computed in float32."""
    y, x = np.meshgrid(np.arange(res, dtype=np.float32), np.arange(res, dtype=np.float32), indexing="ij")
    step = np.float32(2.0) / np.float32(res)
    r = np.zeros((res * res, 8), np.float32)
    r[:, 0:3] = np.asarray(origin, np.float32)
    r[:, 3] = np.float32(tmin)
    r[:, 4] = -1.0
    r[:, 5] = (np.float32(-1.0) + step * y).reshape(-1)
    r[:, 6] = (np.float32(-1.0) + step * x).reshape(-1)
    r[:, 7] = np.float32(tmax)
    return r


def axis_rays(v: np.ndarray, rng, per_axis: int = 4000) -> np.ndarray:
    """Rays along +-x, +-y, +-z (the other two direction components exactly
    zero) whose origins take the mesh's own vertex coordinates (so they lie
    on the planes of box faces and triangle edges) mixed with random ones."""
    lo, hi = v.min(axis=0), v.max(axis=0)
    out = []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for sgn in (-1.0, 1.0):
            o = np.empty((per_axis, 3), np.float32)
            cb = np.unique(v[:, b])
            cc = np.unique(v[:, c])
            half = per_axis // 2
            o[:half, b] = rng.choice(cb, half)
            o[:half, c] = rng.choice(cc, half)
            o[half:, b] = rng.uniform(lo[b], hi[b], per_axis - half)
            o[half:, c] = rng.uniform(lo[c], hi[c], per_axis - half)
            # Start outside the mesh, or exactly on one of its planes.
            start = np.where(rng.uniform(size=per_axis) < 0.5, lo[a] - 1 if sgn > 0 else hi[a] + 1,
                             rng.choice(np.unique(v[:, a]), per_axis))
            o[:, a] = start
            d = np.zeros((per_axis, 3), np.float32)
            d[:, a] = sgn
            out.append(scenes.make_rays(o, d))
    return np.concatenate(out)


def _check_blas(ctx, dev, v, i, rays, restructure, tag):
    """Build on the GPU, compare nodes and hits (closest + any, full_t and
    instance-id outputs) with the oracle; closest hits also with brute force."""
    import torch

    accel = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=restructure)
    torch.cuda.synchronize()
    buf = accel.buffer.cpu().numpy()
    ref = oracle.build_blas(v, i, restructure)
    ok, msg = nodes_equal(buf, ref)
    assert ok, f"{tag}: {msg}"
    assert oracle.is_valid(buf) and oracle.check_consistency(buf)
    drays = to_dev(rays, dev)
    closest = None
    for query in ("closest", "any"):
        hits = ctx.intersect(accel, drays, query, "full_t")
        torch.cuda.synchronize()
        g = api.decode_hits(hits, "full_t")
        r = oracle.trace(ref, rays, any_hit=(query == "any"))
        ok, msg = hits_equal(g, r)
        assert ok, f"{tag} {query}: {msg}"
        if query == "closest":
            closest = g
            bf = oracle.brute_force(ref, rays)
            ok, msg = hits_equal(g, bf)
            assert ok, f"{tag} brute force: {msg}"
    ids = ctx.intersect(accel, drays, "closest", "instance")
    torch.cuda.synchronize()
    # Single level: the instance-id output is the primitive id (isect.comp:235).
    assert np.array_equal(api.decode_hits(ids, "instance")["id"], oracle.trace(ref, rays)["prim_id"])
    return accel, ref, closest


@pytest.mark.parametrize("scale", [1.0, 20.0])
@pytest.mark.parametrize("restructure", [False, True])
def test_cornell_box_build_obj_fan(ctx, dev, scale, restructure):
    v, i, _, _ = cornell(scale)
    rays = fan_rays(2048)
    say(f"cornell box x{scale:g}, {i.shape[0]} tris, restructure={restructure}, 2048^2 fan rays")
    _, _, g = _check_blas(ctx, dev, v, i, rays, restructure, f"cornell x{scale:g}")
    hit = (g["prim_id"] != RR_INVALID).mean()
    if scale == 1.0:
        assert hit == 0.0  # the fan passes 13 units above the unscaled box
    else:
        assert hit > 0.5, hit


@pytest.mark.parametrize("scale", [1.0, 20.0])
def test_cornell_box_axis_parallel_rays(ctx, dev, scale):
    v, i, _, _ = cornell(scale)
    rng = scenes.rng_for(70, int(scale))
    rays = axis_rays(v, rng)
    say(f"cornell box x{scale:g}: {rays.shape[0]} axis-parallel rays from face planes")
    assert (rays[:, 4:7] == 0).sum(axis=1).min() == 2
    for restructure in (False, True):
        _, _, g = _check_blas(ctx, dev, v, i, rays, restructure, f"axis x{scale:g}")
        assert (g["prim_id"] != RR_INVALID).mean() > 0.2


def test_cornell_box_two_level(ctx, dev):
    """BuildObj2Level: a BLAS per object (fast build), identity instances."""
    import torch

    v, i, obj, names = cornell(20.0)
    say(f"cornell box two-level: {len(names)} objects as identity instances")
    blases, host_bufs = [], []
    for k in range(len(names)):
        tri = i[obj == k]
        used = np.unique(tri)
        remap = np.full(v.shape[0], -1, np.int64)
        remap[used] = np.arange(used.size)
        vk, ik = v[used], remap[tri].astype(np.int32)
        blases.append(ctx.build_geometry(to_dev(vk, dev), to_dev(ik, dev), restructure=False))
        host_bufs.append((vk, ik))
    xf = np.tile(np.eye(3, 4, dtype=np.float32), (len(names), 1, 1))
    tlas = ctx.build_scene(blases, xf)
    torch.cuda.synchronize()
    host = {b.address: b.buffer.cpu().numpy() for b in blases}
    for b, (vk, ik) in zip(blases, host_bufs):
        ok, msg = nodes_equal(host[b.address], oracle.build_blas(vk, ik, False))
        assert ok, msg
    buf = tlas.buffer.cpu().numpy()
    ref = oracle.build_tlas(xf, [b.address for b in blases], [host[b.address] for b in blases])
    ok, msg = nodes_equal(buf, ref)
    assert ok, msg
    ok, msg = tlas_tables_equal(buf, ref)
    assert ok, msg
    rays = fan_rays(2048)
    drays = to_dev(rays, dev)
    for query in ("closest", "any"):
        hits = ctx.intersect(tlas, drays, query, "full_t")
        torch.cuda.synchronize()
        g = api.decode_hits(hits, "full_t")
        ok, msg = hits_equal(g, oracle.trace(ref, rays, any_hit=(query == "any"), blas_map=host))
        assert ok, f"{query}: {msg}"
    # Closest hits as the one-level build of the whole box reports them:
    # the same surfaces; instance = object, primitive = index within it.
    one = oracle.trace(oracle.build_blas(v, i), rays)
    hit = one["prim_id"] != RR_INVALID
    assert np.array_equal(g["inst_id"] != RR_INVALID, hit)
    assert np.array_equal(obj[one["prim_id"][hit]], g["inst_id"][hit])
    assert np.array_equal(g["t"][hit].view(np.uint32), one["t"][hit].view(np.uint32))


def test_single_triangle(ctx, dev):
    """BuildSingleTriangle (basic_test.h:284-292): the one-leaf BVH, traced."""
    v = np.array([[0, -0.7, 0.7], [-0.7, 0.7, 1.0], [0.7, 0.7, 1.0]], np.float32)
    i = np.array([[0, 1, 2]], np.int32)
    say("single triangle")
    rng = scenes.rng_for(71)
    o = rng.uniform(-1, 1, size=(4096, 3)).astype(np.float32)
    o[:, 2] = -1.0
    d = np.tile(np.array([[0, 0, 1]], np.float32), (4096, 1))
    rays = np.concatenate([scenes.make_rays(o, d), fan_rays(64, origin=(0, 0, -2))])
    for restructure in (False, True):
        accel, ref, g = _check_blas(ctx, dev, v, i, rays, restructure, "single triangle")
        assert int(accel.buffer[12:16].cpu().numpy().view(np.uint32)[0]) == 1  # the root is the leaf
        assert (g["prim_id"] == 0).sum() > 100
        # Against FP64: t within 1e-6 relative.
        for k in np.nonzero(g["prim_id"] == 0)[0][:200]:
            ok, t, _, _ = oracle.triangle_fp64(rays[k], v[0], v[1], v[2], eps=1e-6)
            assert ok and abs(g["t"][k] - t) <= 1e-6 * abs(t)


@pytest.mark.parametrize("case", ["update_obj", "update_test"])
def test_cornell_box_refit_as_the_reference_tests(ctx, dev, case):
    """UpdateObj (y -= 40, basic_test.h:1168-1174) and UpdateTest (y += 10,
    hlbvh_test.h:445-448): refit the built box to the moved vertices."""
    import torch

    v, i, _, _ = cornell(20.0)
    dy = np.float32(-40.0 if case == "update_obj" else 10.0)
    say(f"cornell box refit, y += {dy:g}")
    accel = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True)
    torch.cuda.synchronize()
    built = accel.buffer.cpu().numpy()
    moved = v.copy()
    moved[:, 1] += dy
    ctx.update_geometry(accel, to_dev(moved, dev), to_dev(i, dev))
    torch.cuda.synchronize()
    buf = accel.buffer.cpu().numpy()
    ref = oracle.refit(built, moved, i)
    ok, msg = nodes_equal(buf, ref)
    assert ok, msg
    assert oracle.check_consistency(buf)
    rays = fan_rays(640)  # basic_test.h:1191-1208
    hits = ctx.intersect(accel, to_dev(rays, dev), "closest", "full_t")
    torch.cuda.synchronize()
    g = api.decode_hits(hits, "full_t")
    ok, msg = hits_equal(g, oracle.trace(ref, rays))
    assert ok, msg
    ok, msg = hits_equal(g, oracle.brute_force(oracle.build_blas(moved, i), rays))
    assert ok, "brute force on the moved box: " + msg
    assert (g["prim_id"] != RR_INVALID).sum() > 1000


def test_shared_edge_is_watertight_on_gpu(ctx, dev):
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    i = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    say("shared edge watertightness")
    s = np.linspace(0.05, 0.95, 19, dtype=np.float32)
    o = np.stack([s, np.full_like(s, 3.0), s], axis=1)
    rays = scenes.make_rays(o, np.tile(np.array([[0, -1, 0]], np.float32), (19, 1)))
    _, _, g = _check_blas(ctx, dev, v, i, rays, False, "shared edge")
    assert np.isin(g["prim_id"], [0, 1]).all()


def test_heightfield_vertical_rays_never_leak_on_gpu(ctx, dev):
    v, i = scenes.heightfield_mesh(65)
    say("heightfield: vertical rays through grid vertices and edges")
    rng = np.random.default_rng(13)
    n = 20000
    xz = rng.uniform(0, 1, size=(n, 2)).astype(np.float32)
    grid = np.linspace(0, 1, 65, dtype=np.float32)
    xz[:64] = np.stack([grid[1:], grid[:-1]], axis=1)          # through vertices
    xz[64:128] = np.stack([grid[1:], (grid[:-1] + grid[1:]) / 2], axis=1)  # along edges
    down = scenes.make_rays(np.stack([xz[:, 0], np.ones(n, np.float32), xz[:, 1]], axis=1),
                            np.tile(np.array([[0, -1, 0]], np.float32), (n, 1)))
    up = scenes.make_rays(np.stack([xz[:, 0], -np.ones(n, np.float32), xz[:, 1]], axis=1),
                          np.tile(np.array([[0, 1, 0]], np.float32), (n, 1)))
    rays = np.concatenate([down, up])
    for restructure in (False, True):
        _, _, g = _check_blas(ctx, dev, v, i, rays, restructure, "heightfield")
        assert (g["prim_id"] != RR_INVALID).all()


def test_equal_distances_resolve_to_lowest_id_on_gpu(ctx, dev):
    rng = np.random.default_rng(5)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    filler, _ = scenes.triangle_soup(200, rng)
    v = np.concatenate([filler + np.float32([0, 5, 0]), np.tile(base, (6, 1))]).astype(np.float32)
    i = np.arange(v.shape[0], dtype=np.int32).reshape(-1, 3)
    say("coincident triangles: lowest primitive id wins")
    o = np.array([[0.2, 2.0, 0.3], [0.1, 2.0, 0.1]], np.float32)
    rays = scenes.make_rays(o, np.array([[0, -1, 0], [0, -1, 0]], np.float32))
    for restructure in (False, True):
        _, _, g = _check_blas(ctx, dev, v, i, rays, restructure, "ties")
        assert (g["prim_id"] == 200).all()


def test_aabb_known_answers_on_gpu(ctx, dev):
    import torch

    say("AABB leaves: known answers")
    b = np.array([[0, 0, 0, 1, 1, 1], [0, 0, 3, 1, 1, 4]], np.float32)
    accel = ctx.build_aabbs(to_dev(b, dev))
    torch.cuda.synchronize()
    buf = accel.buffer.cpu().numpy()
    ok, msg = nodes_equal(buf, oracle.build_blas_aabbs(b))
    assert ok, msg
    cases = [  # origin, direction, tmin, tmax, (prim, t) or None
        ((0.5, 0.5, -2), (0, 0, 1), 0.0, 1e30, (0, 2.0)),
        ((0.5, 0.5, 2), (0, 0, 1), 0.0, 1e30, (1, 1.0)),     # between the boxes
        ((0.5, 0.5, 0.5), (0, 0, 1), 0.25, 1e30, (0, 0.25)),  # inside: t = tmin
        ((0.5, 0.5, -2), (0, 0, -1), 0.0, 1e30, None),       # pointing away
        ((0.5, 0.5, -2), (0, 0, 1), 0.0, 1.5, None),         # too short
        ((1.0, 0.5, -2), (0, 0, 1), 0.0, 1e30, (0, 2.0)),    # on a face plane, parallel to it
        ((0.0, 0.0, -2), (0, 0, 1), 0.0, 1e30, (0, 2.0)),    # along an edge
    ]
    o = np.array([c[0] for c in cases], np.float32)
    d = np.array([c[1] for c in cases], np.float32)
    rays = scenes.make_rays(o, d)
    rays[:, 3] = [c[2] for c in cases]
    rays[:, 7] = [c[3] for c in cases]
    hits = ctx.intersect(accel, to_dev(rays, dev), "closest", "full_t")
    torch.cuda.synchronize()
    g = api.decode_hits(hits, "full_t")
    ok, msg = hits_equal(g, oracle.trace(buf, rays))
    assert ok, msg
    for k, c in enumerate(cases):
        if c[4] is None:
            assert g["prim_id"][k] == RR_INVALID, k
        else:
            assert (g["prim_id"][k], g["t"][k], g["u"][k], g["v"][k]) == (c[4][0], np.float32(c[4][1]), 0, 0), k
