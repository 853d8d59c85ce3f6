"""CPU tests that pin the oracle.

The reference holds no golden vectors for this path (its tests only check
scan/sort against the STL, BVH consistency and SAH monotonicity, and all of
them load the missing Sponza mesh).  Each oracle function is therefore pinned
by what the mathematics fixes: closed forms, brute force on small inputs,
invariants, FP64, and the reference's own compiled bvh_analyzer where it is
present.
"""
import itertools
import math
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import scenes

INV = 0xFFFFFFFF


# --- a3: Morton codes --------------------------------------------------------
def test_morton_closed_forms():
    # vlk/kernels/common.h:251-268: 10 bits per axis, x in bit 2 of each triple.
    assert oracle.morton3(0.0, 0.0, 0.0) == 0
    assert oracle.morton3(1.0, 1.0, 1.0) == 0x3FFFFFFF          # clamped to 1023 per axis
    assert oracle.morton3(0.5, 0.0, 0.0) == 0x20000000          # x = 512 -> bit 9 -> bit 29
    assert oracle.morton3(0.0, 0.5, 0.0) == 0x10000000
    assert oracle.morton3(0.0, 0.0, 0.5) == 0x08000000
    assert oracle.morton3(-3.0, 7.0, 0.0) == oracle.morton3(0.0, 1.0, 0.0)  # clamp
    # 1/1024 -> coordinate 1 -> lowest bit of its axis.
    assert oracle.morton3(0.0, 0.0, 1.0 / 1024) == 1
    assert oracle.morton3(1.0 / 1024, 0.0, 0.0) == 4


def _interleave(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b)
    return code


def test_morton_matches_bitwise_interleave():
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, size=(500, 3)):
        p = ((x + 0.5) / 1024, (y + 0.5) / 1024, (z + 0.5) / 1024)
        assert oracle.morton3(*p) == _interleave(int(x), int(y), int(z))


def test_morton_codes_of_mesh_use_centroid_of_box():
    # Three triangles whose box centroids sit at known normalised positions.
    v = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0],          # box centre (0,0,0) -> 0
                  [1, 1, 1], [1, 1, 1], [1, 1, 1],          # (1,1,1) -> all ones
                  [0, 0, 0], [1, 0, 0], [0, 0, 0]], np.float32)  # centre x = 0.5
    i = np.arange(9, dtype=np.uint32).reshape(3, 3)
    codes = oracle.morton_codes(v, i)
    assert list(codes) == [0, 0x3FFFFFFF, 0x20000000]
    lo, hi = oracle.mesh_bounds(v, i)
    assert list(lo) == [0, 0, 0] and list(hi) == [1, 1, 1]


def test_morton_flat_axis_maps_to_zero():
    # A zero-extent axis contributes 0 bits instead of 0/0.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.5, 0], [1, 1, 0]], np.float32)
    i = np.arange(6, dtype=np.uint32).reshape(2, 3)
    codes = oracle.morton_codes(v, i)
    assert all(c & 0x09249249 == 0 for c in codes)  # no z bits


# --- a4-a6: sort -------------------------------------------------------------
def test_sort_pairs_stable_like_numpy():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 50, size=10000).astype(np.uint32)
    vals = np.arange(10000, dtype=np.uint32)
    k, v = oracle.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


# --- a7: Karras hierarchy ----------------------------------------------------
def _karras_brute(codes):
    """O(n^2) binary radix tree over augmented keys (code, index): a range's
    split is the last position sharing more than the range's common prefix
    with its first key; internal node of a range is numbered by its split's
    neighbour convention (left child = split, right child = split + 1)."""
    n = len(codes)

    def delta(i, j):
        if codes[i] != codes[j]:
            return 32 - int(codes[i] ^ codes[j]).bit_length()
        return 32 + 32 - int(i ^ j).bit_length()

    child = {}

    def rec(a, b, name):
        common = delta(a, b)
        split = max(k for k in range(a, b) if delta(a, k) > common)
        left = ("leaf", split) if split == a else ("int", split)
        right = ("leaf", split + 1) if split + 1 == b else ("int", split + 1)
        child[name] = (left, right)
        if left[0] == "int":
            rec(a, split, left)
        if right[0] == "int":
            rec(split + 1, b, right)

    if n > 1:
        rec(0, n - 1, ("int", 0))
    return child


@pytest.mark.parametrize("case", ["random", "dups", "all_equal", "pairs", "sorted_bits"])
@pytest.mark.parametrize("n", [2, 3, 5, 17, 64])
def test_karras_matches_brute_force(case, n):
    rng = np.random.default_rng(n * 31 + len(case))
    if case == "random":
        codes = np.sort(rng.integers(0, 1 << 30, size=n)).astype(np.uint32)
    elif case == "dups":
        codes = np.sort(rng.integers(0, 4, size=n)).astype(np.uint32)
    elif case == "all_equal":
        codes = np.full(n, 12345, np.uint32)
    elif case == "pairs":
        codes = np.repeat(np.sort(rng.integers(0, 1 << 30, size=(n + 1) // 2)), 2)[:n].astype(np.uint32)
    else:
        codes = np.arange(n, dtype=np.uint32) << 20
    prims = rng.permutation(n).astype(np.uint32)
    nodes = oracle.emit_hierarchy(codes, prims)
    brute = _karras_brute(codes)
    for k in range(n - 1):
        (lk, li), (rk, ri) = brute[("int", k)]
        want_l = li + n - 1 if lk == "leaf" else li
        want_r = ri + n - 1 if rk == "leaf" else ri
        assert nodes[k, 3] == want_l and nodes[k, 7] == want_r, (k, nodes[k], brute[("int", k)])
        assert nodes[want_l, 11] == k and nodes[want_r, 11] == k
    assert nodes[0, 11] == INV
    for i in range(n):
        assert nodes[n - 1 + i, 3] == INV and nodes[n - 1 + i, 7] == prims[i]


# --- a8/a10: refit -----------------------------------------------------------
def _nodes_f(buf):
    return oracle.nodes_of(buf).view(np.float32)


def test_refit_boxes_are_exact_unions():
    v, i = scenes.triangle_soup(3000, scenes.rng_for(9))
    buf = oracle.build_blas(v, i)
    assert oracle.check_consistency(buf) and oracle.is_valid(buf)
    nodes = oracle.nodes_of(buf)
    nf = nodes.view(np.float32)
    n = 3000
    # Every leaf stores its own triangle.
    for k in range(n - 1, 2 * n - 1):
        prim = nodes[k, 7]
        tri = v[i[prim]]
        assert np.array_equal(nf[k, [0, 1, 2]], tri[0]) and np.array_equal(nf[k, [4, 5, 6]], tri[1])
        assert np.array_equal(nf[k, [8, 9, 10]], tri[2])
    # The root's two child boxes together are the mesh bounds.
    lo, hi = oracle.mesh_bounds(v, i)
    assert np.array_equal(np.minimum(nf[0, [0, 1, 2]], nf[0, [8, 9, 10]]), lo)
    assert np.array_equal(np.maximum(nf[0, [4, 5, 6]], nf[0, [12, 13, 14]]), hi)


def test_refit_after_motion_equals_rebuild_boxes():
    # The reference's UpdateTest shifts every y by +10 (hlbvh_test.h:447).
    v, i = scenes.triangle_soup(2000, scenes.rng_for(10))
    buf = oracle.build_blas(v, i)
    moved = v.copy()
    moved[:, 1] += 10.0
    up = oracle.refit(buf, moved, i)
    assert oracle.check_consistency(up)
    nf = _nodes_f(up)
    assert np.all(np.minimum(nf[0, 1], nf[0, 9]) >= 10.0 - 0.05)


# --- a9: treelet restructuring -------------------------------------------------
def _all_trees(leaves):
    if len(leaves) == 1:
        yield leaves[0]
        return
    first, rest = leaves[0], leaves[1:]
    # Split into two non-empty groups; first leaf always in the left group.
    for r in range(0, len(rest)):
        for combo in itertools.combinations(rest, r):
            left = (first,) + combo
            right = tuple(x for x in rest if x not in combo)
            if not right:
                continue
            for lt in _all_trees(left):
                for rt in _all_trees(right):
                    yield (lt, rt)


def _tree_cost(tree, boxes):
    def leaves(t):
        return [t] if isinstance(t, int) else leaves(t[0]) + leaves(t[1])

    def area(ids):
        lo = boxes[ids, :3].min(axis=0)
        hi = boxes[ids, 3:].max(axis=0)
        e = hi - lo
        return 2.0 * (e[0] * e[2] + e[1] * e[0] + e[2] * e[1])

    def cost(t):
        if isinstance(t, int):
            return 1.2 * area([t])
        return 1.2 * area(leaves(t)) + cost(t[0]) + cost(t[1])

    return cost(tree)


def test_treelet_dp_is_optimal_over_all_trees():
    rng = np.random.default_rng(11)
    trees = list(_all_trees(tuple(range(7))))
    assert len(trees) == 10395  # (2*7-3)!!
    for trial in range(3):
        c = rng.uniform(0, 1, size=(7, 3))
        e = rng.uniform(0.01, 0.3, size=(7, 3))
        boxes = np.concatenate([c - e, c + e], axis=1).astype(np.float32)
        cost, part = oracle.treelet_dp(boxes)
        best = min(_tree_cost(t, boxes.astype(np.float64)) for t in trees)
        assert math.isclose(float(cost[127]), best, rel_tol=1e-5), (cost[127], best)
        # Every stored split is a proper, non-empty partition.
        for s in range(1, 128):
            if bin(s).count("1") > 1:
                p = int(part[s])
                assert p and (p & s) == p and p != s


def test_restructure_keeps_validity_and_lowers_sah():
    for n, seed in ((64, 1), (500, 2), (20000, 3)):
        v, i = scenes.triangle_soup(n, scenes.rng_for(12, seed))
        base = oracle.build_blas(v, i, restructure=False)
        opt = oracle.build_blas(v, i, restructure=True)
        assert oracle.is_valid(opt) and oracle.check_consistency(opt)
        # Reference RestructureTest: SAH after <= SAH before (hlbvh_test.h:417).
        assert oracle.sah(opt) <= oracle.sah(base) * (1 + 1e-6)
        leaves = oracle.nodes_of(opt)[n - 1:]
        assert sorted(leaves[:, 7].tolist()) == list(range(n))  # each primitive in exactly one leaf


@pytest.mark.parametrize("leaves", [5, 9])
def test_restructure_other_treelet_sizes(leaves):
    """Treelets of 5 or 9 leaves (a tuning option; the reference uses 7):
    valid, consistent, every primitive in one leaf, SAH not above the
    unrestructured build, and traversal still equals brute force."""
    for n, seed in ((64, 1), (3000, 4)):
        v, i = scenes.triangle_soup(n, scenes.rng_for(12, seed))
        base = oracle.build_blas(v, i, restructure=False)
        opt = oracle.build_blas(v, i, restructure=True, treelet_leaves=leaves)
        assert oracle.is_valid(opt) and oracle.check_consistency(opt)
        assert oracle.sah(opt) <= oracle.sah(base) * (1 + 1e-6)
        leaf_nodes = oracle.nodes_of(opt)[n - 1:]
        assert sorted(leaf_nodes[:, 7].tolist()) == list(range(n))
    w = scenes.soup_camera(n_tris=5000, width=48, height=48)
    buf = oracle.build_blas(w.vertices, w.indices, True, treelet_leaves=leaves)
    h, bf = oracle.trace(buf, w.rays), oracle.brute_force(buf, w.rays)
    assert np.array_equal(h["prim_id"], bf["prim_id"])
    assert np.array_equal(h["t"].view(np.uint32), bf["t"].view(np.uint32))


# --- a12/a13/a15: traversal ----------------------------------------------------
def test_watertight_known_answer():
    # Ray straight down onto the triangle (0,0,0),(1,0,0),(0,0,1) at x=0.25, z=0.5.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    i = np.array([[0, 1, 2]], np.uint32)
    buf = oracle.build_blas(v, i)
    rays = scenes.make_rays(np.array([[0.25, 2.0, 0.5]], np.float32), np.array([[0, -1, 0]], np.float32))
    h = oracle.trace(buf, rays)[0]
    assert h["prim_id"] == 0 and h["inst_id"] == 0
    assert h["t"] == 2.0 and h["u"] == 0.25 and h["v"] == 0.5   # weights of v1, v2
    # Outside the triangle, and behind the origin: misses.
    rays = scenes.make_rays(np.array([[0.75, 2.0, 0.5], [0.1, -1.0, 0.1]], np.float32),
                            np.array([[0, -1, 0], [0, -1, 0]], np.float32))
    assert (oracle.trace(buf, rays)["prim_id"] == INV).all()


def test_shared_edge_is_watertight():
    # A quad split along its diagonal: rays exactly through the diagonal hit
    # (no gap between the two triangles), at t = 3 up to rounding.
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    i = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    buf = oracle.build_blas(v, i)
    s = np.linspace(0.05, 0.95, 19, dtype=np.float32)
    o = np.stack([s, np.full_like(s, 3.0), s], axis=1)
    rays = scenes.make_rays(o, np.tile(np.array([[0, -1, 0]], np.float32), (19, 1)))
    h = oracle.trace(buf, rays)
    assert np.isin(h["prim_id"], [0, 1]).all()
    assert np.all(np.abs(h["t"] - 3.0) <= 4 * np.spacing(np.float32(3.0)))


def test_equal_distances_resolve_to_lowest_id():
    # Coincident copies of a triangle give bit-identical t; the closest hit
    # must be the lowest primitive id whatever the BVH's shape.
    rng = np.random.default_rng(5)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    filler, _ = scenes.triangle_soup(200, rng)
    v = np.concatenate([filler + np.float32([0, 5, 0]), np.tile(base, (6, 1))]).astype(np.float32)
    i = np.arange(v.shape[0], dtype=np.uint32).reshape(-1, 3)
    for restructure in (False, True):
        buf = oracle.build_blas(v, i, restructure)
        o = np.array([[0.2, 2.0, 0.3], [0.1, 2.0, 0.1]], np.float32)
        rays = scenes.make_rays(o, np.array([[0, -1, 0], [0, -1, 0]], np.float32))
        h = oracle.trace(buf, rays)
        assert (h["prim_id"] == 200).all()


def test_heightfield_downward_rays_never_leak():
    # Watertightness: every vertical ray inside the domain hits exactly one surface.
    v, i = scenes.heightfield_mesh(65)
    buf = oracle.build_blas(v, i, restructure=True)
    rng = np.random.default_rng(13)
    xz = rng.uniform(0, 1, size=(20000, 2)).astype(np.float32)
    # Include rays exactly through grid vertices and edges.
    grid = np.linspace(0, 1, 65, dtype=np.float32)
    xz[:64] = np.stack([grid[1:], grid[:-1]], axis=1)
    o = np.stack([xz[:, 0], np.full(20000, 1.0, np.float32), xz[:, 1]], axis=1)
    rays = scenes.make_rays(o, np.tile(np.array([[0, -1, 0]], np.float32), (20000, 1)))
    h = oracle.trace(buf, rays)
    assert (h["prim_id"] != INV).all()
    # And upward rays from below see it too.
    o2 = o.copy()
    o2[:, 1] = -1.0
    rays2 = scenes.make_rays(o2, np.tile(np.array([[0, 1, 0]], np.float32), (20000, 1)))
    assert (oracle.trace(buf, rays2)["prim_id"] != INV).all()


@pytest.mark.parametrize("restructure", [False, True])
def test_traversal_equals_brute_force(restructure):
    w = scenes.soup_camera(n_tris=20000, width=96, height=96)
    buf = oracle.build_blas(w.vertices, w.indices, restructure)
    h = oracle.trace(buf, w.rays)
    bf = oracle.brute_force(buf, w.rays)
    for f in ("prim_id", "inst_id"):
        assert np.array_equal(h[f], bf[f])
    for f in ("t", "u", "v"):
        assert np.array_equal(h[f].view(np.uint32), bf[f].view(np.uint32))
    assert (h["prim_id"] != INV).mean() > 0.2


def test_fp32_hits_agree_with_fp64():
    w = scenes.soup_camera(n_tris=65536, width=128, height=128)
    buf = oracle.build_blas(w.vertices, w.indices)
    h = oracle.trace(buf, w.rays)
    hit = np.nonzero(h["prim_id"] != INV)[0]
    assert hit.size > 1000
    worst_t = worst_k = 0.0
    for k in hit:
        tri = w.vertices[w.indices[h["prim_id"][k]]]
        ok, t, u, v = oracle.triangle_fp64(w.rays[k], tri[0], tri[1], tri[2], eps=1e-5)
        assert ok
        worst_t = max(worst_t, abs(h["t"][k] - t) / abs(t))
        # The watertight test forms edge functions from vertex offsets to the
        # ray origin (absolute error ~2^-24 * D, D = largest vertex distance)
        # over edge functions of size ~l^2 (l^2 = area projected along the
        # ray), so barycentrics carry error ~2^-24 * D / l.
        t64 = tri.astype(np.float64)
        o, d = w.rays[k, :3].astype(np.float64), w.rays[k, 4:7].astype(np.float64)
        dist = np.max(np.linalg.norm(t64 - o, axis=1))
        proj = abs(np.cross(t64[1] - t64[0], t64[2] - t64[0]) @ d) / np.abs(d).max()
        bound = 2.0 ** -24 * dist / math.sqrt(proj)
        worst_k = max(worst_k, max(abs(h["u"][k] - u), abs(h["v"][k] - v)) / bound)
    # Tolerances: t within 1e-6 relative; barycentrics within 8 * 2^-24 * D / l
    # (measured worst: 7.3e-7 and 2.8x on this workload).
    assert worst_t < 1e-6, worst_t
    assert worst_k < 8.0, worst_k


def test_any_hit_is_a_valid_hit_and_agrees_on_hit_or_miss():
    w = scenes.soup_camera(n_tris=20000, width=96, height=96)
    buf = oracle.build_blas(w.vertices, w.indices, True)
    closest = oracle.trace(buf, w.rays)
    anyh = oracle.trace(buf, w.rays, any_hit=True)
    assert np.array_equal(closest["prim_id"] != INV, anyh["prim_id"] != INV)
    k = np.nonzero(anyh["prim_id"] != INV)[0]
    assert np.all(anyh["t"][k] >= closest["t"][k])


def test_tlas_traversal_equals_brute_force_and_inverse():
    rng = scenes.rng_for(2, 77)
    meshes = [scenes.triangle_soup(int(c), rng, sigma=0.08) for c in (300, 1000, 50)]
    blas = [oracle.build_blas(v, i, True) for v, i in meshes]
    addrs = [0x1000 * (k + 1) for k in range(3)]
    pick = rng.integers(0, 3, size=40)
    xf = scenes.random_transforms(rng, 40, trans=4.0)
    tlas = oracle.build_tlas(xf, [addrs[k] for k in pick], [blas[k] for k in pick])
    assert oracle.is_valid(tlas) and oracle.check_consistency(tlas)
    bmap = dict(zip(addrs, blas))
    o = rng.uniform(-6, 6, size=(8000, 3)).astype(np.float32)
    rays = scenes.make_rays(o, scenes.unit_sphere(8000, rng))
    h = oracle.trace(tlas, rays, blas_map=bmap)
    bf = oracle.brute_force(tlas, rays, blas_map=bmap)
    assert np.array_equal(h["prim_id"], bf["prim_id"]) and np.array_equal(h["inst_id"], bf["inst_id"])
    assert (h["inst_id"] != INV).mean() > 0.02
    # Stored inverse transforms: M * M^-1 = I.
    off = int(tlas[24:32].view(np.uint64)[0])
    x = tlas[off:off + 96 * 40].view(np.float32).reshape(40, 2, 3, 4)
    for k in range(40):
        inv = np.vstack([x[k, 0], [0, 0, 0, 1]]).astype(np.float64)
        fwd = np.vstack([x[k, 1], [0, 0, 0, 1]]).astype(np.float64)
        assert np.allclose(fwd @ inv, np.eye(4), atol=1e-5)


# --- the reference's own analyzer (built from its sources into oracle/_ref) ----
@pytest.mark.skipif(not oracle.REF_ANALYZER.exists(), reason="oracle/_ref/bvh_analyzer not built")
def test_reference_bvh_analyzer_agrees_on_validity_and_sah():
    w = scenes.soup_camera(n_tris=4096, width=32, height=32)
    for restructure in (False, True):
        buf = oracle.build_blas(w.vertices, w.indices, restructure)
        with tempfile.TemporaryDirectory() as d:
            d = Path(d)
            oracle.write_analyzer_dump(buf, w.rays, 32, 32, d / "bvh.bin", d / "rays.bin", d / "cfg.txt")
            out = subprocess.run([str(oracle.REF_ANALYZER), str(d / "cfg.txt")], cwd=d, capture_output=True,
                                 text=True, timeout=60, env={**os.environ, "OMP_NUM_THREADS": "2"})
        assert out.returncode == 0, out.stderr
        stats = dict(line.split(": ", 1) for line in out.stdout.splitlines() if ": " in line)
        assert stats["is_valid"].strip() == "1"
        assert math.isclose(float(stats["sah"]), oracle.sah(buf), rel_tol=1e-5)


# --- AABB-list primitives ------------------------------------------------------
def _ray(o, d, tmin=0.0, tmax=1e30):
    return scenes.make_rays(np.asarray([o], np.float32), np.asarray([d], np.float32), tmin, tmax)


def test_aabb_leaf_known_answers():
    b = np.array([[0, 0, 0, 1, 1, 1], [0, 0, 3, 1, 1, 4]], np.float32)
    buf = oracle.build_blas_aabbs(b)
    assert oracle.is_valid(buf)
    h = oracle.trace(buf, _ray((0.5, 0.5, -2), (0, 0, 1)))[0]
    assert (h["prim_id"], h["t"], h["u"], h["v"]) == (0, 2.0, 0.0, 0.0)
    h = oracle.trace(buf, _ray((0.5, 0.5, 2), (0, 0, 1)))[0]  # between the boxes: the far one at t = 1
    assert (h["prim_id"], h["t"]) == (1, 1.0)
    h = oracle.trace(buf, _ray((0.5, 0.5, 0.5), (0, 0, 1), tmin=0.25))[0]  # inside: t = tmin
    assert (h["prim_id"], h["t"]) == (0, 0.25)
    h = oracle.trace(buf, _ray((0.5, 0.5, -2), (0, 0, -1)))[0]  # pointing away
    assert h["prim_id"] == 0xFFFFFFFF
    h = oracle.trace(buf, _ray((0.5, 0.5, -2), (0, 0, 1), tmax=1.5))[0]  # too short
    assert h["prim_id"] == 0xFFFFFFFF
    h = oracle.trace(buf, _ray((1.0, 0.5, -2), (0, 0, 1)))[0]  # on a face plane, parallel to it
    assert (h["prim_id"], h["t"]) == (0, 2.0)


def _boxes(n, rng):
    c = rng.uniform(0, 1, size=(n, 3))
    e = rng.uniform(0.001, 0.03, size=(n, 3))
    return np.concatenate([c - e, c + e], axis=1).astype(np.float32)


@pytest.mark.parametrize("restructure", [False, True])
def test_aabb_traversal_equals_fp64_brute_force(restructure):
    rng = scenes.rng_for(60)
    b = _boxes(3000, rng)
    buf = oracle.build_blas_aabbs(b, restructure)
    assert oracle.is_valid(buf) and oracle.check_consistency(buf)
    o = rng.uniform(-0.5, 1.5, size=(4000, 3)).astype(np.float32)
    rays = scenes.make_rays(o, scenes.unit_sphere(4000, rng))
    got = oracle.trace(buf, rays)
    # Float64 slab test over every box, independent of the BVH.
    O = rays[:, 0:3].astype(np.float64)[:, None, :]
    D = rays[:, 4:7].astype(np.float64)[:, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (b[None, :, 0:3] - O) / D
        t1 = (b[None, :, 3:6] - O) / D
    near = np.nanmax(np.minimum(t0, t1), axis=2)
    far = np.nanmin(np.maximum(t0, t1), axis=2)
    near = np.maximum(near, 0.0)
    hit = near <= far
    tt = np.where(hit, near, np.inf)
    best = tt.argmin(axis=1)
    bt = tt[np.arange(len(rays)), best]
    anyhit = np.isfinite(bt)
    assert np.array_equal(got["prim_id"] != 0xFFFFFFFF, anyhit)
    with np.errstate(invalid="ignore"):
        sel = anyhit & (np.sort(tt, axis=1)[:, 1] - bt > 1e-5 * np.maximum(bt, 1))  # no near-ties
    assert np.array_equal(got["prim_id"][sel], best[sel])
    assert np.allclose(got["t"][anyhit], bt[anyhit], rtol=1e-6, atol=1e-7)
    assert (got["u"] == 0).all() and (got["v"] == 0).all()
    bf = oracle.brute_force(buf, rays)
    for f in ("prim_id", "t"):
        assert np.array_equal(got[f], bf[f]), f
