"""GPU test of the traversal autotuner: every combination of the kernel's
run-time knobs must return the same hits (the tuner raises otherwise), and
the fastest combination is left set."""
import pytest

from radeonrays_sdk_amd import api, autotune, scenes
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu


def test_tuning_sweep_keeps_results(ctx, dev):
    import torch

    say("autotune sweep on a 256x256 heightfield, 200k rays")
    w = scenes.heightfield(res=256, n_rays=200000)
    accel = ctx.build_geometry(to_dev(w.vertices, dev), to_dev(w.indices, dev), restructure=True)
    rays = to_dev(w.rays, dev)
    torch.cuda.synchronize()
    start = autotune.current(ctx.lib)
    try:
        out = autotune.tune_traversal(ctx, accel, rays, reps=2, min_busy=(16, 32, 48))
        assert len(out["table"]) == 3 * 3 * 3
        best = out["best"]
        assert autotune.current(ctx.lib) == {k: best[k] for k in ("rays_per_fetch", "lds_stack", "min_busy")}
    finally:
        autotune.apply(ctx.lib, start)


@pytest.mark.parametrize("n", [2 << 20, 4 << 20])
def test_sort_variant_sweep_keeps_results(ctx, dev, n):
    say(f"autotune sort variants on {n} keys")
    lib = ctx.lib
    start = (autotune.get(lib, autotune.rr.RRX_TUNING_SORT_DIGIT_BITS),
             autotune.get(lib, autotune.rr.RRX_TUNING_SORT_HYBRID))
    try:
        out = autotune.tune_sort(ctx, n, reps=2)
        want = [(4, 0), (8, 0), (10, 0)] + ([(8, 1)] if n >= 1 << 22 else [])
        assert [(r["digit_bits"], r["hybrid"]) for r in out["table"]] == want
        assert autotune.get(lib, autotune.rr.RRX_TUNING_SORT_DIGIT_BITS) == out["best"]["digit_bits"]
        assert autotune.get(lib, autotune.rr.RRX_TUNING_SORT_HYBRID) == out["best"]["hybrid"]
    finally:
        autotune.set_(lib, autotune.rr.RRX_TUNING_SORT_DIGIT_BITS, start[0])
        autotune.set_(lib, autotune.rr.RRX_TUNING_SORT_HYBRID, start[1])


@pytest.mark.parametrize("leaves", [5, 9])
def test_treelet_sizes_match_oracle(ctx, dev, leaves):
    """Restructuring with 5- or 9-leaf treelets equals the oracle's
    restatement node for node, and its hits equal the oracle's."""
    import numpy as np
    import torch

    import oracle
    from radeonrays_sdk_amd import api
    from tests.gpu_util import hits_equal, nodes_equal

    say(f"restructure with {leaves}-leaf treelets vs oracle")
    lib = ctx.lib
    start = autotune.get(lib, autotune.rr.RRX_TUNING_TREELET_LEAVES)
    try:
        autotune.set_(lib, autotune.rr.RRX_TUNING_TREELET_LEAVES, leaves)
        for v, i in (scenes.triangle_soup(20000, scenes.rng_for(40, leaves)),
                     scenes.heightfield_mesh(90, seed=scenes.SEED_BASE + leaves)):
            i = i.reshape(-1, 3)
            accel = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True)
            torch.cuda.synchronize()
            buf = accel.buffer.cpu().numpy()
            ok, msg = nodes_equal(buf, oracle.build_blas(v, i, True, treelet_leaves=leaves))
            assert ok, msg
            lo, hi = v.min(axis=0), v.max(axis=0)
            rng = np.random.default_rng(leaves)
            o = rng.uniform(lo, hi, size=(20000, 3)).astype(np.float32)
            rays = scenes.make_rays(o, scenes.unit_sphere(20000, rng))
            for q in ("closest", "any"):
                hits = ctx.intersect(accel, to_dev(rays, dev), q, "full_t")
                torch.cuda.synchronize()
                ok, msg = hits_equal(api.decode_hits(hits, "full_t"), oracle.trace(buf, rays, any_hit=(q == "any")))
                assert ok, f"{q}: {msg}"
    finally:
        autotune.set_(lib, autotune.rr.RRX_TUNING_TREELET_LEAVES, start)


def test_treelet_sweep_keeps_closest_hits(ctx, dev):
    say("autotune treelet size on a 256x256 heightfield")
    w = scenes.heightfield(res=256, n_rays=100000)
    lib = ctx.lib
    start = autotune.get(lib, autotune.rr.RRX_TUNING_TREELET_LEAVES)
    try:
        out = autotune.tune_treelet(ctx, to_dev(w.vertices, dev), to_dev(w.indices, dev), to_dev(w.rays, dev), reps=2)
        assert [r["treelet_leaves"] for r in out["table"]] == [5, 7, 9]
        assert autotune.get(lib, autotune.rr.RRX_TUNING_TREELET_LEAVES) == out["best"]["treelet_leaves"]
    finally:
        autotune.set_(lib, autotune.rr.RRX_TUNING_TREELET_LEAVES, start)


def test_query_grid_size_keeps_results(ctx, dev):
    """RRX_TUNING_TRACE_RAYS_PER_LANE changes only how many workgroups a
    query's persistent launch gets (from one per CU up to the resident
    grid): closest and any hits must equal the plain one-lane-per-ray
    kernel's at every setting, for both levels and for ray counts below,
    inside and above the range the setting shapes."""
    import numpy as np
    import torch

    say("query grid sizes: 256x256 heightfield and a 12-instance scene, rays per lane 0/1/2/3/64")
    lib = ctx.lib
    w = scenes.heightfield(res=256, n_rays=600000)
    blas = ctx.build_geometry(to_dev(w.vertices, dev), to_dev(w.indices, dev), restructure=True)
    soups = [scenes.triangle_soup(n, scenes.rng_for(11, n), sigma=0.2) for n in (3000, 20000, 50000)]
    blases = [ctx.build_geometry(to_dev(v, dev), to_dev(i.reshape(-1, 3), dev), restructure=True) for v, i in soups]
    xf = scenes.random_transforms(scenes.rng_for(11, 99), 12, trans=1.0)
    tlas = ctx.build_scene([blases[k % 3] for k in range(12)], xf)
    rng = scenes.rng_for(11, 7)
    o = rng.uniform(-3, 3, size=(600000, 3)).astype(np.float32)
    scene_rays = to_dev(scenes.make_rays(o, scenes.unit_sphere(600000, rng)), dev)
    rays = to_dev(w.rays, dev)
    torch.cuda.synchronize()
    start = autotune.get(lib, autotune.rr.RRX_TUNING_TRACE_RAYS_PER_LANE)
    refs = {}
    try:
        for accel, r, what in ((blas, rays, "single level"), (tlas, scene_rays, "two level")):
            for n in (1000, 70000, 300000, 600000):
                sub = r[:n].contiguous()
                for query in ("closest", "any"):
                    with autotune.plain_traversal(lib):
                        ref = ctx.intersect(accel, sub, query, "full_t")
                    refs[(what, n, query)] = ref
                    for k in (0, 1, 2, 3, 64):
                        autotune.set_(lib, autotune.rr.RRX_TUNING_TRACE_RAYS_PER_LANE, k)
                        got = ctx.intersect(accel, sub, query, "full_t")
                        torch.cuda.synchronize()
                        assert torch.equal(got, ref), f"{what}, {n} rays, {query}, rays per lane {k}"
                    hit = api.decode_hits(ref, "full_t")["prim_id"] != 0xFFFFFFFF
                    assert int(hit.sum()) > n // 100, f"{what}, {n} rays, {query}: almost no hits"
        # The batch kernels size their grids by the same rule.
        mix = [(blas, rays, "single level", 70000, "closest"), (tlas, scene_rays, "two level", 300000, "any"),
               (blas, rays, "single level", 600000, "any"), (tlas, scene_rays, "two level", 1000, "closest")]
        for k in (0, 1, 2, 64):
            autotune.set_(lib, autotune.rr.RRX_TUNING_TRACE_RAYS_PER_LANE, k)
            outs = ctx.intersect_batch([(a, r[:n].contiguous(), q, "full_t") for a, r, _, n, q in mix])
            torch.cuda.synchronize()
            for (_, _, what, n, q), got in zip(mix, outs):
                assert torch.equal(got, refs[(what, n, q)]), f"batch: {what}, {n} rays, {q}, rays per lane {k}"
    finally:
        autotune.set_(lib, autotune.rr.RRX_TUNING_TRACE_RAYS_PER_LANE, start)


def test_two_level_knobs_keep_results(ctx, dev):
    """RRX_TUNING_TRACE_2L_MIN_BUSY / _2L_RAYS_PER_FETCH tune only the
    persistent two-level kernel's refills: its hits must equal the plain
    kernel's at every setting, and the single-level knobs must not move them."""
    import numpy as np
    import torch

    say("two-level refill knobs on a 24-instance scene, 200k rays")
    lib = ctx.lib
    soups = [scenes.triangle_soup(n, scenes.rng_for(12, n), sigma=0.2) for n in (5000, 40000)]
    blases = [ctx.build_geometry(to_dev(v, dev), to_dev(i.reshape(-1, 3), dev), restructure=True) for v, i in soups]
    tlas = ctx.build_scene([blases[k % 2] for k in range(24)], scenes.random_transforms(scenes.rng_for(12, 5), 24,
                                                                                         trans=1.5))
    rng = scenes.rng_for(12, 6)
    o = rng.uniform(-3, 3, size=(200000, 3)).astype(np.float32)
    rays = to_dev(scenes.make_rays(o, scenes.unit_sphere(200000, rng)), dev)
    torch.cuda.synchronize()
    keys = (autotune.rr.RRX_TUNING_TRACE_2L_MIN_BUSY, autotune.rr.RRX_TUNING_TRACE_2L_RAYS_PER_FETCH)
    start = [autotune.get(lib, k) for k in keys]
    try:
        for query in ("closest", "any"):
            with autotune.plain_traversal(lib):
                ref = ctx.intersect(tlas, rays, query, "full_t")
            for mb, f in ((1, 32), (16, 128), (32, 64), (64, 64), (48, 32)):
                autotune.set_(lib, keys[0], mb)
                autotune.set_(lib, keys[1], f)
                got = ctx.intersect(tlas, rays, query, "full_t")
                # The batch kernel's scene requests take the same threshold.
                parts = ctx.intersect_batch([(tlas, rays[:70000], query, "full_t"),
                                             (tlas, rays[70000:], query, "full_t")])
                torch.cuda.synchronize()
                assert torch.equal(got, ref), f"{query}, min busy {mb}, fetch {f}"
                assert torch.equal(torch.cat(parts), ref), f"batch: {query}, min busy {mb}"
            hit = api.decode_hits(ref, "full_t")["prim_id"] != 0xFFFFFFFF
            assert int(hit.sum()) > 2000, f"{query}: almost no hits"
    finally:
        for k, v in zip(keys, start):
            autotune.set_(lib, k, v)


def test_batch_refill_knob_keeps_results(ctx, dev):
    """RRX_TUNING_TRACE_BATCH_MIN_BUSY moves only when the batch kernel's
    waves refill: the geometry requests' hits must equal the plain kernel's
    at every setting."""
    import torch

    say("batch refill threshold 1/12/24/64 on a 256x256 heightfield, 6 requests")
    lib = ctx.lib
    w = scenes.heightfield(res=256, n_rays=300000)
    blas = ctx.build_geometry(to_dev(w.vertices, dev), to_dev(w.indices, dev), restructure=True)
    rays = to_dev(w.rays, dev)
    cuts = [0, 1, 700, 50000, 50001, 200000, 300000]
    torch.cuda.synchronize()
    refs = []
    with autotune.plain_traversal(lib):
        for k in range(len(cuts) - 1):
            q = "closest" if k % 2 else "any"
            refs.append(ctx.intersect(blas, rays[cuts[k]:cuts[k + 1]].contiguous(), q, "full_t"))
    start = autotune.get(lib, autotune.rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY)
    try:
        for mb in (1, 12, 24, 64):
            autotune.set_(lib, autotune.rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, mb)
            outs = ctx.intersect_batch([(blas, rays[cuts[k]:cuts[k + 1]].contiguous(), "closest" if k % 2 else "any",
                                         "full_t") for k in range(len(cuts) - 1)])
            torch.cuda.synchronize()
            for k, (got, ref) in enumerate(zip(outs, refs)):
                assert torch.equal(got, ref), f"request {k}, min busy {mb}"
    finally:
        autotune.set_(lib, autotune.rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, start)


def test_two_level_and_batch_tuners(ctx, dev):
    """tune_two_level and tune_batch sweep their knobs on the caller's own
    work, raise if any setting changes a hit, and leave the fastest set."""
    import numpy as np
    import torch

    say("tune_two_level on a 16-instance scene, tune_batch on 40 requests")
    lib = ctx.lib
    v, i = scenes.triangle_soup(20000, scenes.rng_for(13, 1), sigma=0.2)
    blas = ctx.build_geometry(to_dev(v, dev), to_dev(i.reshape(-1, 3), dev), restructure=True)
    tlas = ctx.build_scene([blas] * 16, scenes.random_transforms(scenes.rng_for(13, 2), 16, trans=1.5))
    rng = scenes.rng_for(13, 3)
    o = rng.uniform(-3, 3, size=(100000, 3)).astype(np.float32)
    rays = to_dev(scenes.make_rays(o, scenes.unit_sphere(100000, rng)), dev)
    torch.cuda.synchronize()
    keys2 = (autotune.rr.RRX_TUNING_TRACE_2L_RAYS_PER_FETCH, autotune.rr.RRX_TUNING_TRACE_2L_MIN_BUSY)
    kb = autotune.rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY
    start = [autotune.get(lib, k) for k in keys2] + [autotune.get(lib, kb)]
    try:
        t = autotune.tune_two_level(ctx, tlas, rays, queries=("closest", "any"), reps=2, min_busy=(16, 32))
        assert len(t["table"]) == 3 * 2 and t["default"] is not None
        assert [autotune.get(lib, k) for k in keys2] == [t["best"]["rays_per_fetch"], t["best"]["min_busy"]]
        reqs = [(blas, rays[k * 2500:(k + 1) * 2500], "any" if k % 3 else "closest", "full") for k in range(40)]
        b = autotune.tune_batch(ctx, reqs, reps=2, min_busy=(8, 24, 48))
        assert [r["min_busy"] for r in b["table"]] == [8, 24, 48] and b["default"]["min_busy"] == 24
        assert autotune.get(lib, kb) == b["best"]["min_busy"]
    finally:
        for k, val in zip(keys2 + (kb,), start):
            autotune.set_(lib, k, val)
