"""GPU test of the traversal autotuner: every combination of the kernel's
run-time knobs must return the same hits (the tuner raises otherwise), and
the fastest combination is left set."""
import pytest

from radeonrays_sdk_amd import autotune, scenes
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


def test_sort_digit_width_sweep_keeps_results(ctx, dev):
    say("autotune sort digit width on 2M keys")
    lib = ctx.lib
    start = autotune.get(lib, autotune.rr.RRX_TUNING_SORT_DIGIT_BITS)
    try:
        out = autotune.tune_sort(ctx, 2 << 20, reps=2)
        assert [r["digit_bits"] for r in out["table"]] == [8, 10]
        assert autotune.get(lib, autotune.rr.RRX_TUNING_SORT_DIGIT_BITS) == out["best"]["digit_bits"]
    finally:
        autotune.set_(lib, autotune.rr.RRX_TUNING_SORT_DIGIT_BITS, start)
