"""rrxCmdSortPairsNoAlloc, the sort entry point that allocates nothing (its
scratch is a static device array per GPU), against rrxCmdSortPairs (the
same sort with a caller-provided scratch buffer) and numpy's stable
argsort: two calls in a row on one GPU, the second with a smaller input,
so the second runs on what the first left in the shared array; keys of 30
bits (four 8-bit passes: the result ends in keys) and of 24 bits (three
passes: it ends in the tmp arrays and is copied back).  A size above the
maximum is refused before anything is enqueued.
"""
import numpy as np
import pytest

from radeonrays_sdk_amd import rr
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 24


def _no_alloc_sort(ctx, keys, vals, key_bits):
    import torch

    kt, vt = torch.empty_like(keys), torch.empty_like(vals)
    with ctx.ptrs(keys.data_ptr(), vals.data_ptr(), kt.data_ptr(), vt.data_ptr()) as (k, v, a, b):
        rr.check(ctx.lib.rrxCmdSortPairsNoAlloc(ctx.ctx, k, v, a, b, int(keys.numel()), key_bits, ctx.cs),
                 "rrxCmdSortPairsNoAlloc")
    return kt, vt


@pytest.mark.parametrize("key_bits", [30, 24])
def test_no_alloc_sort_equals_sort_pairs(ctx, dev, key_bits):
    import torch

    for n in (5_000_003, 700_001):  # the second call is smaller than the first
        say(f"no-allocation sort n={n} key_bits={key_bits}")
        rng = np.random.default_rng(n + key_bits)
        keys = rng.integers(0, 1 << key_bits, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 5] = keys[3]  # duplicates: stability matters
        order = np.argsort(keys, kind="stable")
        ka = to_dev(keys.view(np.int32), dev)
        va = torch.arange(n, dtype=torch.int32, device=dev)
        kb, vb = ka.clone(), va.clone()
        ctx.sort_pairs(ka, va, key_bits)
        keep = _no_alloc_sort(ctx, kb, vb, key_bits)
        torch.cuda.synchronize()
        assert torch.equal(ka, kb) and torch.equal(va, vb), "differs from rrxCmdSortPairs"
        assert np.array_equal(vb.cpu().numpy().view(np.uint32), order.astype(np.uint32))
        assert np.array_equal(kb.cpu().numpy().view(np.uint32), keys[order])
        del keep


def test_no_alloc_sort_refuses_sizes_above_the_maximum(ctx, dev):
    import torch

    say("no-allocation sort: size above the maximum")
    k = torch.zeros(1, dtype=torch.int32, device=dev)
    with ctx.ptrs(k.data_ptr(), k.data_ptr(), k.data_ptr(), k.data_ptr()) as (a, b, c, d):
        assert ctx.lib.rrxCmdSortPairsNoAlloc(ctx.ctx, a, b, c, d, MAX_PAIRS + 1, 30, ctx.cs) == \
            rr.RR_ERROR_INVALID_PARAMETER
        # Nothing was enqueued: the stream still runs normally.
        assert ctx.lib.rrxCmdSortPairsNoAlloc(ctx.ctx, a, b, c, d, 1, 30, ctx.cs) == rr.RR_SUCCESS
    torch.cuda.synchronize()
