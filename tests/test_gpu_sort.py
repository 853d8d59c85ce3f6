"""GPU tests of the hybrid sort rrxCmdSortPairs takes on 2^22..2^24 keys:
one most-significant-digit pass (up to 10 bits, onesweep look-back), then
every bucket sorted by one workgroup, in LDS when it holds at most 24,576
keys and in chunks through memory when larger.

Each case is compared with numpy's stable argsort (keys and values, so
stability is checked too) and with the LSD passes (hybrid switched off).
The distributions cover uniform keys (every bucket in LDS), partial last
tiles, 31- and 32-bit keys, keys whose low part is 2 bits or empty, the
Morton codes of a heightfield (a surface: buckets of very uneven size,
some sorted in chunks), a few hot buckets of 10^5 keys, and all keys
equal (one bucket holds everything).
"""
import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import autotune, rr, scenes
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu


def _keys(n: int, key_bits: int, dist: str) -> np.ndarray:
    rng = np.random.default_rng(n + key_bits)
    if dist == "uniform":
        k = rng.integers(0, 1 << key_bits, size=n, dtype=np.uint64).astype(np.uint32)
        k[: n // 8] = k[5]  # duplicates too
        return k
    if dist == "hot":
        k = rng.integers(0, 1 << key_bits, size=n, dtype=np.uint64).astype(np.uint32)
        for b in range(5):  # five buckets of about 150,000 keys each
            sel = rng.random(n) < 0.035
            hot = np.uint32(rng.integers(0, 1 << 10) << (key_bits - 10))
            k[sel] = hot | (k[sel] & np.uint32((1 << (key_bits - 10)) - 1))
        return k
    if dist == "equal":
        return np.full(n, 0x2ABCDEF & ((1 << key_bits) - 1), np.uint32)
    if dist == "morton":
        v, i = scenes.heightfield_mesh(res=1449)  # 4,190,208 triangles
        k = oracle.morton_codes(v, i)
        return k[:n] if n <= k.shape[0] else np.concatenate([k, k[: n - k.shape[0]]])
    raise ValueError(dist)


def _sort(ctx, dev, keys, key_bits, hybrid):
    import torch

    lib = ctx.lib
    start = autotune.get(lib, rr.RRX_TUNING_SORT_HYBRID)
    try:
        autotune.set_(lib, rr.RRX_TUNING_SORT_HYBRID, hybrid)
        k = to_dev(keys.view(np.int32), dev)
        v = torch.arange(keys.shape[0], dtype=torch.int32, device=dev)
        ctx.sort_pairs(k, v, key_bits)
        torch.cuda.synchronize()
        return k.cpu().numpy().view(np.uint32), v.cpu().numpy().view(np.uint32)
    finally:
        autotune.set_(lib, rr.RRX_TUNING_SORT_HYBRID, start)


@pytest.mark.parametrize("n,key_bits,dist", [
    (1 << 22, 30, "uniform"),
    ((1 << 22) + 12345, 30, "uniform"),
    (5_000_003, 32, "uniform"),
    (9_999_999, 31, "uniform"),
    (1 << 24, 30, "uniform"),
    ((1 << 22) + 1, 12, "uniform"),
    ((1 << 22) + 3, 10, "uniform"),
    (4_190_208, 30, "morton"),
    ((1 << 22) + 77, 30, "hot"),
    (1 << 22, 30, "equal"),
])
def test_hybrid_sort_matches_numpy_and_lsd(ctx, dev, n, key_bits, dist):
    say(f"hybrid sort n={n} key_bits={key_bits} {dist}")
    keys = _keys(n, key_bits, dist)
    order = np.argsort(keys, kind="stable")
    hk, hv = _sort(ctx, dev, keys, key_bits, 1)
    assert np.array_equal(hv, order.astype(np.uint32)), "values (stable order) differ from numpy"
    assert np.array_equal(hk, keys[order]), "keys differ from numpy"
    lk, lv = _sort(ctx, dev, keys, key_bits, 0)
    assert np.array_equal(lk, hk) and np.array_equal(lv, hv), "hybrid and LSD passes differ"
    assert _device_errors(ctx) == 0


def _device_errors(ctx) -> int:
    from ctypes import byref, c_uint32

    e = c_uint32()
    rr.check(ctx.lib.rrxGetDeviceErrors(ctx.ctx, byref(e)), "rrxGetDeviceErrors")
    return e.value
