"""Helpers shared by the GPU parity tests."""
from __future__ import annotations

import numpy as np

RR_INVALID = 0xFFFFFFFF


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def to_dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def nodes_equal(gpu_buf: np.ndarray, ref_buf: np.ndarray) -> tuple[bool, str]:
    """Compare two BLAS/TLAS buffers node by node.  Float fields compare with
    == (so +0 and -0 are equal); the integer fields (links and the `update`
    counter, 0 between refits) must be equal."""
    import oracle

    g, r = oracle.nodes_of(gpu_buf), oracle.nodes_of(ref_buf)
    if g.shape != r.shape:
        return False, f"node count {g.shape} vs {r.shape}"
    if g.size == 0:
        return True, ""
    ints = [3, 7, 11, 15]
    floats = [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14]
    bad_i = np.nonzero((g[:, ints] != r[:, ints]).any(axis=1))[0]
    if bad_i.size:
        k = bad_i[0]
        return False, f"{bad_i.size} nodes differ in links; first {k}: {g[k]} vs {r[k]}"
    gf, rf = g[:, floats].view(np.float32), r[:, floats].view(np.float32)
    bad_f = np.nonzero((gf != rf).any(axis=1))[0]
    if bad_f.size:
        k = bad_f[0]
        return False, f"{bad_f.size} nodes differ in boxes; first {k}: {gf[k]} vs {rf[k]}"
    return True, ""


def hits_equal(gpu: dict, ref: np.ndarray, fields=("prim_id", "inst_id", "t", "u", "v")) -> tuple[bool, str]:
    """Bit-exact comparison of decoded GPU hits with oracle hits."""
    for f in fields:
        a = gpu[f]
        b = ref[f]
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), np.ascontiguousarray(b).view(np.uint32)
        bad = np.nonzero(a != b)[0]
        if bad.size:
            k = bad[0]
            return False, f"{f}: {bad.size} rays differ; first ray {k}: gpu {gpu[f][k]} oracle {ref[f][k]}"
    return True, ""


def tlas_tables_equal(gpu_buf: np.ndarray, ref_buf: np.ndarray) -> tuple[bool, str]:
    """Bit-exact comparison of a TLAS's per-instance transforms (inverse at
    2i, forward at 2i+1, 48 bytes each) and BLAS address table (8 bytes per
    instance), located through each buffer's own header."""
    n = int(ref_buf[8:12].view(np.uint32)[0])
    if int(gpu_buf[8:12].view(np.uint32)[0]) != n:
        return False, "instance counts differ"
    for name, field, per in (("transforms", slice(24, 32), 96), ("BLAS table", slice(32, 40), 8)):
        go, ro = int(gpu_buf[field].view(np.uint64)[0]), int(ref_buf[field].view(np.uint64)[0])
        if go != ro:
            return False, f"{name} offset {go} vs {ro}"
        g, r = gpu_buf[go:go + per * n], ref_buf[ro:ro + per * n]
        bad = np.nonzero(g != r)[0]
        if bad.size:
            inst = bad[0] // per
            return False, f"{name}: {np.unique(bad // per).size} instances differ; first {inst}"
    return True, ""
