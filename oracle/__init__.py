"""CPU oracle (test infrastructure only): ctypes binding of liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  See oracle.cpp for what each
function restates and which reference file it follows.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# RR_ORACLE_LIB names another build (the host-sanitizer build, liboracle_asan.so).
LIB = Path(os.environ.get("RR_ORACLE_LIB") or HERE / "liboracle.so")
REF_ANALYZER = HERE / "_ref" / "bvh_analyzer"

HIT_DTYPE = np.dtype([("u", "<f4"), ("v", "<f4"), ("inst_id", "<u4"), ("prim_id", "<u4"), ("t", "<f4"),
                      ("pad", "<u4", 3)])
STATS_DTYPE = np.dtype([("internal_nodes", "<u4"), ("leaves", "<u4"), ("box_tests", "<u4"), ("tri_tests", "<u4")])

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        lib = ctypes.CDLL(str(LIB))
        P = c_void_p
        sig = {
            "orc_mesh_bounds": ([P, c_uint32, P, c_uint32, P, P], None),
            "orc_morton_codes": ([P, c_uint32, P, c_uint32, P], None),
            "orc_morton3": ([c_float, c_float, c_float], c_uint32),
            "orc_sort_pairs": ([P, P, c_uint32], None),
            "orc_set_threads": ([c_int], None),
            "orc_emit_hierarchy": ([P, P, c_uint32, P], None),
            "orc_refit": ([P, c_uint32, P, c_uint32, P], None),
            "orc_restructure": ([c_uint32, P], None),
            "orc_restructure_leaves": ([c_uint32, P, c_int], None),
            "orc_blas_bytes": ([c_uint32], c_size_t),
            "orc_build_blas": ([P, c_uint32, P, c_uint32, c_int, P], None),
            "orc_build_blas_leaves": ([P, c_uint32, P, c_uint32, c_int, c_int, P], None),
            "orc_tlas_bytes": ([c_uint32], c_size_t),
            "orc_build_tlas": ([P, P, P, c_uint32, P], None),
            "orc_trace": ([P, P, P, c_uint32, c_int, P, P, c_int], None),
            "orc_brute_force": ([P, P, P, c_uint32, P, c_int], None),
            "orc_triangle_fp64": ([P, P, P, P, c_double, POINTER(c_double), POINTER(c_double), POINTER(c_double)], c_int),
            "orc_treelet_dp": ([P, P, P], None),
            "orc_is_valid": ([P], c_int),
            "orc_check_consistency": ([P], c_int),
            "orc_sah": ([P], c_float),
            "orc_height": ([P], c_uint32),
            "orc_write_analyzer_dump": ([P, P, c_uint32, c_uint32, c_char_p, c_char_p, c_char_p], c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def check_accel(buf: np.ndarray) -> None:
    """Raise unless `buf` holds a well-formed acceleration-structure header
    whose node array (and TLAS tables) lie inside the buffer."""
    if buf.dtype != np.uint8 or buf.size < 128:
        raise ValueError("accel buffer must be uint8 with a 128-byte header")
    hdr = buf[:128].view(np.uint32)
    magic, kind, prims, count = (int(x) for x in hdr[:4])
    offs = buf[16:40].view(np.uint64)
    if magic != 0x53415252 or kind not in (1, 2):
        raise ValueError(f"bad header magic/kind {magic:#x}/{kind}")
    if count != (2 * prims - 1 if prims else 0):
        raise ValueError(f"node_count {count} inconsistent with prim_count {prims}")
    if int(offs[0]) + 64 * count > buf.size:
        raise ValueError("node array runs past the buffer")
    if kind == 2 and (int(offs[1]) + 96 * prims > buf.size or int(offs[2]) + 8 * prims > buf.size):
        raise ValueError("TLAS tables run past the buffer")


def _p(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def morton3(x, y, z) -> int:
    return int(load().orc_morton3(x, y, z))


def mesh_bounds(v, i):
    v, i = _c(v, np.float32), _c(i, np.uint32)
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    load().orc_mesh_bounds(_p(v), v.shape[1], _p(i), i.shape[0], _p(lo), _p(hi))
    return lo, hi


def morton_codes(v, i):
    v, i = _c(v, np.float32), _c(i, np.uint32)
    out = np.zeros(i.shape[0], np.uint32)
    load().orc_morton_codes(_p(v), v.shape[1], _p(i), i.shape[0], _p(out))
    return out


def sort_pairs(keys, vals):
    k, v = _c(keys, np.uint32).copy(), _c(vals, np.uint32).copy()
    load().orc_sort_pairs(_p(k), _p(v), k.size)
    return k, v


def set_threads(threads: int) -> None:
    """OpenMP threads of the parallel build stages (results do not depend on it)."""
    load().orc_set_threads(int(threads))


def emit_hierarchy(sorted_codes, sorted_prims):
    n = len(sorted_codes)
    nodes = np.zeros((max(2 * n - 1, 0), 16), np.uint32)
    load().orc_emit_hierarchy(_p(_c(sorted_codes, np.uint32)), _p(_c(sorted_prims, np.uint32)), n, _p(nodes))
    return nodes


def build_blas(v, i, restructure=False, treelet_leaves=7) -> np.ndarray:
    v, i = _c(v, np.float32), _c(i, np.uint32)
    n = i.shape[0]
    buf = np.zeros(load().orc_blas_bytes(n), np.uint8)
    load().orc_build_blas_leaves(_p(v), v.shape[1], _p(i), n, int(restructure), int(treelet_leaves), _p(buf))
    return buf


def build_blas_aabbs(boxes, restructure=False) -> np.ndarray:
    """BLAS over an AABB list [n, 6] (min xyz, max xyz): the LBVH of the
    unindexed triangles (min, max, min), whose bounds, centroids and leaf
    boxes are the boxes', flagged as an AABB BLAS (header flag 2)."""
    b = _c(boxes, np.float32).reshape(-1, 6)
    n = b.shape[0]
    v = np.stack([b[:, 0:3], b[:, 3:6], b[:, 0:3]], axis=1).reshape(-1, 3)
    buf = build_blas(v, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3), restructure)
    hdr = buf[:128].view(np.uint32)
    hdr[10] |= 2
    return buf


def restructure(buf: np.ndarray, tris: int, treelet_leaves=7) -> np.ndarray:
    out = buf.copy()
    load().orc_restructure_leaves(tris, _p(out[128:]), int(treelet_leaves))
    return out


def refit(buf: np.ndarray, v, i) -> np.ndarray:
    v, i = _c(v, np.float32), _c(i, np.uint32)
    out = buf.copy()
    load().orc_refit(_p(v), v.shape[1], _p(i), i.shape[0], _p(out[128:]))
    return out


def build_tlas(xforms, blas_addrs, blas_bufs) -> np.ndarray:
    n = len(blas_bufs)
    x = _c(xforms, np.float32).reshape(n, 12) if n else np.zeros((1, 12), np.float32)
    a = np.asarray(blas_addrs, dtype=np.uint64) if n else np.zeros(1, np.uint64)
    ptrs = (c_void_p * max(n, 1))(*[b.ctypes.data for b in blas_bufs])
    buf = np.zeros(load().orc_tlas_bytes(n), np.uint8)
    load().orc_build_tlas(_p(x), _p(a), ptrs, n, _p(buf))
    return buf


def _blas_ptrs(accel: np.ndarray, blas_map):
    """Array of host pointers, one per TLAS instance, from {device address: host buffer}."""
    hdr = accel[:128].view(np.uint32)
    if hdr[1] != 2:
        return None, ()
    n = int(hdr[2])
    off = int(accel[32:40].view(np.uint64)[0])  # blas_table_offset
    addrs = accel[off:off + 8 * n].view(np.uint64)
    bufs = [blas_map[int(a)] for a in addrs]
    for b in bufs:
        check_accel(b)
    return (c_void_p * max(n, 1))(*[b.ctypes.data for b in bufs]), bufs


def trace(accel: np.ndarray, rays, any_hit=False, blas_map=None, stats=False, threads=0):
    check_accel(accel)
    rays = _c(rays, np.float32)
    n = rays.shape[0]
    hits = np.zeros(n, HIT_DTYPE)
    st = np.zeros(n, STATS_DTYPE) if stats else None
    ptrs, keep = _blas_ptrs(accel, blas_map or {})
    load().orc_trace(_p(accel), ptrs, _p(rays), n, int(any_hit), _p(hits), _p(st), threads)
    return (hits, st) if stats else hits


def brute_force(accel: np.ndarray, rays, blas_map=None, threads=0):
    check_accel(accel)
    rays = _c(rays, np.float32)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    ptrs, keep = _blas_ptrs(accel, blas_map or {})
    load().orc_brute_force(_p(accel), ptrs, _p(rays), rays.shape[0], _p(hits), threads)
    return hits


def triangle_fp64(ray, v0, v1, v2, eps=1e-9):
    t, u, v = c_double(), c_double(), c_double()
    ok = load().orc_triangle_fp64(_p(_c(ray, np.float32)), _p(_c(v0, np.float32)), _p(_c(v1, np.float32)),
                                  _p(_c(v2, np.float32)), eps, ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return bool(ok), t.value, u.value, v.value


def treelet_dp(boxes):
    """(cost[128], part[128]) of the treelet dynamic programme over 7 leaf boxes [7, 6]."""
    b = _c(boxes, np.float32).reshape(7, 6)
    cost, part = np.zeros(128, np.float32), np.zeros(128, np.uint32)
    load().orc_treelet_dp(_p(b), _p(cost), _p(part))
    return cost, part


def is_valid(accel) -> bool:
    check_accel(accel)
    return bool(load().orc_is_valid(_p(accel)))


def check_consistency(accel) -> bool:
    check_accel(accel)
    return bool(load().orc_check_consistency(_p(accel)))


def sah(accel) -> float:
    check_accel(accel)
    return float(load().orc_sah(_p(accel)))


def height(accel) -> int:
    check_accel(accel)
    return int(load().orc_height(_p(accel)))


def write_analyzer_dump(accel, rays, width, height_, bvh_path, rays_path, cfg_path) -> None:
    rc = load().orc_write_analyzer_dump(_p(accel), _p(_c(rays, np.float32)), width, height_, str(bvh_path).encode(),
                                        str(rays_path).encode(), str(cfg_path).encode())
    if rc:
        raise RuntimeError("orc_write_analyzer_dump failed")


def nodes_of(accel: np.ndarray) -> np.ndarray:
    """Node array of a buffer as uint32 [count, 16] (float fields viewable via .view(np.float32))."""
    check_accel(accel)
    hdr = accel[:128].view(np.uint32)
    off = int(accel[16:24].view(np.uint64)[0])
    count = int(hdr[3])
    return accel[off:off + 64 * count].view(np.uint32).reshape(count, 16)
