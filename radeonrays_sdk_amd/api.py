"""Pythonic wrapper over the C ABI (see rr.py) for tests, benchmarks and the
serving loop.  torch supplies device memory and the stream; every operation
goes through the rr* entry points of libradeonrays.so.
"""
from __future__ import annotations

import ctypes
from contextlib import contextmanager
from ctypes import byref, c_size_t, c_void_p
from dataclasses import dataclass

import numpy as np

from . import rr

_QUERY = {"closest": rr.RR_INTERSECT_QUERY_CLOSEST, "any": rr.RR_INTERSECT_QUERY_ANY}
_OUTPUT = {"full": rr.RR_INTERSECT_QUERY_OUTPUT_FULL_HIT, "instance": rr.RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID,
           "full_t": rr.RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T}
_HIT_BYTES = {"full": 16, "instance": 4, "full_t": 32}


def _nbytes(t) -> int:
    return int(t.numel()) * int(t.element_size())


def _check_rays(rays, where: str) -> int:
    """Ray count of a [R, 8] float32 device tensor (RRRay records); anything
    else is refused here, before a kernel could read past the buffer."""
    if rays.dim() != 2 or int(rays.shape[1]) != 8 or rays.element_size() != 4 or not rays.is_contiguous():
        raise ValueError(f"{where}: rays must be a contiguous float32 [R, 8] tensor, got {tuple(rays.shape)}")
    return int(rays.shape[0])


def _check_size(buf, need: int, what: str, where: str) -> None:
    if _nbytes(buf) < need:
        raise ValueError(f"{where}: {what} holds {_nbytes(buf)} bytes, {need} needed")


def _torch():
    import torch
    return torch


@dataclass
class Accel:
    """A built BLAS or TLAS: its result buffer (uint8 tensor) and kind."""
    buffer: "object"
    kind: str              # "blas" or "tlas"
    prim_count: int
    children: tuple = ()   # BLASes a TLAS refers to (kept alive with it)

    @property
    def address(self) -> int:
        return self.buffer.data_ptr()


class Context:
    """An RRContext on one HIP device whose submissions run on torch's current stream."""

    def __init__(self, device: int = 0):
        torch = _torch()
        self.lib = rr.load()
        self.device = device
        self.torch_device = torch.device("cuda", device)
        with torch.cuda.device(device):
            self.stream = torch.cuda.current_stream(device)
        ctx = c_void_p()
        rr.check(self.lib.rrCreateContextHip(rr.RR_API_VERSION, device, c_void_p(self.stream.cuda_stream), byref(ctx)),
                 "rrCreateContextHip")
        self.ctx = ctx
        cs = c_void_p()
        rr.check(self.lib.rrGetCommandStreamFromHipStream(self.ctx, c_void_p(self.stream.cuda_stream), byref(cs)),
                 "rrGetCommandStreamFromHipStream")
        self.cs = cs  # immediate stream: rrCmd* enqueue on torch's stream

    # -- handles ------------------------------------------------------------
    @contextmanager
    def ptrs(self, *addrs):
        """RRDevicePtr handles for device addresses (None stays NULL), released on exit."""
        handles = []
        try:
            for a in addrs:
                if not a:  # None, or the null address of an empty tensor
                    handles.append(None)
                    continue
                p = c_void_p()
                rr.check(self.lib.rrGetDevicePtrFromHipPtr(self.ctx, c_void_p(int(a)), 0, byref(p)),
                         "rrGetDevicePtrFromHipPtr")
                handles.append(p)
            yield handles
        finally:
            for p in handles:
                if p is not None:
                    self.lib.rrReleaseDevicePtr(self.ctx, p)

    @contextmanager
    def on_stream(self, stream):
        """Enqueue on `stream` (a torch stream) instead of the context's stream
        inside the block, e.g. while that stream is captured into a graph."""
        cs = c_void_p()
        rr.check(self.lib.rrGetCommandStreamFromHipStream(self.ctx, c_void_p(stream.cuda_stream), byref(cs)),
                 "rrGetCommandStreamFromHipStream")
        saved, self.cs = self.cs, cs
        try:
            yield self
        finally:
            self.cs = saved
            self.lib.rrReleaseExternalCommandStream(self.ctx, cs)

    def empty(self, nbytes: int):
        torch = _torch()
        return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=self.torch_device)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.rrReleaseExternalCommandStream(self.ctx, self.cs)
            self.lib.rrDestroyContext(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- builds ---------------------------------------------------------------
    @staticmethod
    def _mesh(vertices, indices, vptr, iptr):
        if vertices.dim() != 2 or int(vertices.shape[1]) < 3 or vertices.element_size() != 4 or vertices.stride(1) != 1:
            raise ValueError(f"vertices must be float32 [V, >=3] with unit column stride, got {tuple(vertices.shape)}")
        if indices.dim() != 2 or int(indices.shape[1]) != 3 or indices.element_size() not in (2, 4):
            raise ValueError(f"indices must be int32/uint16 [T, 3], got {tuple(indices.shape)}")
        v_stride = vertices.stride(0) * vertices.element_size()
        prim = rr.RRTriangleMeshPrimitive(vptr, int(vertices.shape[0]), int(v_stride), iptr, int(indices.shape[0]),
                                          rr.RR_INDEX_TYPE_UINT16 if indices.element_size() == 2 else rr.RR_INDEX_TYPE_UINT32)
        inp = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, 1, ctypes.cast(ctypes.pointer(prim), c_void_p))
        return prim, inp

    @staticmethod
    def _options(restructure: bool):
        return rr.RRBuildOptions(0 if restructure else rr.RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD, None)

    def geometry_requirements(self, tri_count: int, vertex_count: int = 3, restructure: bool = False):
        prim = rr.RRTriangleMeshPrimitive(None, vertex_count, 12, None, tri_count, rr.RR_INDEX_TYPE_UINT32)
        inp = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, 1, ctypes.cast(ctypes.pointer(prim), c_void_p))
        req = rr.RRMemoryRequirements()
        opts = self._options(restructure)
        rr.check(self.lib.rrGetGeometryBuildMemoryRequirements(self.ctx, byref(inp), byref(opts), byref(req)),
                 "rrGetGeometryBuildMemoryRequirements")
        return req

    def build_geometry(self, vertices, indices, restructure: bool = False, scratch=None, result=None,
                       extra_flags: int = 0) -> Accel:
        """LBVH (+ treelet restructuring) of a mesh: vertices float32 [V, >=3], indices int32/uint16 [T, 3]."""
        torch = _torch()
        assert vertices.is_cuda and indices.is_cuda and vertices.dtype == torch.float32
        indices = indices.contiguous()
        tris = int(indices.shape[0])
        req = self.geometry_requirements(tris, max(int(vertices.shape[0]), 1), restructure)
        result = result if result is not None else self.empty(req.result_buffer_size)
        scratch = scratch if scratch is not None else self.empty(req.temporary_build_buffer_size)
        _check_size(result, req.result_buffer_size, "result", "build_geometry")
        _check_size(scratch, req.temporary_build_buffer_size, "scratch", "build_geometry")
        with self.ptrs(vertices.data_ptr(), indices.data_ptr(), scratch.data_ptr(), result.data_ptr()) as (v, i, s, r):
            prim, inp = self._mesh(vertices, indices, v, i)
            opts = self._options(restructure)
            opts.build_flags |= extra_flags
            rr.check(self.lib.rrCmdBuildGeometry(self.ctx, rr.RR_BUILD_OPERATION_BUILD, byref(inp), byref(opts), s, r,
                                                 self.cs), "rrCmdBuildGeometry")
        return Accel(result, "blas", tris)

    def update_geometry(self, accel: Accel, vertices, indices) -> None:
        """Refit `accel` to moved vertices (same triangles)."""
        indices = indices.contiguous()
        if int(indices.shape[0]) != accel.prim_count:
            raise ValueError(f"update_geometry: {int(indices.shape[0])} triangles, the BVH has {accel.prim_count}")
        with self.ptrs(vertices.data_ptr(), indices.data_ptr(), accel.address) as (v, i, r):
            prim, inp = self._mesh(vertices, indices, v, i)
            rr.check(self.lib.rrCmdBuildGeometry(self.ctx, rr.RR_BUILD_OPERATION_UPDATE, byref(inp), None, None, r,
                                                 self.cs), "rrCmdBuildGeometry(update)")

    def build_aabbs(self, boxes, restructure: bool = False, update: Accel = None) -> Accel:
        """BLAS over an AABB list: boxes float32 [n, 6] (min xyz, max xyz) on the
        device (RR_PRIMITIVE_TYPE_AABB_LIST); with update=<accel>, refit it."""
        boxes = boxes.contiguous()
        n = int(boxes.shape[0])
        prim0 = rr.RRAABBListPrimitive(None, n, 24)
        inp0 = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_AABB_LIST, 1, c_void_p(ctypes.addressof(prim0)))
        req = rr.RRMemoryRequirements()
        opts = self._options(restructure)
        rr.check(self.lib.rrGetGeometryBuildMemoryRequirements(self.ctx, byref(inp0), byref(opts), byref(req)),
                 "rrGetGeometryBuildMemoryRequirements")
        result = update.buffer if update is not None else self.empty(req.result_buffer_size)
        scratch = self.empty(req.temporary_update_buffer_size if update is not None else req.temporary_build_buffer_size)
        with self.ptrs(boxes.data_ptr() if n else None, scratch.data_ptr(), result.data_ptr()) as (b, s, r):
            prim = rr.RRAABBListPrimitive(b, n, 24)
            inp = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_AABB_LIST, 1, c_void_p(ctypes.addressof(prim)))
            op = rr.RR_BUILD_OPERATION_UPDATE if update is not None else rr.RR_BUILD_OPERATION_BUILD
            rr.check(self.lib.rrCmdBuildGeometry(self.ctx, op, byref(inp), byref(opts), s, r, self.cs),
                     "rrCmdBuildGeometry(aabbs)")
        return Accel(result, "blas", n)

    def build_geometry_multi(self, meshes, restructure: bool = False, update: Accel = None) -> Accel:
        """One BLAS over several meshes [(vertices, indices)] (one build input with
        primitive_count = len(meshes)); with update=<accel>, refit it instead."""
        meshes = [(v, i.contiguous()) for v, i in meshes]
        n = len(meshes)
        prims0, _ = self._batch_inputs(meshes, [None] * (2 * n))
        inp0 = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, n, c_void_p(ctypes.addressof(prims0)))
        req = rr.RRMemoryRequirements()
        opts = self._options(restructure)
        rr.check(self.lib.rrGetGeometryBuildMemoryRequirements(self.ctx, byref(inp0), byref(opts), byref(req)),
                 "rrGetGeometryBuildMemoryRequirements")
        result = update.buffer if update is not None else self.empty(req.result_buffer_size)
        scratch = self.empty(req.temporary_update_buffer_size if update is not None else req.temporary_build_buffer_size)
        addrs = [a for v, i in meshes for a in (v.data_ptr(), i.data_ptr())]
        with self.ptrs(*addrs, scratch.data_ptr(), result.data_ptr()) as hs:
            prims, _ = self._batch_inputs(meshes, hs[:2 * n])
            inp = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, n, c_void_p(ctypes.addressof(prims)))
            op = rr.RR_BUILD_OPERATION_UPDATE if update is not None else rr.RR_BUILD_OPERATION_BUILD
            rr.check(self.lib.rrCmdBuildGeometry(self.ctx, op, byref(inp), byref(opts), hs[-2], hs[-1], self.cs),
                     "rrCmdBuildGeometry(multi)")
        return Accel(result, "blas", sum(int(i.shape[0]) for _, i in meshes))

    def _batch_inputs(self, meshes, handles):
        """ctypes build inputs for [(vertices, indices)] with their RRDevicePtr handles."""
        n = len(meshes)
        prims = (rr.RRTriangleMeshPrimitive * max(n, 1))()
        inputs = (rr.RRGeometryBuildInput * max(n, 1))()
        for k, (v, i) in enumerate(meshes):
            vs = v.stride(0) * v.element_size()
            prims[k] = rr.RRTriangleMeshPrimitive(handles[2 * k], int(v.shape[0]), int(vs), handles[2 * k + 1],
                                                  int(i.shape[0]),
                                                  rr.RR_INDEX_TYPE_UINT16 if i.element_size() == 2 else rr.RR_INDEX_TYPE_UINT32)
            inputs[k] = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, 1,
                                                c_void_p(ctypes.addressof(prims) + k * ctypes.sizeof(prims[0])))
        return prims, inputs

    def batch_build_requirements(self, meshes, restructure: bool = False) -> int:
        _, inputs = self._batch_inputs(meshes, [None] * (2 * len(meshes)))
        size = c_size_t()
        opts = self._options(restructure)
        rr.check(self.lib.rrxGetGeometryBatchBuildMemoryRequirements(self.ctx, inputs, len(meshes), byref(opts),
                                                                    byref(size)),
                 "rrxGetGeometryBatchBuildMemoryRequirements")
        return size.value

    def result_arena(self, sizes, align: int = 256) -> list:
        """Result buffers for several builds carved from one device allocation
        (each slice `align`-aligned), for acceleration structures traversed
        together: separate allocations each get their own address range, and
        a two-level traversal over many of them pays for the scattered
        ranges in address translation (64 BLASes of the instanced scene,
        secondary rays: 5.4 ms from separate allocations against 4.1 ms from
        one range).  Returns one uint8 tensor view per size; the views keep
        the allocation alive."""
        offs, total = [], 0
        for n in sizes:
            offs.append(total)
            total += (int(n) + align - 1) // align * align
        base = self.empty(max(total, 1))
        return [base[o:o + int(n)] for o, n in zip(offs, sizes)]

    def build_geometry_batch(self, meshes, restructure: bool = False, scratch=None, results=None, update=False):
        """rrxCmdBuildGeometryBatch over [(vertices, indices)] (cuda tensors); returns one Accel per mesh."""
        meshes = [(v, i.contiguous()) for v, i in meshes]
        if results is None:
            results = [self.empty(self.geometry_requirements(int(i.shape[0]), max(int(v.shape[0]), 1),
                                                             restructure).result_buffer_size) for v, i in meshes]
        scratch = scratch if scratch is not None else self.empty(self.batch_build_requirements(meshes, restructure))
        addrs = [a for v, i in meshes for a in (v.data_ptr(), i.data_ptr())]
        with self.ptrs(*addrs, *[r.data_ptr() for r in results], scratch.data_ptr()) as hs:
            n = len(meshes)
            prims, inputs = self._batch_inputs(meshes, hs[:2 * n])
            outs = (c_void_p * max(n, 1))(*[h.value if h is not None else None for h in hs[2 * n:3 * n]])
            opts = self._options(restructure)
            op = rr.RR_BUILD_OPERATION_UPDATE if update else rr.RR_BUILD_OPERATION_BUILD
            rr.check(self.lib.rrxCmdBuildGeometryBatch(self.ctx, op, inputs, n, byref(opts), hs[-1], outs, self.cs),
                     "rrxCmdBuildGeometryBatch")
        return [Accel(r, "blas", int(i.shape[0])) for r, (v, i) in zip(results, meshes)]

    def update_geometry_batch(self, accels, meshes, scratch=None):
        """Refit every accel to its mesh's moved vertices (rrxCmdBuildGeometryBatch, UPDATE)."""
        return self.build_geometry_batch(meshes, scratch=scratch, results=[a.buffer for a in accels], update=True)

    def scene_requirements(self, count: int):
        inp = rr.RRSceneBuildInput(None, count)
        req = rr.RRMemoryRequirements()
        rr.check(self.lib.rrGetSceneBuildMemoryRequirements(self.ctx, byref(inp), None, byref(req)),
                 "rrGetSceneBuildMemoryRequirements")
        return req

    def build_scene(self, blases, transforms, update: Accel = None) -> Accel:
        """TLAS over instances: blases[i] (Accel) placed by transforms[i] (3x4 row-major).
        With update=<scene>, refit that scene to the new transforms instead (rrxCmdUpdateScene)."""
        transforms = np.asarray(transforms, dtype=np.float32).reshape(len(blases), 3, 4)
        n = len(blases)
        req = self.scene_requirements(n)
        result = update.buffer if update is not None else self.empty(req.result_buffer_size)
        scratch = self.empty(req.temporary_update_buffer_size if update is not None else req.temporary_build_buffer_size)
        uniq = list(dict.fromkeys(b.address for b in blases))  # one handle per distinct BLAS
        with self.ptrs(*uniq, scratch.data_ptr(), result.data_ptr()) as hs:
            handle = {addr: h.value for addr, h in zip(uniq, hs[:len(uniq)])}
            s, r = hs[len(uniq)], hs[len(uniq) + 1]
            arr = (rr.RRInstance * max(n, 1))()
            view = np.frombuffer(arr, dtype=_INSTANCE_DTYPE)  # the same 56-byte records
            view["geometry"][:n] = [handle[b.address] for b in blases]
            view["transform"][:n] = transforms
            inp = rr.RRSceneBuildInput(arr, n)
            if update is not None:
                rr.check(self.lib.rrxCmdUpdateScene(self.ctx, byref(inp), s, r, self.cs), "rrxCmdUpdateScene")
            else:
                rr.check(self.lib.rrCmdBuildScene(self.ctx, byref(inp), None, s, r, self.cs), "rrCmdBuildScene")
        return Accel(result, "tlas", n, tuple(blases))

    # -- queries --------------------------------------------------------------
    def trace_scratch(self, ray_count: int):
        size = c_size_t()
        rr.check(self.lib.rrGetTraceMemoryRequirements(self.ctx, max(ray_count, 1), byref(size)),
                 "rrGetTraceMemoryRequirements")
        return self.empty(size.value)

    def intersect(self, accel: Accel, rays, query: str = "closest", output: str = "full", indirect=None, hits=None,
                  scratch=None):
        """rrCmdIntersect: rays float32 [R, 8] on the device; returns the hit buffer (uint8 tensor)."""
        n = _check_rays(rays, "intersect")
        hits = hits if hits is not None else self.empty(n * _HIT_BYTES[output])
        scratch = scratch if scratch is not None else self.trace_scratch(n)
        _check_size(hits, n * _HIT_BYTES[output], "hits", "intersect")
        if n:
            size = c_size_t()
            rr.check(self.lib.rrGetTraceMemoryRequirements(self.ctx, n, byref(size)), "rrGetTraceMemoryRequirements")
            _check_size(scratch, size.value, "scratch", "intersect")
        ind = indirect.data_ptr() if indirect is not None else None
        with self.ptrs(accel.address, rays.data_ptr(), ind, hits.data_ptr(), scratch.data_ptr()) as (a, r, c, h, s):
            rr.check(self.lib.rrCmdIntersect(self.ctx, a, _QUERY[query], r, n, c, _OUTPUT[output], h, s, self.cs),
                     "rrCmdIntersect")
        return hits

    def intersect_with_stats(self, accel: Accel, rays, query: str = "closest"):
        torch = _torch()
        n = _check_rays(rays, "intersect_with_stats")
        hits = self.empty(n * 32)
        stats = torch.empty((n, 4), dtype=torch.int32, device=self.torch_device)
        scratch = self.trace_scratch(n)
        with self.ptrs(accel.address, rays.data_ptr(), hits.data_ptr(), scratch.data_ptr(), stats.data_ptr()) as (
                a, r, h, s, st):
            rr.check(self.lib.rrxCmdIntersectWithStats(self.ctx, a, _QUERY[query], r, n, h, s, st, self.cs),
                     "rrxCmdIntersectWithStats")
        return hits, stats

    def write_analyzer_dump(self, accel: Accel, bvh_path, config_path, rays=None, width: int = 0, height: int = 0,
                            rays_path=None):
        """rrxWriteAnalyzerDump: the BLAS as a bvh_analyzer dump + 7-line config (rays optional)."""
        rays_path = rays_path or (str(bvh_path) + ".rays")
        with self.ptrs(accel.address, rays.data_ptr() if rays is not None else None) as (a, r):
            rr.check(self.lib.rrxWriteAnalyzerDump(self.ctx, a, r, width, height, str(bvh_path).encode(),
                                                    str(rays_path).encode(), str(config_path).encode()),
                     "rrxWriteAnalyzerDump")

    def batch_scratch(self, request_count: int, total_rays: int, max_rays: int):
        size = c_size_t()
        rr.check(self.lib.rrxGetIntersectBatchMemoryRequirements(self.ctx, request_count, total_rays, max_rays,
                                                                 byref(size)),
                 "rrxGetIntersectBatchMemoryRequirements")
        return self.empty(size.value)

    def intersect_batch(self, requests, scratch=None):
        """rrxCmdIntersectBatch over requests = [(accel, rays, query, output[, hits])]; returns the hit buffers."""
        outs, addrs = [], []
        for req in requests:
            accel, rays, query, output = req[:4]
            n = _check_rays(rays, "intersect_batch")
            hits = req[4] if len(req) > 4 and req[4] is not None else self.empty(n * _HIT_BYTES[output])
            _check_size(hits, n * _HIT_BYTES[output], "hits", "intersect_batch")
            outs.append(hits)
            addrs += [accel.address, rays.data_ptr() if n else None, hits.data_ptr()]
        total = sum(int(r[1].shape[0]) for r in requests)
        largest = max([int(r[1].shape[0]) for r in requests] + [1])
        if scratch is None:
            scratch = self.batch_scratch(len(requests), total, largest)
        else:
            size = c_size_t()
            rr.check(self.lib.rrxGetIntersectBatchMemoryRequirements(self.ctx, len(requests), total, largest,
                                                                     byref(size)),
                     "rrxGetIntersectBatchMemoryRequirements")
            _check_size(scratch, size.value, "scratch", "intersect_batch")
        with self.ptrs(*addrs, scratch.data_ptr()) as hs:
            arr = (rr.RRxIntersectRequest * max(len(requests), 1))()
            for k, req in enumerate(requests):
                a, r, h = hs[3 * k: 3 * k + 3]
                arr[k] = rr.RRxIntersectRequest(a, r, h, int(req[1].shape[0]), _QUERY[req[2]], _OUTPUT[req[3]], 0)
            rr.check(self.lib.rrxCmdIntersectBatch(self.ctx, arr, len(requests), hs[-1], self.cs),
                     "rrxCmdIntersectBatch")
        return outs

    def sort_pairs(self, keys, values, key_bits: int = 32, tmp=None, scratch=None):
        """Stable radix sort of uint32 (stored as int32) key/value tensors in place."""
        torch = _torch()
        n = int(keys.numel())
        size = c_size_t()
        rr.check(self.lib.rrxGetSortMemoryRequirements(self.ctx, n, byref(size)), "rrxGetSortMemoryRequirements")
        kt, vt = tmp if tmp is not None else (torch.empty_like(keys), torch.empty_like(values))
        scratch = scratch if scratch is not None else self.empty(size.value)
        with self.ptrs(keys.data_ptr(), values.data_ptr(), kt.data_ptr(), vt.data_ptr(), scratch.data_ptr()) as (
                k, v, a, b, s):
            rr.check(self.lib.rrxCmdSortPairs(self.ctx, k, v, a, b, n, key_bits, s, self.cs), "rrxCmdSortPairs")


# RRInstance (radeonrays.h:228-232) as a numpy record: geometry handle, 3x4 transform.
_INSTANCE_DTYPE = np.dtype([("geometry", "<u8"), ("transform", "<f4", (3, 4))])
assert _INSTANCE_DTYPE.itemsize == ctypes.sizeof(rr.RRInstance)


# -- decoding helpers ------------------------------------------------------------
def decode_hits(buf, output: str = "full") -> dict:
    """Host view of a hit buffer: dict of numpy arrays (u, v, inst_id, prim_id[, t])."""
    a = buf.cpu().numpy() if hasattr(buf, "cpu") else np.asarray(buf)
    if output == "instance":
        return {"id": a.view(np.uint32)}
    rec = _HIT_BYTES[output] // 4
    w = a.view(np.uint32).reshape(-1, rec)
    f = a.view(np.float32).reshape(-1, rec)
    out = {"u": f[:, 0].copy(), "v": f[:, 1].copy(), "inst_id": w[:, 2].copy(), "prim_id": w[:, 3].copy()}
    if output == "full_t":
        out["t"] = f[:, 4].copy()
    return out
