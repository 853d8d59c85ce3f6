"""Python face of the serving loop (rrxCreateServer and friends).

The loop itself is native (csrc/serve.cpp): a host thread releases queued
requests at their arrival times, admits them under an HBM budget for their
ray and hit buffers, and dispatches them.  In "batched" mode every dispatch
is one rrxCmdIntersectBatch over everything admitted so far, on whichever of
`max_inflight` streams is idle, so requests that arrive while the GPU is busy
are coalesced into the next persistent-traversal launch and launches overlap.
In "plain" mode each request is one rrCmdIntersect + rrSumbitCommandStream +
rrWaitEvent, in order: the per-call path a client of the reference API runs.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_uint64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import rr
from .api import _HIT_BYTES, _OUTPUT, _QUERY, Accel, Context

TIMING_DTYPE = np.dtype([("arrival_ns", "<u8"), ("dispatch_ns", "<u8"), ("complete_ns", "<u8"),
                         ("batch_id", "<u4"), ("batch_size", "<u4")])


@dataclass
class Request:
    """One ray query: `count` rays starting at byte `ray_offset` of `rays`
    (a device tensor of RRRay records), hits written to `hits`."""
    accel: Accel
    rays: object
    ray_offset: int
    count: int
    query: str
    output: str
    hits: object


@dataclass
class HostRequest:
    """One ray query whose rays and hits are host tensors (pinned memory
    copies fastest): `rays` float32 [count, 8], `hits` uint8 of at least
    count * record bytes.  The server stages them through its device arena."""
    accel: Accel
    rays: object
    query: str
    output: str
    hits: object


class Server:
    def __init__(self, ctx: Context, mode: str = "batched", max_inflight: int = 2, max_batch_requests: int = 256,
                 max_batch_rays: int = 4 << 20, memory_budget: int = 200 << 30, staging_bytes: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        opts = rr.RRxServerOptions(rr.RRX_SERVE_PLAIN if mode == "plain" else rr.RRX_SERVE_BATCHED, max_inflight,
                                   max_batch_requests, max_batch_rays, memory_budget, staging_bytes)
        h = c_void_p()
        rr.check(self.lib.rrxCreateServer(ctx.ctx, byref(opts), byref(h)), "rrxCreateServer")
        self.handle = h
        self._ptrs = []  # RRDevicePtr handles of submitted requests, released on close

    def now(self) -> int:
        return int(self.lib.rrxServerNow(self.handle))

    def _ptr(self, addr, offset=0):
        if not addr:
            return None
        p = c_void_p()
        rr.check(self.lib.rrGetDevicePtrFromHipPtr(self.ctx.ctx, c_void_p(int(addr)), offset, byref(p)),
                 "rrGetDevicePtrFromHipPtr")
        self._ptrs.append(p)
        return p

    def submit(self, requests, release_ns=None) -> int:
        """Queue requests (released at release_ns[i] on the server clock, or now); returns the first ticket."""
        n = len(requests)
        arr = (rr.RRxIntersectRequest * max(n, 1))()
        accel_ptrs = {}
        for k, q in enumerate(requests):
            if q.accel.address not in accel_ptrs:
                accel_ptrs[q.accel.address] = self._ptr(q.accel.address)
            if q.hits.numel() * q.hits.element_size() < q.count * _HIT_BYTES[q.output]:
                raise ValueError(f"request {k}: hit buffer too small")
            if q.ray_offset + 32 * q.count > q.rays.numel() * q.rays.element_size():
                raise ValueError(f"request {k}: rays run past the ray buffer")
            arr[k] = rr.RRxIntersectRequest(accel_ptrs[q.accel.address], self._ptr(q.rays.data_ptr(), q.ray_offset),
                                            self._ptr(q.hits.data_ptr()), q.count, _QUERY[q.query],
                                            _OUTPUT[q.output], 0)
        rel = None
        if release_ns is not None:
            rel = (c_uint64 * max(n, 1))(*[int(x) for x in release_ns])
        first = c_uint64()
        rr.check(self.lib.rrxServerSubmit(self.handle, arr, rel, n, byref(first)), "rrxServerSubmit")
        return first.value

    def submit_host(self, requests, release_ns=None) -> int:
        """Queue host-memory requests (rrxServerSubmitHost); returns the first
        ticket.  The tensors must stay alive until the requests complete."""
        n = len(requests)
        arr = (rr.RRxHostIntersectRequest * max(n, 1))()
        accel_ptrs = {}
        for k, q in enumerate(requests):
            if q.rays.is_cuda or q.hits.is_cuda:
                raise ValueError(f"request {k}: host requests take host tensors")
            count = int(q.rays.shape[0])
            if q.rays.dim() != 2 or int(q.rays.shape[1]) != 8 or q.rays.element_size() != 4 or not q.rays.is_contiguous():
                raise ValueError(f"request {k}: rays must be a contiguous float32 [R, 8] tensor")
            if q.hits.numel() * q.hits.element_size() < count * _HIT_BYTES[q.output] or not q.hits.is_contiguous():
                raise ValueError(f"request {k}: hit buffer too small")
            if q.accel.address not in accel_ptrs:
                accel_ptrs[q.accel.address] = self._ptr(q.accel.address)
            arr[k] = rr.RRxHostIntersectRequest(accel_ptrs[q.accel.address], c_void_p(q.rays.data_ptr()),
                                                c_void_p(q.hits.data_ptr()), count, _QUERY[q.query],
                                                _OUTPUT[q.output], 0)
        rel = None
        if release_ns is not None:
            rel = (c_uint64 * max(n, 1))(*[int(x) for x in release_ns])
        first = c_uint64()
        rr.check(self.lib.rrxServerSubmitHost(self.handle, arr, rel, n, byref(first)), "rrxServerSubmitHost")
        return first.value

    def drain(self) -> None:
        rr.check(self.lib.rrxServerDrain(self.handle), "rrxServerDrain")

    def timings(self, first: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=TIMING_DTYPE)
        buf = (rr.RRxServedTiming * max(count, 1)).from_buffer(out) if count else None
        rr.check(self.lib.rrxServerGetTimings(self.handle, first, count, buf), "rrxServerGetTimings")
        return out

    def peak_admitted(self) -> int:
        v = c_uint64()
        rr.check(self.lib.rrxServerGetPeakAdmitted(self.handle, byref(v)), "rrxServerGetPeakAdmitted")
        return v.value

    def close(self) -> None:
        if self.handle:
            self.lib.rrxDestroyServer(self.handle)
            self.handle = None
            for p in self._ptrs:
                self.lib.rrReleaseDevicePtr(self.ctx.ctx, p)
            self._ptrs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def latency_summary(t: np.ndarray, ray_counts) -> dict:
    """p50/p99/mean latency (ms) and throughput of one served run."""
    lat = (t["complete_ns"].astype(np.int64) - t["arrival_ns"].astype(np.int64)) / 1e6
    span = (int(t["complete_ns"].max()) - int(t["arrival_ns"].min())) / 1e9
    rays = int(np.sum(ray_counts))
    queue = (t["dispatch_ns"].astype(np.int64) - t["arrival_ns"].astype(np.int64)) / 1e6
    service = (t["complete_ns"].astype(np.int64) - t["dispatch_ns"].astype(np.int64)) / 1e6
    return {"requests": int(t.size), "p50_ms": round(float(np.percentile(lat, 50)), 4),
            "queue_p50_ms": round(float(np.percentile(queue, 50)), 4),
            "service_p50_ms": round(float(np.percentile(service, 50)), 4),
            "p99_ms": round(float(np.percentile(lat, 99)), 4), "mean_ms": round(float(lat.mean()), 4),
            "mrays_s": round(rays / span / 1e6, 2), "span_s": round(span, 4),
            "mean_batch": round(float(t["batch_size"].mean()), 2)}
