"""ctypes binding of the C ABI in include/radeonrays.h, radeonrays_hip.h and
radeonrays_ext.h.

This is the host-side mirror of the reference interface: the same entry
points, argument meaning and RRError codes (src/core/include/radeonrays.h),
callable from Python.  Device memory and streams are plain HIP addresses;
the helpers in radeonrays_sdk_amd.api take them from torch tensors.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

# RADEONRAYS_LIB names another build of the library (the host-sanitizer
# build, libradeonrays_asan.so, in the sanitizer test).
LIB_PATH = Path(os.environ.get("RADEONRAYS_LIB") or Path(__file__).resolve().parent / "libradeonrays.so")

# --- enums (radeonrays.h) ---------------------------------------------------
RR_SUCCESS = 0
RR_ERROR_NOT_IMPLEMENTED = 1
RR_ERROR_INTERNAL = 2
RR_ERROR_OUT_OF_HOST_MEMORY = 3
RR_ERROR_OUT_OF_DEVICE_MEMORY = 4
RR_ERROR_INVALID_API_VERSION = 5
RR_ERROR_INVALID_PARAMETER = 6
RR_ERROR_UNSUPPORTED_API = 7
RR_ERROR_UNSUPPORTED_INTEROP = 8

RR_API_DX = 1
RR_API_VK = 2
RR_API_HIP = 3

RR_LOG_LEVEL_DEBUG, RR_LOG_LEVEL_INFO, RR_LOG_LEVEL_WARN, RR_LOG_LEVEL_ERROR, RR_LOG_LEVEL_OFF = 1, 2, 3, 4, 5

RR_BUILD_OPERATION_BUILD = 1
RR_BUILD_OPERATION_UPDATE = 2

RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD = 1
RR_BUILD_FLAG_BITS_ALLOW_UPDATE = 2

RR_PRIMITIVE_TYPE_TRIANGLE_MESH = 0
RR_PRIMITIVE_TYPE_AABB_LIST = 1

RR_INDEX_TYPE_UINT32 = 0
RR_INDEX_TYPE_UINT16 = 1

RR_INTERSECT_QUERY_CLOSEST = 0
RR_INTERSECT_QUERY_ANY = 1

RR_INTERSECT_QUERY_OUTPUT_FULL_HIT = 0
RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID = 1
RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T = 2

RRX_TUNING_TRACE_MIN_BUSY = 0
RRX_TUNING_TRACE_RAYS_PER_FETCH = 1
RRX_TUNING_TRACE_LDS_STACK = 2
RRX_TUNING_TRACE_PLAIN = 3
RRX_TUNING_SORT_DIGIT_BITS = 4
RRX_TUNING_TREELET_LEAVES = 5
RRX_TUNING_SORT_SPIN_LIMIT = 6
RRX_TUNING_SORT_HYBRID = 7
RRX_TUNING_TRACE_RAYS_PER_LANE = 8
RRX_TUNING_TRACE_2L_MIN_BUSY = 9
RRX_TUNING_TRACE_2L_RAYS_PER_FETCH = 10
RRX_TUNING_TRACE_BATCH_MIN_BUSY = 11

RRX_DEVICE_ERROR_STACK_OVERFLOW = 1 << 0
RRX_DEVICE_ERROR_SORT_LOOKBACK = 1 << 1
RRX_DEVICE_ERROR_INVALID_BUFFER = 1 << 2

RRX_SERVE_BATCHED = 0
RRX_SERVE_PLAIN = 1

RR_API_VERSION = 0 * 1000000 + 1 * 1000 + 1
RR_INVALID = 0xFFFFFFFF

ERROR_NAMES = {
    0: "RR_SUCCESS", 1: "RR_ERROR_NOT_IMPLEMENTED", 2: "RR_ERROR_INTERNAL", 3: "RR_ERROR_OUT_OF_HOST_MEMORY",
    4: "RR_ERROR_OUT_OF_DEVICE_MEMORY", 5: "RR_ERROR_INVALID_API_VERSION", 6: "RR_ERROR_INVALID_PARAMETER",
    7: "RR_ERROR_UNSUPPORTED_API", 8: "RR_ERROR_UNSUPPORTED_INTEROP",
}


class RRError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where} failed: {ERROR_NAMES.get(code, code)}")
        self.code = code


# --- structs (radeonrays.h) --------------------------------------------------
class RRTriangleMeshPrimitive(Structure):
    _fields_ = [("vertices", c_void_p), ("vertex_count", c_uint32), ("vertex_stride", c_uint32),
                ("triangle_indices", c_void_p), ("triangle_count", c_uint32), ("index_type", c_int)]


class RRAABBListPrimitive(Structure):
    _fields_ = [("aabbs", c_void_p), ("aabb_count", c_uint32), ("aabb_stride", c_uint32)]


class RRGeometryBuildInput(Structure):
    _fields_ = [("primitive_type", c_int), ("primitive_count", c_uint32), ("primitives", c_void_p)]


class RRBuildOptions(Structure):
    _fields_ = [("build_flags", c_uint32), ("backend_specific_info", c_void_p)]


class RRInstance(Structure):
    _fields_ = [("geometry", c_void_p), ("transform", (c_float * 4) * 3)]


class RRSceneBuildInput(Structure):
    _fields_ = [("instances", POINTER(RRInstance)), ("instance_count", c_uint32)]


class RRMemoryRequirements(Structure):
    _fields_ = [("temporary_build_buffer_size", c_size_t), ("temporary_update_buffer_size", c_size_t),
                ("result_buffer_size", c_size_t)]


class RRxIntersectRequest(Structure):
    _fields_ = [("scene_buffer", c_void_p), ("rays", c_void_p), ("hits", c_void_p), ("ray_count", c_uint32),
                ("query", c_int), ("query_output", c_int), ("padding", c_uint32)]


class RRxServerOptions(Structure):
    _fields_ = [("mode", c_int), ("max_inflight", c_uint32), ("max_batch_requests", c_uint32),
                ("max_batch_rays", c_uint32), ("memory_budget", c_uint64), ("staging_bytes", c_uint64)]


class RRxHostIntersectRequest(Structure):
    _fields_ = [("scene_buffer", c_void_p), ("rays", c_void_p), ("hits", c_void_p), ("ray_count", c_uint32),
                ("query", c_int), ("query_output", c_int), ("padding", c_uint32)]


class RRxServedTiming(Structure):
    _fields_ = [("arrival_ns", c_uint64), ("dispatch_ns", c_uint64), ("complete_ns", c_uint64),
                ("batch_id", c_uint32), ("batch_size", c_uint32)]


# Every exported entry point with its argument types; tests check that the
# library exports exactly what the headers declare.
_SIGNATURES = {
    "rrCreateContext": [c_uint32, c_int, POINTER(c_void_p)],
    "rrDestroyContext": [c_void_p],
    "rrSetLogLevel": [c_int],
    "rrSetLogFile": [c_char_p],
    "rrCmdBuildGeometry": [c_void_p, c_int, POINTER(RRGeometryBuildInput), POINTER(RRBuildOptions), c_void_p,
                           c_void_p, c_void_p],
    "rrGetGeometryBuildMemoryRequirements": [c_void_p, POINTER(RRGeometryBuildInput), POINTER(RRBuildOptions),
                                             POINTER(RRMemoryRequirements)],
    "rrCmdBuildScene": [c_void_p, POINTER(RRSceneBuildInput), POINTER(RRBuildOptions), c_void_p, c_void_p, c_void_p],
    "rrGetSceneBuildMemoryRequirements": [c_void_p, POINTER(RRSceneBuildInput), POINTER(RRBuildOptions),
                                          POINTER(RRMemoryRequirements)],
    "rrCmdIntersect": [c_void_p, c_void_p, c_int, c_void_p, c_uint32, c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "rrGetTraceMemoryRequirements": [c_void_p, c_uint32, POINTER(c_size_t)],
    "rrAllocateCommandStream": [c_void_p, POINTER(c_void_p)],
    "rrReleaseCommandStream": [c_void_p, c_void_p],
    "rrSumbitCommandStream": [c_void_p, c_void_p, c_void_p, POINTER(c_void_p)],
    "rrReleaseEvent": [c_void_p, c_void_p],
    "rrWaitEvent": [c_void_p, c_void_p],
    "rrReleaseDevicePtr": [c_void_p, c_void_p],
    "rrReleaseExternalCommandStream": [c_void_p, c_void_p],
    # radeonrays_hip.h
    "rrCreateContextHip": [c_uint32, c_int, c_void_p, POINTER(c_void_p)],
    "rrGetDevicePtrFromHipPtr": [c_void_p, c_void_p, c_size_t, POINTER(c_void_p)],
    "rrGetCommandStreamFromHipStream": [c_void_p, c_void_p, POINTER(c_void_p)],
    "rrAllocateDeviceBuffer": [c_void_p, c_size_t, POINTER(c_void_p)],
    "rrMapDevicePtr": [c_void_p, c_void_p, POINTER(c_void_p)],
    "rrUnmapDevicePtr": [c_void_p, c_void_p, POINTER(c_void_p)],
    "rrGetHipAddress": [c_void_p],
    "rrGetHipStream": [c_void_p],
    # radeonrays_ext.h
    "rrxGetSortMemoryRequirements": [c_void_p, c_uint32, POINTER(c_size_t)],
    "rrxCmdSortPairs": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, c_void_p],
    "rrxCmdSortPairsNoAlloc": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p],
    "rrxCmdIntersectWithStats": [c_void_p, c_void_p, c_int, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p],
    "rrxCmdUpdateScene": [c_void_p, POINTER(RRSceneBuildInput), c_void_p, c_void_p, c_void_p],
    "rrxSetTuning": [c_int, c_uint32],
    "rrxGetTuning": [c_int, POINTER(c_uint32)],
    "rrxGetDeviceErrors": [c_void_p, POINTER(c_uint32)],
    "rrxGetGeometryBatchBuildMemoryRequirements": [c_void_p, c_void_p, c_uint32, POINTER(RRBuildOptions),
                                                   POINTER(c_size_t)],
    "rrxCmdBuildGeometryBatch": [c_void_p, c_int, c_void_p, c_uint32, POINTER(RRBuildOptions), c_void_p, c_void_p,
                                 c_void_p],
    "rrxWriteAnalyzerDump": [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_char_p, c_char_p, c_char_p],
    "rrxGetIntersectBatchMemoryRequirements": [c_void_p, c_uint32, c_uint64, c_uint32, POINTER(c_size_t)],
    "rrxCmdIntersectBatch": [c_void_p, POINTER(RRxIntersectRequest), c_uint32, c_void_p, c_void_p],
    "rrxCreateServer": [c_void_p, POINTER(RRxServerOptions), POINTER(c_void_p)],
    "rrxDestroyServer": [c_void_p],
    "rrxServerNow": [c_void_p],
    "rrxServerSubmit": [c_void_p, POINTER(RRxIntersectRequest), POINTER(c_uint64), c_uint32, POINTER(c_uint64)],
    "rrxServerSubmitHost": [c_void_p, POINTER(RRxHostIntersectRequest), POINTER(c_uint64), c_uint32,
                            POINTER(c_uint64)],
    "rrxServerDrain": [c_void_p],
    "rrxServerGetTimings": [c_void_p, c_uint64, c_uint32, POINTER(RRxServedTiming)],
    "rrxServerGetPeakAdmitted": [c_void_p, POINTER(c_uint64)],
}
_RESTYPES = {"rrGetHipAddress": c_void_p, "rrGetHipStream": c_void_p, "rrxServerNow": c_uint64}

_lib = None


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libradeonrays.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(str(p))
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    if path is None:
        _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != RR_SUCCESS:
        raise RRError(code, where)
