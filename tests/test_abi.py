"""CPU checks of the C ABI: the library loads, exports exactly what the
headers declare, and rejects bad arguments before touching the GPU (the
reference's shim checks pointers first, radeonrays.cpp:252-256,424-428)."""
import ctypes
import re
from ctypes import byref, c_size_t, c_void_p
from pathlib import Path

import pytest

from radeonrays_sdk_amd import rr

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("radeonrays*.h"))
# A context handle the library never handed out: size queries accept it (they
# read nothing from the context), every other entry point must reject it
# without dereferencing it.
FAKE_CTX = c_void_p(0x1000)


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        names |= set(re.findall(r"RR_API\s+[\w\s\*]+?\b(rrx?\w+)\s*\(", text))
    return names


def test_headers_declare_the_reference_entry_points():
    names = declared_symbols()
    reference_abi = {  # radeonrays.h:276-473
        "rrCreateContext", "rrDestroyContext", "rrSetLogLevel", "rrSetLogFile", "rrCmdBuildGeometry",
        "rrGetGeometryBuildMemoryRequirements", "rrCmdBuildScene", "rrGetSceneBuildMemoryRequirements",
        "rrCmdIntersect", "rrGetTraceMemoryRequirements", "rrAllocateCommandStream", "rrReleaseCommandStream",
        "rrSumbitCommandStream", "rrReleaseEvent", "rrWaitEvent", "rrReleaseDevicePtr",
        "rrReleaseExternalCommandStream"}
    assert reference_abi <= names


def test_library_exports_every_declared_symbol():
    lib = rr.load()
    names = declared_symbols()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(rr._SIGNATURES) == names  # the binding covers exactly the headers


def test_struct_layouts_match_the_abi():
    # Sizes as compiled from include/radeonrays.h by gcc on x86-64.
    assert ctypes.sizeof(rr.RRTriangleMeshPrimitive) == 32
    assert ctypes.sizeof(rr.RRGeometryBuildInput) == 16
    assert ctypes.sizeof(rr.RRBuildOptions) == 16
    assert ctypes.sizeof(rr.RRInstance) == 56
    assert ctypes.sizeof(rr.RRSceneBuildInput) == 16
    assert ctypes.sizeof(rr.RRMemoryRequirements) == 24
    # include/radeonrays_ext.h
    assert ctypes.sizeof(rr.RRxIntersectRequest) == 40
    assert ctypes.sizeof(rr.RRxServerOptions) == 32
    assert ctypes.sizeof(rr.RRxHostIntersectRequest) == 40
    assert ctypes.sizeof(rr.RRxServedTiming) == 32


def _mesh_input(tris, ptype=rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, count=1):
    prim = rr.RRTriangleMeshPrimitive(None, 3 * tris, 12, None, tris, rr.RR_INDEX_TYPE_UINT32)
    inp = rr.RRGeometryBuildInput(ptype, count, ctypes.cast(ctypes.pointer(prim), c_void_p))
    return prim, inp


def test_null_arguments_are_invalid_parameter():
    lib = rr.load()
    prim, inp = _mesh_input(10)
    req = rr.RRMemoryRequirements()
    E = rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrCreateContext(rr.RR_API_VERSION, rr.RR_API_HIP, None) == E
    assert lib.rrDestroyContext(None) == E
    assert lib.rrGetGeometryBuildMemoryRequirements(None, byref(inp), None, byref(req)) == E
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, None, None, byref(req)) == E
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(inp), None, None) == E
    assert lib.rrCmdBuildGeometry(FAKE_CTX, rr.RR_BUILD_OPERATION_BUILD, None, None, None, None, c_void_p(1)) == E
    assert lib.rrCmdBuildGeometry(FAKE_CTX, rr.RR_BUILD_OPERATION_BUILD, byref(inp), None, None, None, None) == E
    assert lib.rrCmdBuildScene(FAKE_CTX, None, None, None, None, c_void_p(1)) == E
    assert lib.rrCmdIntersect(FAKE_CTX, None, 0, None, 1, None, 0, None, None, None) == E
    size = c_size_t()
    # radeonrays.cpp:741: zero rays is an invalid parameter too.
    assert lib.rrGetTraceMemoryRequirements(FAKE_CTX, 0, byref(size)) == E
    assert lib.rrGetTraceMemoryRequirements(FAKE_CTX, 1, None) == E
    assert lib.rrAllocateCommandStream(None, None) == E
    assert lib.rrSumbitCommandStream(FAKE_CTX, None, None, None) == E
    assert lib.rrReleaseEvent(FAKE_CTX, None) == E
    assert lib.rrWaitEvent(FAKE_CTX, None) == E
    assert lib.rrReleaseDevicePtr(FAKE_CTX, None) == E
    assert lib.rrReleaseExternalCommandStream(FAKE_CTX, None) == E
    assert lib.rrGetDevicePtrFromHipPtr(FAKE_CTX, None, 0, None) == E
    assert lib.rrSetLogFile(None) == E
    # Extensions.
    assert lib.rrxCmdIntersectBatch(FAKE_CTX, None, 1, c_void_p(1), c_void_p(1)) == E
    assert lib.rrxCmdIntersectBatch(FAKE_CTX, None, 0, None, c_void_p(1)) == E
    bad = (rr.RRxIntersectRequest * 1)(rr.RRxIntersectRequest(None, None, None, 5, 0, 0, 0))
    assert lib.rrxCmdIntersectBatch(FAKE_CTX, bad, 1, c_void_p(1), c_void_p(1)) == E
    bad[0] = rr.RRxIntersectRequest(c_void_p(1), c_void_p(1), c_void_p(1), 5, 7, 0, 0)  # unknown query
    assert lib.rrxCmdIntersectBatch(FAKE_CTX, bad, 1, c_void_p(1), c_void_p(1)) == E
    assert lib.rrxGetIntersectBatchMemoryRequirements(FAKE_CTX, 1, 1, 1, None) == E
    opts = rr.RRxServerOptions(rr.RRX_SERVE_BATCHED, 0, 16, 1 << 20, 1 << 30)  # no streams
    h = c_void_p()
    assert lib.rrxCreateServer(FAKE_CTX, byref(opts), byref(h)) == E
    opts = rr.RRxServerOptions(5, 1, 16, 1 << 20, 1 << 30)  # unknown mode
    assert lib.rrxCreateServer(FAKE_CTX, byref(opts), byref(h)) == E
    assert lib.rrxCreateServer(None, byref(opts), byref(h)) == E
    assert lib.rrxDestroyServer(None) == E
    assert lib.rrxServerDrain(None) == E
    assert lib.rrxServerSubmit(None, None, None, 0, None) == E
    assert lib.rrxServerGetTimings(None, 0, 0, None) == E
    assert lib.rrxServerGetPeakAdmitted(None, None) == E
    assert lib.rrxServerNow(None) == 0
    assert lib.rrxWriteAnalyzerDump(FAKE_CTX, None, None, 0, 0, b"a", b"b", b"c") == E
    assert lib.rrxWriteAnalyzerDump(FAKE_CTX, c_void_p(1), None, 0, 0, None, b"b", b"c") == E


def test_tuning_validation():
    lib = rr.load()
    from ctypes import c_uint32
    v = c_uint32()
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_FETCH, byref(v)) == rr.RR_SUCCESS and v.value == 128
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_LDS_STACK, byref(v)) == rr.RR_SUCCESS and v.value == 16
    for key, bad in ((rr.RRX_TUNING_TRACE_RAYS_PER_FETCH, 48), (rr.RRX_TUNING_TRACE_LDS_STACK, 12),
                     (rr.RRX_TUNING_TRACE_MIN_BUSY, 0), (rr.RRX_TUNING_TRACE_MIN_BUSY, 65),
                     (rr.RRX_TUNING_SORT_HYBRID, 2), (rr.RRX_TUNING_TRACE_RAYS_PER_LANE, 65),
                     (rr.RRX_TUNING_TRACE_2L_MIN_BUSY, 0), (rr.RRX_TUNING_TRACE_2L_MIN_BUSY, 65),
                     (rr.RRX_TUNING_TRACE_2L_RAYS_PER_FETCH, 48),
                     (rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, 0), (rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, 65), (12, 1)):
        assert lib.rrxSetTuning(key, bad) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_LDS_STACK, None) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_LDS_STACK, 32) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_LDS_STACK, byref(v)) == rr.RR_SUCCESS and v.value == 32
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_LDS_STACK, 16) == rr.RR_SUCCESS
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_PLAIN, 2) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_PLAIN, byref(v)) == rr.RR_SUCCESS
    before = v.value
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_PLAIN, 1) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_PLAIN, byref(v)) == rr.RR_SUCCESS and v.value == 1
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_PLAIN, before) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_HYBRID, byref(v)) == rr.RR_SUCCESS and v.value == 0
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_HYBRID, 1) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_HYBRID, byref(v)) == rr.RR_SUCCESS and v.value == 1
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_HYBRID, 0) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_LANE, byref(v)) == rr.RR_SUCCESS and v.value == 2
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_LANE, 0) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_LANE, byref(v)) == rr.RR_SUCCESS and v.value == 0
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_LANE, 2) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_2L_MIN_BUSY, byref(v)) == rr.RR_SUCCESS and v.value == 32
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_2L_RAYS_PER_FETCH, byref(v)) == rr.RR_SUCCESS and v.value == 32
    # The single-level knobs leave the two-level ones alone.
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_MIN_BUSY, 40) == rr.RR_SUCCESS
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_FETCH, 32) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_2L_MIN_BUSY, byref(v)) == rr.RR_SUCCESS and v.value == 32
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_2L_RAYS_PER_FETCH, byref(v)) == rr.RR_SUCCESS and v.value == 32
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_MIN_BUSY, 24) == rr.RR_SUCCESS
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_RAYS_PER_FETCH, 128) == rr.RR_SUCCESS
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_2L_MIN_BUSY, 16) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_2L_MIN_BUSY, byref(v)) == rr.RR_SUCCESS and v.value == 16
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_2L_MIN_BUSY, 32) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, byref(v)) == rr.RR_SUCCESS and v.value == 24
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, 40) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TRACE_MIN_BUSY, byref(v)) == rr.RR_SUCCESS and v.value == 24
    assert lib.rrxSetTuning(rr.RRX_TUNING_TRACE_BATCH_MIN_BUSY, 24) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, byref(v)) == rr.RR_SUCCESS and v.value == 8
    for bad in (0, 6, 9, 16):
        assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, bad) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, 4) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, byref(v)) == rr.RR_SUCCESS and v.value == 4
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, 10) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, byref(v)) == rr.RR_SUCCESS and v.value == 10
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_DIGIT_BITS, 8) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TREELET_LEAVES, byref(v)) == rr.RR_SUCCESS and v.value == 7
    for bad in (6, 8, 11):
        assert lib.rrxSetTuning(rr.RRX_TUNING_TREELET_LEAVES, bad) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrxSetTuning(rr.RRX_TUNING_TREELET_LEAVES, 9) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_TREELET_LEAVES, byref(v)) == rr.RR_SUCCESS and v.value == 9
    assert lib.rrxSetTuning(rr.RRX_TUNING_TREELET_LEAVES, 7) == rr.RR_SUCCESS


def test_aabb_list_requirements_and_unknown_primitive_type():
    # AABB lists are built (the reference answers RR_ERROR_NOT_IMPLEMENTED,
    # radeonrays.cpp:286-289); an unknown primitive type still is not.
    lib = rr.load()
    boxes = rr.RRAABBListPrimitive(None, 10, 24)
    inp = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_AABB_LIST, 1, ctypes.cast(ctypes.pointer(boxes), c_void_p))
    req, tri = rr.RRMemoryRequirements(), rr.RRMemoryRequirements()
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(inp), None, byref(req)) == rr.RR_SUCCESS
    p10, i10 = _mesh_input(10)
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(i10), None, byref(tri)) == rr.RR_SUCCESS
    assert req.result_buffer_size == tri.result_buffer_size
    assert req.temporary_update_buffer_size >= 10 * 36
    prim, bad = _mesh_input(10, ptype=7)
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(bad), None, byref(req)) == rr.RR_ERROR_NOT_IMPLEMENTED


def test_log_level_validation():
    lib = rr.load()
    assert lib.rrSetLogLevel(0) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrSetLogLevel(9) == rr.RR_ERROR_INVALID_PARAMETER
    assert lib.rrSetLogLevel(rr.RR_LOG_LEVEL_WARN) == rr.RR_SUCCESS


def test_vulkan_and_dx_contexts_are_unsupported():
    lib = rr.load()
    ctx = c_void_p()
    assert lib.rrCreateContext(rr.RR_API_VERSION, rr.RR_API_VK, byref(ctx)) == rr.RR_ERROR_UNSUPPORTED_API
    assert lib.rrCreateContext(rr.RR_API_VERSION, rr.RR_API_DX, byref(ctx)) == rr.RR_ERROR_UNSUPPORTED_API


@pytest.mark.parametrize("tris", [0, 1, 2, 1000, 1045458])
def test_memory_requirements(tris):
    lib = rr.load()
    for flags, restructure in ((rr.RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD, False), (0, True)):
        prim, inp = _mesh_input(tris)
        opts = rr.RRBuildOptions(flags, None)
        req = rr.RRMemoryRequirements()
        assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(inp), byref(opts), byref(req)) == rr.RR_SUCCESS
        nodes = 2 * tris - 1 if tris else 0
        # 128-byte header + (2N-1) 64-byte nodes (hlbvh_builder.cpp:376-377), 256-byte aligned.
        assert req.result_buffer_size == 128 + (nodes * 64 + 255) // 256 * 256
        assert req.temporary_update_buffer_size == 0
        assert req.temporary_build_buffer_size >= 16 * tris
    for inst in (1, 37, 4096):
        scene = rr.RRSceneBuildInput(None, inst)
        req = rr.RRMemoryRequirements()
        assert lib.rrGetSceneBuildMemoryRequirements(FAKE_CTX, byref(scene), None, byref(req)) == rr.RR_SUCCESS
        assert req.result_buffer_size >= 128 + (2 * inst - 1) * 64 + 2 * inst * 48 + inst * 8
        # A scene refit stages 64 bytes per instance and a 32-byte world box
        # per instance after a 256-byte head.
        assert req.temporary_update_buffer_size >= 96 * inst + 256
    size = c_size_t()
    assert lib.rrGetTraceMemoryRequirements(FAKE_CTX, 1000, byref(size)) == rr.RR_SUCCESS
    assert size.value >= 1000 * 256


def test_multiple_meshes_per_build_input():
    # Two 10-triangle meshes build one 20-triangle BLAS (the reference asserts
    # a single mesh, intersector.cpp:110,146); refit needs scratch for them.
    lib = rr.load()
    prims = (rr.RRTriangleMeshPrimitive * 2)(*[rr.RRTriangleMeshPrimitive(None, 30, 12, None, 10, rr.RR_INDEX_TYPE_UINT32)
                                               for _ in range(2)])
    inp = rr.RRGeometryBuildInput(rr.RR_PRIMITIVE_TYPE_TRIANGLE_MESH, 2, ctypes.cast(prims, c_void_p))
    req, one = rr.RRMemoryRequirements(), rr.RRMemoryRequirements()
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(inp), None, byref(req)) == rr.RR_SUCCESS
    p20, i20 = _mesh_input(20)
    assert lib.rrGetGeometryBuildMemoryRequirements(FAKE_CTX, byref(i20), None, byref(one)) == rr.RR_SUCCESS
    assert req.result_buffer_size == one.result_buffer_size
    assert req.temporary_build_buffer_size > one.temporary_build_buffer_size
    assert req.temporary_update_buffer_size >= 20 * 36


def test_foreign_handles_are_rejected_without_dereference():
    """Handles the library never handed out (or already released) give
    RR_ERROR_INVALID_PARAMETER: none of these calls may touch the memory the
    handle points at (1 and 0x1000 are unmapped addresses)."""
    lib = rr.load()
    E = rr.RR_ERROR_INVALID_PARAMETER
    bogus = c_void_p(1)
    prim, inp = _mesh_input(10)
    assert lib.rrDestroyContext(FAKE_CTX) == E
    assert lib.rrCmdBuildGeometry(FAKE_CTX, rr.RR_BUILD_OPERATION_BUILD, byref(inp), None, bogus, bogus, bogus) == E
    prim, bad = _mesh_input(10, ptype=7)
    assert lib.rrCmdBuildGeometry(FAKE_CTX, rr.RR_BUILD_OPERATION_BUILD, byref(bad), None, bogus, bogus, bogus) == E
    scene = rr.RRSceneBuildInput(None, 0)
    assert lib.rrCmdBuildScene(FAKE_CTX, byref(scene), None, bogus, bogus, bogus) == E
    assert lib.rrCmdIntersect(FAKE_CTX, bogus, 0, bogus, 1, None, 0, bogus, bogus, bogus) == E
    out = c_void_p()
    assert lib.rrAllocateCommandStream(FAKE_CTX, byref(out)) == E
    assert lib.rrReleaseCommandStream(FAKE_CTX, bogus) == E
    assert lib.rrSumbitCommandStream(FAKE_CTX, bogus, None, byref(out)) == E
    assert lib.rrReleaseEvent(FAKE_CTX, bogus) == E
    assert lib.rrWaitEvent(FAKE_CTX, bogus) == E
    assert lib.rrReleaseDevicePtr(FAKE_CTX, bogus) == E
    assert lib.rrReleaseExternalCommandStream(FAKE_CTX, bogus) == E
    assert lib.rrGetDevicePtrFromHipPtr(FAKE_CTX, bogus, 0, byref(out)) == E
    assert lib.rrGetCommandStreamFromHipStream(FAKE_CTX, None, byref(out)) == E
    assert lib.rrAllocateDeviceBuffer(FAKE_CTX, 64, byref(out)) == E
    assert lib.rrMapDevicePtr(FAKE_CTX, bogus, byref(out)) == E
    assert lib.rrUnmapDevicePtr(FAKE_CTX, bogus, byref(out)) == E
    assert lib.rrGetHipAddress(bogus) is None
    assert lib.rrGetHipStream(bogus) is None
    assert lib.rrxCmdSortPairs(FAKE_CTX, bogus, bogus, bogus, bogus, 10, 32, bogus, bogus) == E
    assert lib.rrxCmdIntersectWithStats(FAKE_CTX, bogus, 0, bogus, 1, bogus, bogus, bogus, bogus) == E
    assert lib.rrxCmdUpdateScene(FAKE_CTX, byref(scene), bogus, bogus, bogus) == E
    assert lib.rrxCmdBuildGeometryBatch(FAKE_CTX, rr.RR_BUILD_OPERATION_BUILD, None, 0, None, bogus, None, bogus) == E
    assert lib.rrxWriteAnalyzerDump(FAKE_CTX, bogus, None, 0, 0, b"a", b"b", b"c") == E
    assert lib.rrxGetIntersectBatchMemoryRequirements(FAKE_CTX, 1, 1, 1, byref(c_size_t())) == E
    from ctypes import c_uint32
    assert lib.rrxGetDeviceErrors(FAKE_CTX, byref(c_uint32())) == E
    opts = rr.RRxServerOptions(rr.RRX_SERVE_BATCHED, 1, 16, 1 << 20, 1 << 30)
    assert lib.rrxCreateServer(FAKE_CTX, byref(opts), byref(out)) == E
    assert lib.rrxDestroyServer(bogus) == E
    assert lib.rrxServerDrain(bogus) == E
    assert lib.rrxServerNow(bogus) == 0
    assert lib.rrxServerGetPeakAdmitted(bogus, byref(ctypes.c_uint64())) == E


def test_sort_spin_limit_tuning_round_trip():
    lib = rr.load()
    from ctypes import c_uint32
    v = c_uint32()
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_SPIN_LIMIT, byref(v)) == rr.RR_SUCCESS and v.value == 1 << 22
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_SPIN_LIMIT, 0) == rr.RR_SUCCESS
    assert lib.rrxGetTuning(rr.RRX_TUNING_SORT_SPIN_LIMIT, byref(v)) == rr.RR_SUCCESS and v.value == 0
    assert lib.rrxSetTuning(rr.RRX_TUNING_SORT_SPIN_LIMIT, 1 << 22) == rr.RR_SUCCESS


def test_only_the_c_abi_is_exported():
    """The shared library exports the rr*/rrx* entry points and nothing else
    (no C++ internals, kernel stubs or template instantiations)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(rr.LIB_PATH)], capture_output=True, text=True, check=True)
    names = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    assert names == declared_symbols(), sorted(names ^ declared_symbols())
