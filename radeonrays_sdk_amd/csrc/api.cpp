// api.cpp — the C ABI (radeonrays.h, radeonrays_hip.h, radeonrays_ext.h).
//
// Mirrors the reference shim src/core/src/radeonrays.cpp: every entry point
// validates its pointers first (RR_ERROR_INVALID_PARAMETER), unsupported
// primitive types give RR_ERROR_NOT_IMPLEMENTED, and no C++ exception
// crosses the ABI (RR_ERROR_INTERNAL).  Behind it there is one backend: HIP
// on the current device.
//
// Command streams: a stream from rrAllocateCommandStream records closures
// that enqueue kernels; rrSumbitCommandStream runs them, in order, on the
// context's HIP stream and returns an event (the reference records into a
// Vulkan command buffer and submits it, vlk/device.cpp:221-240).  A stream
// wrapped from a client hipStream_t enqueues immediately on that stream.
//
// Handles: every context, device pointer, command stream and event the
// library hands out is kept in a registry until it is released, and each
// entry point checks the handles it is given against it, so a stale or
// foreign handle gives RR_ERROR_INVALID_PARAMETER instead of being
// dereferenced.  No entry point that records work reads device memory: the
// level of a buffer the host has not seen built is decided on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/radeonrays.h"
#include "../../include/radeonrays_ext.h"
#include "../../include/radeonrays_hip.h"
#include "api_internal.h"
#include "batch.h"
#include "build.h"
#include "device_errors.h"
#include "rr_layout.h"
#include "sort.h"
#include "trace.h"

namespace {

// ---------------------------------------------------------------------------
// Logging (the reference's spdlog singleton, utils/logger.h:33-79, reduced to
// a level and a sink).  Default level WARN; RR_LOG_LEVEL=1..5 overrides.
struct Logger {
    std::mutex mu;
    int level = RR_LOG_LEVEL_WARN;
    FILE* sink = stderr;
    Logger() {
        if (const char* e = std::getenv("RR_LOG_LEVEL")) {
            const int v = std::atoi(e);
            if (v >= RR_LOG_LEVEL_DEBUG && v <= RR_LOG_LEVEL_OFF) level = v;
        }
    }
    void Log(int lvl, const char* fmt, ...) {
        if (lvl < level || level == RR_LOG_LEVEL_OFF) return;
        static const char* names[] = {"", "debug", "info", "warn", "error", "off"};
        std::lock_guard<std::mutex> g(mu);
        std::fprintf(sink, "[radeonrays] [%s] ", names[lvl]);
        va_list ap;
        va_start(ap, fmt);
        std::vfprintf(sink, fmt, ap);
        va_end(ap);
        std::fputc('\n', sink);
        std::fflush(sink);
    }
};
Logger& Log() {
    static Logger l;
    return l;
}
#define RR_LOG(lvl, ...) Log().Log(lvl, __VA_ARGS__)

// A handle that is not live (never handed out, or already released).
struct BadHandle : std::runtime_error {
    BadHandle() : std::runtime_error("handle is not live") {}
};

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};
void Check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(e, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

// ---------------------------------------------------------------------------
// Handles.

// Live handles of one kind.
class HandleSet {
public:
    void Add(const void* p) {
        std::lock_guard<std::mutex> g(mu_);
        set_.insert(p);
    }
    bool Remove(const void* p) {
        std::lock_guard<std::mutex> g(mu_);
        return set_.erase(p) != 0;
    }
    bool Has(const void* p) const {
        if (!p) return false;
        std::lock_guard<std::mutex> g(mu_);
        return set_.count(p) != 0;
    }

private:
    mutable std::mutex mu_;
    std::unordered_set<const void*> set_;
};
HandleSet& LiveContexts() { static HandleSet s; return s; }
HandleSet& LivePtrs() { static HandleSet s; return s; }
HandleSet& LiveStreams() { static HandleSet s; return s; }
HandleSet& LiveEvents() { static HandleSet s; return s; }
bool LiveCtx(RRContext c) { return LiveContexts().Has(c); }
bool LivePtr(RRDevicePtr p) { return LivePtrs().Has(p); }
bool OptPtr(RRDevicePtr p) { return !p || LivePtrs().Has(p); }  // null allowed
bool LiveStream(RRCommandStream s) { return LiveStreams().Has(s); }
bool LiveEvent(RREvent e) { return LiveEvents().Has(e); }

struct DevicePtr {
    void*  base = nullptr;
    size_t offset = 0;
    bool   owned = false;   // from rrAllocateDeviceBuffer
    size_t size = 0;
    void*  mapped = nullptr;
    void*  addr() const { return static_cast<uint8_t*>(base) + offset; }
};

using Command = std::function<void(hipStream_t)>;

struct Event {
    hipEvent_t ev = nullptr;
    bool pooled = false;  // released by the client, kept for reuse
};

struct Context;

struct CommandStream {
    Context* ctx = nullptr;
    bool external = false;
    hipStream_t stream = nullptr;      // external streams only
    std::vector<Command> commands;     // internal streams only
    std::vector<void*> staging;        // pinned uploads owned by recorded commands
};

struct PendingFree {
    hipEvent_t ev;
    void* host;
};

// What the context knows of a result buffer it recorded a build into.
struct AccelInfo {
    rr::AccelKind kind;
    uint32_t prims;  // triangles or instances
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    std::mutex mu;
    std::unordered_map<uintptr_t, AccelInfo> kinds;  // result buffer -> kind and size
    std::vector<PendingFree> pending;                      // pinned staging to free once done
    std::vector<Event*> event_pool;
    // Pinned staging blocks kept for reuse (hipHostMalloc costs far more than
    // the small uploads it serves), at most kPinnedPoolBytes in all.
    static constexpr size_t kPinnedPoolBytes = size_t(64) << 20;
    std::unordered_map<void*, size_t> pinned_size;
    std::vector<void*> pinned_free;
    size_t pinned_free_bytes = 0;
    // Staging read by commands captured into a graph: every replay reads it
    // again, so it lives until the context is destroyed.  Captured commands
    // take it from an arena reserved by the first uncaptured staging
    // request, so that recording under capture makes no HIP call that could
    // invalidate the capture; only when the arena is full is a block
    // allocated (and kept in graph_staging).
    static constexpr size_t kGraphArenaBytes = size_t(16) << 20;
    std::vector<void*> graph_staging;
    uint8_t* graph_arena = nullptr;
    size_t graph_arena_used = 0;

    bool InArena(const void* p) const {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        return graph_arena && b >= graph_arena && b < graph_arena + kGraphArenaBytes;
    }

    // Pinned host block of at least `bytes` (caller holds mu).
    void* AcquirePinned(size_t bytes) {
        Reap(false);
        for (size_t k = 0; k < pinned_free.size(); ++k) {
            void* p = pinned_free[k];
            const size_t sz = pinned_size[p];
            if (sz >= bytes) {
                pinned_free.erase(pinned_free.begin() + k);
                pinned_free_bytes -= sz;
                return p;
            }
        }
        const size_t sz = std::max<size_t>(bytes, 4096);
        void* p = nullptr;
        Check(hipHostMalloc(&p, sz, hipHostMallocDefault), "hipHostMalloc");
        pinned_size[p] = sz;
        return p;
    }
    void ReleasePinned(void* p) {
        auto it = pinned_size.find(p);
        if (it == pinned_size.end() || pinned_free_bytes + it->second > kPinnedPoolBytes) {
            if (it != pinned_size.end()) pinned_size.erase(it);
            (void)hipHostFree(p);
            return;
        }
        pinned_free.push_back(p);
        pinned_free_bytes += it->second;
    }
    void FreePool() {
        for (void* p : pinned_free) (void)hipHostFree(p);
        pinned_free.clear();
        pinned_free_bytes = 0;
        for (void* p : graph_staging) (void)hipHostFree(p);
        graph_staging.clear();
        if (graph_arena) (void)hipHostFree(graph_arena);
        graph_arena = nullptr;
    }

    void Reap(bool wait_all) {
        std::vector<PendingFree> keep;
        for (auto& p : pending) {
            const hipError_t q = wait_all ? hipEventSynchronize(p.ev) : hipEventQuery(p.ev);
            if (q == hipSuccess) {
                ReleasePinned(p.host);
                (void)hipEventDestroy(p.ev);
            } else {
                keep.push_back(p);
            }
        }
        pending.swap(keep);
    }
    // Free `host` once everything enqueued on `s` so far has finished.
    void FreeAfter(hipStream_t s, void* host) {
        hipEvent_t ev;
        Check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        Check(hipEventRecord(ev, s), "hipEventRecord");
        pending.push_back({ev, host});
    }
};

Context* AsCtx(RRContext c) { return reinterpret_cast<Context*>(c); }
DevicePtr* AsPtr(RRDevicePtr p) { return reinterpret_cast<DevicePtr*>(p); }
CommandStream* AsStream(RRCommandStream s) { return reinterpret_cast<CommandStream*>(s); }
Event* AsEvent(RREvent e) { return reinterpret_cast<Event*>(e); }
// Device address behind a handle (null stays null); throws BadHandle for a
// handle that is not live, which Guarded turns into RR_ERROR_INVALID_PARAMETER.
void* Addr(RRDevicePtr p) {
    if (!p) return nullptr;
    if (!LivePtr(p)) throw BadHandle();
    return AsPtr(p)->addr();
}

void RecordAccel(Context* c, void* result, rr::AccelKind kind, uint32_t prims) {
    std::lock_guard<std::mutex> lk(c->mu);
    c->kinds[reinterpret_cast<uintptr_t>(result)] = AccelInfo{kind, prims};
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Enqueue now (external stream) or record for submission (internal stream).
void Record(CommandStream* cs, Command cmd) {
    if (cs->external) {
        cmd(cs->stream);
        Check(hipGetLastError(), "kernel launch");
    } else {
        cs->commands.push_back(std::move(cmd));
    }
}

bool IsCapturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Pinned staging for a command about to be recorded on `cs`.
void* AcquireStaging(Context* c, CommandStream* cs, size_t bytes) {
    std::lock_guard<std::mutex> lk(c->mu);
    if (cs->external && IsCapturing(cs->stream)) {
        const size_t need = rr::AlignUp(bytes, 256);
        if (c->graph_arena && c->graph_arena_used + need <= Context::kGraphArenaBytes) {
            void* p = c->graph_arena + c->graph_arena_used;
            c->graph_arena_used += need;
            return p;
        }
        void* p = nullptr;
        Check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
        c->graph_staging.push_back(p);
        return p;
    }
    if (!c->graph_arena) Check(hipHostMalloc(reinterpret_cast<void**>(&c->graph_arena), Context::kGraphArenaBytes,
                                             hipHostMallocDefault), "hipHostMalloc");
    return c->AcquirePinned(bytes);
}

// Hand pinned staging read by a just-recorded command to its owner: freed
// once the work has run (external stream), or with the command stream's
// submission (internal stream); kept for the context's life when the stream
// is being captured into a graph.
void KeepStaging(Context* c, CommandStream* cs, void* host) {
    if (!host) return;
    if (!cs->external) {
        cs->staging.push_back(host);
        return;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    // Captured staging is already owned by the arena or graph_staging.
    if (c->InArena(host) || IsCapturing(cs->stream)) return;
    c->FreeAfter(cs->stream, host);
}

template <typename F>
RRError Guarded(const char* name, F&& f) {
    try {
        return f();
    } catch (const BadHandle&) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "%s: invalid parameter (a device pointer handle is not live)", name);
        return RR_ERROR_INVALID_PARAMETER;
    } catch (const HipError& e) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "%s: %s", name, e.what());
        return (e.code == hipErrorOutOfMemory) ? RR_ERROR_OUT_OF_DEVICE_MEMORY : RR_ERROR_INTERNAL;
    } catch (const std::bad_alloc&) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "%s: out of host memory", name);
        return RR_ERROR_OUT_OF_HOST_MEMORY;
    } catch (const std::exception& e) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "%s: %s", name, e.what());
        return RR_ERROR_INTERNAL;
    } catch (...) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "%s: unknown exception", name);
        return RR_ERROR_INTERNAL;
    }
}

#define RR_REQUIRE(cond, name)                                                 \
    do {                                                                       \
        if (!(cond)) {                                                         \
            RR_LOG(RR_LOG_LEVEL_ERROR, "%s: invalid parameter (%s)", name, #cond); \
            return RR_ERROR_INVALID_PARAMETER;                                 \
        }                                                                      \
    } while (0)

// Mesh description from a build input (GetTriangleMeshBuildInfo, radeonrays.cpp:59-74).
// With resolve == false only the counts are read (memory requirements never
// look at the device pointers, as in the reference).
RRError MeshFromInput(const RRGeometryBuildInput* in, rr::MeshDesc& m, const char* name, bool resolve) {
    if (rr::Raw(in->primitive_type) != RR_PRIMITIVE_TYPE_TRIANGLE_MESH) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "%s: build input type not supported", name);
        return RR_ERROR_NOT_IMPLEMENTED;
    }
    RR_REQUIRE(in->primitive_count == 1 && in->triangle_mesh_primitives, name);
    const RRTriangleMeshPrimitive& p = in->triangle_mesh_primitives[0];
    RR_REQUIRE(rr::Raw(p.index_type) == RR_INDEX_TYPE_UINT32 || rr::Raw(p.index_type) == RR_INDEX_TYPE_UINT16, name);
    RR_REQUIRE(p.triangle_count < (1u << 30), name);
    m.tri_count = p.triangle_count;
    if (!resolve) return RR_SUCCESS;
    if (p.triangle_count > 0) {
        RR_REQUIRE(p.vertices && p.triangle_indices && p.vertex_count > 0, name);
        RR_REQUIRE(p.vertex_stride >= 12 && p.vertex_stride % 4 == 0, name);
    }
    m.vertices = Addr(p.vertices);
    m.vertex_count = p.vertex_count;
    m.vertex_stride = p.vertex_stride;
    m.indices = Addr(p.triangle_indices);
    m.index16 = rr::Raw(p.index_type) == RR_INDEX_TYPE_UINT16 ? 1u : 0u;
    m.tri_count = p.triangle_count;
    return RR_SUCCESS;
}

// How long rrWaitEvent polls before it blocks (RR_WAIT_SPIN_US, default 2000).
long WaitSpinMicros() {
    static const long us = [] {
        const char* e = std::getenv("RR_WAIT_SPIN_US");
        return e ? std::max(0L, std::atol(e)) : 2000L;
    }();
    return us;
}

// An AABB list (radeonrays.h:201-206; the reference declares it and answers
// RR_ERROR_NOT_IMPLEMENTED, radeonrays.cpp:286-289).  One list per input.
bool IsAabbList(const RRGeometryBuildInput* in) { return rr::Raw(in->primitive_type) == RR_PRIMITIVE_TYPE_AABB_LIST; }

RRError AabbsFromInput(const RRGeometryBuildInput* in, rr::AabbDesc& a, const char* name, bool resolve) {
    RR_REQUIRE(in->primitive_count == 1 && in->aabb_primitives, name);
    const RRAABBListPrimitive& p = in->aabb_primitives[0];
    RR_REQUIRE(p.aabb_count < (1u << 30), name);
    a.count = p.aabb_count;
    if (!resolve) return RR_SUCCESS;
    if (p.aabb_count > 0) RR_REQUIRE(p.aabbs && p.aabb_stride >= 24 && p.aabb_stride % 4 == 0, name);
    a.aabbs = Addr(p.aabbs);
    a.stride = p.aabb_stride;
    return RR_SUCCESS;
}

// Several triangle meshes in one build input (primitive_count > 1; the
// reference asserts a single mesh, intersector.cpp:110,146).
bool IsMultiMesh(const RRGeometryBuildInput* in) {
    return rr::Raw(in->primitive_type) == RR_PRIMITIVE_TYPE_TRIANGLE_MESH && in->primitive_count > 1;
}

RRError SubMeshesFromInput(const RRGeometryBuildInput* in, std::vector<rr::SubMesh>& out, uint32_t& tris,
                           const char* name, bool resolve) {
    RR_REQUIRE(in->triangle_mesh_primitives, name);
    out.clear();
    uint64_t total = 0;
    for (uint32_t k = 0; k < in->primitive_count; ++k) {
        RRGeometryBuildInput one = *in;
        one.primitive_count = 1;
        one.triangle_mesh_primitives = in->triangle_mesh_primitives + k;
        rr::MeshDesc m;
        const RRError e = MeshFromInput(&one, m, name, resolve);
        if (e != RR_SUCCESS) return e;
        rr::SubMesh sm{};
        sm.vertices = reinterpret_cast<uint64_t>(m.vertices);
        sm.indices = reinterpret_cast<uint64_t>(m.indices);
        sm.vertex_count = m.vertex_count;
        sm.vertex_stride = m.vertex_stride;
        sm.index16 = m.index16;
        sm.tri_count = m.tri_count;
        sm.first_tri = (uint32_t)total;
        total += m.tri_count;
        RR_REQUIRE(total < (1u << 30), name);
        out.push_back(sm);
    }
    tris = (uint32_t)total;
    return RR_SUCCESS;
}

bool WantsRestructure(const RRBuildOptions* o) {
    // intersector.cpp:116,170: restructure only with options that do not ask for a fast build.
    return o && (o->build_flags & RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD) == 0;
}

RRError CreateContext(int device, hipStream_t stream, RRContext* out) {
    return Guarded("rrCreateContext", [&]() -> RRError {
        int count = 0;
        Check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) return RR_ERROR_INVALID_PARAMETER;
        auto* c = new Context;
        c->device = device;
        DeviceGuard g(device);
        if (stream) {
            c->stream = stream;
            c->own_stream = false;
        } else {
            Check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        }
        rr::InitDeviceErrors();
        // The no-allocation sort's scratch on this device (sort.hip).
        if (!rr::InitSortNoAllocScratch())
            RR_LOG(RR_LOG_LEVEL_WARN, "rrCreateContext: no-allocation sort scratch not found on device %d", device);
        LiveContexts().Add(c);
        *out = reinterpret_cast<RRContext>(c);
        return RR_SUCCESS;
    });
}

// Report device-side failures seen since the last report (device_errors.h).
RRError DeviceErrorsAsResult(const char* name) {
    const uint32_t e = rr::TakeDeviceErrors();
    if (!e) return RR_SUCCESS;
    RR_LOG(RR_LOG_LEVEL_ERROR, "%s: device reported%s%s%s", name, (e >> rr::kErrStackOverflow) & 1 ? " traversal stack overflow" : "",
           (e >> rr::kErrLookbackTimeout) & 1 ? " sort look-back timeout" : "",
           (e >> rr::kErrBadAccel) & 1 ? " a query on a buffer without a valid header" : "");
    return RR_ERROR_INTERNAL;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device error words (device_errors.h): pinned, coherent, mapped host memory.

namespace {
uint32_t* g_err_host = nullptr;    // host view
uint32_t* g_err_device = nullptr;  // the address kernels store to
std::once_flag g_err_once;
}  // namespace

void rr::InitDeviceErrors() {
    std::call_once(g_err_once, [] {
        void* p = nullptr;
        Check(hipHostMalloc(&p, 256, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
        std::memset(p, 0, 256);
        void* d = nullptr;
        Check(hipHostGetDevicePointer(&d, p, 0), "hipHostGetDevicePointer");
        g_err_host = static_cast<uint32_t*>(p);
        g_err_device = static_cast<uint32_t*>(d);
    });
}

uint32_t* rr::DeviceErrorWords() { return g_err_device; }

uint32_t rr::TakeDeviceErrors() {
    if (!g_err_host) return 0;
    uint32_t bits = 0;
    for (uint32_t k = 0; k < rr::kErrKinds; ++k) {
        volatile uint32_t* w = g_err_host + k;
        if (*w) {
            bits |= 1u << k;
            *w = 0;
        }
    }
    return bits;
}

extern "C" {

// ---------------------------------------------------------------------------
// Context and logging

RRError rrCreateContext(uint32_t api_version, RRApi api, RRContext* context) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrCreateContext(%u)", api_version);
    RR_REQUIRE(context, "rrCreateContext");
    if (rr::Raw(api) != RR_API_HIP) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "rrCreateContext: only RR_API_HIP is available");
        return RR_ERROR_UNSUPPORTED_API;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    return CreateContext(dev, nullptr, context);
}

RRError rrCreateContextHip(uint32_t api_version, int device, void* hip_stream, RRContext* context) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrCreateContextHip(%u, %d)", api_version, device);
    RR_REQUIRE(context, "rrCreateContextHip");
    return CreateContext(device, static_cast<hipStream_t>(hip_stream), context);
}

RRError rrDestroyContext(RRContext context) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrDestroyContext");
    RR_REQUIRE(LiveCtx(context), "rrDestroyContext");
    return Guarded("rrDestroyContext", [&]() -> RRError {
        Context* c = AsCtx(context);
        LiveContexts().Remove(c);
        DeviceGuard g(c->device);
        (void)hipStreamSynchronize(c->stream);
        c->Reap(true);
        c->FreePool();
        for (Event* e : c->event_pool) {
            LiveEvents().Remove(e);
            (void)hipEventDestroy(e->ev);
            delete e;
        }
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return RR_SUCCESS;
    });
}

RRError rrSetLogLevel(RRLogLevel log_level) {
    const uint32_t level = rr::Raw(log_level);
    if (level < RR_LOG_LEVEL_DEBUG || level > RR_LOG_LEVEL_OFF) return RR_ERROR_INVALID_PARAMETER;
    Log().level = (int)level;
    RR_LOG(RR_LOG_LEVEL_INFO, "rrSetLogLevel(%d)", (int)level);
    return RR_SUCCESS;
}

RRError rrSetLogFile(char const* filename) {
    RR_REQUIRE(filename, "rrSetLogFile");
    FILE* f = std::fopen(filename, "a");
    if (!f) return RR_ERROR_INTERNAL;
    {
        std::lock_guard<std::mutex> g(Log().mu);
        if (Log().sink && Log().sink != stderr && Log().sink != stdout) std::fclose(Log().sink);
        Log().sink = f;
    }
    RR_LOG(RR_LOG_LEVEL_INFO, "rrSetLogFile(%s)", filename);
    return RR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Memory requirements.  Sizes are functions of the build input alone: these
// queries read nothing from the context (as long as it is not null), so
// they also answer where no device is present.

RRError rrGetGeometryBuildMemoryRequirements(RRContext context, const RRGeometryBuildInput* build_input,
                                             const RRBuildOptions* build_options,
                                             RRMemoryRequirements* memory_requirements) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrGetGeometryBuildMemoryRequirements");
    RR_REQUIRE(context && build_input && memory_requirements, "rrGetGeometryBuildMemoryRequirements");
    if (IsAabbList(build_input)) {
        rr::AabbDesc a;
        const RRError e = AabbsFromInput(build_input, a, "rrGetGeometryBuildMemoryRequirements", false);
        if (e != RR_SUCCESS) return e;
        memory_requirements->result_buffer_size = rr::BlasResultSize(a.count);
        memory_requirements->temporary_build_buffer_size = rr::AabbScratchBytes(a.count, true, WantsRestructure(build_options));
        memory_requirements->temporary_update_buffer_size = rr::AabbScratchBytes(a.count, false, false);
        return RR_SUCCESS;
    }
    if (IsMultiMesh(build_input)) {
        std::vector<rr::SubMesh> subs;
        uint32_t tris = 0;
        const RRError e = SubMeshesFromInput(build_input, subs, tris, "rrGetGeometryBuildMemoryRequirements", false);
        if (e != RR_SUCCESS) return e;
        memory_requirements->result_buffer_size = rr::BlasResultSize(tris);
        memory_requirements->temporary_build_buffer_size =
            rr::MultiMeshScratchBytes((uint32_t)subs.size(), tris, true, WantsRestructure(build_options));
        // Refitting several meshes gathers them again, into scratch.
        memory_requirements->temporary_update_buffer_size = rr::MultiMeshScratchBytes((uint32_t)subs.size(), tris, false, false);
        return RR_SUCCESS;
    }
    rr::MeshDesc m;
    const RRError e = MeshFromInput(build_input, m, "rrGetGeometryBuildMemoryRequirements", false);
    if (e != RR_SUCCESS) return e;
    memory_requirements->result_buffer_size = rr::BlasResultSize(m.tri_count);
    memory_requirements->temporary_build_buffer_size = rr::BlasBuildScratchBytes(m.tri_count, WantsRestructure(build_options));
    memory_requirements->temporary_update_buffer_size = 0;  // refit needs no scratch
    return RR_SUCCESS;
}

RRError rrGetSceneBuildMemoryRequirements(RRContext context, const RRSceneBuildInput* build_input,
                                          const RRBuildOptions*, RRMemoryRequirements* memory_requirements) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrGetSceneBuildMemoryRequirements");
    RR_REQUIRE(context && build_input && memory_requirements, "rrGetSceneBuildMemoryRequirements");
    RR_REQUIRE(build_input->instance_count < (1u << 30), "rrGetSceneBuildMemoryRequirements");
    memory_requirements->result_buffer_size = rr::TlasResultSize(build_input->instance_count);
    memory_requirements->temporary_build_buffer_size = rr::TlasBuildScratchBytes(build_input->instance_count);
    // A scene refit (rrxCmdUpdateScene) stages the instances and their world
    // boxes in scratch.
    memory_requirements->temporary_update_buffer_size = rr::TlasUpdateScratchBytes(build_input->instance_count);
    return RR_SUCCESS;
}

RRError rrGetTraceMemoryRequirements(RRContext context, uint32_t ray_count, size_t* scratch_size) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrGetTraceMemoryRequirements");
    // radeonrays.cpp:741 rejects ray_count == 0 as well.
    RR_REQUIRE(context && scratch_size && ray_count, "rrGetTraceMemoryRequirements");
    *scratch_size = rr::TraceScratchBytes(ray_count);
    return RR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Builds and queries

RRError rrCmdBuildGeometry(RRContext context, RRBuildOperation build_operation, const RRGeometryBuildInput* build_input,
                           const RRBuildOptions* build_options, RRDevicePtr temporary_buffer,
                           RRDevicePtr geometry_buffer, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrCmdBuildGeometry");
    RR_REQUIRE(LiveCtx(context) && LiveStream(command_stream) && build_input, "rrCmdBuildGeometry");
    RR_REQUIRE(rr::Raw(build_operation) == RR_BUILD_OPERATION_BUILD || rr::Raw(build_operation) == RR_BUILD_OPERATION_UPDATE,
               "rrCmdBuildGeometry");
    RR_REQUIRE(OptPtr(temporary_buffer) && OptPtr(geometry_buffer), "rrCmdBuildGeometry");
    return Guarded("rrCmdBuildGeometry", [&]() -> RRError {
        Context* c = AsCtx(context);
        CommandStream* cs = AsStream(command_stream);
        const bool restructure = WantsRestructure(build_options);
        const bool update = rr::Raw(build_operation) == RR_BUILD_OPERATION_UPDATE;
        if (IsAabbList(build_input)) {
            rr::AabbDesc a;
            const RRError e = AabbsFromInput(build_input, a, "rrCmdBuildGeometry", true);
            if (e != RR_SUCCESS) return e;
            RR_REQUIRE(geometry_buffer && temporary_buffer, "rrCmdBuildGeometry");
            DeviceGuard g(c->device);
            void* result = Addr(geometry_buffer);
            void* scratch = Addr(temporary_buffer);
            RecordAccel(c, result, rr::kAccelBlas, a.count);
            if (!update)
                Record(cs, [=](hipStream_t s) { rr::BuildAabbs(a, restructure, scratch, result, s); });
            else
                Record(cs, [=](hipStream_t s) { rr::UpdateAabbs(a, scratch, result, s); });
            return RR_SUCCESS;
        }
        if (IsMultiMesh(build_input)) {
            std::vector<rr::SubMesh> subs;
            uint32_t tris = 0;
            const RRError e = SubMeshesFromInput(build_input, subs, tris, "rrCmdBuildGeometry", true);
            if (e != RR_SUCCESS) return e;
            RR_REQUIRE(geometry_buffer && temporary_buffer, "rrCmdBuildGeometry");
            DeviceGuard g(c->device);
            void* result = Addr(geometry_buffer);
            void* scratch = Addr(temporary_buffer);
            RecordAccel(c, result, rr::kAccelBlas, tris);
            const uint32_t count = (uint32_t)subs.size();
            const size_t bytes = sizeof(rr::SubMesh) * count;
            void* host = AcquireStaging(c, cs, bytes);
            std::memcpy(host, subs.data(), bytes);
            Record(cs, [=](hipStream_t s) {
                auto* table = reinterpret_cast<rr::SubMesh*>(static_cast<uint8_t*>(scratch) + rr::MultiMeshTableOffset(tris));
                Check(hipMemcpyAsync(table, host, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
                if (update) rr::UpdateMultiMesh(table, count, tris, scratch, result, s);
                else rr::BuildMultiMesh(table, count, tris, restructure, scratch, result, s);
            });
            KeepStaging(c, cs, host);
            return RR_SUCCESS;
        }
        rr::MeshDesc m;
        const RRError e = MeshFromInput(build_input, m, "rrCmdBuildGeometry", true);
        if (e != RR_SUCCESS) return e;
        RR_REQUIRE(geometry_buffer, "rrCmdBuildGeometry");
        if (!update) RR_REQUIRE(temporary_buffer, "rrCmdBuildGeometry");
        DeviceGuard g(c->device);
        void* result = Addr(geometry_buffer);
        void* scratch = Addr(temporary_buffer);
        RecordAccel(c, result, rr::kAccelBlas, m.tri_count);
        if (!update)
            Record(cs, [=](hipStream_t s) { rr::BuildBlas(m, restructure, scratch, result, s); });
        else
            Record(cs, [=](hipStream_t s) { rr::UpdateBlas(m, result, s); });
        return RR_SUCCESS;
    });
}

// Build (or refit) a scene over the instances, as rrCmdBuildScene.
static RRError SceneCommand(bool update, RRContext context, const RRSceneBuildInput* build_input,
                            RRDevicePtr temporary_buffer, RRDevicePtr scene_buffer, RRCommandStream command_stream) {
    RR_REQUIRE(LiveCtx(context) && LiveStream(command_stream) && build_input, "rrCmdBuildScene");
    RR_REQUIRE(LivePtr(scene_buffer) && LivePtr(temporary_buffer), "rrCmdBuildScene");
    RR_REQUIRE(build_input->instance_count == 0 || build_input->instances, "rrCmdBuildScene");
    RR_REQUIRE(build_input->instance_count < (1u << 30), "rrCmdBuildScene");
    for (uint32_t i = 0; i < build_input->instance_count; ++i)
        RR_REQUIRE(LivePtr(build_input->instances[i].geometry), "rrCmdBuildScene");
    return Guarded("rrCmdBuildScene", [&]() -> RRError {
        Context* c = AsCtx(context);
        CommandStream* cs = AsStream(command_stream);
        DeviceGuard g(c->device);
        const uint32_t count = build_input->instance_count;
        void* result = Addr(scene_buffer);
        void* scratch = Addr(temporary_buffer);
        if (update) {
            // A refit keeps the scene's layout, which depends on the instance
            // count: refuse another count for a scene this context built.
            std::lock_guard<std::mutex> lk(c->mu);
            auto it = c->kinds.find(reinterpret_cast<uintptr_t>(result));
            if (it != c->kinds.end() && (it->second.kind != rr::kAccelTlas || it->second.prims != count)) {
                RR_LOG(RR_LOG_LEVEL_ERROR, "rrxCmdUpdateScene: invalid parameter (instance count %u, scene built with %u)",
                       count, it->second.prims);
                return RR_ERROR_INVALID_PARAMETER;
            }
        }
        // Pack the instances now (the caller's array may go away after this
        // call, as in intersector.cpp:218-247) into pinned memory that the
        // recorded upload reads at execution time.
        rr::InstanceUpload* host = nullptr;
        if (count) {
            host = static_cast<rr::InstanceUpload*>(AcquireStaging(c, cs, sizeof(rr::InstanceUpload) * count));
            for (uint32_t i = 0; i < count; ++i) {
                std::memcpy(host[i].xform.m, build_input->instances[i].transform, sizeof(rr::Transform));
                host[i].blas = reinterpret_cast<uint64_t>(Addr(build_input->instances[i].geometry));
                host[i].pad = 0;
            }
        }
        RecordAccel(c, result, rr::kAccelTlas, count);
        auto cmd = [=](hipStream_t s) {
            auto* dev_inst = reinterpret_cast<rr::InstanceUpload*>(static_cast<uint8_t*>(scratch) + rr::TlasUploadOffset());
            if (count)
                Check(hipMemcpyAsync(dev_inst, host, sizeof(rr::InstanceUpload) * count, hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
            if (update) rr::UpdateTlas(dev_inst, count, scratch, result, s);
            else rr::BuildTlas(dev_inst, count, scratch, result, s);
        };
        Record(cs, cmd);
        KeepStaging(c, cs, host);
        return RR_SUCCESS;
    });
}

RRError rrCmdBuildScene(RRContext context, const RRSceneBuildInput* build_input, const RRBuildOptions*,
                        RRDevicePtr temporary_buffer, RRDevicePtr scene_buffer, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrCmdBuildScene");
    return SceneCommand(false, context, build_input, temporary_buffer, scene_buffer, command_stream);
}

RRError rrxCmdUpdateScene(RRContext context, const RRSceneBuildInput* build_input, RRDevicePtr temporary_buffer,
                          RRDevicePtr scene_buffer, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrxCmdUpdateScene");
    return SceneCommand(true, context, build_input, temporary_buffer, scene_buffer, command_stream);
}

// Kind of a result buffer as this context recorded it, or kAccelUnknown for
// a buffer built elsewhere (another context, a copy): traversal then reads
// the kind from the buffer's header on the device, so recording never waits
// for the device (the reference decides through a host map,
// vlk/intersector.cpp:289-291).
static rr::AccelKind KindOf(Context* c, void* accel) {
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->kinds.find(reinterpret_cast<uintptr_t>(accel));
    return it != c->kinds.end() ? it->second.kind : rr::kAccelUnknown;
}

RRError rrCmdIntersect(RRContext context, RRDevicePtr scene_buffer, RRIntersectQuery query, RRDevicePtr rays,
                       uint32_t ray_count, RRDevicePtr indirect_ray_count, RRIntersectQueryOutput query_output,
                       RRDevicePtr hits, RRDevicePtr scratch, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrCmdIntersect");
    RR_REQUIRE(LiveCtx(context) && LivePtr(scene_buffer) && LivePtr(rays) && LivePtr(hits) && LivePtr(scratch) &&
                   LiveStream(command_stream) && OptPtr(indirect_ray_count),
               "rrCmdIntersect");
    RR_REQUIRE(rr::Raw(query) == RR_INTERSECT_QUERY_CLOSEST || rr::Raw(query) == RR_INTERSECT_QUERY_ANY, "rrCmdIntersect");
    RR_REQUIRE(rr::Raw(query_output) == RR_INTERSECT_QUERY_OUTPUT_FULL_HIT ||
                   rr::Raw(query_output) == RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID ||
                   rr::Raw(query_output) == RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T,
               "rrCmdIntersect");
    return Guarded("rrCmdIntersect", [&]() -> RRError {
        Context* c = AsCtx(context);
        DeviceGuard g(c->device);
        void* accel = Addr(scene_buffer);
        const rr::AccelKind kind = KindOf(c, accel);
        const void* r = Addr(rays);
        void* h = Addr(hits);
        void* sc = Addr(scratch);
        const uint32_t* ind = static_cast<const uint32_t*>(Addr(indirect_ray_count));
        const auto q = query == RR_INTERSECT_QUERY_ANY ? rr::kQueryAny : rr::kQueryClosest;
        const auto o = static_cast<rr::OutputKind>(query_output);
        Record(AsStream(command_stream), [=](hipStream_t s) { rr::Trace(accel, kind, q, o, r, ray_count, ind, h, sc, s); });
        return RR_SUCCESS;
    });
}

// ---------------------------------------------------------------------------
// Command streams and events

RRError rrAllocateCommandStream(RRContext context, RRCommandStream* command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrAllocateCommandStream");
    RR_REQUIRE(LiveCtx(context) && command_stream, "rrAllocateCommandStream");
    return Guarded("rrAllocateCommandStream", [&]() -> RRError {
        auto* cs = new CommandStream;
        cs->ctx = AsCtx(context);
        LiveStreams().Add(cs);
        *command_stream = reinterpret_cast<RRCommandStream>(cs);
        return RR_SUCCESS;
    });
}

static void ReleaseStaging(Context* c, CommandStream* cs) {
    if (cs->staging.empty()) return;
    std::lock_guard<std::mutex> lk(c->mu);
    const bool capturing = IsCapturing(c->stream);
    for (void* p : cs->staging) {
        if (capturing) c->graph_staging.push_back(p);
        else c->FreeAfter(c->stream, p);
    }
    cs->staging.clear();
}

RRError rrReleaseCommandStream(RRContext context, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrReleaseCommandStream");
    RR_REQUIRE(LiveCtx(context) && LiveStream(command_stream) && !AsStream(command_stream)->external,
               "rrReleaseCommandStream");
    return Guarded("rrReleaseCommandStream", [&]() -> RRError {
        Context* c = AsCtx(context);
        CommandStream* cs = AsStream(command_stream);
        LiveStreams().Remove(cs);
        DeviceGuard g(c->device);
        ReleaseStaging(c, cs);
        delete cs;
        return RR_SUCCESS;
    });
}

RRError rrSumbitCommandStream(RRContext context, RRCommandStream command_stream, RREvent wait_event, RREvent* out_event) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrSumbitCommandStream");
    RR_REQUIRE(LiveCtx(context) && LiveStream(command_stream) && out_event, "rrSumbitCommandStream");
    RR_REQUIRE(!wait_event || LiveEvent(wait_event), "rrSumbitCommandStream");
    return Guarded("rrSumbitCommandStream", [&]() -> RRError {
        Context* c = AsCtx(context);
        CommandStream* cs = AsStream(command_stream);
        DeviceGuard g(c->device);
        const hipStream_t s = cs->external ? cs->stream : c->stream;
        if (wait_event) Check(hipStreamWaitEvent(s, AsEvent(wait_event)->ev, 0), "hipStreamWaitEvent");
        for (auto& cmd : cs->commands) cmd(s);
        cs->commands.clear();
        Check(hipGetLastError(), "kernel launch");
        // Staging read by this submission is freed once it has run.
        ReleaseStaging(c, cs);
        Event* ev = nullptr;
        {
            std::lock_guard<std::mutex> lk(c->mu);
            c->Reap(false);
            if (!c->event_pool.empty()) {
                ev = c->event_pool.back();
                c->event_pool.pop_back();
            }
        }
        if (!ev) {
            ev = new Event;
            Check(hipEventCreateWithFlags(&ev->ev, hipEventDisableTiming), "hipEventCreate");
            LiveEvents().Add(ev);
        }
        ev->pooled = false;
        Check(hipEventRecord(ev->ev, s), "hipEventRecord");
        *out_event = reinterpret_cast<RREvent>(ev);
        return RR_SUCCESS;
    });
}

RRError rrReleaseEvent(RRContext context, RREvent event) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrReleaseEvent");
    RR_REQUIRE(LiveCtx(context) && LiveEvent(event), "rrReleaseEvent");
    Context* c = AsCtx(context);
    std::lock_guard<std::mutex> lk(c->mu);
    Event* e = AsEvent(event);
    if (e->pooled) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "rrReleaseEvent: invalid parameter (event already released)");
        return RR_ERROR_INVALID_PARAMETER;
    }
    e->pooled = true;
    c->event_pool.push_back(e);
    return RR_SUCCESS;
}

RRError rrWaitEvent(RRContext context, RREvent event) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrWaitEvent");
    RR_REQUIRE(LiveCtx(context) && LiveEvent(event) && !AsEvent(event)->pooled, "rrWaitEvent");
    return Guarded("rrWaitEvent", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        // Poll first (short queries finish in tens of microseconds, less than
        // a blocking wait takes to wake up), then block.  Failures the
        // device reported while the work ran make the wait fail.
        const hipEvent_t ev = AsEvent(event)->ev;
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(WaitSpinMicros());
        for (;;) {
            const hipError_t q = hipEventQuery(ev);
            if (q == hipSuccess) return DeviceErrorsAsResult("rrWaitEvent");
            if (q != hipErrorNotReady) Check(q, "hipEventQuery");
            if (std::chrono::steady_clock::now() >= until) break;
            std::this_thread::yield();
        }
        Check(hipEventSynchronize(ev), "hipEventSynchronize");
        return DeviceErrorsAsResult("rrWaitEvent");
    });
}

RRError rrGetCommandStreamFromHipStream(RRContext context, void* hip_stream, RRCommandStream* command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrGetCommandStreamFromHipStream");
    RR_REQUIRE(LiveCtx(context) && command_stream, "rrGetCommandStreamFromHipStream");
    return Guarded("rrGetCommandStreamFromHipStream", [&]() -> RRError {
        auto* cs = new CommandStream;
        cs->ctx = AsCtx(context);
        cs->external = true;
        cs->stream = static_cast<hipStream_t>(hip_stream);
        LiveStreams().Add(cs);
        *command_stream = reinterpret_cast<RRCommandStream>(cs);
        return RR_SUCCESS;
    });
}

RRError rrReleaseExternalCommandStream(RRContext context, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrReleaseExternalCommandStream");
    RR_REQUIRE(LiveCtx(context) && LiveStream(command_stream) && AsStream(command_stream)->external,
               "rrReleaseExternalCommandStream");
    LiveStreams().Remove(command_stream);
    delete AsStream(command_stream);
    return RR_SUCCESS;
}

void* rrGetHipStream(RRCommandStream command_stream) {
    if (!LiveStream(command_stream)) return nullptr;
    CommandStream* cs = AsStream(command_stream);
    return cs->external ? cs->stream : cs->ctx->stream;
}

// ---------------------------------------------------------------------------
// Device pointers

RRError rrGetDevicePtrFromHipPtr(RRContext context, void* device_address, size_t offset, RRDevicePtr* device_ptr) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrGetDevicePtrFromHipPtr");
    RR_REQUIRE(LiveCtx(context) && device_address && device_ptr, "rrGetDevicePtrFromHipPtr");
    return Guarded("rrGetDevicePtrFromHipPtr", [&]() -> RRError {
        auto* p = new DevicePtr;
        p->base = device_address;
        p->offset = offset;
        LivePtrs().Add(p);
        *device_ptr = reinterpret_cast<RRDevicePtr>(p);
        return RR_SUCCESS;
    });
}

RRError rrReleaseDevicePtr(RRContext context, RRDevicePtr ptr) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrReleaseDevicePtr");
    RR_REQUIRE(LiveCtx(context) && LivePtr(ptr), "rrReleaseDevicePtr");
    return Guarded("rrReleaseDevicePtr", [&]() -> RRError {
        DevicePtr* p = AsPtr(ptr);
        LivePtrs().Remove(p);
        if (p->owned) {
            Context* c = AsCtx(context);
            DeviceGuard g(c->device);
            (void)hipStreamSynchronize(c->stream);
            if (p->mapped) std::free(p->mapped);
            (void)hipFree(p->base);
            // Forget the levels recorded for results built into this memory:
            // the next allocation may reuse the addresses for anything.
            std::lock_guard<std::mutex> lk(c->mu);
            const uintptr_t lo = reinterpret_cast<uintptr_t>(p->base), hi = lo + p->size;
            for (auto it = c->kinds.begin(); it != c->kinds.end();) {
                if (it->first >= lo && it->first < hi) it = c->kinds.erase(it);
                else ++it;
            }
        }
        delete p;
        return RR_SUCCESS;
    });
}

RRError rrAllocateDeviceBuffer(RRContext context, size_t size, RRDevicePtr* device_ptr) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrAllocateDeviceBuffer(%zu)", size);
    RR_REQUIRE(LiveCtx(context) && size && device_ptr, "rrAllocateDeviceBuffer");
    return Guarded("rrAllocateDeviceBuffer", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        void* mem = nullptr;
        Check(hipMalloc(&mem, size), "hipMalloc");
        auto* p = new DevicePtr;
        p->base = mem;
        p->owned = true;
        p->size = size;
        LivePtrs().Add(p);
        *device_ptr = reinterpret_cast<RRDevicePtr>(p);
        return RR_SUCCESS;
    });
}

RRError rrMapDevicePtr(RRContext context, RRDevicePtr device_ptr, void** mapping_ptr) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrMapDevicePtr");
    RR_REQUIRE(LiveCtx(context) && LivePtr(device_ptr) && mapping_ptr, "rrMapDevicePtr");
    DevicePtr* p = AsPtr(device_ptr);
    RR_REQUIRE(p->owned, "rrMapDevicePtr");
    return Guarded("rrMapDevicePtr", [&]() -> RRError {
        Context* c = AsCtx(context);
        DeviceGuard g(c->device);
        if (!p->mapped) {
            p->mapped = std::malloc(p->size);
            if (!p->mapped) return RR_ERROR_OUT_OF_HOST_MEMORY;
        }
        Check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        Check(hipMemcpy(p->mapped, p->base, p->size, hipMemcpyDeviceToHost), "hipMemcpy");
        *mapping_ptr = p->mapped;
        return RR_SUCCESS;
    });
}

RRError rrUnmapDevicePtr(RRContext context, RRDevicePtr device_ptr, void** mapping_ptr) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrUnmapDevicePtr");
    RR_REQUIRE(LiveCtx(context) && LivePtr(device_ptr) && mapping_ptr, "rrUnmapDevicePtr");
    DevicePtr* p = AsPtr(device_ptr);
    RR_REQUIRE(p->owned && p->mapped, "rrUnmapDevicePtr");
    return Guarded("rrUnmapDevicePtr", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        Check(hipMemcpy(p->base, p->mapped, p->size, hipMemcpyHostToDevice), "hipMemcpy");
        std::free(p->mapped);
        p->mapped = nullptr;
        *mapping_ptr = nullptr;
        return RR_SUCCESS;
    });
}

void* rrGetHipAddress(RRDevicePtr device_ptr) { return LivePtr(device_ptr) ? AsPtr(device_ptr)->addr() : nullptr; }

// ---------------------------------------------------------------------------
// Extensions

RRError rrxGetSortMemoryRequirements(RRContext context, uint32_t n, size_t* scratch_size) {
    RR_REQUIRE(context && scratch_size, "rrxGetSortMemoryRequirements");
    RR_REQUIRE(n < (1u << 30), "rrxGetSortMemoryRequirements");
    *scratch_size = rr::SortScratchBytes(n);
    return RR_SUCCESS;
}

RRError rrxCmdSortPairs(RRContext context, RRDevicePtr keys, RRDevicePtr values, RRDevicePtr keys_tmp,
                        RRDevicePtr values_tmp, uint32_t n, uint32_t key_bits, RRDevicePtr scratch,
                        RRCommandStream command_stream) {
    RR_REQUIRE(LiveCtx(context) && LivePtr(keys) && LivePtr(values) && LivePtr(keys_tmp) && LivePtr(values_tmp) &&
                   LivePtr(scratch) && LiveStream(command_stream),
               "rrxCmdSortPairs");
    RR_REQUIRE(n < (1u << 30) && key_bits >= 1 && key_bits <= 32, "rrxCmdSortPairs");
    return Guarded("rrxCmdSortPairs", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        auto* k = static_cast<uint32_t*>(Addr(keys));
        auto* v = static_cast<uint32_t*>(Addr(values));
        auto* kt = static_cast<uint32_t*>(Addr(keys_tmp));
        auto* vt = static_cast<uint32_t*>(Addr(values_tmp));
        void* sc = Addr(scratch);
        Record(AsStream(command_stream), [=](hipStream_t s) {
            const rr::SortScratch ss = rr::CarveSortScratch(sc, n);
            const bool in_place = rr::RadixSortPairs(k, v, kt, vt, n, (int)key_bits, ss, false, s);
            if (!in_place) {
                Check(hipMemcpyAsync(k, kt, size_t(n) * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                Check(hipMemcpyAsync(v, vt, size_t(n) * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
            }
        });
        return RR_SUCCESS;
    });
}

RRError rrxCmdSortPairsNoAlloc(RRContext context, RRDevicePtr keys, RRDevicePtr values, RRDevicePtr keys_tmp,
                               RRDevicePtr values_tmp, uint32_t n, uint32_t key_bits, RRCommandStream command_stream) {
    RR_REQUIRE(LiveCtx(context) && LivePtr(keys) && LivePtr(values) && LivePtr(keys_tmp) && LivePtr(values_tmp) &&
                   LiveStream(command_stream),
               "rrxCmdSortPairsNoAlloc");
    RR_REQUIRE(key_bits >= 1 && key_bits <= 32, "rrxCmdSortPairsNoAlloc");
    if (n > rr::kSortNoAllocMaxKeys) {
        RR_LOG(RR_LOG_LEVEL_ERROR, "rrxCmdSortPairsNoAlloc: invalid parameter (%u pairs; at most %u without allocation)",
               n, rr::kSortNoAllocMaxKeys);
        return RR_ERROR_INVALID_PARAMETER;
    }
    return Guarded("rrxCmdSortPairsNoAlloc", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        auto* k = static_cast<uint32_t*>(Addr(keys));
        auto* v = static_cast<uint32_t*>(Addr(values));
        auto* kt = static_cast<uint32_t*>(Addr(keys_tmp));
        auto* vt = static_cast<uint32_t*>(Addr(values_tmp));
        Record(AsStream(command_stream), [=](hipStream_t s) {
            const int r = rr::SortPairsNoAlloc(k, v, kt, vt, n, (int)key_bits, s);
            if (r < 0) throw std::runtime_error("no-allocation sort scratch unavailable on this device");
            if (r == 0) {
                Check(hipMemcpyAsync(k, kt, size_t(n) * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                Check(hipMemcpyAsync(v, vt, size_t(n) * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
            }
        });
        return RR_SUCCESS;
    });
}

RRError rrxCmdIntersectWithStats(RRContext context, RRDevicePtr scene_buffer, RRIntersectQuery query, RRDevicePtr rays,
                                 uint32_t ray_count, RRDevicePtr hits, RRDevicePtr scratch, RRDevicePtr stats,
                                 RRCommandStream command_stream) {
    RR_REQUIRE(LiveCtx(context) && LivePtr(scene_buffer) && LivePtr(rays) && LivePtr(hits) && LivePtr(scratch) &&
                   LivePtr(stats) && LiveStream(command_stream),
               "rrxCmdIntersectWithStats");
    return Guarded("rrxCmdIntersectWithStats", [&]() -> RRError {
        Context* c = AsCtx(context);
        DeviceGuard g(c->device);
        void* accel = Addr(scene_buffer);
        const rr::AccelKind kind = KindOf(c, accel);
        const void* r = Addr(rays);
        void* h = Addr(hits);
        void* sc = Addr(scratch);
        auto* st = static_cast<uint32_t*>(Addr(stats));
        const auto q = query == RR_INTERSECT_QUERY_ANY ? rr::kQueryAny : rr::kQueryClosest;
        Record(AsStream(command_stream), [=](hipStream_t s) { rr::TraceWithStats(accel, kind, q, r, ray_count, h, sc, st, s); });
        return RR_SUCCESS;
    });
}

RRError rrxGetIntersectBatchMemoryRequirements(RRContext context, uint32_t request_count, uint64_t total_ray_count,
                                               uint32_t max_request_rays, size_t* scratch_size) {
    RR_REQUIRE(LiveCtx(context) && scratch_size, "rrxGetIntersectBatchMemoryRequirements");
    RR_REQUIRE(total_ray_count < (uint64_t(1) << 37), "rrxGetIntersectBatchMemoryRequirements");
    return Guarded("rrxGetIntersectBatchMemoryRequirements", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        // Every request shares the two batch launches (one per level).  The
        // size never shrinks below what a single query of max_request_rays
        // needs, which earlier versions ran scene requests with.
        *scratch_size = std::max(rr::TraceBatchScratchBytes(request_count, total_ray_count),
                                 rr::TraceScratchBytes(max_request_rays));
        return RR_SUCCESS;
    });
}

RRError rrxCmdIntersectBatch(RRContext context, const RRxIntersectRequest* requests, uint32_t request_count,
                             RRDevicePtr scratch, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrxCmdIntersectBatch");
    RR_REQUIRE(LiveCtx(context) && LivePtr(scratch) && LiveStream(command_stream), "rrxCmdIntersectBatch");
    RR_REQUIRE(request_count == 0 || requests, "rrxCmdIntersectBatch");
    for (uint32_t i = 0; i < request_count; ++i) {
        const RRxIntersectRequest& q = requests[i];
        RR_REQUIRE(LivePtr(q.scene_buffer) && (q.ray_count == 0 || (LivePtr(q.rays) && LivePtr(q.hits))),
                   "rrxCmdIntersectBatch");
        RR_REQUIRE(rr::Raw(q.query) == RR_INTERSECT_QUERY_CLOSEST || rr::Raw(q.query) == RR_INTERSECT_QUERY_ANY, "rrxCmdIntersectBatch");
        RR_REQUIRE(rr::Raw(q.query_output) == RR_INTERSECT_QUERY_OUTPUT_FULL_HIT ||
                       rr::Raw(q.query_output) == RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID ||
                       rr::Raw(q.query_output) == RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T,
                   "rrxCmdIntersectBatch");
    }
    return Guarded("rrxCmdIntersectBatch", [&]() -> RRError {
        Context* c = AsCtx(context);
        CommandStream* cs = AsStream(command_stream);
        DeviceGuard g(c->device);
        // One table, ordered as: requests whose address this context
        // recorded as a geometry, requests it has no record of (the kernels
        // read their level from the header), requests recorded as a scene.
        // The geometry kernel takes the first two groups, the two-level
        // kernel the last two.
        std::vector<rr::BatchRequest> group[3];
        for (uint32_t i = 0; i < request_count; ++i) {
            const RRxIntersectRequest& q = requests[i];
            if (q.ray_count == 0) continue;
            void* accel = Addr(q.scene_buffer);
            const rr::AccelKind kind = KindOf(c, accel);
            const auto qk = q.query == RR_INTERSECT_QUERY_ANY ? rr::kQueryAny : rr::kQueryClosest;
            const auto ok = static_cast<rr::OutputKind>(q.query_output);
            rr::BatchRequest b{};
            b.accel = reinterpret_cast<uint64_t>(accel);
            b.rays = reinterpret_cast<uint64_t>(Addr(q.rays));
            b.hits = reinterpret_cast<uint64_t>(Addr(q.hits));
            b.ray_count = q.ray_count;
            b.flags = (qk == rr::kQueryAny ? 1u : 0u) | (uint32_t(ok) << 1);
            const int g = kind == rr::kAccelBlas ? 0 : kind == rr::kAccelTlas ? 2 : 1;
            if (g == 1) b.flags |= rr::kBatchFlagCheckLevel;
            group[g].push_back(b);
        }
        std::vector<rr::BatchRequest> table;
        uint64_t total_rays = 0;
        for (int g = 0; g < 3; ++g)
            for (const rr::BatchRequest& b : group[g]) total_rays += b.ray_count;
        const uint32_t chunk = rr::BatchChunkFor(total_rays, request_count);
        uint64_t chunks = 0, group_chunks[3] = {0, 0, 0};
        for (int g = 0; g < 3; ++g) {
            for (rr::BatchRequest& b : group[g]) {
                b.first_chunk = uint32_t(chunks);
                const uint64_t n = (b.ray_count + chunk - 1) / chunk;
                chunks += n;
                group_chunks[g] += n;
                RR_REQUIRE(chunks < (uint64_t(1) << 32), "rrxCmdIntersectBatch");
                table.push_back(b);
            }
        }
        const uint32_t n_req = uint32_t(table.size());
        const uint32_t c_geo = uint32_t(group_chunks[0]), c_unrec = uint32_t(group_chunks[1]),
                       c_scene = uint32_t(group_chunks[2]);
        void* sc = Addr(scratch);
        void* host = nullptr;
        const size_t table_bytes = sizeof(rr::BatchRequest) * n_req;
        if (n_req) host = AcquireStaging(c, cs, table_bytes);
        if (host) std::memcpy(host, table.data(), table_bytes);
        auto cmd = [=](hipStream_t s) {
            if (n_req) {
                auto* dev_table =
                    reinterpret_cast<rr::BatchRequest*>(static_cast<uint8_t*>(sc) + rr::kBatchTableOffset);
                Check(hipMemcpyAsync(dev_table, host, table_bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
                rr::TraceBatch(dev_table, n_req, chunk, c_geo, c_unrec, c_scene, sc, s);
            }
        };
        Record(cs, cmd);
        KeepStaging(c, cs, host);
        return RR_SUCCESS;
    });
}

// Batched geometry builds.  Scratch: queue + mesh table, then a work area
// per batched mesh; meshes above kBatchMaxTris are built one at a time from
// the start of the scratch after the batch.
struct BatchPlan {
    std::vector<rr::MeshDesc> meshes;
    std::vector<uint32_t> small, large;  // indices into meshes; small sorted by size, largest first
    std::vector<size_t> work_offset;     // per small mesh
    size_t head = 0, total = 0;
    size_t roots_offset = 0;             // treelet-root list of the batched restructuring
    uint32_t max_roots[3] = {0, 0, 0};   // its bound per round
};

static RRError PlanBatch(const RRGeometryBuildInput* inputs, uint32_t count, bool restructure, bool resolve,
                         BatchPlan& plan, const char* name) {
    plan.meshes.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        const RRError e = MeshFromInput(&inputs[i], plan.meshes[i], name, resolve);
        if (e != RR_SUCCESS) return e;
        (plan.meshes[i].tri_count <= rr::kBatchMaxTris ? plan.small : plan.large).push_back(i);
    }
    std::stable_sort(plan.small.begin(), plan.small.end(),
                     [&](uint32_t a, uint32_t b) { return plan.meshes[a].tri_count > plan.meshes[b].tri_count; });
    plan.head = rr::BatchHeadBytes((uint32_t)plan.small.size());
    size_t o = plan.head;
    plan.work_offset.resize(plan.small.size());
    for (size_t k = 0; k < plan.small.size(); ++k) {
        plan.work_offset[k] = o;
        o += rr::BatchWorkBytes(plan.meshes[plan.small[k]].tri_count);
    }
    if (restructure) {
        // Every treelet root of round r covers at least 64 << r primitives.
        for (uint32_t k : plan.small)
            for (int r = 0; r < 3; ++r) plan.max_roots[r] += plan.meshes[k].tri_count / (64u << r);
        plan.roots_offset = o;
        o += rr::AlignUp(size_t(plan.max_roots[0]) * 8, 256);
    }
    plan.total = o;
    for (uint32_t i : plan.large) plan.total = std::max(plan.total, rr::BlasBuildScratchBytes(plan.meshes[i].tri_count, restructure));
    return RR_SUCCESS;
}

RRError rrxGetGeometryBatchBuildMemoryRequirements(RRContext context, const RRGeometryBuildInput* build_inputs,
                                                   uint32_t count, const RRBuildOptions* build_options,
                                                   size_t* scratch_size) {
    // A size query: reads nothing from the context.
    RR_REQUIRE(context && scratch_size && (count == 0 || build_inputs), "rrxGetGeometryBatchBuildMemoryRequirements");
    BatchPlan plan;
    const RRError e = PlanBatch(build_inputs, count, WantsRestructure(build_options), false, plan,
                                "rrxGetGeometryBatchBuildMemoryRequirements");
    if (e != RR_SUCCESS) return e;
    *scratch_size = std::max<size_t>(plan.total, 256);
    return RR_SUCCESS;
}

RRError rrxCmdBuildGeometryBatch(RRContext context, RRBuildOperation build_operation,
                                 const RRGeometryBuildInput* build_inputs, uint32_t count,
                                 const RRBuildOptions* build_options, RRDevicePtr temporary_buffer,
                                 const RRDevicePtr* geometry_buffers, RRCommandStream command_stream) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrxCmdBuildGeometryBatch");
    RR_REQUIRE(LiveCtx(context) && LiveStream(command_stream) && LivePtr(temporary_buffer), "rrxCmdBuildGeometryBatch");
    RR_REQUIRE(count == 0 || (build_inputs && geometry_buffers), "rrxCmdBuildGeometryBatch");
    RR_REQUIRE(rr::Raw(build_operation) == RR_BUILD_OPERATION_BUILD || rr::Raw(build_operation) == RR_BUILD_OPERATION_UPDATE,
               "rrxCmdBuildGeometryBatch");
    for (uint32_t i = 0; i < count; ++i) RR_REQUIRE(LivePtr(geometry_buffers[i]), "rrxCmdBuildGeometryBatch");
    return Guarded("rrxCmdBuildGeometryBatch", [&]() -> RRError {
        Context* c = AsCtx(context);
        CommandStream* cs = AsStream(command_stream);
        DeviceGuard g(c->device);
        const bool restructure = WantsRestructure(build_options);
        BatchPlan plan;
        const RRError e = PlanBatch(build_inputs, count, restructure, true, plan, "rrxCmdBuildGeometryBatch");
        if (e != RR_SUCCESS) return e;
        void* scratch = Addr(temporary_buffer);
        std::vector<void*> results(count);
        for (uint32_t i = 0; i < count; ++i) results[i] = Addr(geometry_buffers[i]);
        const bool update = rr::Raw(build_operation) == RR_BUILD_OPERATION_UPDATE;
        if (!update)
            for (uint32_t i = 0; i < count; ++i) RecordAccel(c, results[i], rr::kAccelBlas, plan.meshes[i].tri_count);
        // The mesh table, packed now into pinned memory that the recorded
        // upload reads at execution time.
        const uint32_t n_small = (uint32_t)plan.small.size();
        rr::BatchMesh* host = nullptr;
        if (n_small) host = static_cast<rr::BatchMesh*>(AcquireStaging(c, cs, sizeof(rr::BatchMesh) * n_small));
        uint32_t chunks = 0, big = 0, small = 0;
        for (uint32_t k = 0; k < n_small; ++k) {
            const rr::MeshDesc& m = plan.meshes[plan.small[k]];
            // Prefixes: the table is ordered largest first.
            if (m.tri_count > rr::kBatchLdsMaxTris) ++big;
            if (m.tri_count > rr::kBatchLdsSmallMaxTris) ++small;
            rr::BatchMesh b{};
            b.vertices = reinterpret_cast<uint64_t>(m.vertices);
            b.indices = reinterpret_cast<uint64_t>(m.indices);
            b.result = reinterpret_cast<uint64_t>(results[plan.small[k]]);
            b.work = reinterpret_cast<uint64_t>(static_cast<uint8_t*>(scratch) + plan.work_offset[k]);
            b.vertex_count = m.vertex_count;
            b.vertex_stride = m.vertex_stride;
            b.index16 = m.index16;
            b.tri_count = m.tri_count;
            b.flags = (restructure && m.tri_count >= 64) ? rr::kAccelFlagRestructured : 0u;
            b.first_chunk = chunks;
            chunks += rr::RefitChunks(m.tri_count);
            host[k] = b;
        }
        const uint32_t big_chunks = big < n_small ? host[big].first_chunk : chunks;  // chunks of the big meshes
        const size_t roots_offset = plan.roots_offset;
        const uint32_t max_roots[3] = {plan.max_roots[0], plan.max_roots[1], plan.max_roots[2]};
        auto cmd = [=](hipStream_t s) {
            if (n_small) {
                auto* table = reinterpret_cast<rr::BatchMesh*>(static_cast<uint8_t*>(scratch) + rr::kMeshTableOffset);
                Check(hipMemcpyAsync(table, host, sizeof(rr::BatchMesh) * n_small, hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
                if (update) {
                    rr::BatchRefit(table, n_small, chunks, scratch, s);
                } else {
                    rr::BatchBuild(table, n_small, big, small, big_chunks, chunks, scratch, s);
                    // Treelet restructuring of all meshes at once, each in its own work area.
                    if (restructure)
                        rr::BatchRestructure(table, n_small, chunks, max_roots, scratch,
                                             static_cast<uint8_t*>(scratch) + roots_offset, s);
                }
            }
            for (uint32_t i : plan.large) {
                if (update) rr::UpdateBlas(plan.meshes[i], results[i], s);
                else rr::BuildBlas(plan.meshes[i], restructure, scratch, results[i], s);
            }
        };
        Record(cs, cmd);
        KeepStaging(c, cs, host);
        return RR_SUCCESS;
    });
}

RRError rrxSetTuning(RRxTuning key, uint32_t value) {
    switch (rr::Raw(key)) {
        case RRX_TUNING_TRACE_MIN_BUSY:
            RR_REQUIRE(value >= 1 && value <= 64, "rrxSetTuning");
            rr::SetMinBusyLanes(value);
            return RR_SUCCESS;
        case RRX_TUNING_TRACE_RAYS_PER_FETCH:
            return rr::SetRaysPerFetch(value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_TRACE_LDS_STACK:
            return rr::SetLdsStackDepth(value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_TRACE_PLAIN:
            RR_REQUIRE(value <= 1, "rrxSetTuning");
            rr::SetTraceMode(value ? rr::kTraceModePlain : rr::kTraceModePersistent);
            return RR_SUCCESS;
        case RRX_TUNING_SORT_DIGIT_BITS:
            return rr::SetSortDigitBits((int)value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_TREELET_LEAVES:
            return rr::SetTreeletLeaves((int)value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_SORT_SPIN_LIMIT:
            rr::SetSortSpinLimit(value);
            return RR_SUCCESS;
        case RRX_TUNING_SORT_HYBRID:
            RR_REQUIRE(value <= 1, "rrxSetTuning");
            rr::SetSortHybrid(value != 0);
            return RR_SUCCESS;
        case RRX_TUNING_TRACE_RAYS_PER_LANE:
            return rr::SetRaysPerLane(value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_TRACE_2L_MIN_BUSY:
            return rr::SetMinBusyLanes2L(value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_TRACE_2L_RAYS_PER_FETCH:
            return rr::SetRaysPerFetch2L(value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
        case RRX_TUNING_TRACE_BATCH_MIN_BUSY:
            return rr::SetMinBusyLanesBatch(value) ? RR_SUCCESS : RR_ERROR_INVALID_PARAMETER;
    }
    return RR_ERROR_INVALID_PARAMETER;
}

RRError rrxGetTuning(RRxTuning key, uint32_t* value) {
    RR_REQUIRE(value, "rrxGetTuning");
    switch (rr::Raw(key)) {
        case RRX_TUNING_TRACE_MIN_BUSY: *value = rr::MinBusyLanes(); return RR_SUCCESS;
        case RRX_TUNING_TRACE_RAYS_PER_FETCH: *value = rr::RaysPerFetch(); return RR_SUCCESS;
        case RRX_TUNING_TRACE_LDS_STACK: *value = rr::LdsStackDepth(); return RR_SUCCESS;
        case RRX_TUNING_TRACE_PLAIN: *value = rr::TraceMode() == rr::kTraceModePlain ? 1u : 0u; return RR_SUCCESS;
        case RRX_TUNING_SORT_DIGIT_BITS: *value = (uint32_t)rr::SortDigitBits(); return RR_SUCCESS;
        case RRX_TUNING_TREELET_LEAVES: *value = (uint32_t)rr::TreeletLeaves(); return RR_SUCCESS;
        case RRX_TUNING_SORT_SPIN_LIMIT: *value = rr::SortSpinLimit(); return RR_SUCCESS;
        case RRX_TUNING_SORT_HYBRID: *value = rr::SortHybrid() ? 1u : 0u; return RR_SUCCESS;
        case RRX_TUNING_TRACE_RAYS_PER_LANE: *value = rr::RaysPerLane(); return RR_SUCCESS;
        case RRX_TUNING_TRACE_2L_MIN_BUSY: *value = rr::MinBusyLanes2L(); return RR_SUCCESS;
        case RRX_TUNING_TRACE_2L_RAYS_PER_FETCH: *value = rr::RaysPerFetch2L(); return RR_SUCCESS;
        case RRX_TUNING_TRACE_BATCH_MIN_BUSY: *value = rr::MinBusyLanesBatch(); return RR_SUCCESS;
    }
    return RR_ERROR_INVALID_PARAMETER;
}

RRError rrxWriteAnalyzerDump(RRContext context, RRDevicePtr geometry, RRDevicePtr rays, uint32_t width, uint32_t height,
                             const char* bvh_path, const char* rays_path, const char* config_path) {
    RR_LOG(RR_LOG_LEVEL_INFO, "rrxWriteAnalyzerDump");
    RR_REQUIRE(LiveCtx(context) && LivePtr(geometry) && OptPtr(rays) && bvh_path && rays_path && config_path,
               "rrxWriteAnalyzerDump");
    RR_REQUIRE(!rays || (uint64_t)width * height > 0, "rrxWriteAnalyzerDump");
    return Guarded("rrxWriteAnalyzerDump", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        // Everything enqueued so far (the build) must have finished.
        Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const uint8_t* accel = static_cast<const uint8_t*>(Addr(geometry));
        rr::AccelHeader h;
        Check(hipMemcpy(&h, accel, sizeof(h), hipMemcpyDeviceToHost), "hipMemcpy");
        if (h.magic != rr::kAccelMagic) return RR_ERROR_INVALID_PARAMETER;
        // The analyzer reads single-level BVHs only (bvh_analyzer/main.cpp:77).
        if (h.kind != rr::kAccelBlas) return RR_ERROR_NOT_IMPLEMENTED;
        std::vector<uint8_t> nodes(size_t(h.node_count) * sizeof(rr::BvhNode));
        if (!nodes.empty())
            Check(hipMemcpy(nodes.data(), accel + h.nodes_offset, nodes.size(), hipMemcpyDeviceToHost), "hipMemcpy");
        auto write = [](const char* path, const void* data, size_t bytes) {
            FILE* f = std::fopen(path, "wb");
            if (!f) return false;
            const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
            return std::fclose(f) == 0 && ok;
        };
        if (!write(bvh_path, nodes.data(), nodes.size())) return RR_ERROR_INTERNAL;
        if (rays) {
            std::vector<uint8_t> r(size_t(width) * height * 32);
            Check(hipMemcpy(r.data(), Addr(rays), r.size(), hipMemcpyDeviceToHost), "hipMemcpy");
            if (!write(rays_path, r.data(), r.size())) return RR_ERROR_INTERNAL;
        }
        // bvh_analyzer/config.h:46-68: bvh path, type, internal node count,
        // triangle count, rays path, ray width, ray height.
        FILE* f = std::fopen(config_path, "w");
        if (!f) return RR_ERROR_INTERNAL;
        std::fprintf(f, "%s\nvkbvh2\n%u\n%u\n%s\n%u\n%u\n", bvh_path, h.prim_count ? h.prim_count - 1 : 0, h.prim_count,
                     rays_path, width, height);
        return std::fclose(f) == 0 ? RR_SUCCESS : RR_ERROR_INTERNAL;
    });
}

RRError rrxGetDeviceErrors(RRContext context, uint32_t* errors) {
    RR_REQUIRE(LiveCtx(context) && errors, "rrxGetDeviceErrors");
    return Guarded("rrxGetDeviceErrors", [&]() -> RRError {
        DeviceGuard g(AsCtx(context)->device);
        Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        *errors = rr::TakeDeviceErrors();
        return RR_SUCCESS;
    });
}

}  // extern "C"

int rr::ContextDevice(RRContext context) { return AsCtx(context)->device; }
bool rr::IsLiveContext(RRContext context) { return LiveCtx(context); }
