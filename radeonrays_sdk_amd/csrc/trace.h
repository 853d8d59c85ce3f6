// trace.h — host interface of the ray traversal kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "rr_layout.h"

namespace rr {

// Per-lane traversal stack of kStackCapacity entries: a ring of
// kLdsStackEntries (or the persistent kernel's tuned depth) in LDS, the rest
// spilled to the lane's kSpillEntries slots of trace scratch (the reference
// keeps 16 in LDS and 64 per ray in scratch, vlk/kernels/isect.comp:26-27,
// geometry_trace.cpp:169).  The capacity is the same for every LDS depth, so
// tuning never changes which rays a too-deep tree overflows; such a ray
// ends early and sets the stack-overflow error word (device_errors.h).
constexpr int kStackCapacity = 80;
constexpr int kLdsStackEntries = 16;
constexpr int kMinLdsStackEntries = 8;
constexpr int kSpillEntries = kStackCapacity - kMinLdsStackEntries;

size_t TraceScratchBytes(uint32_t ray_count);

// Queries run the persistent kernels (single- or two-level) unless the mode
// is plain (one lane per ray, the exact depth-first loop); RR_TRACE_MODE=plain
// selects it at start-up.
enum TraceModeKind : int { kTraceModePersistent = 0, kTraceModePlain = 1 };
int TraceMode();
void SetTraceMode(int mode);

// Persistent kernel: a wave goes back to fetch rays for its idle lanes once
// fewer than this many lanes are still traversing (RR_TRACE_MIN_BUSY).
constexpr uint32_t kDefaultMinBusy = 24;
uint32_t MinBusyLanes();
void SetMinBusyLanes(uint32_t lanes);

// Tuning of the persistent kernel (process-wide): ray indices a wave takes
// per queue atomic (32, 64 or 128) and LDS ring entries per lane (8, 16 or
// 32).  Setters reject other values.  Defaults: the sweep's best on the
// 1M-triangle heightfield with 4M incoherent rays.
constexpr uint32_t kDefaultFetch = 128;
constexpr uint32_t kDefaultLdsStack = 16;
uint32_t RaysPerFetch();
bool SetRaysPerFetch(uint32_t rays);
// The same two knobs for the two-level kernels (the persistent one and the
// batch kernel's scene requests, which take the refill threshold), whose
// rays are longer: the sweep's best on the 64-BLAS, 4,096-instance scene.
constexpr uint32_t kDefaultMinBusy2L = 32;
constexpr uint32_t kDefaultFetch2L = 32;
uint32_t MinBusyLanes2L();
bool SetMinBusyLanes2L(uint32_t lanes);
uint32_t RaysPerFetch2L();
bool SetRaysPerFetch2L(uint32_t rays);
// Refill threshold of the batch kernel's geometry requests (64-ray chunks).
// 24 suits the serving mix (BLASes of 10k-2M triangles); batches of many
// small BLASes with short rays run faster at 12.
constexpr uint32_t kDefaultMinBusyBatch = 24;
uint32_t MinBusyLanesBatch();
bool SetMinBusyLanesBatch(uint32_t lanes);
uint32_t LdsStackDepth();
bool SetLdsStackDepth(uint32_t entries);

// Rays per lane k that size a single query's persistent grid (0..64; setter
// rejects others): at most four workgroups per CU unless 2 x kBlock x k rays
// per workgroup need more; below that a workgroup per kBlock x k rays, at
// least one per CU (single level) or one ray per lane (two level); never
// more than one ray per lane or the resident grid.  0: the resident grid
// whenever the rays fill it.
constexpr uint32_t kDefaultRaysPerLane = 2;
uint32_t RaysPerLane();
bool SetRaysPerLane(uint32_t rays);

enum QueryKind : int { kQueryClosest = 0, kQueryAny = 1 };
enum OutputKind : int { kOutFullHit = 0, kOutInstanceId = 1, kOutFullHitT = 2 };

// Trace ray_count rays (or min(ray_count, *indirect) when indirect is not
// null) against the BLAS or TLAS at `accel`.  With kind kAccelUnknown both
// levels' kernels are launched and the buffer's header selects the one that
// runs, so the host never reads the buffer.  Failures are reported through
// the device error words (device_errors.h).
void Trace(const void* accel, AccelKind kind, QueryKind query, OutputKind output, const void* rays, uint32_t ray_count,
           const uint32_t* indirect, void* hits, void* scratch, hipStream_t stream);

// One request of a batched (coalesced) query: a BLAS or a scene, its rays
// and its hit buffer.  flags: bit 0 any-hit, bits 1-2 OutputKind.
struct BatchRequest {
    uint64_t accel;
    uint64_t rays;
    uint64_t hits;
    uint32_t ray_count;
    uint32_t flags;
    uint32_t first_chunk;  // chunks of the batch's chunk size before this request
    uint32_t pad[3];
};
static_assert(sizeof(BatchRequest) == 48, "BatchRequest must be 48 bytes");
constexpr uint32_t kBatchChunk = 64;
// Rays per chunk of one batch: kBatchChunk, or kBatchChunkLarge when its
// requests average at least kBatchLargeChunkRays rays (half the chunk
// atomics and table searches for large requests; small requests keep more,
// finer pieces of work).
constexpr uint32_t kBatchChunkLarge = 128;
constexpr uint64_t kBatchLargeChunkRays = 2048;
inline uint32_t BatchChunkFor(uint64_t total_rays, uint32_t requests) {
    return requests && total_rays >= (uint64_t)requests * kBatchLargeChunkRays ? kBatchChunkLarge : kBatchChunk;
}

// Scratch for a batch of `requests` requests and `total_rays` rays: a
// 2 KB head (the two kernels' chunk counters, eight each), the request
// table (uploaded to byte kBatchTableOffset, BatchTableBytes long), then
// spill slots.
size_t TraceBatchScratchBytes(uint32_t requests, uint64_t total_rays);
constexpr size_t kBatchTableOffset = 2048;
inline size_t BatchTableBytes(uint32_t requests) {
    return ((size_t)requests * sizeof(BatchRequest) + 255) & ~(size_t)255;
}
// BatchRequest.flags bit: the context has no record of this address's level,
// so the kernels read it from the header.
constexpr uint32_t kBatchFlagCheckLevel = 8;

// Trace every request of the table (already in device memory), ordered as
// recorded geometries (geo_chunks chunks), unrecorded requests (flag
// kBatchFlagCheckLevel), recorded scenes: one persistent launch per level
// with requests present; requests may name different BVHs, queries and
// outputs.
void TraceBatch(const BatchRequest* table, uint32_t requests, uint32_t chunk, uint32_t geo_chunks,
                uint32_t unrecorded_chunks, uint32_t scene_chunks, void* scratch, hipStream_t stream);

// Same traversal with per-ray visit counters written to `stats`
// (uint32 x4 per ray: internal nodes visited, leaves visited, box tests,
// triangle tests), for the roofline's algorithmic byte count.
void TraceWithStats(const void* accel, AccelKind kind, QueryKind query, const void* rays, uint32_t ray_count, void* hits,
                    void* scratch, uint32_t* stats, hipStream_t stream);

}  // namespace rr
