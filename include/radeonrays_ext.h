/*
 * radeonrays_ext.h — extensions of the MI355X library with no counterpart
 * in the reference API.  They expose pieces of the hot path on their own
 * (for measurement and for callers that batch work across many
 * acceleration structures); none is needed to use the rr* API.
 */
#ifndef RADEONRAYS_EXT_H
#define RADEONRAYS_EXT_H

#include "radeonrays.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scratch bytes for rrxCmdSortPairs on n pairs. */
RR_API RRError rrxGetSortMemoryRequirements(RRContext context, uint32_t n, size_t* scratch_size);

/* Stable radix sort of n (uint32 key, uint32 value) pairs whose keys are
 * below 2^key_bits (bits above key_bits must be zero), in place in
 * keys/values, using keys_tmp / values_tmp (n uint32 each) and scratch.
 * The onesweep LSD sort the BVH builder runs on Morton codes, or on
 * 2^22..2^24 keys one most-significant-digit pass followed by a per-bucket
 * sort (RRX_TUNING_SORT_HYBRID); the order is the same.  Reference
 * counterpart: the internal RadixSortKeyValue of vlk/radix_sort.cpp, which
 * the public API does not expose. */
RR_API RRError rrxCmdSortPairs(RRContext       context,
                               RRDevicePtr     keys,
                               RRDevicePtr     values,
                               RRDevicePtr     keys_tmp,
                               RRDevicePtr     values_tmp,
                               uint32_t        n,
                               uint32_t        key_bits,
                               RRDevicePtr     scratch,
                               RRCommandStream command_stream);

/* rrxCmdSortPairs without a scratch argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * must not allocate).  No counterpart in the reference.  Its scratch
 * (histograms, tickets, look-back words: 4,329,984 bytes) is a static device
 * array that every GPU the library runs on holds until the process ends;
 * the call allocates nothing.  At most 16,777,216 pairs (more gives
 * RR_ERROR_INVALID_PARAMETER); the same stable order as rrxCmdSortPairs, with
 * its default 8-bit LSD passes whatever the sort tuning.  All such sorts on
 * one GPU share the array: calls that overlap in time on one GPU race
 * (enqueue them on one stream, or order their streams).  The context must
 * have been created on the GPU the call runs on. */
RR_API RRError rrxCmdSortPairsNoAlloc(RRContext       context,
                                      RRDevicePtr     keys,
                                      RRDevicePtr     values,
                                      RRDevicePtr     keys_tmp,
                                      RRDevicePtr     values_tmp,
                                      uint32_t        n,
                                      uint32_t        key_bits,
                                      RRCommandStream command_stream);

/* Closest- or any-hit query like rrCmdIntersect with FULL_HIT_T output that
 * also writes per-ray traversal counters to `stats` (4 uint32 per ray: inner
 * nodes visited, leaves visited, box tests, triangle tests), the figures
 * bvh_analyzer reports as TraversalStats (bvh_analyzer/intersection_primitives.h:70-100). */
RR_API RRError rrxCmdIntersectWithStats(RRContext        context,
                                        RRDevicePtr      scene_buffer,
                                        RRIntersectQuery query,
                                        RRDevicePtr      rays,
                                        uint32_t         ray_count,
                                        RRDevicePtr      hits,
                                        RRDevicePtr      scratch,
                                        RRDevicePtr      stats,
                                        RRCommandStream  command_stream);

/* Process-wide tuning of the persistent traversal kernel and of the sort
 * (the autotuner in radeonrays_sdk_amd/autotune.py sweeps these).  Results
 * do not depend on them, only speed does. */
typedef enum RRxTuning
{
    RRX_TUNING_TRACE_MIN_BUSY = 0,        /* refill a wave once fewer lanes traverse (1..64; default 24) */
    RRX_TUNING_TRACE_RAYS_PER_FETCH = 1,  /* ray indices per queue atomic (32, 64, 128; default 128) */
    RRX_TUNING_TRACE_LDS_STACK = 2,       /* LDS stack entries per lane (8, 16, 32; default 16) */
    RRX_TUNING_TRACE_PLAIN = 3,           /* 1: one lane per ray for its whole traversal, no
                                             persistent kernels (default 0, or 1 when the
                                             environment sets RR_TRACE_MODE=plain) */
    RRX_TUNING_SORT_DIGIT_BITS = 4,       /* digit width of rrxCmdSortPairs on >= 2^20 keys:
                                             8 (default), 10 (three passes for 30-bit keys)
                                             or 4 (the reference's radix, eight passes) */
    RRX_TUNING_TREELET_LEAVES = 5,        /* leaves per treelet of the restructuring that builds
                                             without RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD run:
                                             5, 7 (default, the reference's) or 9; changes the
                                             tree (so any-hit may report another hit), not the
                                             closest hits */
    RRX_TUNING_SORT_SPIN_LIMIT = 6,       /* test hook, not a speed setting: bound of a sort
                                             tile's wait for its predecessors, in polls that find
                                             one unpublished (default 2^22; 0 gives up at the
                                             first).  A tile that gives up sorts wrongly and the
                                             failure is reported (rrxGetDeviceErrors) */
    RRX_TUNING_SORT_HYBRID = 7,           /* 1: rrxCmdSortPairs on 2^22..2^24 keys takes one
                                             10-bit most-significant-digit pass and then sorts each
                                             bucket in one workgroup's LDS (two passes over the data
                                             instead of four); a bucket of more than 20,480 keys is
                                             sorted by its workgroup through memory, so keys
                                             concentrated in a few buckets sort slower.  0 (default):
                                             the LSD passes.  Same order either way */
    RRX_TUNING_TRACE_RAYS_PER_LANE = 8,   /* k sizing a single query's persistent grid (0..64;
                                             default 2): at most four 256-lane workgroups per CU
                                             unless 512 x k rays per workgroup need more; below
                                             that a workgroup per 256 x k rays, at least one per CU
                                             (single level) or one ray per lane (two level); never
                                             more than one ray per lane or the resident grid.
                                             Queries that do not fill the resident grid finish
                                             sooner on fewer waves per SIMD.  0: the resident grid
                                             whenever the rays fill it */
    RRX_TUNING_TRACE_2L_MIN_BUSY = 9,     /* RRX_TUNING_TRACE_MIN_BUSY for the two-level kernels
                                             (rrCmdIntersect on a scene, scene requests of
                                             rrxCmdIntersectBatch; 1..64; default 32);
                                             RRX_TUNING_TRACE_MIN_BUSY sets the persistent
                                             single-level kernel's */
    RRX_TUNING_TRACE_2L_RAYS_PER_FETCH = 10, /* RRX_TUNING_TRACE_RAYS_PER_FETCH for the persistent
                                             two-level kernel (32, 64, 128; default 32) */
    RRX_TUNING_TRACE_BATCH_MIN_BUSY = 11  /* RRX_TUNING_TRACE_MIN_BUSY for the geometry requests
                                             of rrxCmdIntersectBatch (1..64; default 24; batches of
                                             many small BLASes run faster at 12) */
} RRxTuning;

RR_API RRError rrxSetTuning(RRxTuning key, uint32_t value);
RR_API RRError rrxGetTuning(RRxTuning key, uint32_t* value);

/* Failures the device reported since they were last read (bit k set for
 * kind k), for work enqueued on client streams, where no rrWaitEvent call
 * sees them; rrWaitEvent reports the same failures as RR_ERROR_INTERNAL.
 * Waits for the device, then clears what it returns.  The reference has no
 * counterpart (its only failure mode on the device is a wait that never
 * ends, vlk/device.cpp:249-262). */
#define RRX_DEVICE_ERROR_STACK_OVERFLOW   (1u << 0) /* a ray needed more than 80 stack entries; it ended early */
#define RRX_DEVICE_ERROR_SORT_LOOKBACK    (1u << 1) /* a sort tile gave up waiting; that sort's order is wrong */
#define RRX_DEVICE_ERROR_INVALID_BUFFER   (1u << 2) /* a query named a buffer without a valid header
                                                       (rrCmdIntersect then writes no hits), or a batched
                                                       request named one whose header disagrees with the
                                                       level the context recorded for that address
                                                       (rrxCmdIntersectBatch answers it with misses) */
RR_API RRError rrxGetDeviceErrors(RRContext context, uint32_t* errors);

/* Scratch bytes for rrxCmdBuildGeometryBatch over `count` build inputs
 * (each one triangle mesh, as for rrGetGeometryBuildMemoryRequirements).
 * Each geometry's result buffer is sized by rrGetGeometryBuildMemoryRequirements
 * for that input alone. */
RR_API RRError rrxGetGeometryBatchBuildMemoryRequirements(RRContext                   context,
                                                          RRGeometryBuildInput const* build_inputs,
                                                          uint32_t                    count,
                                                          RRBuildOptions const*       build_options,
                                                          size_t*                     scratch_size);

/* Build (or, with RR_BUILD_OPERATION_UPDATE, refit) `count` geometries as one
 * command: meshes of up to 65,536 triangles are built by one persistent
 * launch in which each workgroup builds whole meshes taken from a queue;
 * larger meshes follow one at a time.  geometry_buffers[i] receives exactly
 * what rrCmdBuildGeometry (radeonrays.h:322-327) writes for build_inputs[i]
 * with the same options.  No counterpart in the reference, which builds
 * one geometry per call (vlk/intersector.cpp:124-182). */
RR_API RRError rrxCmdBuildGeometryBatch(RRContext                   context,
                                        RRBuildOperation            build_operation,
                                        RRGeometryBuildInput const* build_inputs,
                                        uint32_t                    count,
                                        RRBuildOptions const*       build_options,
                                        RRDevicePtr                 temporary_buffer,
                                        RRDevicePtr const*          geometry_buffers,
                                        RRCommandStream             command_stream);

/* Refit a scene built by rrCmdBuildScene (radeonrays.h:357-362) to new
 * instance transforms: build_input must name the same instance count and
 * geometries as the build; their transforms may change.  The scene keeps its
 * topology and gets the new transforms and refitted boxes, so hits are those
 * of the moved instances.  temporary_buffer needs
 * temporary_update_buffer_size bytes from rrGetSceneBuildMemoryRequirements.
 * No counterpart in the reference, which rebuilds scenes only. */
RR_API RRError rrxCmdUpdateScene(RRContext                context,
                                 RRSceneBuildInput const* build_input,
                                 RRDevicePtr              temporary_buffer,
                                 RRDevicePtr              scene_buffer,
                                 RRCommandStream          command_stream);

/* Write a geometry (BLAS) in the dump format of the reference's bvh_analyzer:
 * the VkBvhNode array to bvh_path, width x height RRRay records from `rays`
 * (may be NULL: no rays file is written) to rays_path, and the analyzer's
 * 7-line config (bvh_analyzer/config.h:46-68) to config_path.  Waits for the
 * device first.  The reference writes the same files from its tests
 * (test/test_vk/basic_test.h:519-555); scenes (TLAS) are not readable by the
 * analyzer and give RR_ERROR_NOT_IMPLEMENTED. */
RR_API RRError rrxWriteAnalyzerDump(RRContext   context,
                                    RRDevicePtr geometry,
                                    RRDevicePtr rays,
                                    uint32_t    width,
                                    uint32_t    height,
                                    const char* bvh_path,
                                    const char* rays_path,
                                    const char* config_path);

/* One query of a batch: the arguments of one rrCmdIntersect call
 * (radeonrays.h:391-400) without the indirect ray count. */
typedef struct RRxIntersectRequest
{
    RRDevicePtr            scene_buffer;  /* geometry (BLAS) or scene (TLAS) result buffer */
    RRDevicePtr            rays;          /* ray_count RRRay records */
    RRDevicePtr            hits;          /* ray_count records of query_output's type */
    uint32_t               ray_count;
    RRIntersectQuery       query;
    RRIntersectQueryOutput query_output;
    uint32_t               padding;
} RRxIntersectRequest;

/* Scratch bytes for rrxCmdIntersectBatch over request_count requests holding
 * total_ray_count rays, none with more than max_request_rays rays. */
RR_API RRError rrxGetIntersectBatchMemoryRequirements(RRContext context,
                                                      uint32_t  request_count,
                                                      uint64_t  total_ray_count,
                                                      uint32_t  max_request_rays,
                                                      size_t*   scratch_size);

/* Run request_count queries as one submission.  Every geometry (BLAS)
 * request shares one persistent traversal launch whose waves take 64-ray
 * chunks of any request, so small requests fill the GPU together, and every
 * scene (TLAS) request shares a second, two-level one.  The level of a
 * buffer this context built comes from the context's record of it, of any
 * other buffer from its header; a buffer whose header disagrees with the
 * record (another buffer put at a freed result's address) is reported as
 * RRX_DEVICE_ERROR_INVALID_BUFFER and answered with misses.  Each request's
 * hits are those an rrCmdIntersect call with the same arguments writes.  No counterpart in the
 * reference API (its rrCmdIntersect runs one query per dispatch,
 * vlk/intersector.cpp:269-339).  The request array is copied during the
 * call. */
RR_API RRError rrxCmdIntersectBatch(RRContext                  context,
                                    const RRxIntersectRequest* requests,
                                    uint32_t                   request_count,
                                    RRDevicePtr                scratch,
                                    RRCommandStream            command_stream);

/* ------------------------------------------------------------------------
 * Serving loop: a host thread that admits queued requests under a memory
 * budget and dispatches them to the GPU.  In batched mode every dispatch is
 * one rrxCmdIntersectBatch over all admitted requests (up to the batch
 * limits) on one of max_inflight streams, so requests that arrive while
 * others run are coalesced into the next launch and launches overlap.  In
 * plain mode every request is one rrCmdIntersect + rrSumbitCommandStream +
 * rrWaitEvent, one after another: the per-call path, for comparison.
 * Times are nanoseconds of the server's steady clock (rrxServerNow). */
typedef struct _RRxServer* RRxServer;

typedef enum RRxServeMode
{
    RRX_SERVE_BATCHED = 0,
    RRX_SERVE_PLAIN = 1
} RRxServeMode;

typedef struct RRxServerOptions
{
    RRxServeMode mode;
    uint32_t     max_inflight;        /* concurrent launches (batched mode), >= 1 */
    uint32_t     max_batch_requests;  /* requests per launch */
    uint32_t     max_batch_rays;      /* rays per launch; also the largest request */
    uint64_t     memory_budget;       /* bytes of admitted ray + hit buffers (accounting) */
    uint64_t     staging_bytes;       /* device arena for host-memory requests (rrxServerSubmitHost),
                                         allocated with the server; 0: host requests are refused */
} RRxServerOptions;

typedef struct RRxServedTiming
{
    uint64_t arrival_ns;   /* max(release time, submission time) */
    uint64_t dispatch_ns;  /* launch enqueued */
    uint64_t complete_ns;  /* completion seen by the server */
    uint32_t batch_id;     /* dispatch the request went out in */
    uint32_t batch_size;   /* requests in that dispatch */
} RRxServedTiming;

RR_API RRError rrxCreateServer(RRContext context, const RRxServerOptions* options, RRxServer* server);
RR_API RRError rrxDestroyServer(RRxServer server);
RR_API uint64_t rrxServerNow(RRxServer server);

/* Queue count requests; request i is admitted no earlier than release_ns[i]
 * (null: now).  Tickets first_ticket .. first_ticket + count - 1 name them.
 * Requests larger than max_batch_rays or the memory budget are rejected. */
RR_API RRError rrxServerSubmit(RRxServer                  server,
                               const RRxIntersectRequest* requests,
                               const uint64_t*            release_ns,
                               uint32_t                   count,
                               uint64_t*                  first_ticket);

/* A request whose rays and hits are in host memory (pinned memory copies
 * fastest).  The server admits it only when space for its rays and hits is
 * free in its device staging arena (staging_bytes), so the arena bounds the
 * device memory of host requests in flight: at dispatch its rays are copied
 * into the arena, traced there with the rest of the launch, and its hits
 * copied back to `hits`; the space is freed when the hits are on the host. */
typedef struct RRxHostIntersectRequest
{
    RRDevicePtr            scene_buffer;  /* geometry (BLAS) or scene (TLAS) result buffer */
    const void*            rays;          /* ray_count RRRay records, host memory */
    void*                  hits;          /* ray_count records of query_output's type, host memory */
    uint32_t               ray_count;
    RRIntersectQuery       query;
    RRIntersectQueryOutput query_output;
    uint32_t               padding;
} RRxHostIntersectRequest;

/* Queue host-memory requests, as rrxServerSubmit.  Requests larger than
 * max_batch_rays or than the staging arena are rejected. */
RR_API RRError rrxServerSubmitHost(RRxServer                      server,
                                   const RRxHostIntersectRequest* requests,
                                   const uint64_t*                release_ns,
                                   uint32_t                       count,
                                   uint64_t*                      first_ticket);

/* Block until every request submitted so far has completed. */
RR_API RRError rrxServerDrain(RRxServer server);

/* Timings of tickets [first_ticket, first_ticket + count), completed ones only. */
RR_API RRError rrxServerGetTimings(RRxServer server, uint64_t first_ticket, uint32_t count, RRxServedTiming* timings);

/* Peak bytes admitted at once since creation. */
RR_API RRError rrxServerGetPeakAdmitted(RRxServer server, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* RADEONRAYS_EXT_H */
