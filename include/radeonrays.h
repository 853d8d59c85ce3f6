/*
 * radeonrays.h — public C ABI of the MI355X build-and-intersect library.
 *
 * This header keeps the binary interface of RadeonRays SDK 4.1
 * (src/core/include/radeonrays.h) so that existing clients recompile and link
 * against libradeonrays.so unchanged: every type name, enumerator value,
 * struct layout and entry-point signature below matches that header.  The
 * text is written for this library.  Two enumerators are additive and marked
 * as such (RR_API_HIP, RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T); a HIP interop
 * header (radeonrays_hip.h) plays the role of radeonrays_vlk.h / _dx.h.
 *
 * Execution model (same contract as the reference, radeonrays.h:267-270 and
 * radeonrays.cpp:504-529): rrCmd* calls record work into a command stream;
 * work recorded into a stream from rrAllocateCommandStream runs when the stream
 * is submitted with rrSumbitCommandStream; work recorded into a stream wrapped
 * from a client hipStream_t is enqueued on that stream immediately.
 */
#ifndef RADEONRAYS_H
#define RADEONRAYS_H

/* radeonrays.h:22-25 */
#define RR_API_MAJOR_VERSION 0x000000
#define RR_API_MINOR_VERSION 0x000001
#define RR_API_PATCH_VERSION 0x000001
#define RR_API_VERSION RR_API_MAJOR_VERSION * 1000000 + RR_API_MINOR_VERSION * 1000 + RR_API_PATCH_VERSION

/* radeonrays.h:27-35 */
#ifdef _WIN32
#ifdef RR_EXPORT_API
#define RR_API __declspec(dllexport)
#else
#define RR_API __declspec(dllimport)
#endif
#else
#define RR_API __attribute__((visibility("default")))
#endif

#include <stddef.h>
#include <stdint.h>

/* Opaque handles (radeonrays.h:37-47). */
struct _RRDevicePtr;
struct _RREvent;
struct _RRContext;
struct _RRCommandStream;

typedef uint32_t                 RRBuildFlags;
typedef uint32_t                 RRRayMask;
typedef struct _RRDevicePtr*     RRDevicePtr;
typedef struct _RRContext*       RRContext;
typedef struct _RREvent*         RREvent;
typedef struct _RRCommandStream* RRCommandStream;

/* radeonrays.h:49-52 (the enumerator's spelling is part of the ABI). */
enum
{
    RR_INAVLID_VALUE = ~0u
};

/* radeonrays.h:57-68 */
typedef enum
{
    RR_SUCCESS                    = 0,
    RR_ERROR_NOT_IMPLEMENTED      = 1,
    RR_ERROR_INTERNAL             = 2,
    RR_ERROR_OUT_OF_HOST_MEMORY   = 3,
    RR_ERROR_OUT_OF_DEVICE_MEMORY = 4,
    RR_ERROR_INVALID_API_VERSION  = 5,
    RR_ERROR_INVALID_PARAMETER    = 6,
    RR_ERROR_UNSUPPORTED_API      = 7,
    RR_ERROR_UNSUPPORTED_INTEROP  = 8
} RRError;

/* radeonrays.h:70-74, plus RR_API_HIP (additive: the only backend here). */
typedef enum
{
    RR_API_DX  = 1,
    RR_API_VK  = 2,
    RR_API_HIP = 3
} RRApi;

/* radeonrays.h:76-83 */
typedef enum
{
    RR_LOG_LEVEL_DEBUG = 1,
    RR_LOG_LEVEL_INFO  = 2,
    RR_LOG_LEVEL_WARN  = 3,
    RR_LOG_LEVEL_ERROR = 4,
    RR_LOG_LEVEL_OFF   = 5
} RRLogLevel;

/* radeonrays.h:85-94: build a new structure or refit an existing one. */
typedef enum
{
    RR_BUILD_OPERATION_BUILD  = 1,
    RR_BUILD_OPERATION_UPDATE = 2
} RRBuildOperation;

/* radeonrays.h:96-105.  PREFER_FAST_BUILD skips treelet restructuring. */
typedef enum
{
    RR_BUILD_FLAG_BITS_PREFER_FAST_BUILD = 1,
    RR_BUILD_FLAG_BITS_ALLOW_UPDATE      = 2
} RRBuildFlagBits;

/* radeonrays.h:107-117 */
typedef enum
{
    RR_PRIMITIVE_TYPE_TRIANGLE_MESH,
    RR_PRIMITIVE_TYPE_AABB_LIST,
} RRPrimitiveType;

/* radeonrays.h:119-125 */
typedef enum
{
    RR_INDEX_TYPE_UINT32,
    RR_INDEX_TYPE_UINT16
} RRIndexType;

/* radeonrays.h:127-133 */
typedef enum
{
    RR_INTERSECT_QUERY_CLOSEST = 0,
    RR_INTERSECT_QUERY_ANY     = 1
} RRIntersectQuery;

/* radeonrays.h:135-142, plus FULL_HIT_T (additive): the 32-byte RRHitT
 * record below, which carries the hit distance that RRHit lacks. */
typedef enum
{
    RR_INTERSECT_QUERY_OUTPUT_FULL_HIT,
    RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID,
    RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T
} RRIntersectQueryOutput;

/* radeonrays.h:147-153 — 32 bytes. */
typedef struct
{
    float origin[3];
    float min_t;
    float direction[3];
    float max_t;
} RRRay;

/* radeonrays.h:157-162 — 16 bytes.  uv are the barycentric weights of the
 * hit triangle's second and third vertices. */
typedef struct
{
    float    uv[2];
    uint32_t inst_id;
    uint32_t prim_id;
} RRHit;

/* Additive 32-byte hit record for RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T. */
typedef struct
{
    float    uv[2];
    uint32_t inst_id;
    uint32_t prim_id;
    float    t;
    uint32_t padding[3];
} RRHitT;

/* radeonrays.h:166-173 */
typedef struct
{
    RRBuildFlags build_flags;
    void*        backend_specific_info;
} RRBuildOptions;

/* radeonrays.h:182-192: indexed triangle mesh; each vertex starts with
 * three floats (x, y, z) and vertices are vertex_stride bytes apart. */
typedef struct
{
    RRDevicePtr vertices;
    uint32_t    vertex_count;
    uint32_t    vertex_stride;

    RRDevicePtr triangle_indices;
    uint32_t    triangle_count;

    RRIndexType index_type;
} RRTriangleMeshPrimitive;

/* radeonrays.h:201-206: boxes are (xmin, ymin, zmin, pad), (xmax, ymax, zmax, pad). */
typedef struct
{
    RRDevicePtr aabbs;
    uint32_t    aabb_count;
    uint32_t    aabb_stride;
} RRAABBListPrimitive;

/* radeonrays.h:213-226 */
typedef struct
{
    RRPrimitiveType primitive_type;
    uint32_t        primitive_count;
    union
    {
        RRTriangleMeshPrimitive* triangle_mesh_primitives;
        RRAABBListPrimitive*     aabb_primitives;
    };
} RRGeometryBuildInput;

/* radeonrays.h:228-232: transform is a row-major 3x4 object-to-world matrix. */
typedef struct
{
    RRDevicePtr geometry;
    float       transform[3][4];
} RRInstance;

/* radeonrays.h:247-253 */
typedef struct
{
    const RRInstance* instances;
    uint32_t          instance_count;
} RRSceneBuildInput;

/* radeonrays.h:254-259 */
typedef struct
{
    size_t temporary_build_buffer_size;
    size_t temporary_update_buffer_size;
    size_t result_buffer_size;
} RRMemoryRequirements;

#ifdef __cplusplus
extern "C" {
#endif

/* Create a context (radeonrays.h:276).  api must be RR_API_HIP; the context
 * uses the current HIP device and a stream of its own. */
RR_API RRError rrCreateContext(uint32_t api_version, RRApi api, RRContext* context);

/* Destroy a context and every pooled stream/event it owns (radeonrays.h:286). */
RR_API RRError rrDestroyContext(RRContext context);

/* Process-wide log level (radeonrays.h:295). */
RR_API RRError rrSetLogLevel(RRLogLevel log_level);

/* Send the log to a file instead of stderr (radeonrays.h:304). */
RR_API RRError rrSetLogFile(char const* filename);

/* Record a BLAS build (LBVH, then treelet restructuring unless build_options
 * is NULL or asks for PREFER_FAST_BUILD) or an update (refit with the same
 * topology) into command_stream (radeonrays.h:322). */
RR_API RRError rrCmdBuildGeometry(RRContext                   context,
                                  RRBuildOperation            build_operation,
                                  const RRGeometryBuildInput* build_input,
                                  const RRBuildOptions*       build_options,
                                  RRDevicePtr                 temporary_buffer,
                                  RRDevicePtr                 geometry_buffer,
                                  RRCommandStream             command_stream);

/* Sizes of the scratch and result buffers for rrCmdBuildGeometry (radeonrays.h:338). */
RR_API RRError rrGetGeometryBuildMemoryRequirements(RRContext                   context,
                                                    const RRGeometryBuildInput* build_input,
                                                    const RRBuildOptions*       build_options,
                                                    RRMemoryRequirements*       memory_requirements);

/* Record a two-level scene (instance TLAS) build (radeonrays.h:357).  The
 * geometry buffers named by the instances must outlive the scene. */
RR_API RRError rrCmdBuildScene(RRContext                context,
                               const RRSceneBuildInput* build_input,
                               const RRBuildOptions*    build_options,
                               RRDevicePtr              temporary_buffer,
                               RRDevicePtr              scene_buffer,
                               RRCommandStream          command_stream);

/* Sizes of the scratch and result buffers for rrCmdBuildScene (radeonrays.h:372). */
RR_API RRError rrGetSceneBuildMemoryRequirements(RRContext                context,
                                                 const RRSceneBuildInput* build_input,
                                                 const RRBuildOptions*    build_options,
                                                 RRMemoryRequirements*    memory_requirements);

/* Record a ray query against a geometry (BLAS) or scene (TLAS) buffer
 * (radeonrays.h:391).  When indirect_ray_count is not NULL it points to a
 * device uint32 and min(*indirect_ray_count, ray_count) rays are traced. */
RR_API RRError rrCmdIntersect(RRContext              context,
                              RRDevicePtr            scene_buffer,
                              RRIntersectQuery       query,
                              RRDevicePtr            rays,
                              uint32_t               ray_count,
                              RRDevicePtr            indirect_ray_count,
                              RRIntersectQueryOutput query_output,
                              RRDevicePtr            hits,
                              RRDevicePtr            scratch,
                              RRCommandStream        command_stream);

/* Scratch size for rrCmdIntersect with up to ray_count rays (radeonrays.h:409). */
RR_API RRError rrGetTraceMemoryRequirements(RRContext context, uint32_t ray_count, size_t* scratch_size);

/* Command streams (radeonrays.h:418-427). */
RR_API RRError rrAllocateCommandStream(RRContext context, RRCommandStream* command_stream);
RR_API RRError rrReleaseCommandStream(RRContext context, RRCommandStream command_stream);

/* Run the recorded work after wait_event (if any) and return an event that
 * signals its completion (radeonrays.h:438; the spelling is the ABI's). */
RR_API RRError rrSumbitCommandStream(RRContext       context,
                                     RRCommandStream command_stream,
                                     RREvent         wait_event,
                                     RREvent*        out_event);

/* Events (radeonrays.h:448-456). */
RR_API RRError rrReleaseEvent(RRContext context, RREvent event);
RR_API RRError rrWaitEvent(RRContext context, RREvent event);

/* Release a device pointer handle (not the memory it names) (radeonrays.h:464). */
RR_API RRError rrReleaseDevicePtr(RRContext context, RRDevicePtr ptr);

/* Release a stream wrapped from a client hipStream_t (radeonrays.h:472). */
RR_API RRError rrReleaseExternalCommandStream(RRContext context, RRCommandStream command_stream);

#ifdef __cplusplus
}
#endif

#endif /* RADEONRAYS_H */
