// build.h — host launchers for acceleration-structure builds.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "rr_layout.h"

namespace rr {

// A triangle mesh in device memory (RRTriangleMeshPrimitive, radeonrays.h:182-192).
struct MeshDesc {
    const void* vertices = nullptr;
    uint32_t    vertex_count = 0;
    uint32_t    vertex_stride = 12;
    const void* indices = nullptr;
    uint32_t    index16 = 0;      // 1: uint16 indices, 0: uint32
    uint32_t    tri_count = 0;
    uint32_t    accel_flags = 0;  // extra AccelHeader flags of the result (kAccelFlagAabbs)
};

// An AABB list in device memory (RRAABBListPrimitive, radeonrays.h:201-206):
// count boxes of 6 floats (min xyz, max xyz) every `stride` bytes.
struct AabbDesc {
    const void* aabbs = nullptr;
    uint32_t    count = 0;
    uint32_t    stride = 24;
};
// An AABB list is built as the unindexed "triangles" (min, max, min), whose
// bounds, centroids, leaf boxes and treelet boxes are exactly the boxes';
// the result is flagged kAccelFlagAabbs so traversal tests leaf boxes.
// Scratch: the gathered corners (36 bytes per box), then the mesh build's.
size_t AabbScratchBytes(uint32_t count, bool build, bool restructure);
void BuildAabbs(const AabbDesc& boxes, bool restructure, void* scratch, void* result, hipStream_t stream);
void UpdateAabbs(const AabbDesc& boxes, void* scratch, void* result, hipStream_t stream);

// One instance as uploaded for a TLAS build: forward object-to-world
// transform and the device address of the instance's BLAS buffer.
struct InstanceUpload {
    Transform xform;
    uint64_t  blas;
    uint64_t  pad;
};
static_assert(sizeof(InstanceUpload) == 64, "InstanceUpload must be 64 bytes");

size_t BlasBuildScratchBytes(uint32_t tris, bool restructure);
// Scratch RestructureBlas needs for a BLAS of `tris` triangles.
size_t RestructureScratchBytes(uint32_t tris);
size_t TlasBuildScratchBytes(uint32_t instances);

// LBVH build of one mesh into `result` (BlasResultSize bytes), optionally
// followed by treelet restructuring.  All work is enqueued on `stream`.
void BuildBlas(const MeshDesc& mesh, bool restructure, void* scratch, void* result, hipStream_t stream);

// Refit of a BLAS built from a mesh with the same triangles and indices.
void UpdateBlas(const MeshDesc& mesh, void* result, hipStream_t stream);

// Several meshes built as one BLAS (RRGeometryBuildInput with
// primitive_count > 1): their triangles are numbered consecutively, mesh 0
// first, and gathered into an unindexed triangle list in scratch, which is
// then built (or refitted) as one mesh.
struct SubMesh {
    uint64_t vertices;
    uint64_t indices;
    uint32_t vertex_count;
    uint32_t vertex_stride;
    uint32_t index16;
    uint32_t tri_count;
    uint32_t first_tri;
    uint32_t pad[3];
};
static_assert(sizeof(SubMesh) == 48, "SubMesh must be 48 bytes");
// Scratch: gathered vertices (36 bytes per triangle), the sub-mesh table,
// then the single-mesh build scratch (none for an update).
size_t MultiMeshScratchBytes(uint32_t meshes, uint32_t tris, bool build, bool restructure);
size_t MultiMeshTableOffset(uint32_t tris);
void BuildMultiMesh(const SubMesh* table, uint32_t meshes, uint32_t tris, bool restructure, void* scratch,
                    void* result, hipStream_t stream);
void UpdateMultiMesh(const SubMesh* table, uint32_t meshes, uint32_t tris, void* scratch, void* result,
                     hipStream_t stream);

// TLAS build over `count` instances whose descriptions are already in
// device memory at `instances` (typically the head of `scratch`).
void BuildTlas(const InstanceUpload* instances, uint32_t count, void* scratch, void* result, hipStream_t stream);

// Refit of a TLAS to new instance transforms (same instance count and
// BLASes): instance boxes, transforms and the BLAS table are rewritten and
// the node boxes refitted; the topology is kept.
size_t TlasUpdateScratchBytes(uint32_t instances);
void UpdateTlas(const InstanceUpload* instances, uint32_t count, void* scratch, void* result, hipStream_t stream);

// Byte offset inside the TLAS scratch where the instance upload goes.
inline size_t TlasUploadOffset() { return 0; }

// Treelet restructuring of a built BLAS in place (three rounds, minimum
// treelet-root sizes 64, 128, 256 primitives).
void RestructureBlas(uint32_t tris, void* scratch, void* result, hipStream_t stream);

// Leaves per treelet of the restructuring (5, 7 or 9; default 7, the
// reference's).  Process-wide; setters reject other values.
int TreeletLeaves();
bool SetTreeletLeaves(int leaves);

}  // namespace rr
