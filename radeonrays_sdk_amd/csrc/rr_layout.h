// rr_layout.h — memory layout of acceleration-structure buffers in HBM.
//
// Shared by host launchers and device kernels.  Everything here is plain
// data: no HIP types, so the same definitions describe what the C ABI
// hands out (opaque result buffers sized by rrGet*MemoryRequirements).
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace rr {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;       // radeonrays.h:49-52 / vlk kernels/common.h:21
constexpr uint32_t kTopLevelSentinel = 0xFFFFFFFEu; // isect_2l.comp:28

// BVH2 node, 64 bytes, in the field order of vlk/kernels/bvh2.h:25-35 so
// that a dump of the node array is readable by bvh_analyzer's VkBvhNode
// reader (bvh_analyzer/transform.h:31-41).
//  internal node: aabb0 = box of child0, aabb1 = box of child1;
//  leaf:          child0 = kInvalid, child1 = primitive id, v0/v1/v2 = the
//                 triangle's vertices (vlk/kernels/bvh2.h:22-23).
//  `update` of an internal node counts refit arrivals modulo 2: 1 tells the
//  second child that the first one is done, and the second's increment
//  wraps it back to 0 (it is 0 between refits).
// Internal nodes occupy [0, N-1), leaves [N-1, 2N-1); the root is node 0
// (lbvh_emit_hierarchy_mesh.comp:31-32).  With N == 1 the root is a leaf.
struct BvhNode {
    float    a0_min_or_v0[3];
    uint32_t child0;
    float    a0_max_or_v1[3];
    uint32_t child1;
    float    a1_min_or_v2[3];
    uint32_t parent;
    float    a1_max_or_v3[3];
    uint32_t update;
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be 64 bytes");

// Header at offset 0 of every result buffer.  The reference keeps a host map
// from scene buffer to BLAS list (vlk/intersector.cpp:263,289-291); here the
// buffer describes itself, so traversal needs no host state and a TLAS
// carries its own table of BLAS addresses (no 2048-instance cap).
// kAccelUnknown is never stored: it asks traversal to read the kind from the
// header on the device (a buffer the host has not seen built).
enum AccelKind : uint32_t { kAccelUnknown = 0, kAccelBlas = 1, kAccelTlas = 2 };
constexpr uint32_t kAccelMagic = 0x53415252u;  // "RRAS"

struct AccelHeader {
    uint32_t magic;
    uint32_t kind;          // AccelKind
    uint32_t prim_count;    // triangles (BLAS) or instances (TLAS)
    uint32_t node_count;    // 2*prim_count - 1, or 0 when empty
    uint64_t nodes_offset;  // byte offset of BvhNode[node_count]
    uint64_t transforms_offset;  // TLAS: Transform[2*I], inverse at 2i, forward at 2i+1
    uint64_t blas_table_offset;  // TLAS: uint64 device address of each instance's BLAS buffer
    uint32_t flags;         // kAccelFlag*
    uint32_t build_serial;  // increments on every build/refit of this buffer
    uint32_t reserved[20];
};
static_assert(sizeof(AccelHeader) == 128, "AccelHeader must be 128 bytes");

constexpr uint32_t kAccelFlagRestructured = 1u;
// The BLAS's primitives are axis-aligned boxes (RR_PRIMITIVE_TYPE_AABB_LIST):
// each leaf stores its box's min corner in v0 and max corner in v1 and v2.
constexpr uint32_t kAccelFlagAabbs = 2u;

// Row-major 3x4 affine transform, 48 bytes (vlk/kernels/common.h:45-50).
struct Transform {
    float m[3][4];
};
static_assert(sizeof(Transform) == 48, "Transform must be 48 bytes");

__host__ __device__ inline size_t AlignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr uint32_t NodeCount(uint32_t prims) { return prims ? 2u * prims - 1u : 0u; }

// Result buffer sizes (cf. hlbvh_builder.cpp:376-377 and
// hlbvh_top_level_builder.cpp:378-381, plus the 128-byte header and, for a
// TLAS, the BLAS address table).
inline size_t BlasResultSize(uint32_t tris) {
    return sizeof(AccelHeader) + AlignUp(size_t(NodeCount(tris)) * sizeof(BvhNode), 256);
}
inline size_t TlasNodesBytes(uint32_t inst) { return AlignUp(size_t(NodeCount(inst)) * sizeof(BvhNode), 256); }
inline size_t TlasTransformsBytes(uint32_t inst) { return AlignUp(size_t(inst) * 2 * sizeof(Transform), 256); }
inline size_t TlasResultSize(uint32_t inst) {
    return sizeof(AccelHeader) + TlasNodesBytes(inst) + TlasTransformsBytes(inst) + AlignUp(size_t(inst) * 8, 256);
}

}  // namespace rr
