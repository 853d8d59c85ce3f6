// batch.h — host interface of the batched BLAS build and refit: many small
// meshes built (or refitted) by one persistent launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "build.h"

namespace rr {

// Meshes up to this many triangles go through the batched kernels; larger
// ones are built one at a time by BuildBlas.
constexpr uint32_t kBatchMaxTris = 1u << 16;

// One mesh of a batch as uploaded to the device (64 bytes).
struct BatchMesh {
    uint64_t vertices;
    uint64_t indices;
    uint64_t result;   // BLAS result buffer
    uint64_t work;     // this mesh's key/value work area (BatchWorkBytes)
    uint32_t vertex_count;
    uint32_t vertex_stride;
    uint32_t index16;
    uint32_t tri_count;
    uint32_t flags;        // header flags of the result (kAccelFlag*)
    uint32_t first_chunk;  // refit chunks (kRefitChunk leaves) of the meshes before this one
    uint32_t pad[2];
};
static_assert(sizeof(BatchMesh) == 64, "BatchMesh must be 64 bytes");

// Bytes of a mesh's work area: keys, values and their ping-pong copies, or
// the treelet-restructuring scratch if that is larger.
size_t BatchWorkBytes(uint32_t tris);

// Scratch head of a batch: queue counter, then the mesh table.
constexpr size_t kMeshTableOffset = 256;
inline size_t BatchHeadBytes(uint32_t meshes) {
    return kMeshTableOffset + AlignUp(size_t(meshes) * sizeof(BatchMesh), 256);
}

// Leaves per refit queue entry.
constexpr uint32_t kRefitChunk = 64;
inline uint32_t RefitChunks(uint32_t tris) { return (tris + kRefitChunk - 1) / kRefitChunk; }

// Meshes of at most this many triangles are built with their whole sort in
// LDS (one 1024-thread workgroup per mesh; 256 threads for meshes of at
// most kBatchLdsSmallMaxTris).
constexpr uint32_t kBatchLdsMaxTris = 16384;
constexpr uint32_t kBatchLdsSmallMaxTris = 4096;

// LBVH build of every mesh in the table (already in device memory at
// `table`, inside `scratch`).  The first `big` meshes (those above
// kBatchLdsMaxTris triangles; the table is ordered largest first): each
// workgroup takes the next mesh from a queue and runs bounds, Morton codes
// and a stable four-pass radix sort through the mesh's work area; then one
// thread per node of these meshes (in chunks of kRefitChunk leaves, the
// first big_chunks chunks) emits the hierarchy.  Every other mesh is sorted
// by one workgroup in LDS, which then emits its hierarchy from the
// LDS-resident codes; meshes from index `small` on (at most
// kBatchLdsSmallMaxTris triangles) by 256-thread workgroups.  Last,
// BatchRefit over all total_chunks chunks.  The results equal BuildBlas
// with restructure off, mesh by mesh.
void BatchBuild(const BatchMesh* table, uint32_t meshes, uint32_t big, uint32_t small, uint32_t big_chunks,
                uint32_t total_chunks, void* scratch, hipStream_t stream);

// Refit of every mesh in the table (the topology of each BLAS is kept), as
// UpdateBlas mesh by mesh (one launch over the refit chunks of all meshes).
void BatchRefit(const BatchMesh* table, uint32_t meshes, uint32_t total_chunks, void* scratch, hipStream_t stream);

// Treelet restructuring of every mesh in the table with at least 64
// triangles, as RestructureBlas mesh by mesh: per round one launch finds the
// treelet roots of all meshes and one wave per root restructures.  Each
// mesh's work area holds its arrival counts; `roots` holds max_roots[0]
// (mesh, node) pairs, max_roots[r] bounding the roots of round r (the sum
// over meshes of tris / (64 << r)); its counter is scratch word
// kBatchRootsCounterWord.
constexpr uint32_t kBatchRootsCounterWord = 1;
constexpr uint32_t kBatchLdsQueueWord = 2;       // queue of the LDS-sorted meshes (1024-thread workgroups)
constexpr uint32_t kBatchLdsSmallQueueWord = 3;  // and of those for 256-thread workgroups
void BatchRestructure(const BatchMesh* table, uint32_t meshes, uint32_t total_chunks, const uint32_t max_roots[3],
                      void* scratch, void* roots, hipStream_t stream);

}  // namespace rr
