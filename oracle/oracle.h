/*
 * oracle.h — CPU oracle for the build-and-intersect path.  TEST
 * INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it as the checker; the library never does.
 *
 * Buffers use the library's accelerations-structure layout: a 128-byte
 * header followed by 64-byte BVH2 nodes (bvh_analyzer/transform.h:31-41
 * VkBvhNode field order); a TLAS adds inverse/forward transforms and a table
 * of BLAS addresses.
 */
#ifndef RR_ORACLE_H
#define RR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-ray result of a query. */
typedef struct {
    float    u, v;
    uint32_t inst_id, prim_id;
    float    t;
    uint32_t pad[3];
} OrcHit;

/* Traversal counters (bvh_analyzer/intersection_primitives.h:70-100). */
typedef struct {
    uint32_t internal_nodes, leaves, box_tests, tri_tests;
} OrcStats;

/* Morton codes of triangle-box centroids and mesh bounds (a1-a3). */
void orc_mesh_bounds(const float* vertices, uint32_t stride_floats, const uint32_t* indices, uint32_t tris,
                     float lo[3], float hi[3]);
void orc_morton_codes(const float* vertices, uint32_t stride_floats, const uint32_t* indices, uint32_t tris,
                      uint32_t* codes);
uint32_t orc_morton3(float x, float y, float z);

/* Stable sort of (key, value) pairs by key (a4-a6). */
void orc_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n);

/* OpenMP threads of the build stages that run in parallel (bounds, Morton
 * codes, sort, emission) and of traversal without a thread count; results
 * do not depend on it. */
void orc_set_threads(int threads);

/* Karras hierarchy over sorted codes into nodes[2n-1] (a7): writes child0,
 * child1, parent, update of internal nodes and child0/child1/parent of leaves. */
void orc_emit_hierarchy(const uint32_t* sorted_codes, const uint32_t* sorted_prims, uint32_t n, void* nodes);

/* Refit (a8/a10): leaves get their triangle's vertices, internal nodes the
 * boxes of their children. */
void orc_refit(const float* vertices, uint32_t stride_floats, const uint32_t* indices, uint32_t tris, void* nodes);

/* Treelet restructuring, three rounds with minimum root sizes 64/128/256 (a9),
 * 7-leaf treelets; orc_restructure_leaves takes the treelet size (5, 7, 9). */
void orc_restructure(uint32_t tris, void* nodes);
void orc_restructure_leaves(uint32_t tris, void* nodes, int leaves);

/* Dynamic programme of one 7-leaf treelet: boxes are 7 x (lo xyz, hi xyz);
 * writes the optimal cost and split of each of the 128 leaf subsets. */
void orc_treelet_dp(const float* boxes, float* cost, uint32_t* part);

/* Full BLAS build into `buffer` (orc_blas_bytes(tris) bytes). */
size_t orc_blas_bytes(uint32_t tris);
void orc_build_blas(const float* vertices, uint32_t stride_floats, const uint32_t* indices, uint32_t tris,
                    int restructure, void* buffer);
void orc_build_blas_leaves(const float* vertices, uint32_t stride_floats, const uint32_t* indices, uint32_t tris,
                           int restructure, int leaves, void* buffer);

/* TLAS over instances (a11): xforms are row-major 3x4 object-to-world,
 * blas_addr the device address recorded for each instance, blas_host the
 * host copy of that BLAS buffer (to read its root box). */
size_t orc_tlas_bytes(uint32_t instances);
void orc_build_tlas(const float* xforms, const uint64_t* blas_addr, const void* const* blas_host, uint32_t instances,
                    void* buffer);

/* Traversal (a12/a13, a15) of a BLAS or TLAS buffer.  For a TLAS,
 * blas_host[i] is the host copy of instance i's BLAS.  any != 0 selects
 * any-hit.  stats may be NULL.  Runs rays in parallel over `threads`
 * OpenMP threads (0: all). */
void orc_trace(const void* accel, const void* const* blas_host, const float* rays, uint32_t ray_count, int any,
               OrcHit* hits, OrcStats* stats, int threads);

/* Brute force over every triangle (every instance's triangles for a TLAS)
 * with the same FP32 watertight test and (t, inst, prim) ordering. */
void orc_brute_force(const void* accel, const void* const* blas_host, const float* rays, uint32_t ray_count,
                     OrcHit* hits, int threads);

/* FP64 reference for one triangle: exact-ish t, u, v of the ray's plane
 * intersection (double precision, Moller-Trumbore).  Returns 1 if the
 * double-precision hit lies inside the triangle (within eps of an edge). */
int orc_triangle_fp64(const float* ray, const float* v0, const float* v1, const float* v2, double eps, double* t,
                      double* u, double* v);

/* Validators over a BLAS/TLAS node array. */
int orc_is_valid(const void* accel);            /* bvh_analyzer/bvh.h:129-184 */
int orc_check_consistency(const void* accel);   /* test/test_vk/hlbvh_test.h:69-100 */
float orc_sah(const void* accel);               /* bvh_analyzer/bvh.h:186-223 */
uint32_t orc_height(const void* accel);

/* Dump a BLAS's nodes in bvh_analyzer's raw VkBvhNode format and write its
 * 7-line config (bvh_analyzer/config.h:46-68).  Returns 0 on success. */
int orc_write_analyzer_dump(const void* accel, const float* rays, uint32_t width, uint32_t height,
                            const char* bvh_path, const char* rays_path, const char* config_path);

#ifdef __cplusplus
}
#endif

#endif
