// build.hip — LBVH construction, refit and TLAS build for gfx950.
//
// Behavioural spec: the reference's Vulkan LBVH path
// (vlk/hlbvh_builder.cpp:159-351 and its kernels lbvh_init_mesh.comp,
// lbvh_calc_mesh_aabb.comp, lbvh_calc_morton_codes_mesh.comp,
// lbvh_emit_hierarchy_mesh.comp, lbvh_fit_aabb_mesh.comp; the scene twins for
// the TLAS).  The launch structure is this library's own:
//   init (header, bounds, sort tickets) -> mesh bounds (wave64 reductions,
//   one set of ordered-uint atomics per block) -> Morton codes fused with the
//   radix histogram -> onesweep sort -> Karras emission -> refit, where each
//   finished node writes its box straight into its parent's child slot and
//   the parity of the parent's `update` counter tells the second arriver to
//   continue (no reset pass is needed before an update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>

#include "batch.h"
#include "build.h"
#include "build_device.h"
#include "rr_device.h"
#include "sort.h"

namespace rr {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxGrid = 4096;
// Bounds blocks: each ends with six device atomics on the same six words,
// so fewer, longer-running blocks contend less.
constexpr uint32_t kBoundsGrid = 1024;

inline uint32_t Blocks(size_t n, int per = kBlock) { return (uint32_t)std::max<size_t>(1, (n + per - 1) / per); }

// Scratch block at the head of every build scratch buffer.
// Ordered-uint bounds (min xyz, pad, max xyz, pad) in kBoundsCopies copies:
// workgroup b reduces into copy b % kBoundsCopies — one per XCD, as
// workgroups go to the XCDs round-robin — so device atomics from different
// XCDs never meet on one address; readers take the union of the copies.
constexpr int kBoundsCopies = 8;
struct BuildHead {
    uint32_t bounds[kBoundsCopies][8];
};
static_assert(sizeof(BuildHead) == 256, "the build head fills its 256-byte scratch slot");

struct BlasScratchLayout {
    size_t head, keys, vals, keys_tmp, vals_tmp, sort, total;
};

BlasScratchLayout LayoutBlas(uint32_t n) {
    BlasScratchLayout l;
    size_t o = 0;
    l.head = o; o += 256;
    l.keys = o; o += AlignUp(size_t(n) * 4, 256);
    l.vals = o; o += AlignUp(size_t(n) * 4, 256);
    l.keys_tmp = o; o += AlignUp(size_t(n) * 4, 256);
    l.vals_tmp = o; o += AlignUp(size_t(n) * 4, 256);
    l.sort = o; o += SortScratchBytes(n);
    l.total = o;
    return l;
}

// Treelet restructuring scratch (reuses the key/value region of the LBVH
// scratch, which is dead once the hierarchy is emitted).
struct RestructureLayout {
    size_t counter, roots, counts, total;
};
__host__ __device__ RestructureLayout LayoutRestructure(uint32_t n) {
    RestructureLayout l;
    size_t o = 0;
    l.counter = o; o += 256;
    l.roots = o; o += AlignUp(size_t(n) * 4, 256);
    l.counts = o; o += AlignUp(size_t(NodeCount(n)) * 4, 256);
    l.total = o;
    return l;
}

template <typename T>
T* At(void* base, size_t off) {
    return reinterpret_cast<T*>(static_cast<uint8_t*>(base) + off);
}

// ---------------------------------------------------------------------------
// Kernels

__global__ __launch_bounds__(kBlock) void InitBuildKernel(AccelHeader* hdr, uint32_t kind, uint32_t prims,
                                                          uint64_t nodes_off, uint64_t xf_off, uint64_t table_off,
                                                          uint32_t flags, BuildHead* head, uint32_t* sort_hist,
                                                          uint32_t* sort_tickets) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        const uint32_t serial = hdr->magic == kAccelMagic ? hdr->build_serial + 1 : 1;
        AccelHeader h = {};
        h.magic = kAccelMagic;
        h.kind = kind;
        h.prim_count = prims;
        h.node_count = NodeCount(prims);
        h.nodes_offset = nodes_off;
        h.transforms_offset = xf_off;
        h.blas_table_offset = table_off;
        h.flags = flags;
        h.build_serial = serial;
        *hdr = h;
    }
    if (t < 8 * kBoundsCopies) (&head->bounds[0][0])[t] = ((t & 7u) < 4) ? 0xFFFFFFFFu : 0u;
    for (uint32_t i = t; i < kSortHistCopies * kSortHistWords; i += kBlock)
        sort_hist[(i / kSortHistWords) * kSortHistStride + i % kSortHistWords] = 0;
    if (t < kSortMaxPasses) sort_tickets[t] = 0;
}

// Block min/max of ordered-uint coordinates, then 6 device atomics per
// block into the block's copy of the bounds (lbvh_calc_mesh_aabb.comp:121-179
// reduces in a 128-wide LDS tree with atomics per loop iteration).
__device__ __forceinline__ void ReduceBoundsToGlobal(uint32_t mn[3], uint32_t mx[3],
                                                     uint32_t (*copies)[8]) {
    uint32_t* bounds = copies[blockIdx.x % kBoundsCopies];
    __shared__ uint32_t s_red[kBlock / 64][6];
    for (int a = 0; a < 3; ++a) {
        mn[a] = WaveMin(mn[a]);
        mx[a] = WaveMax(mx[a]);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 3; ++a) { s_red[wave][a] = mn[a]; s_red[wave][3 + a] = mx[a]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t a = threadIdx.x;
        uint32_t r = s_red[0][a];
        for (int w = 1; w < kBlock / 64; ++w) r = a < 3 ? min(r, s_red[w][a]) : max(r, s_red[w][a]);
        if (a < 3) atomicMin(&bounds[a], r);
        else atomicMax(&bounds[4 + (a - 3)], r);
    }
}

// Each step loads kBoundsTris triangles a grid stride apart before using
// any of them, so a thread waits for one index load and one vertex load
// per step rather than per triangle.  A triangle index past the end is
// replaced by the last triangle, which leaves the bounds unchanged.
constexpr int kBoundsTris = 4;
__global__ __launch_bounds__(kBlock) void MeshBoundsKernel(MeshView m, BuildHead* head) {
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t last = m.tri_count - 1;  // launched only for tri_count >= 1
    for (uint64_t t0 = blockIdx.x * kBlock + threadIdx.x; t0 < m.tri_count; t0 += kBoundsTris * stride) {
        float v[kBoundsTris][9];
#pragma unroll
        for (int j = 0; j < kBoundsTris; ++j) LoadTriangle(m, (uint32_t)min(t0 + j * stride, last), v[j]);
#pragma unroll
        for (int j = 0; j < kBoundsTris; ++j)
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    const uint32_t o = FloatToOrdered(v[j][3 * k + a]);
                    mn[a] = min(mn[a], o);
                    mx[a] = max(mx[a], o);
                }
    }
    ReduceBoundsToGlobal(mn, mx, head->bounds);
}

struct SceneBounds {
    float lo[3], ext[3];
};
__device__ __forceinline__ SceneBounds ReadBounds(const BuildHead* head) {
    SceneBounds b;
    for (int a = 0; a < 3; ++a) {
        uint32_t lo_o = head->bounds[0][a], hi_o = head->bounds[0][4 + a];
        for (int c = 1; c < kBoundsCopies; ++c) {
            lo_o = min(lo_o, head->bounds[c][a]);
            hi_o = max(hi_o, head->bounds[c][4 + a]);
        }
        const float lo = OrderedToFloat(lo_o);
        const float hi = OrderedToFloat(hi_o);
        b.lo[a] = lo;
        b.ext[a] = hi - lo;
    }
    return b;
}

// Morton code of each primitive's box centroid (lbvh_calc_morton_codes_mesh.comp:78-130),
// fused with the sort's digit histogram and the clearing of its look-back words.
template <typename CentroidFn>
__device__ __forceinline__ void MortonAndHistogram(uint32_t n, const BuildHead* head, CentroidFn centroid,
                                                   uint32_t* keys, uint32_t* vals, int passes, uint32_t* sort_hist,
                                                   uint32_t* sort_status, size_t status_words) {
    __shared__ uint32_t s_h[kSortMaxPasses][kSortRadix];
    for (uint32_t i = threadIdx.x; i < kSortMaxPasses * kSortRadix; i += kBlock) (&s_h[0][0])[i] = 0;
    const SceneBounds b = ReadBounds(head);
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0;
    if (valid) {
        float c[3];
        centroid(i, c);
        key = Morton30(UnitCoord(c[0], b.lo[0], b.ext[0]), UnitCoord(c[1], b.lo[1], b.ext[1]),
                       UnitCoord(c[2], b.lo[2], b.ext[2]));
        keys[i] = key;
        vals[i] = i;
    }
    AccumulateDigits(key, valid, s_h, passes);
    __syncthreads();
    uint32_t* copy = HistCopyOfBlock(sort_hist);
    for (uint32_t d = threadIdx.x; d < (uint32_t)passes * kSortRadix; d += kBlock) {
        const uint32_t c = (&s_h[0][0])[d];
        if (c) atomicAdd(&copy[d], c);
    }
    for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < status_words; w += (size_t)gridDim.x * kBlock)
        sort_status[w] = 0;
}

__global__ __launch_bounds__(kBlock) void MeshMortonKernel(MeshView m, const BuildHead* head, uint32_t* keys,
                                                           uint32_t* vals, int passes, uint32_t* sort_hist,
                                                           uint32_t* sort_status, size_t status_words) {
    auto centroid = [&](uint32_t t, float c[3]) { TriangleCentroid(m, t, c); };
    MortonAndHistogram(m.tri_count, head, centroid, keys, vals, passes, sort_hist, sort_status, status_words);
}

// Block b emits nodes [256 b, 256 b + 256).  Its codes and kEmitMargin
// codes on either side are staged in LDS: most nodes have short key
// ranges, so most probes of their searches read LDS; the others read
// memory.  Same codes either way, so the same tree.
constexpr int kEmitMargin = 256;
constexpr int kEmitWindow = kBlock + 2 * kEmitMargin;
__global__ __launch_bounds__(kBlock) void EmitHierarchyKernel(const uint32_t* __restrict__ codes,
                                                              const uint32_t* __restrict__ prims, uint32_t n,
                                                              BvhNode* __restrict__ nodes) {
    __shared__ uint32_t s_codes[kEmitWindow];
    const int lo = (int)(blockIdx.x * kBlock) - kEmitMargin;
    for (int k = threadIdx.x; k < kEmitWindow; k += kBlock) {
        const int g = lo + k;
        s_codes[k] = (g >= 0 && g < (int)n) ? codes[g] : 0u;
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // An LDS read always (clamped index) and a memory read only outside the
    // window, kept as two loads: one select between the two addresses would
    // become a flat load.
    auto code = [&](int k) -> uint32_t {
        const uint32_t o = (uint32_t)(k - lo);
        const bool in = o < (uint32_t)kEmitWindow;
        uint32_t v = s_codes[in ? o : 0u];
        if (!in) v = LdG(codes + k);
        return v;
    };
    EmitNodeWith(code, prims, n, i, nodes);
}

// One thread per leaf; block b's leaves (key positions [256 b, 256 b + 256))
// form one refit window.
__global__ __launch_bounds__(kBlock) void RefitMeshKernel(MeshView m, BvhNode* __restrict__ nodes) {
    __shared__ RefitLds<kBlock> lds;
    const uint32_t n = m.tri_count;
    const uint32_t base = XcdContiguousBlock(blockIdx.x, gridDim.x) * kBlock;
    const uint32_t i = base + threadIdx.x;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    uint32_t parent = kInvalid;
    if (i < n) parent = WriteLeaf(m, nodes, i, lo, hi);
    WindowRefit<kBlock>(nodes, 2 * n - 1, base, kBlock, 0, i < n, i + 1 < n ? i : kInvalid, n - 1 + i, parent, lo, hi,
                        lds);
}

// Refit of the BLAS in `result` (nodes after the header; topology and leaf
// links in place), one thread per leaf.
void LaunchMeshRefit(const MeshView& m, void* result, hipStream_t s) {
    const uint32_t n = m.tri_count;
    if (n == 0) return;
    RefitMeshKernel<<<Blocks(n), kBlock, 0, s>>>(m, At<BvhNode>(result, sizeof(AccelHeader)));
}


// ---------------------------------------------------------------------------
// Treelet restructuring (Karras & Aila, "Fast Parallel Construction of
// High-Quality Bounding Volume Hierarchies", HPG 2013), following
// init_primitive_count.comp, find_treelet_roots.comp and
// restructure_bvh.comp:149-542 with restructure_hlbvh.cpp:133-278.

// Treelet size: 7 leaves in the reference (restructure_bvh.comp); 5 and 9
// are tuning options (process-wide, SetTreeletLeaves).
static std::atomic<int> g_treelet_leaves{7};
constexpr float kCostInternal = 1.2f;  // C_INT, restructure_bvh.comp:38

// The subsets of 7 leaves with `bits` leaves, in increasing order
// (C(7, 3) = C(7, 4) = 35 of them), for the DP's one-subset-per-lane levels.
struct SubsetTable {
    uint8_t mask[35];
};
constexpr SubsetTable SubsetsOf7(int bits) {
    SubsetTable t{};
    int k = 0;
    for (int m = 0; m < 128; ++m) {
        int c = 0;
        for (int b = 0; b < 7; ++b) c += (m >> b) & 1;
        if (c == bits && k < 35) t.mask[k++] = (uint8_t)m;
    }
    return t;
}
__constant__ constexpr SubsetTable kSubsets7Of3 = SubsetsOf7(3);
__constant__ constexpr SubsetTable kSubsets7Of4 = SubsetsOf7(4);

__device__ __forceinline__ float SurfaceArea(const float lo[3], const float hi[3]) {
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float s = ex * ez;
    s = s + ey * ex;
    s = s + ez * ey;
    return 2.0f * s;
}

// Box of a node as the reference's GetNodeAabb (restructure_bvh.comp:71-89):
// the four child-box corners of an internal node, the three vertices of a leaf.
__device__ __forceinline__ void NodeBox(const BvhNode* __restrict__ nodes, uint32_t idx, uint32_t n_internal, float lo[3],
                                        float hi[3]) {
    const BvhNode& nd = nodes[idx];
    for (int a = 0; a < 3; ++a) {
        float l = fminf(fminf(nd.a0_min_or_v0[a], nd.a0_max_or_v1[a]), nd.a1_min_or_v2[a]);
        float h = fmaxf(fmaxf(nd.a0_min_or_v0[a], nd.a0_max_or_v1[a]), nd.a1_min_or_v2[a]);
        if (idx < n_internal) {
            l = fminf(l, nd.a1_max_or_v3[a]);
            h = fmaxf(h, nd.a1_max_or_v3[a]);
        }
        lo[a] = l;
        hi[a] = h;
    }
}

// The two child boxes held in an internal node's slots, as (lo xyz, hi xyz).
__device__ __forceinline__ void SlotBoxes(const float4& q0, const float4& q1, const float4& q2, const float4& q3,
                                          float b0[6], float b1[6]) {
    b0[0] = q0.x; b0[1] = q0.y; b0[2] = q0.z; b0[3] = q1.x; b0[4] = q1.y; b0[5] = q1.z;
    b1[0] = q2.x; b1[1] = q2.y; b1[2] = q2.z; b1[3] = q3.x; b1[4] = q3.y; b1[5] = q3.z;
}

// Agent-scope store of a node link for the restructuring kernel (see step 5).
__device__ __forceinline__ void StU(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBlock) void InitTreeletCountsKernel(uint32_t n, uint32_t* counts, uint32_t* counter) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n - 1) counts[i] = 0;
    if (i == 0) *counter = 0;
}

// Climb from each leaf, accumulating primitive counts; the first node whose
// subtree reaches `min_prims` becomes a treelet root (find_treelet_roots.comp:63-98).
// This is the climb from node p on, `count` primitives arriving: an
// exchange on each node's count, the first arrival stops.  Root found, or
// kInvalid.  Afterwards counts[p] is non-zero exactly on the nodes some
// climb reached, which the restructuring's hand-off relies on.
__device__ __forceinline__ uint32_t ClimbCountsFrom(const BvhNode* __restrict__ nodes, uint32_t n, uint32_t p,
                                                    uint32_t count, uint32_t min_prims, uint32_t* counts) {
    const uint32_t total = 2 * n - 1;
    for (uint32_t level = 0; p < total && level < total; ++level) {
        const uint32_t old = __hip_atomic_exchange(&counts[p], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        count += old;
        if (old == 0) return kInvalid;
        if (count >= min_prims) return p;
        p = nodes[p].parent;
    }
    return kInvalid;
}

__device__ __forceinline__ uint32_t ClimbToTreeletRoot(const BvhNode* __restrict__ nodes, uint32_t n, uint32_t i,
                                                       uint32_t min_prims, uint32_t* counts) {
    return ClimbCountsFrom(nodes, n, nodes[n - 1 + i].parent, 1, min_prims, counts);
}

// The same climbs over a window of leaves, as the windowed refit
// (WindowRefit): a node whose subtree lies inside the workgroup's window
// counts its two arrivals in LDS — one atomic add of 4 * count + 1, so the
// second arrival reads the first one's count with it — and the second
// arrival marks the node in memory with a plain store of its count (no one
// else touches it).  A climb that leaves the window, or that arrived first
// at a node whose other child crosses the window edge (one arrival when
// the workgroup is done), continues with the exchange climb.  The roots
// found and the nodes marked are those of the climb from every leaf with
// exchanges.  Every thread of the workgroup calls this (it holds barriers).
//
// Leaf parents come from the window too: the thread that loads internal
// node i's parent reads i's child links from the same line and records i
// as the parent of each child leaf inside the window, so a leaf reads its
// node (a whole line for one link) only when its parent is outside.
//
// After the barrier the climbs that continue with exchanges (the few at
// window edges) are handed to the first wave through LDS, as in
// WindowRefit, so the workgroup's other waves end at once; a pending climb
// carries its tree (nodes, leaf count, counts array) and its owner tag.
constexpr int kRootPendingWords = 8;  // p, count, n, tag, nodes (2), counts (2)
template <int W>
struct RootsLds {
    uint32_t acc[W];
    uint32_t parent[W];
    uint32_t leaf_parent[W];
    uint32_t pending[W * kRootPendingWords];
    uint32_t npending;
};
// emit(tag, root) records a root found (tag: the caller's owner tag, e.g.
// the mesh); counts_arg is any global-memory pointer of the launch, from
// which handed-over counts arrays are addressed.
template <int W, typename Emit>
__device__ __forceinline__ void WindowTreeletRoot(const BvhNode* __restrict__ nodes, uint32_t n, uint32_t base,
                                                  uint32_t width, uint32_t slot0, bool active, uint32_t own,
                                                  uint32_t leaf, uint32_t min_prims, uint32_t* counts,
                                                  uint32_t* counts_arg, uint32_t tag, RootsLds<W>& s,
                                                  const Emit& emit) {
    s.acc[threadIdx.x] = 0;
    if (threadIdx.x == 0) s.npending = 0;
    s.leaf_parent[threadIdx.x] = kInvalid;
    __syncthreads();
    uint32_t own_parent = kInvalid;
    if (own != kInvalid && own + 1 < n) {
        const BvhNode& nd = nodes[own];
        own_parent = nd.parent;
        const uint32_t c[2] = {nd.child0, nd.child1};
        for (int k = 0; k < 2; ++k) {
            const uint32_t j = c[k] - (n - 1);  // leaf position, if c[k] is a leaf
            if (c[k] >= n - 1 && c[k] < 2 * n - 1 && j - base < width) s.leaf_parent[slot0 + (j - base)] = own;
        }
    }
    s.parent[threadIdx.x] = own_parent;
    __syncthreads();
    uint32_t p = kInvalid;
    if (active) {
        p = s.leaf_parent[slot0 + (leaf - (n - 1) - base)];
        if (p == kInvalid) p = nodes[leaf].parent;
    }
    uint32_t count = 1, l = 0, root = kInvalid;
    int state = active ? 1 : 0;  // 0 done, 1 climbing in LDS, 2 exchange climb from p, 3 first arrival at p
    for (uint32_t level = 0; state == 1; ++level) {
        if (p - base >= width || p >= n - 1) {  // outside the window, or not an internal node
            state = 2;
            break;
        }
        if (level >= width) {  // more levels than a window's subtree has: corrupt links
            state = 0;
            break;
        }
        l = slot0 + (p - base);
        const uint32_t old = atomicAdd(&s.acc[l], 4u * count + 1u);
        if (old == 0) {
            state = 3;
            break;
        }
        count += old >> 2;
        counts[p] = count;  // both children arrived here; no exchange will touch this node
        if (count >= min_prims) {
            root = p;
            state = 0;
            break;
        }
        p = s.parent[l];
    }
    __syncthreads();
    if (state == 3 && (s.acc[l] & 3u) == 1u) state = 2;
    if (root != kInvalid) emit(tag, root);
    if (state == 2) {
        uint32_t* q = s.pending + kRootPendingWords * atomicAdd(&s.npending, 1u);
        const uint64_t nb = reinterpret_cast<uint64_t>(nodes), cb = reinterpret_cast<uint64_t>(counts);
        q[0] = p;
        q[1] = count;
        q[2] = n;
        q[3] = tag;
        q[4] = (uint32_t)nb;
        q[5] = (uint32_t)(nb >> 32);
        q[6] = (uint32_t)cb;
        q[7] = (uint32_t)(cb >> 32);
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    const uint32_t npending = s.npending;
    for (uint32_t e = threadIdx.x; e < npending; e += 64) {
        const uint32_t* q = s.pending + kRootPendingWords * e;
        // Addressed from this thread's own pointers so that they stay global.
        const BvhNode* tree = RebaseOn<const BvhNode>(nodes, (uint64_t)q[4] | ((uint64_t)q[5] << 32));
        uint32_t* tcounts = RebaseOn<uint32_t>(counts_arg, (uint64_t)q[6] | ((uint64_t)q[7] << 32));
        const uint32_t r = ClimbCountsFrom(tree, q[2], q[0], q[1], min_prims, tcounts);
        if (r != kInvalid) emit(q[3], r);
    }
}

// Roots are staged in LDS and appended with one device atomic per
// workgroup (a per-wave append put thousands of atomics on one counter word,
// and those serialise across XCDs); a block that finds more than 2 x kBlock
// roots appends the rest one by one.  Root order does not matter: each
// root's treelet is restructured on its own.
__global__ __launch_bounds__(kBlock) void FindTreeletRootsKernel(const BvhNode* __restrict__ nodes, uint32_t n,
                                                                 uint32_t min_prims, uint32_t* counts, uint32_t* counter,
                                                                 uint32_t* roots) {
    __shared__ RootsLds<kBlock> lds;
    __shared__ uint32_t s_roots[2 * kBlock];
    __shared__ uint32_t s_nroots, s_base;
    if (threadIdx.x == 0) s_nroots = 0;
    const uint32_t base = blockIdx.x * kBlock;
    const uint32_t i = base + threadIdx.x;
    WindowTreeletRoot<kBlock>(nodes, n, base, kBlock, 0, i < n, i + 1 < n ? i : kInvalid, n - 1 + i, min_prims,
                              counts, counts, 0u, lds, [&](uint32_t, uint32_t root) {
                                  const uint32_t k = atomicAdd(&s_nroots, 1u);
                                  if (k < 2 * kBlock) s_roots[k] = root;
                                  else roots[atomicAdd(counter, 1u)] = root;
                              });
    __syncthreads();
    const uint32_t cnt = min(s_nroots, 2u * kBlock);
    if (threadIdx.x == 0 && cnt) s_base = atomicAdd(counter, cnt);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) roots[s_base + k] = s_roots[k];
}

// One wave per treelet root.  Lane 0 does the pointer-chasing parts (forming
// the treelet, rewiring it, climbing); the 128 subset areas and the dynamic
// programme over subsets run across the wave.
// The body runs for one root of one BVH (nodes, n leaves, arrival counts
// of its internal nodes); the single-mesh and batched kernels call it.
template <int LEAVES, bool kFiveLeafGroups>
__device__ __forceinline__ void RestructureFromRoot(BvhNode* nodes, uint32_t n, uint32_t root, uint32_t* counts) {
    constexpr int kTreeletLeaves = LEAVES;
    constexpr uint32_t kTreeletMasks = 1u << LEAVES;
    __shared__ uint32_t s_leaf[kTreeletLeaves];
    __shared__ uint32_t s_inner[kTreeletLeaves - 1];
    __shared__ float s_box[kTreeletLeaves][6];
    __shared__ float s_leaf_area[kTreeletLeaves];
    __shared__ float s_area[kTreeletMasks];
    // Box of each leaf subset (lo xyz, hi xyz), kept from step 2 for the
    // rewiring; not for 9-leaf treelets, whose 512 boxes would take 12 KB
    // of LDS per wave.
    constexpr bool kSubsetBoxes = LEAVES <= 7;
    __shared__ float s_sbox[kSubsetBoxes ? kTreeletMasks : 1][6];
    __shared__ float s_cost[kTreeletMasks];
    __shared__ uint32_t s_part[kTreeletMasks];
    __shared__ uint32_t s_root;
    __shared__ uint32_t s_parent;
    __shared__ uint32_t s_go;
    const uint32_t lane = threadIdx.x;
    const uint32_t n_int = n - 1;
    if (lane == 0) s_root = root;
    __syncthreads();
    // Each pass moves one level up, so a tree of n leaves needs fewer than n passes.
    for (uint32_t pass = 0; pass < n; ++pass) {
        // 1. Form a treelet: repeatedly open the treelet leaf with the largest
        //    surface area (restructure_bvh.comp:170-218).
        //    A node's two child boxes are in its own slots (exact unions,
        //    kept current by refit and by every finished treelet), so opening
        //    a node is one load of that node: its child links and both boxes.
        if (lane == 0) {
            const uint32_t root = s_root;
            float* area = s_leaf_area;  // LDS: indexed at run time (registers would spill)
            float (*box)[6] = s_box;
            s_inner[0] = root;
            float4 q0, q1, q2, q3;
            LoadNodeAgent(nodes + root, q0, q1, q2, q3);
            s_parent = __float_as_uint(q2.w);
            s_leaf[0] = __float_as_uint(q0.w);
            s_leaf[1] = __float_as_uint(q1.w);
            SlotBoxes(q0, q1, q2, q3, box[0], box[1]);
            area[0] = SurfaceArea(box[0], box[0] + 3);
            area[1] = SurfaceArea(box[1], box[1] + 3);
            bool ok = s_leaf[0] < 2 * n - 1 && s_leaf[1] < 2 * n - 1;
            for (int size = 2; ok && size < kTreeletLeaves; ++size) {
                float largest = 0.0f;
                int sel = -1;
                for (int k = 0; k < size; ++k) {
                    if (s_leaf[k] < n_int && (largest == 0.0f || area[k] > largest)) {
                        largest = area[k];
                        sel = k;
                    }
                }
                if (sel < 0) { ok = false; break; }
                const uint32_t opened = s_leaf[sel];
                s_inner[size - 1] = opened;
                LoadNodeAgent(nodes + opened, q0, q1, q2, q3);
                const uint32_t c0 = __float_as_uint(q0.w), c1 = __float_as_uint(q1.w);
                if (c0 >= 2 * n - 1 || c1 >= 2 * n - 1) { ok = false; break; }
                s_leaf[sel] = c0;
                s_leaf[size] = c1;
                SlotBoxes(q0, q1, q2, q3, box[sel], box[size]);
                area[sel] = SurfaceArea(box[sel], box[sel] + 3);
                area[size] = SurfaceArea(box[size], box[size] + 3);
            }
            s_go = ok ? 1u : 0u;
        }
        __syncthreads();
        if (s_go) {
            // 2. Surface area of every subset of the 7 treelet leaves.
            for (uint32_t m = lane; m < kTreeletMasks; m += 64) {
                float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
                for (int k = 0; k < kTreeletLeaves; ++k)
                    if (m & (1u << k))
                        for (int a = 0; a < 3; ++a) {
                            lo[a] = fminf(lo[a], s_box[k][a]);
                            hi[a] = fmaxf(hi[a], s_box[k][3 + a]);
                        }
                s_area[m] = m ? SurfaceArea(lo, hi) : 0.0f;
                if constexpr (kSubsetBoxes)
                    for (int a = 0; a < 3; ++a) {
                        s_sbox[m][a] = lo[a];
                        s_sbox[m][3 + a] = hi[a];
                    }
                s_cost[m] = __popc(m) == 1 ? kCostInternal * s_area[m] : 0.0f;
                s_part[m] = 0;
            }
            __syncthreads();
            // 3. Optimal partition of each subset, by increasing size.  The
            //    candidates of subset s are the non-empty subsets p of s
            //    without its lowest leaf, in increasing order; the first
            //    cheapest wins.  Orientation as restructure_bvh.comp:395-397
            //    (sizes 2-5 flipped for even sizes, 6 and 7 unflipped).
            //    The full set, last, has the most candidates (63 for 7
            //    leaves): they are (x + 1) << 1 in increasing order, split
            //    over the lanes and reduced to the cheapest with the lowest
            //    x, the sequential loop's choice.
            //    For 7 leaves the six-leaf subsets (127 without one leaf g)
            //    get eight lanes each: candidate x + 1 is x + 1's bits
            //    deposited into the subset's bits above its lowest leaf,
            //    and the group reduces to its first cheapest.
            //    And, when kFiveLeafGroups (the single-mesh kernel: the
            //    batched one, held to 64 VGPRs, would spill), the five-leaf
            //    subsets (127 without leaves a < b, 21 of them in (a, b)
            //    order) get three lanes each.
            for (int bits = 2; bits <= kTreeletLeaves; ++bits) {
                if (kFiveLeafGroups && kTreeletLeaves == 7 && (bits == 3 || bits == 4)) {
                    // One subset per lane (35 subsets, lanes 0-34).
                    if (lane < 35) {
                        const uint32_t s = bits == 3 ? kSubsets7Of3.mask[lane] : kSubsets7Of4.mask[lane];
                        const uint32_t delta = (s - 1) & s;
                        uint32_t p = (0u - delta) & s;
                        float best_cost = FLT_MAX;
                        uint32_t best = 0;
                        do {
                            const float c = s_cost[p] + s_cost[s ^ p];
                            if (best == 0 || c < best_cost) {
                                best_cost = c;
                                best = p;
                            }
                            p = (p - delta) & s;
                        } while (p != 0);
                        s_cost[s] = kCostInternal * s_area[s] + best_cost;
                        s_part[s] = (bits <= 5 && (bits & 1) == 0) ? (s ^ best) : best;
                    }
                    __syncthreads();
                    continue;
                }
                if (kFiveLeafGroups && kTreeletLeaves == 7 && bits == 5) {
                    const uint32_t g = lane / 3, j = lane - 3 * g;
                    uint32_t a = 0, b = 1;
                    for (uint32_t r = g < 21 ? g : 0, first = 0; first < 6; ++first) {
                        if (r < 6 - first) {
                            a = first;
                            b = first + 1 + r;
                            break;
                        }
                        r -= 6 - first;
                    }
                    const uint32_t s = (kTreeletMasks - 1) ^ (1u << a) ^ (1u << b);
                    const uint32_t delta = (s - 1) & s;
                    float best_cost = FLT_MAX;
                    uint32_t best_x = 0xFFFFFFFFu, best_p = 0;
                    if (g < 21) {
                        for (uint32_t x = j; x < 15; x += 3) {
                            uint32_t p = 0, m = delta, v = x + 1;
                            for (int k = 0; k < 5; ++k) {
                                const uint32_t low = m & (0u - m);
                                if (v & 1u) p |= low;
                                v >>= 1;
                                m ^= low;
                            }
                            const float c = s_cost[p] + s_cost[s ^ p];
                            if (best_x == 0xFFFFFFFFu || c < best_cost) {
                                best_cost = c;
                                best_x = x;
                                best_p = p;
                            }
                        }
                    }
                    // The group's first cheapest, in candidate order.
                    const uint32_t base = 3 * (g < 21 ? g : 0);
                    float gc = FLT_MAX;
                    uint32_t gx = 0xFFFFFFFFu, gp = 0;
                    for (uint32_t t = 0; t < 3; ++t) {
                        const float oc = __shfl(best_cost, (int)(base + t), 64);
                        const uint32_t ox = __shfl(best_x, (int)(base + t), 64);
                        const uint32_t op = __shfl(best_p, (int)(base + t), 64);
                        if (ox != 0xFFFFFFFFu && (gx == 0xFFFFFFFFu || oc < gc || (oc == gc && ox < gx))) {
                            gc = oc;
                            gx = ox;
                            gp = op;
                        }
                    }
                    if (g < 21 && j == 0) {
                        s_cost[s] = kCostInternal * s_area[s] + gc;
                        s_part[s] = gp;  // five leaves: odd, not flipped
                    }
                    __syncthreads();
                    continue;
                }
                if (kTreeletLeaves == 7 && bits == 6) {
                    const uint32_t g = lane >> 3, j = lane & 7u;
                    const uint32_t s = (kTreeletMasks - 1) ^ (1u << (g < 7 ? g : 0u));
                    const uint32_t delta = (s - 1) & s;
                    float best_cost = FLT_MAX;
                    uint32_t best_x = 0xFFFFFFFFu, best_p = 0;
                    if (g < 7) {
                        for (uint32_t x = j; x < 31; x += 8) {
                            uint32_t p = 0, m = delta, v = x + 1;
                            for (int k = 0; k < 6; ++k) {
                                const uint32_t low = m & (0u - m);
                                if (v & 1u) p |= low;
                                v >>= 1;
                                m ^= low;
                            }
                            const float c = s_cost[p] + s_cost[s ^ p];
                            if (best_x == 0xFFFFFFFFu || c < best_cost) {
                                best_cost = c;
                                best_x = x;
                                best_p = p;
                            }
                        }
                    }
                    for (int off = 4; off >= 1; off >>= 1) {  // within the group of eight
                        const float oc = __shfl_xor(best_cost, off, 64);
                        const uint32_t ox = __shfl_xor(best_x, off, 64);
                        const uint32_t op = __shfl_xor(best_p, off, 64);
                        if (ox != 0xFFFFFFFFu &&
                            (best_x == 0xFFFFFFFFu || oc < best_cost || (oc == best_cost && ox < best_x))) {
                            best_cost = oc;
                            best_x = ox;
                            best_p = op;
                        }
                    }
                    if (g < 7 && j == 0) {
                        s_cost[s] = kCostInternal * s_area[s] + best_cost;
                        s_part[s] = best_p;  // six leaves: not flipped
                    }
                    __syncthreads();
                    continue;
                }
                if (bits == kTreeletLeaves) {
                    constexpr uint32_t s = kTreeletMasks - 1;
                    constexpr uint32_t candidates = kTreeletMasks / 2 - 1;
                    float best_cost = FLT_MAX;
                    uint32_t best_x = 0xFFFFFFFFu;
                    for (uint32_t x = lane; x < candidates; x += 64) {
                        const uint32_t p = (x + 1) << 1;
                        const float c = s_cost[p] + s_cost[s ^ p];
                        if (best_x == 0xFFFFFFFFu || c < best_cost) {
                            best_cost = c;
                            best_x = x;
                        }
                    }
                    for (int off = 32; off >= 1; off >>= 1) {
                        const float oc = __shfl_xor(best_cost, off, 64);
                        const uint32_t ox = __shfl_xor(best_x, off, 64);
                        if (ox != 0xFFFFFFFFu &&
                            (best_x == 0xFFFFFFFFu || oc < best_cost || (oc == best_cost && ox < best_x))) {
                            best_cost = oc;
                            best_x = ox;
                        }
                    }
                    if (lane == 0) {
                        const uint32_t best = (best_x + 1) << 1;
                        s_cost[s] = kCostInternal * s_area[s] + best_cost;
                        s_part[s] = (bits <= 5 && (bits & 1) == 0) ? (s ^ best) : best;
                    }
                    __syncthreads();
                    continue;
                }
                for (uint32_t s = lane; s < kTreeletMasks; s += 64) {
                    if (__popc(s) != bits) continue;
                    const uint32_t delta = (s - 1) & s;
                    uint32_t p = (0u - delta) & s;
                    float best_cost = FLT_MAX;
                    uint32_t best = 0;
                    do {
                        const float c = s_cost[p] + s_cost[s ^ p];
                        if (best == 0 || c < best_cost) {
                            best_cost = c;
                            best = p;
                        }
                        p = (p - delta) & s;
                    } while (p != 0);
                    s_cost[s] = kCostInternal * s_area[s] + best_cost;
                    s_part[s] = (bits <= 5 && (bits & 1) == 0) ? (s ^ best) : best;
                }
                __syncthreads();
            }
            // 4. Rewire the treelet top-down from the optimal partitions and
            //    write every internal node's child boxes (the union of the
            //    leaves on each side) (restructure_bvh.comp:418-494).
            if (lane == 0) {
                uint32_t stack_mask[kTreeletLeaves], stack_node[kTreeletLeaves];
                int sp = 0, allocated = 1;
                stack_mask[sp] = kTreeletMasks - 1;
                stack_node[sp] = s_inner[0];
                ++sp;
                while (sp > 0) {
                    --sp;
                    const uint32_t m = stack_mask[sp], node = stack_node[sp];
                    const uint32_t side[2] = {s_part[m], m ^ s_part[m]};
                    uint32_t child[2];
                    for (int c = 0; c < 2; ++c) {
                        if (__popc(side[c]) > 1) {
                            child[c] = s_inner[allocated++];
                            stack_mask[sp] = side[c];
                            stack_node[sp] = child[c];
                            ++sp;
                        } else {
                            child[c] = s_leaf[__ffs(side[c]) - 1];
                        }
                    }
                    // Each side's box: its leaf subset's, from step 2, or
                    // the same union in the same leaf order.
                    float box[2][6];
                    for (int c = 0; c < 2; ++c) {
                        if constexpr (kSubsetBoxes) {
                            for (int a = 0; a < 6; ++a) box[c][a] = s_sbox[side[c]][a];
                        } else {
                            for (int a = 0; a < 3; ++a) { box[c][a] = FLT_MAX; box[c][3 + a] = -FLT_MAX; }
                            for (int k = 0; k < kTreeletLeaves; ++k)
                                if (side[c] & (1u << k))
                                    for (int a = 0; a < 3; ++a) {
                                        box[c][a] = fminf(box[c][a], s_box[k][a]);
                                        box[c][3 + a] = fmaxf(box[c][3 + a], s_box[k][3 + a]);
                                    }
                        }
                    }
                    BvhNode& nd = nodes[node];
                    StoreBoxAgent(nd.a0_min_or_v0, box[0][0], box[0][1], box[0][2]);
                    StoreBoxAgent(nd.a0_max_or_v1, box[0][3], box[0][4], box[0][5]);
                    StoreBoxAgent(nd.a1_min_or_v2, box[1][0], box[1][1], box[1][2]);
                    StoreBoxAgent(nd.a1_max_or_v3, box[1][3], box[1][4], box[1][5]);
                    StU(&nd.child0, child[0]);
                    StU(&nd.child1, child[1]);
                    StU(&nodes[child[0]].parent, node);
                    StU(&nodes[child[1]].parent, node);
                }
            }
        }
        // 5. Climb: the last of the two subtrees to finish below the parent
        //    restructures there next (restructure_bvh.comp:497-540).  The
        //    reference also refits the parent's two child boxes here; those
        //    are unchanged (a restructured treelet covers the same leaves, so
        //    its root's box, the exact union of them, is bit for bit the one
        //    already in the parent's slot), so this climb only hands off.
        //    Hand-off as in ClimbRefit: every node access of this kernel is
        //    an agent-scope (write-through / coherent) load or store, the
        //    wave waits for its stores before the arrival counter's relaxed
        //    atomic, and the second arriver then reads what the first wrote.
        //    The root's parent link was read with the root in step 1 (no
        //    other wave writes it before this hand-off).
        if (lane == 0) {
            const uint32_t parent = s_parent;
            uint32_t go = 0;
            if (parent < n_int) {
                WaitStores();
                const uint32_t old =
                    __hip_atomic_fetch_add(&counts[parent], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (old != 0) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    s_root = parent;
                    go = 1;
                }
            }
            s_go = go;
        }
        __syncthreads();
        if (!s_go) return;
    }
}

template <int LEAVES>
__global__ __launch_bounds__(64) void RestructureTreeletsKernel(BvhNode* nodes, uint32_t n, const uint32_t* counter,
                                                                const uint32_t* __restrict__ roots, uint32_t* counts) {
    if (blockIdx.x >= *counter) return;
    RestructureFromRoot<LEAVES, true>(nodes, n, roots[blockIdx.x], counts);
}

// ---------------------------------------------------------------------------
// Batched restructuring: each step of each round runs once for all meshes of
// a batch table (threads over 64-leaf chunks of any mesh, one wave per
// treelet root of any mesh).  A mesh's arrival counts live in its work area
// (LayoutRestructure); meshes of fewer than 64 triangles stay as built.

// Mesh and leaf of thread threadIdx.x of block blockIdx.x over the refit
// chunks of the table; false past the chunks or the mesh's leaves.
__device__ __forceinline__ bool BatchChunkLeaf(const BatchMesh* __restrict__ table, uint32_t meshes,
                                               uint32_t total_chunks, uint32_t& mesh, uint32_t& leaf) {
    const uint32_t c = blockIdx.x * (kBlock / kRefitChunk) + threadIdx.x / kRefitChunk;
    if (c >= total_chunks) return false;
    uint32_t lo = 0, hi = meshes;  // last mesh with first_chunk <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (table[mid].first_chunk <= c) lo = mid;
        else hi = mid;
    }
    mesh = lo;
    leaf = (c - table[lo].first_chunk) * kRefitChunk + threadIdx.x % kRefitChunk;
    return c >= table[lo].first_chunk && leaf < table[lo].tri_count;
}

__device__ __forceinline__ uint32_t* BatchCounts(const BatchMesh& bm, uint32_t* arg) {
    return RebaseOn<uint32_t>(arg, bm.work + LayoutRestructure(bm.tri_count).counts);
}

__global__ __launch_bounds__(kBlock) void BatchTreeletInitKernel(const BatchMesh* __restrict__ table, uint32_t meshes,
                                                                 uint32_t total_chunks, uint32_t* roots_counter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *roots_counter = 0;
    uint32_t mesh, leaf;
    if (!BatchChunkLeaf(table, meshes, total_chunks, mesh, leaf)) return;
    const BatchMesh bm = table[mesh];
    if (bm.tri_count < 64 || leaf + 1 >= bm.tri_count) return;  // counts of internal nodes [0, n - 1)
    BatchCounts(bm, roots_counter)[leaf] = 0;
}

// Windows as in BatchRefitKernel: the chunks of one mesh within a block.
__global__ __launch_bounds__(kBlock) void BatchFindTreeletRootsKernel(const BatchMesh* __restrict__ table,
                                                                      uint32_t meshes, uint32_t total_chunks,
                                                                      uint32_t min_prims, uint32_t* roots_counter,
                                                                      uint2* roots) {
    __shared__ RootsLds<kBlock> lds;
    __shared__ uint2 s_roots[2 * kBlock];
    __shared__ uint32_t s_nroots, s_base;
    if (threadIdx.x == 0) s_nroots = 0;
    constexpr uint32_t kChunksPerBlock = kBlock / kRefitChunk;
    const uint32_t cb = blockIdx.x * kChunksPerBlock;
    const uint32_t c = cb + threadIdx.x / kRefitChunk;
    bool active = c < total_chunks;
    uint32_t lo = 0, hi = meshes;  // last mesh with first_chunk <= c (empty meshes own no chunk)
    while (active && hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (table[mid].first_chunk <= c) lo = mid;
        else hi = mid;
    }
    BatchMesh bm{};
    if (active) bm = table[lo];
    active = active && c >= bm.first_chunk && bm.tri_count >= 64;
    const uint32_t i = (c - bm.first_chunk) * kRefitChunk + threadIdx.x % kRefitChunk;
    active = active && i < bm.tri_count;
    const uint32_t first = std::max(cb, bm.first_chunk);
    const uint32_t last = std::min(cb + kChunksPerBlock, bm.first_chunk + (bm.tri_count + kRefitChunk - 1) / kRefitChunk);
    const uint32_t n = bm.tri_count;
    const BvhNode* nodes =
        reinterpret_cast<const BvhNode*>(RebaseOn<uint8_t>(roots_counter, bm.result) + sizeof(AccelHeader));
    uint32_t* counts = active ? BatchCounts(bm, roots_counter) : nullptr;
    WindowTreeletRoot<kBlock>(
        nodes, n, (first - bm.first_chunk) * kRefitChunk, active ? (last - first) * kRefitChunk : 0,
        (first - cb) * kRefitChunk, active, active && i + 1 < n ? i : kInvalid, n - 1 + i, min_prims, counts,
        roots_counter, lo, lds, [&](uint32_t mesh, uint32_t root) {
            const uint32_t k = atomicAdd(&s_nroots, 1u);
            if (k < 2 * kBlock) s_roots[k] = make_uint2(mesh, root);
            else roots[atomicAdd(roots_counter, 1u)] = make_uint2(mesh, root);
        });
    __syncthreads();
    const uint32_t cnt = min(s_nroots, 2u * kBlock);
    if (threadIdx.x == 0 && cnt) s_base = atomicAdd(roots_counter, cnt);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) roots[s_base + k] = s_roots[k];
}

// Eight waves per SIMD (64 VGPRs): the batched rounds run hundreds of
// thousands of short treelet chains at once and are bound by how many are
// in flight (18.8 against 23.5 ms for the 8,192-mesh batch).  The body is
// inlined into each kernel, so this bound applies to this kernel alone (a
// shared out-of-line body takes the largest register budget of its callers,
// and the call itself cost the single-mesh kernel 36 VGPRs and a spill).
template <int LEAVES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void BatchRestructureTreeletsKernel(const BatchMesh* __restrict__ table,
                                                                     uint32_t* roots_counter,
                                                                     const uint2* __restrict__ roots) {
    if (blockIdx.x >= *roots_counter) return;
    const uint2 r = roots[blockIdx.x];
    const BatchMesh bm = table[r.x];
    BvhNode* nodes = reinterpret_cast<BvhNode*>(RebaseOn<uint8_t>(roots_counter, bm.result) + sizeof(AccelHeader));
    RestructureFromRoot<LEAVES, false>(nodes, bm.tri_count, r.y, BatchCounts(bm, roots_counter));
}


// ---------------------------------------------------------------------------
// TLAS (lbvh_init_scene.comp, lbvh_calc_scene_aabb.comp:126-165,
// lbvh_calc_morton_codes_scene.comp:83-114, lbvh_fit_aabb_scene.comp:98-167).

// World boxes of instances are widened by this fraction of the magnitude of
// the terms that produced them, so that rounding in the corner transforms
// and in the ray's object-space transform can never cull a hit instance.
constexpr float kInstanceBoxWiden = 1.52587890625e-05f;  // 2^-16

struct TlasScratchLayout {
    size_t upload, head, boxes, keys, vals, keys_tmp, vals_tmp, sort, total;
};
TlasScratchLayout LayoutTlas(uint32_t n) {
    TlasScratchLayout l;
    size_t o = 0;
    l.upload = o; o += AlignUp(size_t(n) * sizeof(InstanceUpload), 256);
    l.head = o; o += 256;
    l.boxes = o; o += AlignUp(size_t(n) * 32, 256);
    l.keys = o; o += AlignUp(size_t(n) * 4, 256);
    l.vals = o; o += AlignUp(size_t(n) * 4, 256);
    l.keys_tmp = o; o += AlignUp(size_t(n) * 4, 256);
    l.vals_tmp = o; o += AlignUp(size_t(n) * 4, 256);
    l.sort = o; o += SortScratchBytes(n);
    l.total = o;
    return l;
}

// Inverse of a 3x4 affine transform: adjugate of the 3x3 part over its
// determinant, then the translation mapped back.  (The reference inverts the
// padded 4x4 with GLSL inverse(), common.h:326-342.)
__device__ __forceinline__ void InverseAffine(const float (*m)[4], float (*r)[4]) {
    const float c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const float c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const float c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const float det = (m[0][0] * c00 + m[0][1] * c01) + m[0][2] * c02;
    const float inv = 1.0f / det;
    r[0][0] = c00 * inv;
    r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    r[1][0] = c01 * inv;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    r[2][0] = c02 * inv;
    r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
    for (int a = 0; a < 3; ++a) r[a][3] = -((r[a][0] * m[0][3] + r[a][1] * m[1][3]) + r[a][2] * m[2][3]);
}

// Box of a BLAS's root: empty for an empty BLAS, the triangle's box for a
// one-triangle BLAS, else the union of the root's two child boxes.
__device__ __forceinline__ bool BlasRootBox(const uint8_t* blas, float lo[3], float hi[3]) {
    const AccelHeader* h = reinterpret_cast<const AccelHeader*>(blas);
    if (h->node_count == 0) return false;
    const BvhNode& r = *reinterpret_cast<const BvhNode*>(blas + h->nodes_offset);
    for (int a = 0; a < 3; ++a) {
        if (h->node_count == 1) {
            lo[a] = fminf(fminf(r.a0_min_or_v0[a], r.a0_max_or_v1[a]), r.a1_min_or_v2[a]);
            hi[a] = fmaxf(fmaxf(r.a0_min_or_v0[a], r.a0_max_or_v1[a]), r.a1_min_or_v2[a]);
        } else {
            lo[a] = fminf(r.a0_min_or_v0[a], r.a1_min_or_v2[a]);
            hi[a] = fmaxf(r.a0_max_or_v1[a], r.a1_max_or_v3[a]);
        }
    }
    return true;
}

__global__ __launch_bounds__(kBlock) void TlasInstanceKernel(const InstanceUpload* __restrict__ inst, uint32_t n,
                                                             Transform* __restrict__ xforms, uint64_t* __restrict__ table,
                                                             float* __restrict__ boxes, BuildHead* head) {
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const InstanceUpload u = inst[i];
        Transform inv;
        InverseAffine(u.xform.m, inv.m);
        xforms[2 * (size_t)i] = inv;
        xforms[2 * (size_t)i + 1] = u.xform;
        table[i] = u.blas;
        float blo[3], bhi[3];
        float wlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, whi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        if (BlasRootBox(RebaseOn<const uint8_t>(inst, u.blas), blo, bhi)) {
            float mag[3] = {0.0f, 0.0f, 0.0f};
            for (int c = 0; c < 8; ++c) {
                const float x = (c & 4) ? bhi[0] : blo[0];
                const float y = (c & 2) ? bhi[1] : blo[1];
                const float z = (c & 1) ? bhi[2] : blo[2];
                for (int a = 0; a < 3; ++a) {
                    const float* m = u.xform.m[a];
                    const float p = ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
                    const float g = ((fabsf(m[0] * x) + fabsf(m[1] * y)) + fabsf(m[2] * z)) + fabsf(m[3]);
                    wlo[a] = fminf(wlo[a], p);
                    whi[a] = fmaxf(whi[a], p);
                    mag[a] = fmaxf(mag[a], g);
                }
            }
            for (int a = 0; a < 3; ++a) {
                const float e = mag[a] * kInstanceBoxWiden;
                wlo[a] = wlo[a] - e;
                whi[a] = whi[a] + e;
                mn[a] = FloatToOrdered(wlo[a]);
                mx[a] = FloatToOrdered(whi[a]);
            }
        }
        float4* b = reinterpret_cast<float4*>(boxes) + 2 * (size_t)i;
        b[0] = make_float4(wlo[0], wlo[1], wlo[2], 0.0f);
        b[1] = make_float4(whi[0], whi[1], whi[2], 0.0f);
    }
    ReduceBoundsToGlobal(mn, mx, head->bounds);
}

__global__ __launch_bounds__(kBlock) void TlasMortonKernel(const float* __restrict__ boxes, uint32_t n,
                                                           const BuildHead* head, uint32_t* keys, uint32_t* vals,
                                                           int passes, uint32_t* sort_hist, uint32_t* sort_status,
                                                           size_t status_words) {
    auto centroid = [&](uint32_t t, float c[3]) {
        const float4 lo = reinterpret_cast<const float4*>(boxes)[2 * (size_t)t];
        const float4 hi = reinterpret_cast<const float4*>(boxes)[2 * (size_t)t + 1];
        c[0] = (lo.x + hi.x) * 0.5f;
        c[1] = (lo.y + hi.y) * 0.5f;
        c[2] = (lo.z + hi.z) * 0.5f;
    };
    MortonAndHistogram(n, head, centroid, keys, vals, passes, sort_hist, sort_status, status_words);
}

// One thread per instance leaf; block b's leaves form one refit window.
__global__ __launch_bounds__(kBlock) void RefitSceneKernel(const float* __restrict__ boxes, uint32_t n,
                                                           BvhNode* __restrict__ nodes) {
    __shared__ RefitLds<kBlock> lds;
    const uint32_t base = blockIdx.x * kBlock;
    const uint32_t i = base + threadIdx.x;
    const uint32_t leaf = n - 1 + i;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    uint32_t parent = kInvalid;
    if (i < n) {
        const uint32_t id = nodes[leaf].child1;
        parent = nodes[leaf].parent;
        const float4 blo = reinterpret_cast<const float4*>(boxes)[2 * (size_t)id];
        const float4 bhi = reinterpret_cast<const float4*>(boxes)[2 * (size_t)id + 1];
        // A TLAS leaf stores the instance's world box twice (lbvh_fit_aabb_scene.comp:125-128).
        float4* q = reinterpret_cast<float4*>(nodes + leaf);
        q[0] = make_float4(blo.x, blo.y, blo.z, __uint_as_float(kInvalid));
        q[1] = make_float4(bhi.x, bhi.y, bhi.z, __uint_as_float(id));
        q[2] = make_float4(blo.x, blo.y, blo.z, __uint_as_float(parent));
        q[3] = make_float4(bhi.x, bhi.y, bhi.z, __uint_as_float(0u));
        lo[0] = blo.x; lo[1] = blo.y; lo[2] = blo.z;
        hi[0] = bhi.x; hi[1] = bhi.y; hi[2] = bhi.z;
    }
    WindowRefit<kBlock>(nodes, 2 * n - 1, base, kBlock, 0, i < n, i + 1 < n ? i : kInvalid, leaf, parent, lo, hi,
                        lds);
}

}  // namespace

// Gather the triangles of several meshes into one unindexed list: thread t
// finds its mesh by binary search over first_tri.
__global__ __launch_bounds__(kBlock) void GatherSubMeshesKernel(const SubMesh* __restrict__ table, uint32_t meshes,
                                                                uint32_t tris, float* __restrict__ out) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= tris) return;
    uint32_t lo = 0, hi = meshes;  // last mesh with first_tri <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (table[mid].first_tri <= t) lo = mid;
        else hi = mid;
    }
    const SubMesh& sm = table[lo];
    const MeshView m{RebaseOn<const uint8_t>(out, sm.vertices), RebaseOn<const void>(out, sm.indices),
                     sm.vertex_stride, sm.index16, sm.vertex_count, sm.tri_count};
    float v[9];
    LoadTriangle(m, t - sm.first_tri, v);
    float* o = out + 9 * (size_t)t;
    for (int k = 0; k < 9; ++k) o[k] = v[k];
}

// Box i of an AABB list as the unindexed triangle (min, max, min).
__global__ __launch_bounds__(kBlock) void GatherAabbsKernel(const uint8_t* __restrict__ aabbs, uint32_t count,
                                                            uint32_t stride, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const float* b = reinterpret_cast<const float*>(aabbs + (size_t)i * stride);
    float* o = out + 9 * (size_t)i;
    const float x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4], z1 = b[5];
    o[0] = x0; o[1] = y0; o[2] = z0;
    o[3] = x1; o[4] = y1; o[5] = z1;
    o[6] = x0; o[7] = y0; o[8] = z0;
}

// ---------------------------------------------------------------------------
// Host launchers

size_t RestructureScratchBytes(uint32_t tris) { return LayoutRestructure(tris).total; }

size_t BlasBuildScratchBytes(uint32_t tris, bool restructure) {
    const size_t lbvh = LayoutBlas(tris).total;
    const size_t rs = restructure ? LayoutRestructure(tris).total : 0;
    return std::max(lbvh, rs);
}

void BuildBlas(const MeshDesc& mesh, bool restructure, void* scratch, void* result, hipStream_t s) {
    const uint32_t n = mesh.tri_count;
    AccelHeader* hdr = static_cast<AccelHeader*>(result);
    BvhNode* nodes = At<BvhNode>(result, sizeof(AccelHeader));
    const BlasScratchLayout L = LayoutBlas(n);
    BuildHead* head = At<BuildHead>(scratch, L.head);
    const SortScratch sc = CarveSortScratch(At<void>(scratch, L.sort), n);
    const bool do_restructure = restructure && n >= 64;
    InitBuildKernel<<<1, kBlock, 0, s>>>(hdr, kAccelBlas, n, sizeof(AccelHeader), 0, 0,
                                         (do_restructure ? kAccelFlagRestructured : 0u) | mesh.accel_flags, head,
                                         sc.hist, sc.tickets);
    if (n == 0) return;
    const MeshView m = ViewOf(mesh);
    MeshBoundsKernel<<<std::min(kBoundsGrid, (Blocks(n) + kBoundsTris - 1) / kBoundsTris), kBlock, 0, s>>>(m, head);
    uint32_t* keys = At<uint32_t>(scratch, L.keys);
    uint32_t* vals = At<uint32_t>(scratch, L.vals);
    const int passes = SortPasses(30);
    MeshMortonKernel<<<Blocks(n), kBlock, 0, s>>>(m, head, keys, vals, passes, sc.hist, sc.status, sc.status_words);
    const bool in_place = RadixSortPairs(keys, vals, At<uint32_t>(scratch, L.keys_tmp), At<uint32_t>(scratch, L.vals_tmp),
                                         n, 30, sc, true, s);
    const uint32_t* sk = in_place ? keys : At<uint32_t>(scratch, L.keys_tmp);
    const uint32_t* sv = in_place ? vals : At<uint32_t>(scratch, L.vals_tmp);
    EmitHierarchyKernel<<<Blocks(n), kBlock, 0, s>>>(sk, sv, n, nodes);
    LaunchMeshRefit(m, result, s);
    if (do_restructure) RestructureBlas(n, scratch, result, s);
}

size_t MultiMeshTableOffset(uint32_t tris) { return AlignUp(size_t(tris) * 36, 256); }

size_t MultiMeshScratchBytes(uint32_t meshes, uint32_t tris, bool build, bool restructure) {
    const size_t head = MultiMeshTableOffset(tris) + AlignUp(size_t(meshes) * sizeof(SubMesh), 256);
    return head + (build ? BlasBuildScratchBytes(tris, restructure) : 0);
}

static MeshDesc GatheredMesh(void* scratch, uint32_t tris) {
    MeshDesc m;
    m.vertices = scratch;
    m.vertex_count = 3 * tris;
    m.vertex_stride = 12;
    m.indices = nullptr;
    m.tri_count = tris;
    return m;
}

void BuildMultiMesh(const SubMesh* table, uint32_t meshes, uint32_t tris, bool restructure, void* scratch,
                    void* result, hipStream_t s) {
    if (tris) GatherSubMeshesKernel<<<Blocks(tris), kBlock, 0, s>>>(table, meshes, tris, static_cast<float*>(scratch));
    void* rest = static_cast<uint8_t*>(scratch) + MultiMeshScratchBytes(meshes, tris, false, false);
    BuildBlas(GatheredMesh(scratch, tris), restructure, rest, result, s);
}

size_t AabbScratchBytes(uint32_t count, bool build, bool restructure) {
    return AlignUp(size_t(count) * 36, 256) + (build ? BlasBuildScratchBytes(count, restructure) : 0);
}

void BuildAabbs(const AabbDesc& boxes, bool restructure, void* scratch, void* result, hipStream_t s) {
    const uint32_t n = boxes.count;
    if (n)
        GatherAabbsKernel<<<Blocks(n), kBlock, 0, s>>>(static_cast<const uint8_t*>(boxes.aabbs), n, boxes.stride,
                                                       static_cast<float*>(scratch));
    MeshDesc m = GatheredMesh(scratch, n);
    m.accel_flags = kAccelFlagAabbs;
    BuildBlas(m, restructure, static_cast<uint8_t*>(scratch) + AabbScratchBytes(n, false, false), result, s);
}

void UpdateAabbs(const AabbDesc& boxes, void* scratch, void* result, hipStream_t s) {
    const uint32_t n = boxes.count;
    if (!n) return;
    GatherAabbsKernel<<<Blocks(n), kBlock, 0, s>>>(static_cast<const uint8_t*>(boxes.aabbs), n, boxes.stride,
                                                   static_cast<float*>(scratch));
    UpdateBlas(GatheredMesh(scratch, n), result, s);
}

void UpdateMultiMesh(const SubMesh* table, uint32_t meshes, uint32_t tris, void* scratch, void* result,
                     hipStream_t s) {
    if (!tris) return;
    GatherSubMeshesKernel<<<Blocks(tris), kBlock, 0, s>>>(table, meshes, tris, static_cast<float*>(scratch));
    UpdateBlas(GatheredMesh(scratch, tris), result, s);
}

void UpdateBlas(const MeshDesc& mesh, void* result, hipStream_t s) {
    const uint32_t n = mesh.tri_count;
    if (n == 0) return;
    LaunchMeshRefit(ViewOf(mesh), result, s);
}

int TreeletLeaves() { return g_treelet_leaves.load(std::memory_order_relaxed); }
bool SetTreeletLeaves(int leaves) {
    if (leaves != 5 && leaves != 7 && leaves != 9) return false;
    g_treelet_leaves.store(leaves, std::memory_order_relaxed);
    return true;
}

void RestructureBlas(uint32_t n, void* scratch, void* result, hipStream_t s) {
    if (n < 64) return;
    BvhNode* nodes = At<BvhNode>(result, sizeof(AccelHeader));
    const RestructureLayout L = LayoutRestructure(n);
    uint32_t* counter = At<uint32_t>(scratch, L.counter);
    uint32_t* roots = At<uint32_t>(scratch, L.roots);
    uint32_t* counts = At<uint32_t>(scratch, L.counts);
    uint32_t min_prims = 64;  // restructure_hlbvh.cpp:34-35,151,276: 64, 128, 256
    for (int round = 0; round < 3; ++round, min_prims *= 2) {
        InitTreeletCountsKernel<<<Blocks(n), kBlock, 0, s>>>(n, counts, counter);
        FindTreeletRootsKernel<<<Blocks(n), kBlock, 0, s>>>(nodes, n, min_prims, counts, counter, roots);
        const dim3 grid(Blocks(n, (int)min_prims));
        switch (TreeletLeaves()) {
            case 5: RestructureTreeletsKernel<5><<<grid, 64, 0, s>>>(nodes, n, counter, roots, counts); break;
            case 9: RestructureTreeletsKernel<9><<<grid, 64, 0, s>>>(nodes, n, counter, roots, counts); break;
            default: RestructureTreeletsKernel<7><<<grid, 64, 0, s>>>(nodes, n, counter, roots, counts); break;
        }
    }
}

void BatchRestructure(const BatchMesh* table, uint32_t meshes, uint32_t total_chunks, const uint32_t max_roots[3],
                      void* scratch, void* roots, hipStream_t s) {
    if (meshes == 0 || total_chunks == 0) return;
    uint32_t* counter = static_cast<uint32_t*>(scratch) + kBatchRootsCounterWord;
    uint2* list = static_cast<uint2*>(roots);
    const uint32_t blocks = (uint32_t)(((uint64_t)total_chunks * kRefitChunk + kBlock - 1) / kBlock);
    uint32_t min_prims = 64;  // restructure_hlbvh.cpp:34-35,151,276: 64, 128, 256
    for (int round = 0; round < 3; ++round, min_prims *= 2) {
        BatchTreeletInitKernel<<<blocks, kBlock, 0, s>>>(table, meshes, total_chunks, counter);
        BatchFindTreeletRootsKernel<<<blocks, kBlock, 0, s>>>(table, meshes, total_chunks, min_prims, counter, list);
        if (max_roots[round] == 0) continue;
        const dim3 grid(max_roots[round]);
        switch (TreeletLeaves()) {
            case 5: BatchRestructureTreeletsKernel<5><<<grid, 64, 0, s>>>(table, counter, list); break;
            case 9: BatchRestructureTreeletsKernel<9><<<grid, 64, 0, s>>>(table, counter, list); break;
            default: BatchRestructureTreeletsKernel<7><<<grid, 64, 0, s>>>(table, counter, list); break;
        }
    }
}

size_t TlasBuildScratchBytes(uint32_t instances) { return LayoutTlas(instances).total; }

size_t TlasUpdateScratchBytes(uint32_t instances) { return LayoutTlas(instances).keys; }

void UpdateTlas(const InstanceUpload* instances, uint32_t n, void* scratch, void* result, hipStream_t s) {
    if (n == 0) return;
    const uint64_t nodes_off = sizeof(AccelHeader);
    const uint64_t xf_off = nodes_off + TlasNodesBytes(n);
    const uint64_t table_off = xf_off + TlasTransformsBytes(n);
    const TlasScratchLayout L = LayoutTlas(n);
    // The scene bounds the instance kernel accumulates are not used by a refit.
    float* boxes = At<float>(scratch, L.boxes);
    TlasInstanceKernel<<<Blocks(n), kBlock, 0, s>>>(instances, n, At<Transform>(result, xf_off),
                                                    At<uint64_t>(result, table_off), boxes, At<BuildHead>(scratch, L.head));
    RefitSceneKernel<<<Blocks(n), kBlock, 0, s>>>(boxes, n, At<BvhNode>(result, nodes_off));
}

void BuildTlas(const InstanceUpload* instances, uint32_t n, void* scratch, void* result, hipStream_t s) {
    AccelHeader* hdr = static_cast<AccelHeader*>(result);
    const uint64_t nodes_off = sizeof(AccelHeader);
    const uint64_t xf_off = nodes_off + TlasNodesBytes(n);
    const uint64_t table_off = xf_off + TlasTransformsBytes(n);
    BvhNode* nodes = At<BvhNode>(result, nodes_off);
    const TlasScratchLayout L = LayoutTlas(n);
    BuildHead* head = At<BuildHead>(scratch, L.head);
    const SortScratch sc = CarveSortScratch(At<void>(scratch, L.sort), n);
    InitBuildKernel<<<1, kBlock, 0, s>>>(hdr, kAccelTlas, n, nodes_off, xf_off, table_off, 0u, head, sc.hist, sc.tickets);
    if (n == 0) return;
    float* boxes = At<float>(scratch, L.boxes);
    TlasInstanceKernel<<<Blocks(n), kBlock, 0, s>>>(instances, n, At<Transform>(result, xf_off),
                                                    At<uint64_t>(result, table_off), boxes, head);
    uint32_t* keys = At<uint32_t>(scratch, L.keys);
    uint32_t* vals = At<uint32_t>(scratch, L.vals);
    TlasMortonKernel<<<Blocks(n), kBlock, 0, s>>>(boxes, n, head, keys, vals, SortPasses(30), sc.hist, sc.status,
                                                  sc.status_words);
    const bool in_place = RadixSortPairs(keys, vals, At<uint32_t>(scratch, L.keys_tmp), At<uint32_t>(scratch, L.vals_tmp),
                                         n, 30, sc, true, s);
    const uint32_t* sk = in_place ? keys : At<uint32_t>(scratch, L.keys_tmp);
    const uint32_t* sv = in_place ? vals : At<uint32_t>(scratch, L.vals_tmp);
    EmitHierarchyKernel<<<Blocks(n), kBlock, 0, s>>>(sk, sv, n, nodes);
    RefitSceneKernel<<<Blocks(n), kBlock, 0, s>>>(boxes, n, nodes);
}

}  // namespace rr
