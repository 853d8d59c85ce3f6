// build_device.h — device helpers shared by the single-mesh builder
// (build.hip) and the batched builder (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "build.h"
#include "rr_device.h"

namespace rr {

struct MeshView {
    const uint8_t* vertices;
    const void*    indices;
    uint32_t       stride;
    uint32_t       index16;
    uint32_t       vertex_count;
    uint32_t       tri_count;
};

inline MeshView ViewOf(const MeshDesc& m) {
    return MeshView{static_cast<const uint8_t*>(m.vertices), m.indices, m.vertex_stride, m.index16, m.vertex_count,
                    m.tri_count};
}

// Fetch triangle t.  Indices past the vertex array are clamped to its last
// vertex so that a bad index buffer cannot read out of bounds.
__device__ __forceinline__ void LoadTriangle(const MeshView& m, uint32_t t, float v[9]) {
    uint32_t i0, i1, i2;
    if (!m.indices) {  // unindexed: triangle t is vertices 3t, 3t+1, 3t+2
        i0 = 3 * t; i1 = 3 * t + 1; i2 = 3 * t + 2;
    } else if (m.index16) {
        const uint16_t* ix = static_cast<const uint16_t*>(m.indices) + 3 * (size_t)t;
        i0 = ix[0]; i1 = ix[1]; i2 = ix[2];
    } else {
        const uint32_t* ix = static_cast<const uint32_t*>(m.indices) + 3 * (size_t)t;
        i0 = ix[0]; i1 = ix[1]; i2 = ix[2];
    }
    const uint32_t last = m.vertex_count - 1;
    i0 = min(i0, last); i1 = min(i1, last); i2 = min(i2, last);
    const float* p0 = reinterpret_cast<const float*>(m.vertices + (size_t)i0 * m.stride);
    const float* p1 = reinterpret_cast<const float*>(m.vertices + (size_t)i1 * m.stride);
    const float* p2 = reinterpret_cast<const float*>(m.vertices + (size_t)i2 * m.stride);
    v[0] = p0[0]; v[1] = p0[1]; v[2] = p0[2];
    v[3] = p1[0]; v[4] = p1[1]; v[5] = p1[2];
    v[6] = p2[0]; v[7] = p2[1]; v[8] = p2[2];
}

// Centroid of triangle t's box (lbvh_calc_morton_codes_mesh.comp:106-112).
__device__ __forceinline__ void TriangleCentroid(const MeshView& m, uint32_t t, float c[3]) {
    float v[9];
    LoadTriangle(m, t, v);
    for (int a = 0; a < 3; ++a) {
        const float lo = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
        const float hi = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
        c[a] = (lo + hi) * 0.5f;
    }
}

// Karras 2012 binary radix tree over sorted codes
// (lbvh_emit_hierarchy_mesh.comp:85-217).  delta(i, j) is the common prefix
// length of codes i and j, with 32 + clz(i ^ j) for equal codes and -1 for
// j outside [0, n).  `code(k)` returns sorted code k (from memory, or from a
// copy in LDS); ci = code(i).
template <typename Code>
__device__ __forceinline__ int DeltaOf(const Code& code, int n, int i, uint32_t ci, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t b = code(j);
    return ci != b ? __clz(ci ^ b) : 32 + __clz((uint32_t)(i ^ j));
}

// Node i's topology (and leaf i's primitive) from the sorted codes: Karras
// 2012.  A leaf's other words (its vertices or box, child0 = kInvalid,
// update = 0) are written by the refit that follows every emission
// (WriteLeaf, RefitSceneKernel), which reads child1 and parent from here.
template <typename Code>
__device__ __forceinline__ void EmitNodeWith(const Code& code, const uint32_t* __restrict__ prims, uint32_t n,
                                             uint32_t i, BvhNode* __restrict__ nodes) {
    const uint32_t leaf = n - 1 + i;
    nodes[leaf].child1 = prims[i];
    if (n == 1) {
        nodes[0].parent = kInvalid;
        return;
    }
    if (i >= n - 1) return;
    const int N = (int)n, I = (int)i;
    const uint32_t ci = code(I);
    // Direction and far end of node i's key range.
    const int d = (DeltaOf(code, N, I, ci, I + 1) - DeltaOf(code, N, I, ci, I - 1)) > 0 ? 1 : -1;
    const int dmin = DeltaOf(code, N, I, ci, I - d);
    int lmax = 2;
    while (DeltaOf(code, N, I, ci, I + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (DeltaOf(code, N, I, ci, I + (l + t) * d) > dmin) l += t;
    const int j = I + l * d;
    const int first = min(I, j), last = max(I, j);
    // Split: last position that shares more than the range's common prefix with `first`.
    const uint32_t cf = code(first);
    const int common = DeltaOf(code, N, first, cf, last);
    int lo = first, hi = last;
    while (hi > lo + 1) {
        const int mid = (lo + hi) / 2;
        if (DeltaOf(code, N, first, cf, mid) > common) lo = mid;
        else hi = mid;
    }
    const int split = lo;
    const uint32_t left = (split == first) ? (n - 1 + (uint32_t)split) : (uint32_t)split;
    const uint32_t right = (split + 1 == last) ? (n - 1 + (uint32_t)split + 1) : (uint32_t)split + 1;
    nodes[i].child0 = left;
    nodes[i].child1 = right;
    nodes[i].update = 0;
    nodes[left].parent = i;
    nodes[right].parent = i;
    if (i == 0) nodes[0].parent = kInvalid;
}

__device__ __forceinline__ void EmitNode(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ prims,
                                         uint32_t n, uint32_t i, BvhNode* __restrict__ nodes) {
    EmitNodeWith([codes](int k) { return codes[k]; }, prims, n, i, nodes);
}

// Bottom-up refit shared by BLAS build/update and TLAS build
// (lbvh_fit_aabb_mesh.comp:122-205, lbvh_fit_aabb_scene.comp:98-167).  Each
// finished node stores its box into its parent's child slot, then bumps the
// parent's arrival counter (wrapping from 1 back to 0); a previous value of
// 1 means the sibling already stored its box, so this thread merges both
// boxes and climbs on.
//
// Cross-workgroup hand-off without an L2 write-back per step: the slot is
// written with agent-scope (write-through) stores, the wave waits for them to
// complete (vmcnt(0)) before the counter's atomic add, and the second
// arriver reads the sibling's slot with agent-scope loads, which do not hit
// stale cache lines.  Both threads' accesses to a slot carry the agent-scope
// cache policy, and the counter's atomic orders them.
__device__ __forceinline__ void WaitStores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A child box side (three floats, 16-byte aligned) as one 12-byte vector
// store or load with the cache policy of agent-scope relaxed atomic accesses
// (sc1: written through, read past stale lines), one memory transaction per
// side instead of three.  The loads wait for their data before returning.
typedef float RrBox3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ void StoreBoxAgent(float* p, float x, float y, float z) {
    RrBox3 v;
    v.x = x; v.y = y; v.z = z;
    asm volatile("global_store_dwordx3 %0, %1, off sc1" : : "v"(p), "v"(v) : "memory");
}
// A whole node (four 16-byte quads), agent-scope, waited for.
__device__ __forceinline__ void LoadNodeAgent(const BvhNode* nd, float4& q0, float4& q1, float4& q2, float4& q3) {
    RrFloat4V a, b, c, d;
    asm volatile(
        "global_load_dwordx4 %0, %4, off sc1\n\tglobal_load_dwordx4 %1, %4, off offset:16 sc1\n\t"
        "global_load_dwordx4 %2, %4, off offset:32 sc1\n\tglobal_load_dwordx4 %3, %4, off offset:48 sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d)
        : "v"(nd)
        : "memory");
    q0 = make_float4(a.x, a.y, a.z, a.w);
    q1 = make_float4(b.x, b.y, b.z, b.w);
    q2 = make_float4(c.x, c.y, c.z, c.w);
    q3 = make_float4(d.x, d.y, d.z, d.w);
}
__device__ __forceinline__ void LoadBoxesAgent(const float* a, const float* b, float lo[3], float hi[3]) {
    RrBox3 x, y;
    asm volatile("global_load_dwordx3 %0, %2, off sc1\n\tglobal_load_dwordx3 %1, %3, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=&v"(x), "=&v"(y)
                 : "v"(a), "v"(b)
                 : "memory");
    lo[0] = x.x; lo[1] = x.y; lo[2] = x.z;
    hi[0] = y.x; hi[1] = y.y; hi[2] = y.z;
}

// A node's links (child0, parent), which no refit changes, loaded together.
__device__ __forceinline__ void LoadLinks(const BvhNode* nd, uint32_t& c0, uint32_t& up) {
    c0 = nd->child0;
    up = nd->parent;
}

// Three dependent memory round trips per level: the slot stores and their
// completion, the counter atomic, and the sibling's boxes loaded together
// with the next node's links.
__device__ __forceinline__ void ClimbRefit(BvhNode* __restrict__ nodes, uint32_t total, uint32_t child, uint32_t p,
                                           float lo[3], float hi[3]) {
    // Guards: a parent index outside the node array or more levels than any
    // tree of `total` nodes can have stops the climb instead of faulting.
    if (p == kInvalid || p >= total) return;
    uint32_t c0, up;
    LoadLinks(nodes + p, c0, up);
    for (uint32_t level = 0; level < total; ++level) {
        BvhNode* pn = nodes + p;
        float* slot_lo = c0 == child ? pn->a0_min_or_v0 : pn->a1_min_or_v2;
        float* slot_hi = c0 == child ? pn->a0_max_or_v1 : pn->a1_max_or_v3;
        const float* other_lo = c0 == child ? pn->a1_min_or_v2 : pn->a0_min_or_v0;
        const float* other_hi = c0 == child ? pn->a1_max_or_v3 : pn->a0_max_or_v1;
        StoreBoxAgent(slot_lo, lo[0], lo[1], lo[2]);
        StoreBoxAgent(slot_hi, hi[0], hi[1], hi[2]);
        WaitStores();
        // Wrapping increment (old >= 1 ? 0 : old + 1): the first arrival
        // reads 0 and leaves 1, the second reads 1 and leaves 0, so `update`
        // is 0 between refits whichever nodes took this hand-off.
        const uint32_t old = __builtin_amdgcn_atomic_inc32(&pn->update, 1u, __ATOMIC_RELAXED, "agent");
        if (old == 0) return;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const bool next = up != kInvalid && up < total;
        uint32_t nc0 = kInvalid, nup = kInvalid;
        if (next) LoadLinks(nodes + up, nc0, nup);  // in flight with the box loads below
        float olo[3], ohi[3];
        LoadBoxesAgent(other_lo, other_hi, olo, ohi);
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], olo[a]);
            hi[a] = fmaxf(hi[a], ohi[a]);
        }
        if (!next) return;
        child = p;
        p = up;
        c0 = nc0;
        up = nup;
    }
}

// Leaf i: write its triangle (and links) into the leaf node; returns the
// leaf's parent and the triangle's box.
__device__ __forceinline__ uint32_t WriteLeaf(const MeshView& m, BvhNode* __restrict__ nodes, uint32_t i, float lo[3],
                                              float hi[3]) {
    const uint32_t n = m.tri_count;
    const uint32_t leaf = n - 1 + i;
    const uint32_t prim = nodes[leaf].child1;
    const uint32_t parent = nodes[leaf].parent;
    float v[9];
    LoadTriangle(m, prim, v);
    float4* q = reinterpret_cast<float4*>(nodes + leaf);
    q[0] = make_float4(v[0], v[1], v[2], __uint_as_float(kInvalid));
    q[1] = make_float4(v[3], v[4], v[5], __uint_as_float(prim));
    q[2] = make_float4(v[6], v[7], v[8], __uint_as_float(parent));
    q[3] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
        hi[a] = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
    }
    return parent;
}

// Windowed refit.  A workgroup of W threads refits leaves whose key
// positions form windows of consecutive positions [base, base + width) and
// keeps in LDS, for every internal node whose index lies in a window, an
// arrival counter, the node's links and its two child boxes.  In the Karras layout an
// internal node's index is one end of its key range, so every node whose
// subtree lies inside a window has its index in that window: those nodes are
// merged in LDS, with no global atomic and no wait for a global store.  A
// climb stops at the first node outside its window, or at an in-window node
// where it arrived first; after the workgroup barrier, a stop whose node saw
// only that one local arrival (the sibling subtree crosses the window edge)
// and every out-of-window stop go on with ClimbRefit's global hand-off, which
// every arrival at a node whose subtree crosses a window edge uses.  Boxes
// are min/max, so the result does not depend on which thread merges them.
// The `update` counter is 0 on every node between refits: nodes merged in
// LDS never touch it, and the global hand-off clears it.
// The window's links (child0, child1, parent of each in-window internal
// node) are read into LDS once, one node per thread, so the climb inside
// the window makes no global load, and the second arrival at an in-window
// node writes the whole node (both boxes and its links, 60 of its 64 bytes;
// `update` is left alone) with plain full-quad stores — the first arrival
// stores nothing unless the node turns out to cross the window edge.  No
// other thread writes an in-window node that both children reach inside
// the window.
//
// After the barrier the climbs that continue with the global hand-off (a
// few per workgroup) are handed to the first wave through LDS, in the box
// slots, which nobody reads any more: the other waves end at once instead
// of staying resident behind one lane's chain of memory round trips.
constexpr int kPendingWords = 11;  // child, p, lo[3], hi[3], node count, nodes (2 words)
template <int W>
struct RefitLds {
    uint32_t count[W];
    uint32_t child0[W];
    uint32_t child1[W];
    uint32_t parent[W];
    float box[2][6][W];  // [child side][lo xyz, hi xyz][slot]; then the pending climbs
    uint32_t pending;
};

// Every thread of the workgroup calls this (inactive ones with
// active = false): it holds a barrier.  Thread t owns LDS slot t, which is
// the slot of internal node `own` (its leaf's key position; pass kInvalid
// when the thread has no in-window node); slot0 = LDS slot of the window's
// first position; `total` = node count of the tree.
template <int W>
__device__ __forceinline__ void WindowRefit(BvhNode* __restrict__ nodes, uint32_t total, uint32_t base, uint32_t width,
                                            uint32_t slot0, bool active, uint32_t own, uint32_t leaf, uint32_t parent,
                                            float lo[3], float hi[3], RefitLds<W>& s) {
    static_assert(kPendingWords <= 12, "a pending climb per thread fits in the box slots");
    s.count[threadIdx.x] = 0;
    if (threadIdx.x == 0) s.pending = 0;
    uint32_t own_c0 = kInvalid, own_c1 = kInvalid, own_up = kInvalid;
    if (own < total) {
        own_c0 = nodes[own].child0;
        own_c1 = nodes[own].child1;
        own_up = nodes[own].parent;
    }
    s.child0[threadIdx.x] = own_c0;
    s.child1[threadIdx.x] = own_c1;
    s.parent[threadIdx.x] = own_up;
    __syncthreads();
    uint32_t child = leaf, p = parent, l = 0;
    int state = active ? 1 : 0;  // 0 done, 1 climbing in LDS, 2 global climb from (child, p), 3 first arrival at p
    for (uint32_t level = 0; state == 1; ++level) {
        if (p - base >= width || p >= total) {
            state = 2;
            break;
        }
        if (level >= width) {  // more levels than a window's subtree has: corrupt links
            state = 0;
            break;
        }
        l = slot0 + (p - base);
        const uint32_t c0 = s.child0[l];
        const uint32_t up = s.parent[l];
        const int side = c0 == child ? 0 : 1;
        for (int a = 0; a < 3; ++a) {
            s.box[side][a][l] = lo[a];
            s.box[side][3 + a][l] = hi[a];
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // the boxes are in LDS before the count moves
        const uint32_t old = __hip_atomic_fetch_add(&s.count[l], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0) {
            state = 3;
            break;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        float olo[3], ohi[3];
        for (int a = 0; a < 3; ++a) {
            olo[a] = s.box[side ^ 1][a][l];
            ohi[a] = s.box[side ^ 1][3 + a][l];
        }
        // Both children are in: the whole node, child 0's box first.
        const float* b0lo = side ? olo : lo;
        const float* b0hi = side ? ohi : hi;
        const float* b1lo = side ? lo : olo;
        const float* b1hi = side ? hi : ohi;
        float4* q = reinterpret_cast<float4*>(nodes + p);
        q[0] = make_float4(b0lo[0], b0lo[1], b0lo[2], __uint_as_float(c0));
        q[1] = make_float4(b0hi[0], b0hi[1], b0hi[2], __uint_as_float(s.child1[l]));
        q[2] = make_float4(b1lo[0], b1lo[1], b1lo[2], __uint_as_float(up));
        RrBox3 last;
        last.x = b1hi[0]; last.y = b1hi[1]; last.z = b1hi[2];
        *reinterpret_cast<RrBox3*>(&q[3]) = last;
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], olo[a]);
            hi[a] = fmaxf(hi[a], ohi[a]);
        }
        child = p;
        p = up;
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (state == 3 && s.count[l] == 1) state = 2;
    uint32_t* list = reinterpret_cast<uint32_t*>(&s.box[0][0][0]);
    if (state == 2) {
        uint32_t* q = list + kPendingWords * atomicAdd(&s.pending, 1u);
        const uint64_t base = reinterpret_cast<uint64_t>(nodes);
        q[0] = child;
        q[1] = p;
        for (int a = 0; a < 3; ++a) {
            q[2 + a] = __float_as_uint(lo[a]);
            q[5 + a] = __float_as_uint(hi[a]);
        }
        q[8] = total;
        q[9] = (uint32_t)base;
        q[10] = (uint32_t)(base >> 32);
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x >= 64) return;
    const uint32_t pending = s.pending;
    for (uint32_t e = threadIdx.x; e < pending; e += 64) {
        const uint32_t* q = list + kPendingWords * e;
        float plo[3], phi[3];
        for (int a = 0; a < 3; ++a) {
            plo[a] = __uint_as_float(q[2 + a]);
            phi[a] = __uint_as_float(q[5 + a]);
        }
        // The climb's own tree (workgroups of the batched refit hold several
        // meshes), addressed from this thread's `nodes` so that it stays a
        // global-memory pointer.
        BvhNode* tree = RebaseOn<BvhNode>(nodes, (uint64_t)q[9] | ((uint64_t)q[10] << 32));
        ClimbRefit(tree, q[8], q[0], q[1], plo, phi);
    }
}

}  // namespace rr
