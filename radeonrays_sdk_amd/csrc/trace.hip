// trace.hip — BVH2 ray traversal for gfx950.
//
// Behavioural spec: vlk/kernels/isect.comp:88-246 (single level) and
// isect_2l.comp:136-323 (two level): one lane per ray, depth-first stack
// traversal that descends into the nearer of two hit children first
// (child0 on ties) and defers the other; a TLAS leaf transforms the ray by
// the instance's inverse transform, pushes a sentinel and restarts at the
// BLAS root; popping the sentinel restores the world ray.
// Differences from the reference, all deliberate:
//   * triangles use the watertight test (rr_device.h) and closest hits are
//     ordered by (t, instance, primitive), so the result does not depend on
//     the BVH's shape;
//   * slab tests are conservative (widened by a few ulp);
//   * the stack keeps 16 entries per lane in LDS as a ring and spills the
//     oldest entry to scratch one at a time;
//   * a miss writes a complete record (inst_id = prim_id = ~0).
#include <hip/hip_runtime.h>

#include "rr_device.h"
#include "trace.h"

namespace rr {

namespace {

constexpr int kBlock = 256;

struct RayIn {
    float4 a;  // origin.xyz, min_t
    float4 b;  // direction.xyz, max_t
};

__device__ __forceinline__ RayIn LoadRay(const float4* __restrict__ rays, uint32_t i) {
    RayIn r;
    r.a = rays[2 * (size_t)i];
    r.b = rays[2 * (size_t)i + 1];
    return r;
}

// LDS ring + scratch spill stack for one lane.
struct LaneStack {
    uint32_t (*lds)[kBlock];
    uint32_t* spill;
    int sp;     // entries on the stack
    int gbase;  // entries [0, gbase) live in scratch, [gbase, sp) in LDS

    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ void push(uint32_t v) {
        if (sp == kLdsStackEntries + kSpillEntries) return;  // full: drop (tree deeper than the stack)
        if (sp - gbase == kLdsStackEntries) {
            spill[gbase] = lds[gbase & (kLdsStackEntries - 1)][threadIdx.x];
            ++gbase;
        }
        lds[sp & (kLdsStackEntries - 1)][threadIdx.x] = v;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        if (sp >= gbase) return lds[sp & (kLdsStackEntries - 1)][threadIdx.x];
        gbase = sp;
        return spill[sp];
    }
};

template <bool ANY, int OUT, bool TWO_LEVEL, bool STATS>
__global__ __launch_bounds__(kBlock) void TraceKernel(const uint8_t* __restrict__ accel, const float4* __restrict__ rays,
                                                      uint32_t ray_count, const uint32_t* __restrict__ indirect,
                                                      void* __restrict__ hits, uint32_t* __restrict__ scratch,
                                                      uint32_t* __restrict__ stats) {
    __shared__ uint32_t s_stack[kLdsStackEntries][kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n = ray_count;
    if (indirect) n = min(n, *indirect);
    if (i >= n) return;

    const AccelHeader* hdr = reinterpret_cast<const AccelHeader*>(accel);
    const uint32_t top_count = hdr->node_count;
    const BvhNode* top_nodes = reinterpret_cast<const BvhNode*>(accel + hdr->nodes_offset);
    const Transform* xforms = nullptr;
    const uint64_t* blas_table = nullptr;
    if (TWO_LEVEL) {
        xforms = reinterpret_cast<const Transform*>(accel + hdr->transforms_offset);
        blas_table = reinterpret_cast<const uint64_t*>(accel + hdr->blas_table_offset);
    }

    const RayIn ray = LoadRay(rays, i);
    RayState r;
    SetupRay(r, ray.a.x, ray.a.y, ray.a.z, ray.a.w, ray.b.x, ray.b.y, ray.b.z);
    const float tmax = ray.b.w;

    float best_t = tmax;
    uint32_t best_prim = kInvalid, best_inst = kInvalid;
    float bV = 0.0f, bW = 0.0f, bdet = 1.0f;
    float limit = WidenLimit(tmax);
    uint32_t n_internal_visits = 0, n_leaf_visits = 0, n_box_tests = 0, n_tri_tests = 0;

    LaneStack st{s_stack, scratch + (size_t)i * kSpillEntries, 0, 0};
    const BvhNode* nodes = top_nodes;
    uint32_t node_count = top_count;  // of the level being traversed (bounds every node index)
    uint32_t inst = kInvalid;  // kInvalid while in the top level
    uint32_t node = 0;
    bool alive = top_count != 0;
    while (alive) {
        if (node >= node_count) break;  // corrupt buffer: stop this ray rather than read out of bounds
        float4 q0, q1, q2, q3;
        LoadNode(nodes, node, q0, q1, q2, q3);
        const uint32_t c0 = __float_as_uint(q0.w);
        const uint32_t c1 = __float_as_uint(q1.w);
        if (c0 != kInvalid) {
            if (STATS) { ++n_internal_visits; n_box_tests += 2; }
            float ta, tb;
            const bool h0 = SlabTest(r, limit, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, ta);
            const bool h1 = SlabTest(r, limit, q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, tb);
            if (h0 || h1) {
                const bool c1first = h1 && (!h0 || ta > tb);
                if (h0 && h1) st.push(c1first ? c0 : c1);
                node = c1first ? c1 : c0;
                continue;
            }
        } else if (TWO_LEVEL && inst == kInvalid) {
            // Enter instance c1 (isect_2l.comp:222-237).
            if (STATS) ++n_leaf_visits;
            const uint8_t* blas = reinterpret_cast<const uint8_t*>(blas_table[c1]);
            const AccelHeader* bh = reinterpret_cast<const AccelHeader*>(blas);
            if (bh->node_count != 0) {
                const Transform x = xforms[2 * (size_t)c1];
                const float ox = ((x.m[0][0] * ray.a.x + x.m[0][1] * ray.a.y) + x.m[0][2] * ray.a.z) + x.m[0][3];
                const float oy = ((x.m[1][0] * ray.a.x + x.m[1][1] * ray.a.y) + x.m[1][2] * ray.a.z) + x.m[1][3];
                const float oz = ((x.m[2][0] * ray.a.x + x.m[2][1] * ray.a.y) + x.m[2][2] * ray.a.z) + x.m[2][3];
                const float dx = (x.m[0][0] * ray.b.x + x.m[0][1] * ray.b.y) + x.m[0][2] * ray.b.z;
                const float dy = (x.m[1][0] * ray.b.x + x.m[1][1] * ray.b.y) + x.m[1][2] * ray.b.z;
                const float dz = (x.m[2][0] * ray.b.x + x.m[2][1] * ray.b.y) + x.m[2][2] * ray.b.z;
                SetupRay(r, ox, oy, oz, ray.a.w, dx, dy, dz);
                inst = c1;
                nodes = reinterpret_cast<const BvhNode*>(blas + bh->nodes_offset);
                node_count = bh->node_count;
                st.push(kTopLevelSentinel);
                node = 0;
                continue;
            }
        } else {
            if (STATS) { ++n_leaf_visits; ++n_tri_tests; }
            TriHit h;
            if (WatertightTest(r, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, h) && h.t >= r.tmin &&
                h.t <= best_t) {
                const uint32_t hit_inst = TWO_LEVEL ? inst : 0u;
                if (ANY) {
                    best_t = h.t; best_prim = c1; best_inst = hit_inst;
                    bV = h.V; bW = h.W; bdet = h.det;
                    break;
                }
                if (CloserHit(h.t, hit_inst, c1, best_t, best_inst, best_prim)) {
                    best_t = h.t; best_prim = c1; best_inst = hit_inst;
                    bV = h.V; bW = h.W; bdet = h.det;
                    limit = WidenLimit(best_t);
                }
            }
        }
        // Pop the next deferred node.
        for (;;) {
            if (st.empty()) { alive = false; break; }
            node = st.pop();
            if (TWO_LEVEL && node == kTopLevelSentinel) {
                // Leave the instance and restore the world ray (isect_2l.comp:278-286).
                inst = kInvalid;
                nodes = top_nodes;
                node_count = top_count;
                SetupRay(r, ray.a.x, ray.a.y, ray.a.z, ray.a.w, ray.b.x, ray.b.y, ray.b.z);
                continue;
            }
            break;
        }
    }

    if (STATS) {
        uint4* s = reinterpret_cast<uint4*>(stats) + i;
        *s = make_uint4(n_internal_visits, n_leaf_visits, n_box_tests, n_tri_tests);
    }
    const bool hit = best_prim != kInvalid;
    const float u = hit ? bV / bdet : 0.0f;
    const float v = hit ? bW / bdet : 0.0f;
    if (OUT == kOutFullHit) {
        reinterpret_cast<float4*>(hits)[i] = hit ? make_float4(u, v, __uint_as_float(best_inst), __uint_as_float(best_prim))
                                                 : make_float4(0.0f, 0.0f, __uint_as_float(kInvalid), __uint_as_float(kInvalid));
    } else if (OUT == kOutInstanceId) {
        // Single level writes the primitive id (isect.comp:235), two level the instance (isect_2l.comp:311).
        reinterpret_cast<uint32_t*>(hits)[i] = hit ? (TWO_LEVEL ? best_inst : best_prim) : kInvalid;
    } else {
        float4* o = reinterpret_cast<float4*>(hits) + 2 * (size_t)i;
        o[0] = hit ? make_float4(u, v, __uint_as_float(best_inst), __uint_as_float(best_prim))
                   : make_float4(0.0f, 0.0f, __uint_as_float(kInvalid), __uint_as_float(kInvalid));
        o[1] = make_float4(hit ? best_t : tmax, 0.0f, 0.0f, 0.0f);
    }
}

template <bool ANY, bool TWO_LEVEL, bool STATS>
void LaunchOut(OutputKind out, dim3 g, hipStream_t s, const uint8_t* accel, const float4* rays, uint32_t count,
               const uint32_t* indirect, void* hits, uint32_t* scratch, uint32_t* stats) {
    switch (out) {
        case kOutFullHit:
            TraceKernel<ANY, kOutFullHit, TWO_LEVEL, STATS><<<g, kBlock, 0, s>>>(accel, rays, count, indirect, hits, scratch, stats);
            break;
        case kOutInstanceId:
            TraceKernel<ANY, kOutInstanceId, TWO_LEVEL, STATS><<<g, kBlock, 0, s>>>(accel, rays, count, indirect, hits, scratch, stats);
            break;
        default:
            TraceKernel<ANY, kOutFullHitT, TWO_LEVEL, STATS><<<g, kBlock, 0, s>>>(accel, rays, count, indirect, hits, scratch, stats);
            break;
    }
}

}  // namespace

size_t TraceScratchBytes(uint32_t ray_count) { return AlignUp((size_t)ray_count * kSpillEntries * 4, 256); }

void Trace(const void* accel, AccelKind kind, QueryKind query, OutputKind output, const void* rays, uint32_t ray_count,
           const uint32_t* indirect, void* hits, void* scratch, hipStream_t s) {
    if (ray_count == 0) return;
    const dim3 g((ray_count + kBlock - 1) / kBlock);
    const uint8_t* a = static_cast<const uint8_t*>(accel);
    const float4* r = static_cast<const float4*>(rays);
    uint32_t* sc = static_cast<uint32_t*>(scratch);
    const bool two = kind == kAccelTlas;
    if (query == kQueryAny) {
        if (two) LaunchOut<true, true, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
        else LaunchOut<true, false, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
    } else {
        if (two) LaunchOut<false, true, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
        else LaunchOut<false, false, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
    }
}

void TraceWithStats(const void* accel, AccelKind kind, QueryKind query, const void* rays, uint32_t ray_count, void* hits,
                    void* scratch, uint32_t* stats, hipStream_t s) {
    if (ray_count == 0) return;
    const dim3 g((ray_count + kBlock - 1) / kBlock);
    const uint8_t* a = static_cast<const uint8_t*>(accel);
    const float4* r = static_cast<const float4*>(rays);
    uint32_t* sc = static_cast<uint32_t*>(scratch);
    const bool two = kind == kAccelTlas;
    if (query == kQueryAny) {
        if (two) TraceKernel<true, kOutFullHitT, true, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats);
        else TraceKernel<true, kOutFullHitT, false, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats);
    } else {
        if (two) TraceKernel<false, kOutFullHitT, true, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats);
        else TraceKernel<false, kOutFullHitT, false, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats);
    }
}

}  // namespace rr
