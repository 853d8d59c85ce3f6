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
//   * the stack keeps 8, 16 or 32 entries per lane in LDS as a ring (8 by
//     default) and spills the oldest entry to scratch one at a time; a ray
//     that needs more than kStackCapacity entries ends and sets the
//     stack-overflow error word instead of dropping entries silently;
//   * a miss writes a complete record (inst_id = prim_id = ~0);
//   * every kernel checks the buffer's header first: a buffer this library
//     did not write sets the bad-buffer error word, and a kernel for the
//     other level returns at once (rrCmdIntersect on a buffer the host has
//     not seen built launches both levels and lets the header decide).
#include <hip/hip_runtime.h>

#include "device_errors.h"
#include "rr_device.h"
#include "trace.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>

namespace rr {

namespace {

constexpr int kBlock = 256;

// Trace scratch: kQueues ray-queue counters, each on its own 128-byte line,
// then kSpillEntries stack slots per lane.
constexpr uint32_t kQueues = 8;
constexpr uint32_t kQueueStride = 32;
constexpr uint32_t kQueueHeadWords = kQueues * kQueueStride;

struct RayIn {
    float4 a;  // origin.xyz, min_t
    float4 b;  // direction.xyz, max_t
};

__device__ __forceinline__ RayIn LoadRay(const float4* __restrict__ rays, uint32_t i) {
    RayIn r;
    r.a = LdG4(rays + 2 * (size_t)i);
    r.b = LdG4(rays + 2 * (size_t)i + 1);
    return r;
}

// True if the buffer's header names kind `want`.  A buffer this library did
// not write (wrong magic or kind) sets the bad-buffer error word and gives
// false, as does a valid buffer of the other kind (silently).
__device__ __forceinline__ bool HeaderIs(const AccelHeader* h, uint32_t want, uint32_t* err) {
    const uint32_t magic = h->magic, kind = h->kind;
    if (magic == kAccelMagic && (kind == kAccelBlas || kind == kAccelTlas)) return kind == want;
    if (err && blockIdx.x == 0 && threadIdx.x == 0) StG(err + kErrBadAccel, 1u);
    return false;
}

// Per-lane helper state of the persistent closest-hit kernel.
constexpr uint32_t kNoHelp = 0x7Fu, kHelpDone = 0x80u, kWaiting = 0x100u, kHelper1 = 0x10000u;

// A lane that owns a ray it has not finished, or that helps with another
// lane's ray and has not finished that subtree.
__device__ __forceinline__ bool Traversing(uint32_t rid, uint32_t help) {
    return rid != kInvalid ? (help & kWaiting) == 0 : help != kNoHelp && (help & kHelpDone) == 0;
}

// Position of the n-th (from 0) set bit of m; m must have more than n set bits.
__device__ __forceinline__ uint32_t NthSetBit(uint64_t m, uint32_t n) {
    uint32_t pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << w) - 1ull));
        if (n >= c) {
            n -= c;
            m >>= w;
            pos += (uint32_t)w;
        }
    }
    return pos;
}

// LDS ring of DEPTH entries + scratch spill stack for one lane.
template <int DEPTH = kLdsStackEntries>
struct LaneStack {
    static_assert((DEPTH & (DEPTH - 1)) == 0, "ring depth must be a power of two");
    static_assert(DEPTH >= kMinLdsStackEntries && kStackCapacity - DEPTH <= kSpillEntries, "spill slots per lane");
    uint32_t (*lds)[kBlock];
    uint32_t* spill;  // this block's spill slots (uniform): lane t's are [t * kSpillEntries, (t + 1) * kSpillEntries)
    uint32_t* err;    // device error words (may be null)
    int sp;     // entries pushed so far (index of the next push)
    int gbase;  // entries [bot, gbase) live in scratch, [gbase, sp) in LDS
    int bot = 0;  // entries [0, bot) were handed to other lanes (take_bottom)

    // A 32-bit lane offset from the block's base: one register, not a 64-bit pointer per lane.
    __device__ __forceinline__ uint32_t* slot(int k) const { return spill + (threadIdx.x * (uint32_t)kSpillEntries + (uint32_t)k); }

    __device__ __forceinline__ bool empty() const { return sp == bot; }
    __device__ __forceinline__ void clear() { sp = gbase = bot = 0; }
    __device__ __forceinline__ void push(uint32_t v) {
        if (sp == kStackCapacity) {
            // Tree deeper than the stack: report it and end the ray (its
            // deferred nodes, and in a scene the instance sentinel, are
            // discarded, so the lane finishes the current descent and stops).
            if (err) StG(err + kErrStackOverflow, 1u);
            clear();
            return;
        }
        if (sp - gbase == DEPTH) {
            StG(slot(gbase), lds[gbase & (DEPTH - 1)][threadIdx.x]);
            ++gbase;
        }
        lds[sp & (DEPTH - 1)][threadIdx.x] = v;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        // Read the LDS slot unconditionally (a plain ds_read) and replace it
        // from scratch only when the entry was spilled.
        uint32_t v = lds[sp & (DEPTH - 1)][threadIdx.x];
        if (sp < gbase) {
            gbase = sp;
            v = LdG(slot(sp));
        }
        return v;
    }
    // The oldest deferred entry (the largest remaining subtree), removed from
    // the bottom; the stack must not be empty.
    __device__ __forceinline__ uint32_t take_bottom() {
        uint32_t v;
        if (bot < gbase) {
            v = LdG(slot(bot));
        } else {
            v = lds[bot & (DEPTH - 1)][threadIdx.x];
            ++gbase;
        }
        ++bot;
        return v;
    }
};

template <bool ANY, int OUT, bool TWO_LEVEL, bool STATS>
__global__ __launch_bounds__(kBlock) void TraceKernel(const uint8_t* __restrict__ accel, const float4* __restrict__ rays,
                                                      uint32_t ray_count, const uint32_t* __restrict__ indirect,
                                                      void* __restrict__ hits, uint32_t* __restrict__ scratch,
                                                      uint32_t* __restrict__ stats, uint32_t* err) {
    __shared__ uint32_t s_stack[kLdsStackEntries][kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n = ray_count;
    if (indirect) n = min(n, *indirect);
    if (i >= n) return;

    const AccelHeader* hdr = reinterpret_cast<const AccelHeader*>(accel);
    if (!HeaderIs(hdr, TWO_LEVEL ? kAccelTlas : kAccelBlas, err)) return;
    const uint32_t top_count = hdr->node_count;
    const BvhNode* top_nodes = reinterpret_cast<const BvhNode*>(accel + hdr->nodes_offset);
    bool boxes = !TWO_LEVEL && (hdr->flags & kAccelFlagAabbs) != 0;  // leaf kind of the level traversed
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

    LaneStack<> st{s_stack, scratch + kQueueHeadWords + (size_t)blockIdx.x * kBlock * kSpillEntries, err, 0, 0};
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
            const uint8_t* blas = reinterpret_cast<const uint8_t*>(LdG(blas_table + c1));
            const AccelHeader* bh = reinterpret_cast<const AccelHeader*>(blas);
            const uint32_t bn = LdG(&bh->node_count);
            if (bn != 0) {
                const Transform x = xforms[2 * (size_t)c1];
                const float ox = ((x.m[0][0] * ray.a.x + x.m[0][1] * ray.a.y) + x.m[0][2] * ray.a.z) + x.m[0][3];
                const float oy = ((x.m[1][0] * ray.a.x + x.m[1][1] * ray.a.y) + x.m[1][2] * ray.a.z) + x.m[1][3];
                const float oz = ((x.m[2][0] * ray.a.x + x.m[2][1] * ray.a.y) + x.m[2][2] * ray.a.z) + x.m[2][3];
                const float dx = (x.m[0][0] * ray.b.x + x.m[0][1] * ray.b.y) + x.m[0][2] * ray.b.z;
                const float dy = (x.m[1][0] * ray.b.x + x.m[1][1] * ray.b.y) + x.m[1][2] * ray.b.z;
                const float dz = (x.m[2][0] * ray.b.x + x.m[2][1] * ray.b.y) + x.m[2][2] * ray.b.z;
                SetupRay(r, ox, oy, oz, ray.a.w, dx, dy, dz);
                inst = c1;
                nodes = reinterpret_cast<const BvhNode*>(blas + LdG(&bh->nodes_offset));
                node_count = bn;
                boxes = (LdG(&bh->flags) & kAccelFlagAabbs) != 0;
                st.push(kTopLevelSentinel);
                node = 0;
                continue;
            }
        } else {
            if (STATS) { ++n_leaf_visits; ++n_tri_tests; }
            TriHit h;
            if (LeafTest(boxes, r, best_t, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, h) && h.t >= r.tmin &&
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
                boxes = false;
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

// Write the query's output record for ray i.
template <int OUT, bool TWO_LEVEL>
__device__ __forceinline__ void WriteHit(void* __restrict__ hits, uint32_t i, bool hit, float u, float v,
                                         uint32_t inst, uint32_t prim, float t, float tmax) {
    if (OUT == kOutFullHit) {
        StG4(reinterpret_cast<float4*>(hits) + i,
             hit ? make_float4(u, v, __uint_as_float(inst), __uint_as_float(prim))
                 : make_float4(0.0f, 0.0f, __uint_as_float(kInvalid), __uint_as_float(kInvalid)));
    } else if (OUT == kOutInstanceId) {
        StG(reinterpret_cast<uint32_t*>(hits) + i, hit ? (TWO_LEVEL ? inst : prim) : kInvalid);
    } else {
        float4* o = reinterpret_cast<float4*>(hits) + 2 * (size_t)i;
        StG4(o, hit ? make_float4(u, v, __uint_as_float(inst), __uint_as_float(prim))
                    : make_float4(0.0f, 0.0f, __uint_as_float(kInvalid), __uint_as_float(kInvalid)));
        StG4(o + 1, make_float4(hit ? t : tmax, 0.0f, 0.0f, 0.0f));
    }
}

// ---------------------------------------------------------------------------
// Persistent single-level traversal.
//
// The grid holds only as many waves as the GPU keeps resident.  Rays are split
// into kQueues contiguous ranges, one per XCD (workgroups are dispatched to
// the XCDs round-robin), so that neighbouring rays share an L2; each wave
// takes 64 ray indices at a time from its range with one atomic and hands
// them to lanes as their rays finish, moving on to the other ranges once its
// own is empty.  Each lane runs a while-while loop (Aila & Laine, "Understanding
// the Efficiency of Ray Traversal on GPUs", HPG 2009): inner nodes until every
// busy lane holds a postponed leaf, then the leaves.  Leaves are still tested
// in depth-first order, so closest and any hits are those of the plain
// kernel; only the number of boxes tested differs.

__global__ void QueueInitKernel(uint32_t* counters, uint32_t ray_count, const uint32_t* __restrict__ indirect) {
    uint32_t n = ray_count;
    if (indirect) n = min(n, *indirect);
    const uint32_t part = (n + kQueues - 1) / kQueues;
    const uint32_t k = threadIdx.x;
    if (k < kQueues) counters[k * kQueueStride] = min(k * part, n);
}

// The same set-up done by a one-block persistent launch that runs before
// the full one (instead of a QueueInitKernel launch): the fence and barrier
// make it visible to the block's own atomics, stream order to the next launch.
__device__ __forceinline__ void InitQueues(uint32_t* counters, uint32_t n) {
    const uint32_t part = (n + kQueues - 1) / kQueues;
    if (threadIdx.x < kQueues) counters[threadIdx.x * kQueueStride] = min(threadIdx.x * part, n);
    __threadfence();
    __syncthreads();
}

template <bool ANY, int OUT, int DEPTH>
__global__ __launch_bounds__(kBlock) void TracePersistentKernel(const uint8_t* __restrict__ accel,
                                                                const float4* __restrict__ rays, uint32_t ray_count,
                                                                const uint32_t* __restrict__ indirect,
                                                                void* __restrict__ hits, uint32_t* __restrict__ scratch,
                                                                uint32_t min_busy, uint32_t fetch, uint32_t init_queues,
                                                                uint32_t* err) {
    __shared__ uint32_t s_stack[DEPTH][kBlock];
    uint32_t n = ray_count;
    if (indirect) n = min(n, *indirect);
    if (init_queues) InitQueues(scratch, n);
    const AccelHeader* hdr = reinterpret_cast<const AccelHeader*>(accel);
    if (!HeaderIs(hdr, kAccelBlas, err)) return;
    const uint32_t total = hdr->node_count;
    const uint32_t n_int = hdr->prim_count ? hdr->prim_count - 1 : 0;
    const BvhNode* nodes = reinterpret_cast<const BvhNode*>(accel + hdr->nodes_offset);
    const bool boxes = (hdr->flags & kAccelFlagAabbs) != 0;
    uint32_t* counters = scratch;
    const uint32_t lane = __lane_id();
    const uint32_t part = (n + kQueues - 1) / kQueues;

    // Wave-uniform queue state.
    uint32_t q = blockIdx.x % kQueues, tries = 0, q_next = 0, q_end = 0;
    bool drained = n == 0;

    // Lane state.
    uint32_t rid = kInvalid;
    RayState r;
    // best_t starts at the ray's tmax and changes only with best_prim, so on
    // a miss it still holds tmax.  bU, bV: barycentrics of the best hit so
    // far (divided by the determinant when the hit is taken).
    float best_t = 0.0f, bU = 0.0f, bV = 0.0f;
    float any_limit = 0.0f;  // any hit: the culling limit, fixed per ray (the first hit ends the ray)
    uint32_t best_prim = kInvalid;
    uint32_t node = kInvalid, leaf = kInvalid;
    LaneStack<DEPTH> st{s_stack, scratch + kQueueHeadWords + (size_t)blockIdx.x * kBlock * kSpillEntries,
                        err, 0, 0};
    // Closest hit, once the queues are drained: a lane without a ray may
    // help a lane of its wave by traversing a subtree that lane deferred
    // (see the drain step below).  One register: bits 0-6 the lane whose
    // ray this lane helps with (kNoHelp: none), kHelpDone once that subtree
    // is traversed, kWaiting on a lane whose own ray is traversed while
    // helpers still work on it, bits 16-23 how many lanes help with this
    // lane's ray.
    uint32_t help = kNoHelp;

    for (;;) {
        // Refill idle lanes.
        const uint64_t idle = __ballot(rid == kInvalid);
        if (idle && !drained) {
            if (q_next == q_end) {
                for (;;) {
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(&counters[q * kQueueStride], fetch);
                    b = __shfl(b, 0, 64);
                    const uint32_t end = min((q + 1) * part, n);
                    if (b < end) {
                        q_next = b;
                        q_end = min(b + fetch, end);
                        break;
                    }
                    if (++tries == kQueues) {
                        drained = true;
                        break;
                    }
                    q = q + 1 == kQueues ? 0 : q + 1;
                }
            }
            const uint32_t take = min((uint32_t)__popcll(idle), q_end - q_next);
            const uint32_t rank = RankIn(idle);
            if (rid == kInvalid && rank < take) {
                rid = q_next + rank;
                const RayIn ray = LoadRay(rays, rid);
                SetupRay(r, ray.a.x, ray.a.y, ray.a.z, ray.a.w, ray.b.x, ray.b.y, ray.b.z);
                best_t = ray.b.w;
                if (ANY) any_limit = WidenLimit(best_t);
                best_prim = kInvalid;
                bU = 0.0f; bV = 0.0f;
                st.clear();
                node = total ? 0u : kInvalid;
                leaf = kInvalid;
            }
            q_next += take;
        }
        if (!ANY && drained) {
            // Drain step (the whole wave, no lane diverged here).  Closest
            // hits do not depend on the order subtrees are visited in (ties
            // go to the lowest primitive), so a long ray's deferred subtrees
            // can be traversed by idle lanes of its wave, each with the
            // ray's state and its own stack; their best hits are merged into
            // the ray's lane one lane at a time, and the ray's hit is
            // written once it and every helper are done.
            // (a) Merge finished helpers into the lanes whose rays they helped.
            for (uint64_t fin = __ballot((help & kHelpDone) != 0); fin; fin &= fin - 1) {
                const int f = __builtin_ctzll(fin);
                const uint32_t o = (uint32_t)__shfl((int)(help & kNoHelp), f, 64);
                const float ft = __shfl(best_t, f, 64);
                const uint32_t fp = (uint32_t)__shfl((int)best_prim, f, 64);
                const float fu = __shfl(bU, f, 64), fv = __shfl(bV, f, 64);
                if (lane == o) {
                    if (fp != kInvalid && CloserHit(ft, 0u, fp, best_t, 0u, best_prim)) {
                        best_t = ft; best_prim = fp;
                        bU = fu; bV = fv;
                    }
                    help -= kHelper1;
                }
                if (lane == (uint32_t)f) help = kNoHelp;
            }
            // (b) Rays whose lane and helpers are all done.
            if ((help & kWaiting) && help < kHelper1) {
                const bool hit = best_prim != kInvalid;
                WriteHit<OUT, false>(hits, rid, hit, hit ? bU : 0.0f, hit ? bV : 0.0f, 0u, best_prim,
                                     best_t, best_t);
                rid = kInvalid;
                help = kNoHelp;
            }
            // (c) Idle lanes take the oldest deferred subtree of a lane still
            //     traversing (the k-th idle lane from the k-th such lane).
            const bool traversing = Traversing(rid, help);
            const uint64_t idle_lanes = __ballot(rid == kInvalid && help == kNoHelp);
            const uint64_t donors = __ballot(traversing && !st.empty());
            const uint32_t pairs = min((uint32_t)__popcll(idle_lanes), (uint32_t)__popcll(donors));
            if (pairs) {
                uint32_t give = kInvalid, give_owner = 0;
                if (((donors >> lane) & 1u) && RankIn(donors) < pairs) {
                    give = st.take_bottom();
                    give_owner = rid != kInvalid ? lane : (help & kNoHelp);
                }
                const uint32_t irank = RankIn(idle_lanes);
                const bool takes = ((idle_lanes >> lane) & 1u) && irank < pairs;
                const int src = takes ? (int)NthSetBit(donors, irank) : (int)lane;
                const uint32_t node_in = (uint32_t)__shfl((int)give, src, 64);
                const uint32_t owner_in = (uint32_t)__shfl((int)give_owner, src, 64);
                // A lane that takes nothing reads its own values (src == lane).
                r.ox = __shfl(r.ox, src, 64); r.oy = __shfl(r.oy, src, 64); r.oz = __shfl(r.oz, src, 64);
                r.dx = __shfl(r.dx, src, 64); r.dy = __shfl(r.dy, src, 64); r.dz = __shfl(r.dz, src, 64);
                r.tmin = __shfl(r.tmin, src, 64);
                r.ix = __shfl(r.ix, src, 64); r.iy = __shfl(r.iy, src, 64); r.iz = __shfl(r.iz, src, 64);
                r.sx = __shfl(r.sx, src, 64); r.sy = __shfl(r.sy, src, 64); r.sz = __shfl(r.sz, src, 64);
                r.k = (uint32_t)__shfl((int)r.k, src, 64);
                best_t = __shfl(best_t, src, 64);
                if (takes) {
                    help = owner_in;
                    best_prim = kInvalid;
                    bU = 0.0f; bV = 0.0f;
                    st.clear();
                    node = node_in < total ? node_in : kInvalid;
                    leaf = kInvalid;
                }
                // Each ray's lane counts its new helpers.
                for (uint64_t nh = __ballot(takes); nh; nh &= nh - 1) {
                    const uint32_t o = (uint32_t)__shfl((int)help, __builtin_ctzll(nh), 64);
                    if (lane == o) help += kHelper1;
                }
            }
        }
        if (__ballot(rid != kInvalid || help != kNoHelp) == 0) {
            if (drained) break;
            continue;
        }
        if (!Traversing(rid, help)) continue;
        const uint64_t entered = __ballot(1);

        // Traverse until the ray ends or fewer than min_busy lanes are still
        // traversing, then go back and hand new rays to the idle lanes.
        for (;;) {
            // Phase A: inner nodes until every traversing lane has postponed a leaf.
            while (node < n_int) {
                float4 q0, q1, q2, q3;
                LoadNode(nodes, node, q0, q1, q2, q3);
                const uint32_t c0 = __float_as_uint(q0.w);
                const uint32_t c1 = __float_as_uint(q1.w);
                float ta, tb;
                // Closest hit: recomputed (one register fewer across the loop).
                const float limit = ANY ? any_limit : WidenLimit(best_t);
                const bool h0 = SlabTest(r, limit, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, ta);
                const bool h1 = SlabTest(r, limit, q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, tb);
                if (h0 || h1) {
                    const bool c1first = h1 && (!h0 || ta > tb);
                    if (h0 && h1) st.push(c1first ? c0 : c1);
                    node = c1first ? c1 : c0;
                } else {
                    node = st.empty() ? kInvalid : st.pop();
                }
                if (node >= total && node != kInvalid) node = kInvalid;  // corrupt index: end the ray
                if (node != kInvalid && node >= n_int && leaf == kInvalid) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                }
                // Leave for the leaf phase once every descending lane holds a
                // leaf, or once too few lanes are still descending.
                if (!__any(leaf == kInvalid) || (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy)) break;
            }
            // Phase B: the postponed leaf, then any leaves that come next in order.
            if (leaf == kInvalid && node != kInvalid && node >= n_int) {
                leaf = node < total ? node : kInvalid;  // corrupt index: dropped
                node = st.empty() ? kInvalid : st.pop();
            }
            bool finished = false;
            while (leaf != kInvalid) {
                float4 q0, q1, q2, q3;
                LoadNode(nodes, leaf, q0, q1, q2, q3);
                const uint32_t prim = __float_as_uint(q1.w);
                TriHit h;
                if (LeafTest(boxes, r, best_t, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, h) &&
                    h.t >= r.tmin && h.t <= best_t) {
                    if (ANY) {
                        best_t = h.t; best_prim = prim;
                        bU = h.V / h.det; bV = h.W / h.det;
                        finished = true;
                        break;
                    }
                    if (CloserHit(h.t, 0u, prim, best_t, 0u, best_prim)) {
                        best_t = h.t; best_prim = prim;
                        bU = h.V / h.det; bV = h.W / h.det;
                    }
                }
                leaf = kInvalid;
                if (node != kInvalid && node >= n_int && node < total) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                }
            }
            if (finished || (node == kInvalid && leaf == kInvalid)) {
                node = kInvalid;
                leaf = kInvalid;
                if (!ANY && rid == kInvalid) {
                    help |= kHelpDone;  // a helper: merged into the ray's lane by the drain step
                    break;
                }
                if (!ANY && help >= kHelper1) {
                    help |= kWaiting;  // written by the drain step once the helpers are merged
                    break;
                }
                const bool hit = best_prim != kInvalid;
                WriteHit<OUT, false>(hits, rid, hit, hit ? bU : 0.0f, hit ? bV : 0.0f, 0u, best_prim,
                                     best_t, best_t);
                rid = kInvalid;
                break;
            }
            if (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy) break;  // refill (nothing left once drained)
            // Drained: back to the drain step once a lane of the wave has
            // finished (it may help) while another still holds deferred work.
            if (!ANY && drained && __ballot(1) != entered && __any(!st.empty())) break;
        }
    }
}

// ---------------------------------------------------------------------------
// Persistent two-level traversal: the single-level kernel's ray queues and
// while-while loop over a TLAS whose leaves are instances.  A lane that
// reaches an instance leaf enters it at once (world ray to object space,
// sentinel pushed, BLAS root next; isect_2l.comp:222-237).  Triangle leaves
// are postponed as in the single-level kernel, and a lane never pops the
// sentinel (leaves the instance) while it holds a postponed leaf, so every
// leaf is tested against its own instance's ray and nodes, and leaves are
// still tested in depth-first order: closest and any hits are those of the
// plain kernel.  Inner nodes and leaves are told apart by index (inner nodes
// [0, N-1), leaves [N-1, 2N-1) at both levels).

template <bool ANY, int OUT, int DEPTH>
__global__ __launch_bounds__(kBlock) void TracePersistent2LKernel(const uint8_t* __restrict__ accel,
                                                                  const float4* __restrict__ rays, uint32_t ray_count,
                                                                  const uint32_t* __restrict__ indirect,
                                                                  void* __restrict__ hits,
                                                                  uint32_t* __restrict__ scratch, uint32_t min_busy,
                                                                  uint32_t fetch, uint32_t init_queues, uint32_t* err) {
    __shared__ uint32_t s_stack[DEPTH][kBlock];
    // Each lane's world ray: origin xyz, direction xyz and reciprocal
    // direction, restored without divisions when the lane leaves an
    // instance.  The TLAS level only tests boxes (instance leaves are entered
    // with a fresh set-up), so the watertight shear and axis permutation of
    // the world ray are not kept.
    __shared__ float s_world[9][kBlock];
    uint32_t n = ray_count;
    if (indirect) n = min(n, *indirect);
    if (init_queues) InitQueues(scratch, n);
    const AccelHeader* hdr = reinterpret_cast<const AccelHeader*>(accel);
    if (!HeaderIs(hdr, kAccelTlas, err)) return;
    const uint32_t top_total = hdr->node_count;
    const uint32_t top_prims = hdr->prim_count;
    const uint32_t top_int = top_prims ? top_prims - 1 : 0;
    const BvhNode* top_nodes = reinterpret_cast<const BvhNode*>(accel + hdr->nodes_offset);
    const Transform* xforms = reinterpret_cast<const Transform*>(accel + hdr->transforms_offset);
    const uint64_t* blas_table = reinterpret_cast<const uint64_t*>(accel + hdr->blas_table_offset);
    uint32_t* counters = scratch;
    const uint32_t lane = __lane_id();
    const uint32_t part = (n + kQueues - 1) / kQueues;

    // Wave-uniform queue state.
    uint32_t q = blockIdx.x % kQueues, tries = 0, q_next = 0, q_end = 0;
    bool drained = n == 0;

    // Lane state: the ray of the level being traversed (the world ray waits
    // in LDS), the best hit (best_t is the ray's t_max until a hit), and the
    // level (inst == kInvalid at the top).
    uint32_t rid = kInvalid;
    RayState r;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;  // the culling limit is WidenLimit(best_t), recomputed per node
    uint32_t best_prim = kInvalid, best_inst = kInvalid;
    uint32_t node = kInvalid, leaf = kInvalid;
    uint32_t inst = kInvalid;
    const BvhNode* cur = top_nodes;
    uint32_t cur_int = top_int, cur_total = top_total;
    bool boxes = false;
    LaneStack<DEPTH> st{s_stack, scratch + kQueueHeadWords + (size_t)blockIdx.x * kBlock * kSpillEntries,
                        err, 0, 0};

    for (;;) {
        // Refill idle lanes.
        const uint64_t idle = __ballot(rid == kInvalid);
        if (idle && !drained) {
            if (q_next == q_end) {
                for (;;) {
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(&counters[q * kQueueStride], fetch);
                    b = __shfl(b, 0, 64);
                    const uint32_t end = min((q + 1) * part, n);
                    if (b < end) {
                        q_next = b;
                        q_end = min(b + fetch, end);
                        break;
                    }
                    if (++tries == kQueues) {
                        drained = true;
                        break;
                    }
                    q = q + 1 == kQueues ? 0 : q + 1;
                }
            }
            const uint32_t take = min((uint32_t)__popcll(idle), q_end - q_next);
            const uint32_t rank = RankIn(idle);
            if (rid == kInvalid && rank < take) {
                rid = q_next + rank;
                const RayIn ray = LoadRay(rays, rid);
                s_world[0][threadIdx.x] = ray.a.x; s_world[1][threadIdx.x] = ray.a.y; s_world[2][threadIdx.x] = ray.a.z;
                s_world[3][threadIdx.x] = ray.b.x; s_world[4][threadIdx.x] = ray.b.y; s_world[5][threadIdx.x] = ray.b.z;
                SetupRay(r, ray.a.x, ray.a.y, ray.a.z, ray.a.w, ray.b.x, ray.b.y, ray.b.z);
                s_world[6][threadIdx.x] = r.ix; s_world[7][threadIdx.x] = r.iy; s_world[8][threadIdx.x] = r.iz;
                best_t = ray.b.w;
                best_prim = kInvalid;
                best_inst = kInvalid;
                bu = 0.0f; bv = 0.0f;
                st.clear();
                inst = kInvalid;
                cur = top_nodes;
                cur_int = top_int;
                cur_total = top_total;
                boxes = false;
                node = top_total ? 0u : kInvalid;
                leaf = kInvalid;
            }
            q_next += take;
        }
        if (__ballot(rid != kInvalid) == 0) {
            if (drained) break;
            continue;
        }
        if (rid == kInvalid) continue;

        for (;;) {
            // Phase A: inner nodes, instance entries and exits, until every
            // traversing lane has postponed a leaf or too few still traverse.
            for (;;) {
                if (node < cur_int) {
                    float4 q0, q1, q2, q3;
                    LoadNode(cur, node, q0, q1, q2, q3);
                    const uint32_t c0 = __float_as_uint(q0.w);
                    const uint32_t c1 = __float_as_uint(q1.w);
                    float ta, tb;
                    const float limit = WidenLimit(best_t);
                    const bool h0 = SlabTest(r, limit, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, ta);
                    const bool h1 = SlabTest(r, limit, q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, tb);
                    if (h0 || h1) {
                        const bool c1first = h1 && (!h0 || ta > tb);
                        if (h0 && h1) st.push(c1first ? c0 : c1);
                        node = c1first ? c1 : c0;
                    } else {
                        node = st.empty() ? kInvalid : st.pop();
                    }
                } else if (inst == kInvalid && node < cur_total) {
                    // Instance leaf: enter it (isect_2l.comp:222-237).
                    // The transform is fetched with the BLAS address, and the BLAS
                    // header's three fields together: two dependent round trips.
                    const uint32_t id = __float_as_uint(LdG4(&top_nodes[node].a0_max_or_v1).w);
                    const uint8_t* blas = nullptr;
                    float4 x0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), x1 = x0, x2 = x0;
                    uint32_t bn = 0, bflags = 0;
                    uint64_t boff = 0;
                    if (id < top_prims) {
                        const float4* xp = reinterpret_cast<const float4*>(xforms + 2 * (size_t)id);
                        x0 = LdG4(xp); x1 = LdG4(xp + 1); x2 = LdG4(xp + 2);
                        blas = reinterpret_cast<const uint8_t*>(LdG(blas_table + id));
                        if (blas) {
                            const AccelHeader* bh = reinterpret_cast<const AccelHeader*>(blas);
                            bn = LdG(&bh->node_count);
                            boff = LdG(&bh->nodes_offset);
                            bflags = LdG(&bh->flags);
                        }
                    }
                    if (bn != 0) {
                        const float wox = s_world[0][threadIdx.x], woy = s_world[1][threadIdx.x], woz = s_world[2][threadIdx.x];
                        const float wdx = s_world[3][threadIdx.x], wdy = s_world[4][threadIdx.x], wdz = s_world[5][threadIdx.x];
                        const float ox = ((x0.x * wox + x0.y * woy) + x0.z * woz) + x0.w;
                        const float oy = ((x1.x * wox + x1.y * woy) + x1.z * woz) + x1.w;
                        const float oz = ((x2.x * wox + x2.y * woy) + x2.z * woz) + x2.w;
                        const float dx = (x0.x * wdx + x0.y * wdy) + x0.z * wdz;
                        const float dy = (x1.x * wdx + x1.y * wdy) + x1.z * wdz;
                        const float dz = (x2.x * wdx + x2.y * wdy) + x2.z * wdz;
                        SetupRay(r, ox, oy, oz, r.tmin, dx, dy, dz);
                        inst = id;
                        cur = reinterpret_cast<const BvhNode*>(blas + boff);
                        cur_total = bn;
                        cur_int = (bn - 1) / 2;
                        boxes = (bflags & kAccelFlagAabbs) != 0;
                        st.push(kTopLevelSentinel);
                        node = 0;
                    } else {
                        node = st.empty() ? kInvalid : st.pop();
                    }
                } else if (node == kTopLevelSentinel && leaf == kInvalid) {
                    // Leave the instance, restore the world ray (isect_2l.comp:278-286).
                    inst = kInvalid;
                    cur = top_nodes;
                    cur_int = top_int;
                    cur_total = top_total;
                    boxes = false;
                    r.ox = s_world[0][threadIdx.x]; r.oy = s_world[1][threadIdx.x]; r.oz = s_world[2][threadIdx.x];
                    r.ix = s_world[6][threadIdx.x]; r.iy = s_world[7][threadIdx.x]; r.iz = s_world[8][threadIdx.x];
                    node = st.empty() ? kInvalid : st.pop();
                } else if (inst != kInvalid && node < cur_total && leaf == kInvalid) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                } else {
                    break;  // a second leaf or the sentinel behind a postponed leaf, or done
                }
                // An index past the level's nodes (corrupt buffer) ends the ray.
                if (node != kInvalid && node != kTopLevelSentinel && node >= cur_total) node = kInvalid;
                if (!__any(leaf == kInvalid) || (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy)) break;
            }
            // Phase B: the postponed leaf, then the leaves that follow it in order.
            if (leaf == kInvalid && inst != kInvalid && node >= cur_int && node < cur_total) {
                leaf = node;
                node = st.empty() ? kInvalid : st.pop();
                if (node != kInvalid && node != kTopLevelSentinel && node >= cur_total) node = kInvalid;
            }
            bool finished = false;
            while (leaf != kInvalid) {
                float4 q0, q1, q2, q3;
                LoadNode(cur, leaf, q0, q1, q2, q3);
                const uint32_t prim = __float_as_uint(q1.w);
                TriHit h;
                if (LeafTest(boxes, r, best_t, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, h) &&
                    h.t >= r.tmin && h.t <= best_t) {
                    if (ANY) {
                        best_t = h.t; best_prim = prim; best_inst = inst;
                        if (OUT != kOutInstanceId) { bu = h.V / h.det; bv = h.W / h.det; }
                        finished = true;
                        break;
                    }
                    if (CloserHit(h.t, inst, prim, best_t, best_inst, best_prim)) {
                        best_t = h.t; best_prim = prim; best_inst = inst;
                        if (OUT != kOutInstanceId) { bu = h.V / h.det; bv = h.W / h.det; }
                    }
                }
                leaf = kInvalid;
                if (node >= cur_int && node < cur_total) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                    if (node != kInvalid && node != kTopLevelSentinel && node >= cur_total) node = kInvalid;
                }
            }
            if (finished || (node == kInvalid && leaf == kInvalid)) {
                const bool hit = best_prim != kInvalid;
                WriteHit<OUT, true>(hits, rid, hit, bu, bv, best_inst,
                                    best_prim, best_t, best_t);
                rid = kInvalid;
                node = kInvalid;
                leaf = kInvalid;
                break;
            }
            if (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy) break;  // refill (nothing left once drained)
        }
    }
}

template <bool ANY, bool TWO_LEVEL, bool STATS>
void LaunchOut(OutputKind out, dim3 g, hipStream_t s, const uint8_t* accel, const float4* rays, uint32_t count,
               const uint32_t* indirect, void* hits, uint32_t* scratch, uint32_t* stats) {
    uint32_t* err = DeviceErrorWords();
    switch (out) {
        case kOutFullHit:
            TraceKernel<ANY, kOutFullHit, TWO_LEVEL, STATS><<<g, kBlock, 0, s>>>(accel, rays, count, indirect, hits, scratch, stats, err);
            break;
        case kOutInstanceId:
            TraceKernel<ANY, kOutInstanceId, TWO_LEVEL, STATS><<<g, kBlock, 0, s>>>(accel, rays, count, indirect, hits, scratch, stats, err);
            break;
        default:
            TraceKernel<ANY, kOutFullHitT, TWO_LEVEL, STATS><<<g, kBlock, 0, s>>>(accel, rays, count, indirect, hits, scratch, stats, err);
            break;
    }
}

// ---------------------------------------------------------------------------
// Batched traversal: one persistent launch per level over many requests.
// The host orders the request table as: requests whose address the context
// recorded as a geometry, requests it has no record of, requests recorded as
// a scene.  The geometry kernel walks the chunks of the first two groups and
// the two-level kernel those of the last two; waves take 64-ray chunks from
// the kernel's counter, a chunk belongs to one request (found by binary
// search over first_chunk), and each lane carries its own request's BVH,
// query and output, so lanes of a wave may serve different requests.  An
// unrecorded request's level is read from its header when a wave takes one
// of its chunks, and a request of the other level is skipped whole; a
// recorded request whose header names the other level (the caller put
// another buffer at that address) is reported as a bad buffer and answered
// with misses.  No kernel beyond the two traversals and a one-wave init
// runs per batch: with several streams in flight, a wider helper kernel
// waits for another stream's persistent launch to leave the CUs.

// Each batch kernel's chunk range is cut into kQueues contiguous parts, one
// per XCD as the persistent kernels' rays are: a wave starts on part
// blockIdx % kQueues and moves to the next part once its own is drained, so
// the chunk atomics of different XCDs land on different words (one 128-byte
// line each; kBatch2LCounterWord is the two-level kernel's first, all
// inside the table offset).  One counter for the whole range took every
// wave's atomic on the same address.
constexpr uint32_t kBatchCounterWord = 0, kBatch2LCounterWord = kQueues * kQueueStride;
static_assert(2 * kQueues * kQueueStride * 4 <= kBatchTableOffset, "batch counters inside the head");

// Next chunk of this kernel's range [first, first + count) of the table's
// chunks, as [q_next, q_end) of request q_req; false once every part is
// drained.  `part` is the part the wave takes chunks from (counters[part *
// kQueueStride] counts them) and `tries` how many parts it found drained;
// a drained part is never revisited, since its counter only grows.  An
// unrecorded request (kBatchFlagCheckLevel) of the other level moves the
// part's counter past its chunks, up to the part's end, with one atomicMax.
// The loop ends because every counter only grows.  Wave-uniform: every lane
// calls it.
template <bool TWO_LEVEL>
__device__ __forceinline__ bool NextBatchChunk(const BatchRequest* __restrict__ table, uint32_t requests,
                                               uint32_t chunk, uint32_t first, uint32_t count, uint32_t* counters,
                                               uint32_t lane, uint32_t& part, uint32_t& tries, uint32_t& q_req,
                                               uint32_t& q_next, uint32_t& q_end) {
    for (;;) {
        // Part `part` holds chunks [first + p_lo, first + p_hi).
        const uint32_t p_lo = (uint32_t)((uint64_t)count * part / kQueues);
        const uint32_t p_hi = (uint32_t)((uint64_t)count * (part + 1) / kQueues);
        uint32_t* counter = counters + part * kQueueStride;
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(counter, 1u);
        c = __shfl(c, 0, 64);
        if (c >= p_hi - p_lo) {
            if (++tries == kQueues) return false;
            part = part + 1 == kQueues ? 0 : part + 1;
            continue;
        }
        c += first + p_lo;
        uint32_t lo = 0, hi = requests;  // last request with first_chunk <= c
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (table[mid].first_chunk <= c) lo = mid;
            else hi = mid;
        }
        if (table[lo].flags & kBatchFlagCheckLevel) {
            const AccelHeader* hdr = reinterpret_cast<const AccelHeader*>(table[lo].accel);
            const bool scene = LdG(&hdr->magic) == kAccelMagic && LdG(&hdr->kind) == kAccelTlas;
            if (scene != TWO_LEVEL) {
                const uint32_t next = lo + 1 < requests ? table[lo + 1].first_chunk : first + count;
                if (lane == 0) atomicMax(counter, min(next, first + p_hi) - (first + p_lo));
                continue;
            }
        }
        q_req = lo;
        q_next = (c - table[lo].first_chunk) * chunk;
        q_end = min(q_next + chunk, table[lo].ray_count);
        return true;
    }
}

__global__ __launch_bounds__(kBlock) void TraceBatchKernel(const BatchRequest* __restrict__ table, uint32_t requests,
                                                           uint32_t first_chunk, uint32_t chunk_count,
                                                           uint32_t* __restrict__ head, uint32_t* __restrict__ spill,
                                                           uint32_t min_busy, uint32_t chunk, uint32_t* err) {
    __shared__ uint32_t s_stack[kLdsStackEntries][kBlock];
    uint32_t* counter = head + kBatchCounterWord;
    const uint32_t lane = __lane_id();
    const uint64_t lt = LaneMaskLt();

    // Wave-uniform queue: rays [q_next, q_end) of request q_req.
    uint32_t q_req = 0, q_next = 0, q_end = 0;
    uint32_t part = blockIdx.x % kQueues, tries = 0;  // this wave's chunk part, parts found drained
    bool drained = false;

    // Lane state.
    uint32_t rid = kInvalid, lreq = 0, lflags = 0, n_int = 0, total = 0;
    bool boxes = false;
    const BvhNode* nodes = nullptr;
    RayState r;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;  // best_t is t_max until a hit; culling limit WidenLimit(best_t)
    uint32_t best_prim = kInvalid;
    uint32_t node = kInvalid, leaf = kInvalid;
    LaneStack<> st{s_stack, spill + (size_t)blockIdx.x * kBlock * kSpillEntries, err, 0, 0};

    for (;;) {
        const uint64_t idle = __ballot(rid == kInvalid);
        if (idle && !drained) {
            if (q_next == q_end)
                drained = !NextBatchChunk<false>(table, requests, chunk, first_chunk, chunk_count, counter, lane, part, tries,
                                                 q_req, q_next, q_end);
            const uint32_t take = drained ? 0u : min((uint32_t)__popcll(idle), q_end - q_next);
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            if (rid == kInvalid && rank < take) {
                rid = q_next + rank;
                lreq = q_req;
                const BatchRequest& rq = table[q_req];
                lflags = rq.flags;
                const uint8_t* accel = reinterpret_cast<const uint8_t*>(rq.accel);
                const AccelHeader* hdr = reinterpret_cast<const AccelHeader*>(accel);
                total = LdG(&hdr->node_count);
                if (LdG(&hdr->magic) != kAccelMagic || LdG(&hdr->kind) != kAccelBlas) {
                    // Not a geometry (no valid header, or a scene at an
                    // address recorded as a geometry): report it, answer misses.
                    if (err) StG(err + kErrBadAccel, 1u);
                    total = 0;
                }
                const uint32_t prims = LdG(&hdr->prim_count);
                n_int = (prims && total) ? prims - 1 : 0;
                nodes = reinterpret_cast<const BvhNode*>(accel + LdG(&hdr->nodes_offset));
                boxes = (LdG(&hdr->flags) & kAccelFlagAabbs) != 0;
                const RayIn ray = LoadRay(reinterpret_cast<const float4*>(rq.rays), rid);
                SetupRay(r, ray.a.x, ray.a.y, ray.a.z, ray.a.w, ray.b.x, ray.b.y, ray.b.z);
                best_t = ray.b.w;
                best_prim = kInvalid;
                bu = 0.0f; bv = 0.0f;
                st.clear();
                node = total ? 0u : kInvalid;
                leaf = kInvalid;
            }
            q_next += take;
        }
        if (__ballot(rid != kInvalid) == 0) {
            if (drained) break;
            continue;
        }
        if (rid == kInvalid) continue;

        const bool any = (lflags & 1u) != 0;
        for (;;) {
            while (node < n_int) {
                float4 q0, q1, q2, q3;
                LoadNode(nodes, node, q0, q1, q2, q3);
                const uint32_t c0 = __float_as_uint(q0.w);
                const uint32_t c1 = __float_as_uint(q1.w);
                float ta, tb;
                const float limit = WidenLimit(best_t);
                const bool h0 = SlabTest(r, limit, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, ta);
                const bool h1 = SlabTest(r, limit, q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, tb);
                if (h0 || h1) {
                    const bool c1first = h1 && (!h0 || ta > tb);
                    if (h0 && h1) st.push(c1first ? c0 : c1);
                    node = c1first ? c1 : c0;
                } else {
                    node = st.empty() ? kInvalid : st.pop();
                }
                if (node >= total && node != kInvalid) node = kInvalid;
                if (node != kInvalid && node >= n_int && leaf == kInvalid) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                }
                if (!__any(leaf == kInvalid) || (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy)) break;
            }
            if (leaf == kInvalid && node != kInvalid && node >= n_int) {
                leaf = node < total ? node : kInvalid;  // corrupt index: dropped
                node = st.empty() ? kInvalid : st.pop();
            }
            bool finished = false;
            while (leaf != kInvalid) {
                float4 q0, q1, q2, q3;
                LoadNode(nodes, leaf, q0, q1, q2, q3);
                const uint32_t prim = __float_as_uint(q1.w);
                TriHit h;
                if (LeafTest(boxes, r, best_t, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, h) &&
                    h.t >= r.tmin && h.t <= best_t) {
                    if (any) {
                        best_t = h.t; best_prim = prim;
                        bu = h.V / h.det; bv = h.W / h.det;
                        finished = true;
                        break;
                    }
                    if (CloserHit(h.t, 0u, prim, best_t, 0u, best_prim)) {
                        best_t = h.t; best_prim = prim;
                        bu = h.V / h.det; bv = h.W / h.det;
                    }
                }
                leaf = kInvalid;
                if (node != kInvalid && node >= n_int && node < total) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                }
            }
            if (finished || (node == kInvalid && leaf == kInvalid)) {
                const bool hit = best_prim != kInvalid;
                void* out = reinterpret_cast<void*>(table[lreq].hits);
                switch ((lflags >> 1) & 3u) {  // a miss leaves best_t at t_max
                    case kOutFullHit: WriteHit<kOutFullHit, false>(out, rid, hit, bu, bv, 0u, best_prim, best_t, best_t); break;
                    case kOutInstanceId: WriteHit<kOutInstanceId, false>(out, rid, hit, bu, bv, 0u, best_prim, best_t, best_t); break;
                    default: WriteHit<kOutFullHitT, false>(out, rid, hit, bu, bv, 0u, best_prim, best_t, best_t); break;
                }
                rid = kInvalid;
                node = kInvalid;
                leaf = kInvalid;
                break;
            }
            if (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy) break;  // refill (nothing left once drained)
        }
    }
}

// The two-level batch kernel: the scene requests of the table, with the
// persistent two-level kernel's loop (instances entered at once, the world
// ray in LDS, postponed leaves, the sentinel never popped while a leaf is
// postponed), so each request's hits equal a TracePersistent2LKernel launch
// of its own.  Per lane: the request's TLAS (its header is re-read on
// instance entry and exit, which are rare next to node visits), query and
// output.
__global__ __launch_bounds__(kBlock) void TraceBatch2LKernel(const BatchRequest* __restrict__ table, uint32_t requests,
                                                             uint32_t first_chunk, uint32_t chunk_count,
                                                             uint32_t* __restrict__ head, uint32_t* __restrict__ spill,
                                                             uint32_t min_busy, uint32_t chunk, uint32_t* err) {
    __shared__ uint32_t s_stack[kLdsStackEntries][kBlock];
    __shared__ float s_world[9][kBlock];  // world ray: origin, direction, reciprocal direction
    uint32_t* counter = head + kBatch2LCounterWord;
    const uint32_t lane = __lane_id();

    // Wave-uniform queue: rays [q_next, q_end) of request q_req.
    uint32_t q_req = 0, q_next = 0, q_end = 0;
    uint32_t part = blockIdx.x % kQueues, tries = 0;  // this wave's chunk part, parts found drained
    bool drained = false;

    // Lane state.
    uint32_t rid = kInvalid, lreq = 0, lflags = 0;
    const AccelHeader* top = nullptr;  // this lane's scene
    uint32_t top_int = 0, top_total = 0;
    const BvhNode* cur = nullptr;
    uint32_t cur_int = 0, cur_total = 0, inst = kInvalid;
    bool boxes = false;
    RayState r;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;  // best_t is t_max until a hit; culling limit WidenLimit(best_t)
    uint32_t best_prim = kInvalid, best_inst = kInvalid;
    uint32_t node = kInvalid, leaf = kInvalid;
    LaneStack<> st{s_stack, spill + (size_t)blockIdx.x * kBlock * kSpillEntries, err, 0, 0};

    for (;;) {
        const uint64_t idle = __ballot(rid == kInvalid);
        if (idle && !drained) {
            if (q_next == q_end)
                drained = !NextBatchChunk<true>(table, requests, chunk, first_chunk, chunk_count, counter, lane, part, tries,
                                                q_req, q_next, q_end);
            const uint32_t take = drained ? 0u : min((uint32_t)__popcll(idle), q_end - q_next);
            const uint32_t rank = RankIn(idle);
            if (rid == kInvalid && rank < take) {
                rid = q_next + rank;
                lreq = q_req;
                const BatchRequest& rq = table[q_req];
                lflags = rq.flags;
                top = reinterpret_cast<const AccelHeader*>(rq.accel);
                top_total = LdG(&top->node_count);
                if (LdG(&top->magic) != kAccelMagic || LdG(&top->kind) != kAccelTlas) {
                    // A geometry (or no valid header) at an address recorded
                    // as a scene: report it, answer misses.
                    if (err) StG(err + kErrBadAccel, 1u);
                    top_total = 0;
                }
                const uint32_t prims = LdG(&top->prim_count);
                top_int = (prims && top_total) ? prims - 1 : 0;
                const RayIn ray = LoadRay(reinterpret_cast<const float4*>(rq.rays), rid);
                s_world[0][threadIdx.x] = ray.a.x; s_world[1][threadIdx.x] = ray.a.y; s_world[2][threadIdx.x] = ray.a.z;
                s_world[3][threadIdx.x] = ray.b.x; s_world[4][threadIdx.x] = ray.b.y; s_world[5][threadIdx.x] = ray.b.z;
                SetupRay(r, ray.a.x, ray.a.y, ray.a.z, ray.a.w, ray.b.x, ray.b.y, ray.b.z);
                s_world[6][threadIdx.x] = r.ix; s_world[7][threadIdx.x] = r.iy; s_world[8][threadIdx.x] = r.iz;
                best_t = ray.b.w;
                best_prim = kInvalid;
                best_inst = kInvalid;
                bu = 0.0f; bv = 0.0f;
                st.clear();
                inst = kInvalid;
                cur = reinterpret_cast<const BvhNode*>(reinterpret_cast<const uint8_t*>(top) + LdG(&top->nodes_offset));
                cur_int = top_int;
                cur_total = top_total;
                boxes = false;
                node = top_total ? 0u : kInvalid;
                leaf = kInvalid;
            }
            q_next += take;
        }
        if (__ballot(rid != kInvalid) == 0) {
            if (drained) break;
            continue;
        }
        if (rid == kInvalid) continue;

        const bool any = (lflags & 1u) != 0;
        for (;;) {
            // Phase A: inner nodes, instance entries and exits.
            for (;;) {
                if (node < cur_int) {
                    float4 q0, q1, q2, q3;
                    LoadNode(cur, node, q0, q1, q2, q3);
                    const uint32_t c0 = __float_as_uint(q0.w);
                    const uint32_t c1 = __float_as_uint(q1.w);
                    float ta, tb;
                    const float limit = WidenLimit(best_t);
                    const bool h0 = SlabTest(r, limit, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, ta);
                    const bool h1 = SlabTest(r, limit, q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, tb);
                    if (h0 || h1) {
                        const bool c1first = h1 && (!h0 || ta > tb);
                        if (h0 && h1) st.push(c1first ? c0 : c1);
                        node = c1first ? c1 : c0;
                    } else {
                        node = st.empty() ? kInvalid : st.pop();
                    }
                } else if (inst == kInvalid && node < cur_total) {
                    // Instance leaf: enter it (isect_2l.comp:222-237).
                    const uint8_t* base = reinterpret_cast<const uint8_t*>(top);
                    const uint32_t id = __float_as_uint(LdG4(&cur[node].a0_max_or_v1).w);
                    const uint8_t* blas = nullptr;
                    float4 x0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), x1 = x0, x2 = x0;
                    uint32_t bn = 0, bflags = 0;
                    uint64_t boff = 0;
                    if (id < LdG(&top->prim_count)) {
                        const Transform* xforms = reinterpret_cast<const Transform*>(base + LdG(&top->transforms_offset));
                        const uint64_t* blas_table = reinterpret_cast<const uint64_t*>(base + LdG(&top->blas_table_offset));
                        const float4* xp = reinterpret_cast<const float4*>(xforms + 2 * (size_t)id);
                        x0 = LdG4(xp); x1 = LdG4(xp + 1); x2 = LdG4(xp + 2);
                        blas = reinterpret_cast<const uint8_t*>(LdG(blas_table + id));
                        if (blas) {
                            const AccelHeader* bh = reinterpret_cast<const AccelHeader*>(blas);
                            bn = LdG(&bh->node_count);
                            boff = LdG(&bh->nodes_offset);
                            bflags = LdG(&bh->flags);
                        }
                    }
                    if (bn != 0) {
                        const float wox = s_world[0][threadIdx.x], woy = s_world[1][threadIdx.x], woz = s_world[2][threadIdx.x];
                        const float wdx = s_world[3][threadIdx.x], wdy = s_world[4][threadIdx.x], wdz = s_world[5][threadIdx.x];
                        const float ox = ((x0.x * wox + x0.y * woy) + x0.z * woz) + x0.w;
                        const float oy = ((x1.x * wox + x1.y * woy) + x1.z * woz) + x1.w;
                        const float oz = ((x2.x * wox + x2.y * woy) + x2.z * woz) + x2.w;
                        const float dx = (x0.x * wdx + x0.y * wdy) + x0.z * wdz;
                        const float dy = (x1.x * wdx + x1.y * wdy) + x1.z * wdz;
                        const float dz = (x2.x * wdx + x2.y * wdy) + x2.z * wdz;
                        SetupRay(r, ox, oy, oz, r.tmin, dx, dy, dz);
                        inst = id;
                        cur = reinterpret_cast<const BvhNode*>(blas + boff);
                        cur_total = bn;
                        cur_int = (bn - 1) / 2;
                        boxes = (bflags & kAccelFlagAabbs) != 0;
                        st.push(kTopLevelSentinel);
                        node = 0;
                    } else {
                        node = st.empty() ? kInvalid : st.pop();
                    }
                } else if (node == kTopLevelSentinel && leaf == kInvalid) {
                    // Leave the instance, restore the world ray (isect_2l.comp:278-286).
                    inst = kInvalid;
                    cur = reinterpret_cast<const BvhNode*>(reinterpret_cast<const uint8_t*>(top) + LdG(&top->nodes_offset));
                    cur_int = top_int;
                    cur_total = top_total;
                    boxes = false;
                    r.ox = s_world[0][threadIdx.x]; r.oy = s_world[1][threadIdx.x]; r.oz = s_world[2][threadIdx.x];
                    r.ix = s_world[6][threadIdx.x]; r.iy = s_world[7][threadIdx.x]; r.iz = s_world[8][threadIdx.x];
                    node = st.empty() ? kInvalid : st.pop();
                } else if (inst != kInvalid && node < cur_total && leaf == kInvalid) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                } else {
                    break;  // a second leaf or the sentinel behind a postponed leaf, or done
                }
                // An index past the level's nodes (corrupt buffer) ends the ray.
                if (node != kInvalid && node != kTopLevelSentinel && node >= cur_total) node = kInvalid;
                if (!__any(leaf == kInvalid) || (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy)) break;
            }
            // Phase B: the postponed leaf, then the leaves that follow it in order.
            if (leaf == kInvalid && inst != kInvalid && node >= cur_int && node < cur_total) {
                leaf = node;
                node = st.empty() ? kInvalid : st.pop();
                if (node != kInvalid && node != kTopLevelSentinel && node >= cur_total) node = kInvalid;
            }
            bool finished = false;
            while (leaf != kInvalid) {
                float4 q0, q1, q2, q3;
                LoadNode(cur, leaf, q0, q1, q2, q3);
                const uint32_t prim = __float_as_uint(q1.w);
                TriHit h;
                if (LeafTest(boxes, r, best_t, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, h) &&
                    h.t >= r.tmin && h.t <= best_t) {
                    if (any) {
                        best_t = h.t; best_prim = prim; best_inst = inst;
                        bu = h.V / h.det; bv = h.W / h.det;
                        finished = true;
                        break;
                    }
                    if (CloserHit(h.t, inst, prim, best_t, best_inst, best_prim)) {
                        best_t = h.t; best_prim = prim; best_inst = inst;
                        bu = h.V / h.det; bv = h.W / h.det;
                    }
                }
                leaf = kInvalid;
                if (node >= cur_int && node < cur_total) {
                    leaf = node;
                    node = st.empty() ? kInvalid : st.pop();
                    if (node != kInvalid && node != kTopLevelSentinel && node >= cur_total) node = kInvalid;
                }
            }
            if (finished || (node == kInvalid && leaf == kInvalid)) {
                const bool hit = best_prim != kInvalid;
                void* out = reinterpret_cast<void*>(table[lreq].hits);
                switch ((lflags >> 1) & 3u) {  // a miss leaves best_t at t_max
                    case kOutFullHit: WriteHit<kOutFullHit, true>(out, rid, hit, bu, bv, best_inst, best_prim, best_t, best_t); break;
                    case kOutInstanceId: WriteHit<kOutInstanceId, true>(out, rid, hit, bu, bv, best_inst, best_prim, best_t, best_t); break;
                    default: WriteHit<kOutFullHitT, true>(out, rid, hit, bu, bv, best_inst, best_prim, best_t, best_t); break;
                }
                rid = kInvalid;
                node = kInvalid;
                leaf = kInvalid;
                break;
            }
            if (!drained && (uint32_t)__popcll(__ballot(1)) < min_busy) break;  // refill (nothing left once drained)
        }
    }
}


__global__ void BatchInitKernel(uint32_t* head) {
    if (threadIdx.x < kQueues) {
        head[kBatchCounterWord + threadIdx.x * kQueueStride] = 0;
        head[kBatch2LCounterWord + threadIdx.x * kQueueStride] = 0;
    }
}

}  // namespace

// Spill slots for every lane the launch can have: the persistent kernel runs
// ceil(ray_count / kBlock) blocks at most and hands rays to any of their
// lanes, so the slots are indexed by lane (up to ray_count rounded up to
// kBlock), not by ray.
size_t TraceScratchBytes(uint32_t ray_count) {
    return AlignUp((size_t)kQueueHeadWords * 4 + AlignUp((size_t)ray_count, kBlock) * kSpillEntries * 4, 256);
}

static std::atomic<uint32_t> g_min_busy{0};
uint32_t MinBusyLanes() {
    uint32_t m = g_min_busy.load(std::memory_order_relaxed);
    if (!m) {
        const char* e = std::getenv("RR_TRACE_MIN_BUSY");
        const int v = e ? std::atoi(e) : 0;
        uint32_t expected = 0;
        m = (v >= 1 && v <= 64) ? (uint32_t)v : kDefaultMinBusy;
        if (!g_min_busy.compare_exchange_strong(expected, m, std::memory_order_relaxed)) m = expected;
    }
    return m;
}
void SetMinBusyLanes(uint32_t v) { g_min_busy.store((v >= 1 && v <= 64) ? v : kDefaultMinBusy, std::memory_order_relaxed); }

static std::atomic<uint32_t> g_fetch{kDefaultFetch};
static std::atomic<uint32_t> g_depth{kDefaultLdsStack};
uint32_t RaysPerFetch() { return g_fetch.load(std::memory_order_relaxed); }
bool SetRaysPerFetch(uint32_t v) {
    if (v != 32 && v != 64 && v != 128) return false;
    g_fetch.store(v, std::memory_order_relaxed);
    return true;
}

static std::atomic<uint32_t> g_min_busy_2l{kDefaultMinBusy2L};
static std::atomic<uint32_t> g_fetch_2l{kDefaultFetch2L};
uint32_t MinBusyLanes2L() { return g_min_busy_2l.load(std::memory_order_relaxed); }
bool SetMinBusyLanes2L(uint32_t v) {
    if (v < 1 || v > 64) return false;
    g_min_busy_2l.store(v, std::memory_order_relaxed);
    return true;
}
static std::atomic<uint32_t> g_min_busy_batch{kDefaultMinBusyBatch};
uint32_t MinBusyLanesBatch() { return g_min_busy_batch.load(std::memory_order_relaxed); }
bool SetMinBusyLanesBatch(uint32_t v) {
    if (v < 1 || v > 64) return false;
    g_min_busy_batch.store(v, std::memory_order_relaxed);
    return true;
}
uint32_t RaysPerFetch2L() { return g_fetch_2l.load(std::memory_order_relaxed); }
bool SetRaysPerFetch2L(uint32_t v) {
    if (v != 32 && v != 64 && v != 128) return false;
    g_fetch_2l.store(v, std::memory_order_relaxed);
    return true;
}
uint32_t LdsStackDepth() { return g_depth.load(std::memory_order_relaxed); }
bool SetLdsStackDepth(uint32_t v) {
    if (v != 8 && v != 16 && v != 32) return false;
    g_depth.store(v, std::memory_order_relaxed);
    return true;
}

static std::atomic<uint32_t> g_rays_per_lane{kDefaultRaysPerLane};
uint32_t RaysPerLane() { return g_rays_per_lane.load(std::memory_order_relaxed); }
bool SetRaysPerLane(uint32_t v) {
    if (v > 64) return false;
    g_rays_per_lane.store(v, std::memory_order_relaxed);
    return true;
}

static std::atomic<int> g_trace_mode{-1};
int TraceMode() {
    int m = g_trace_mode.load(std::memory_order_relaxed);
    if (m < 0) {
        const char* e = std::getenv("RR_TRACE_MODE");
        m = (e && std::strcmp(e, "plain") == 0) ? kTraceModePlain : kTraceModePersistent;
        int expected = -1;
        if (!g_trace_mode.compare_exchange_strong(expected, m, std::memory_order_relaxed)) m = expected;
    }
    return m;
}
void SetTraceMode(int mode) { g_trace_mode.store(mode, std::memory_order_relaxed); }

namespace {

// Resident waves of the persistent kernel (queried once per device).
// Resident blocks of `kernel` on the current device (cached per device;
// which = 0 for the single-request kernel, 1 for the batch kernel).
uint32_t ResidentBlocks(const void* kernel, int which) {
    static std::atomic<uint32_t> blocks[8][64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    uint32_t b = blocks[which][dev].load(std::memory_order_relaxed);
    if (!b) {
        int per_cu = 0, cus = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        b = (uint32_t)std::max(1, per_cu * std::max(1, cus));
        blocks[which][dev].store(b, std::memory_order_relaxed);
    }
    return b;
}

// Resident blocks of the single- or two-level persistent kernel with the
// current LDS stack depth (the occupancy depends on its LDS).
uint32_t PersistentBlocks(bool two) {
    if (two) {
        switch (LdsStackDepth()) {
            case 8: return ResidentBlocks(reinterpret_cast<const void*>(&TracePersistent2LKernel<false, kOutFullHit, 8>), 4);
            case 32: return ResidentBlocks(reinterpret_cast<const void*>(&TracePersistent2LKernel<false, kOutFullHit, 32>), 5);
            default: return ResidentBlocks(reinterpret_cast<const void*>(&TracePersistent2LKernel<false, kOutFullHit, 16>), 6);
        }
    }
    switch (LdsStackDepth()) {
        case 8: return ResidentBlocks(reinterpret_cast<const void*>(&TracePersistentKernel<false, kOutFullHit, 8>), 2);
        case 32: return ResidentBlocks(reinterpret_cast<const void*>(&TracePersistentKernel<false, kOutFullHit, 32>), 3);
        default: return ResidentBlocks(reinterpret_cast<const void*>(&TracePersistentKernel<false, kOutFullHit, 16>), 0);
    }
}

// Workgroups a single query of n rays asks for (the caller takes the minimum
// with the resident grid).  Never more than one ray per lane and never more
// than four workgroups per CU (four waves per SIMD) unless 512 x k rays per
// workgroup need more, k the tuned rays per lane (default 2).  Below that
// the single-level kernel takes a workgroup per 256 x k rays, at least one
// per CU; the two-level kernel, whose rays are four times longer, keeps one
// ray per lane.  A query that fills fewer lanes than the resident grid ends
// when its longest rays do, and their dependent node fetches slow down as
// more waves share a SIMD and the memory system.  Measured against the full
// grid: heightfield (1M triangles, 10 visits per ray) 512K rays 0.24 against
// 0.31 ms, 1M rays 0.29 against 0.32 ms, 128K-4M never slower by more than
// noise; instanced scene (40 visits per ray) 512K secondary rays 1.13
// against 1.19 ms, while 256K of them take 0.79 ms on one ray per lane and
// 0.98 ms on two.  k = 0: the full grid whenever the rays fill it.
uint32_t QueryBlocks(uint64_t n, bool two) {
    const uint64_t one_per_lane = std::max<uint64_t>(1, (n + kBlock - 1) / kBlock);
    const uint32_t k = RaysPerLane();
    if (k == 0) return (uint32_t)std::min<uint64_t>(one_per_lane, UINT32_MAX);
    static std::atomic<uint32_t> cus_of[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    uint32_t cus = cus_of[dev].load(std::memory_order_relaxed);
    if (!cus) {
        int c = 0;
        (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
        cus = (uint32_t)std::max(1, c);
        cus_of[dev].store(cus, std::memory_order_relaxed);
    }
    const uint64_t lo = two ? one_per_lane : (n + (uint64_t)kBlock * k - 1) / ((uint64_t)kBlock * k);
    const uint64_t hi = (n + (uint64_t)kBlock * 2 * k - 1) / ((uint64_t)kBlock * 2 * k);
    const uint64_t b = std::max<uint64_t>({cus, std::min<uint64_t>(4ull * cus, lo), hi});
    return (uint32_t)std::min<uint64_t>({b, one_per_lane, UINT32_MAX});
}

uint32_t BatchBlocks() { return ResidentBlocks(reinterpret_cast<const void*>(&TraceBatchKernel), 1); }
uint32_t Batch2LBlocks() { return ResidentBlocks(reinterpret_cast<const void*>(&TraceBatch2LKernel), 7); }

// Rays per queue atomic for n rays over `blocks` blocks: the level's tuned
// value, cut (to a multiple of 32, at least 32) when fewer rays than waves x
// fetch would leave waves without any.
uint32_t FetchFor(uint32_t n, uint32_t blocks, bool two) {
    const uint64_t waves = (uint64_t)blocks * (kBlock / 64);
    const uint64_t per = (n + waves - 1) / waves;
    const uint32_t f = (uint32_t)std::max<uint64_t>(32, (per + 31) / 32 * 32);
    return std::min(two ? RaysPerFetch2L() : RaysPerFetch(), f);
}

template <bool ANY, bool TWO, int OUT, int DEPTH>
void LaunchPersistentKernel(uint32_t blocks, hipStream_t s, const uint8_t* a, const float4* r, uint32_t n,
                            const uint32_t* ind, void* hits, uint32_t* sc, uint32_t mb, uint32_t f, uint32_t init) {
    uint32_t* err = DeviceErrorWords();
    if (TWO) TracePersistent2LKernel<ANY, OUT, DEPTH><<<blocks, kBlock, 0, s>>>(a, r, n, ind, hits, sc, mb, f, init, err);
    else TracePersistentKernel<ANY, OUT, DEPTH><<<blocks, kBlock, 0, s>>>(a, r, n, ind, hits, sc, mb, f, init, err);
}

template <bool ANY, bool TWO, int DEPTH>
void LaunchPersistentDepth(OutputKind out, uint32_t blocks, hipStream_t s, const uint8_t* a, const float4* r,
                           uint32_t n, const uint32_t* ind, void* hits, uint32_t* sc, uint32_t init) {
    const uint32_t mb = TWO ? MinBusyLanes2L() : MinBusyLanes(), f = FetchFor(n, blocks, TWO);
    switch (out) {
        case kOutFullHit: LaunchPersistentKernel<ANY, TWO, kOutFullHit, DEPTH>(blocks, s, a, r, n, ind, hits, sc, mb, f, init); break;
        case kOutInstanceId: LaunchPersistentKernel<ANY, TWO, kOutInstanceId, DEPTH>(blocks, s, a, r, n, ind, hits, sc, mb, f, init); break;
        default: LaunchPersistentKernel<ANY, TWO, kOutFullHitT, DEPTH>(blocks, s, a, r, n, ind, hits, sc, mb, f, init); break;
    }
}

// init: this launch (one block) sets the queues up first (see InitQueues).
template <bool ANY, bool TWO>
void LaunchPersistent(OutputKind out, uint32_t blocks, hipStream_t s, const uint8_t* a, const float4* r, uint32_t n,
                      const uint32_t* ind, void* hits, uint32_t* sc, bool init) {
    const uint32_t i = init ? 1u : 0u;
    switch (LdsStackDepth()) {
        case 8: LaunchPersistentDepth<ANY, TWO, 8>(out, blocks, s, a, r, n, ind, hits, sc, i); break;
        case 32: LaunchPersistentDepth<ANY, TWO, 32>(out, blocks, s, a, r, n, ind, hits, sc, i); break;
        default: LaunchPersistentDepth<ANY, TWO, 16>(out, blocks, s, a, r, n, ind, hits, sc, i); break;
    }
}

}  // namespace

void Trace(const void* accel, AccelKind kind, QueryKind query, OutputKind output, const void* rays, uint32_t ray_count,
           const uint32_t* indirect, void* hits, void* scratch, hipStream_t s) {
    if (ray_count == 0) return;
    const uint8_t* a = static_cast<const uint8_t*>(accel);
    const float4* r = static_cast<const float4*>(rays);
    uint32_t* sc = static_cast<uint32_t*>(scratch);
    // Both levels are launched every time: each kernel reads the level from
    // the buffer's header and the one for the other level returns at once.
    // `kind` is the level the context recorded when it built this address;
    // it is only a hint, because the caller may have freed that buffer and
    // put another one (a copy, a build elsewhere) at the same address, so it
    // decides which persistent kernel gets the full grid and which gets one
    // block (all that an exiting kernel needs; a persistent kernel with one
    // block still traces every ray, slowly, if the hint was stale).
    if (TraceMode() != kTraceModePlain) {
        // At most ceil(ray_count / kBlock) blocks: TraceScratchBytes holds
        // spill slots for that many lanes.  The queues are set up once: a
        // kernel of the wrong level returns before touching them.
        // With a recorded level the one-block launch (first) sets the queues
        // up; with none, both grids are full and a QueueInitKernel does.
        const bool recorded = kind != kAccelUnknown;
        if (!recorded) QueueInitKernel<<<1, 64, 0, s>>>(sc, ray_count, indirect);
        const uint32_t b1 = kind == kAccelTlas ? 1u : std::min(PersistentBlocks(false), QueryBlocks(ray_count, false));
        const uint32_t b2 = kind == kAccelBlas ? 1u : std::min(PersistentBlocks(true), QueryBlocks(ray_count, true));
        // The one-block launch goes first: with other streams' persistent
        // launches holding the CUs it waits as long as the full launch
        // behind it would, instead of delaying the query's end.
        const bool single_first = kind == kAccelTlas;
        for (int k = 0; k < 2; ++k) {
            const bool two = (k == 0) != single_first;
            const uint32_t b = two ? b2 : b1;
            const bool init = recorded && k == 0;
            if (query == kQueryAny) {
                if (two) LaunchPersistent<true, true>(output, b, s, a, r, ray_count, indirect, hits, sc, init);
                else LaunchPersistent<true, false>(output, b, s, a, r, ray_count, indirect, hits, sc, init);
            } else {
                if (two) LaunchPersistent<false, true>(output, b, s, a, r, ray_count, indirect, hits, sc, init);
                else LaunchPersistent<false, false>(output, b, s, a, r, ray_count, indirect, hits, sc, init);
            }
        }
        return;
    }
    // One lane per ray: both grids cover every ray.
    const dim3 g((ray_count + kBlock - 1) / kBlock);
    if (query == kQueryAny) {
        LaunchOut<true, false, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
        LaunchOut<true, true, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
    } else {
        LaunchOut<false, false, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
        LaunchOut<false, true, false>(output, g, s, a, r, ray_count, indirect, hits, sc, nullptr);
    }
}

void TraceWithStats(const void* accel, AccelKind kind, QueryKind query, const void* rays, uint32_t ray_count, void* hits,
                    void* scratch, uint32_t* stats, hipStream_t s) {
    if (ray_count == 0) return;
    const dim3 g((ray_count + kBlock - 1) / kBlock);
    const uint8_t* a = static_cast<const uint8_t*>(accel);
    const float4* r = static_cast<const float4*>(rays);
    uint32_t* sc = static_cast<uint32_t*>(scratch);
    (void)kind;  // a hint only (see Trace): both levels run, the other one returns at once
    uint32_t* err = DeviceErrorWords();
    if (query == kQueryAny) {
        TraceKernel<true, kOutFullHitT, false, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats, err);
        TraceKernel<true, kOutFullHitT, true, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats, err);
    } else {
        TraceKernel<false, kOutFullHitT, false, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats, err);
        TraceKernel<false, kOutFullHitT, true, true><<<g, kBlock, 0, s>>>(a, r, ray_count, nullptr, hits, sc, stats, err);
    }
}

size_t TraceBatchScratchBytes(uint32_t requests, uint64_t total_rays) {
    // The launch has at most one lane per chunked ray, and the chunks of at
    // most kBatchChunkLarge rays number fewer than total_rays / chunk + requests.
    const uint64_t lanes = std::min<uint64_t>((uint64_t)std::max(BatchBlocks(), Batch2LBlocks()) * kBlock,
                                              AlignUp(total_rays + (uint64_t)requests * kBatchChunkLarge, kBlock));
    return kBatchTableOffset + BatchTableBytes(requests) + AlignUp((size_t)lanes * kSpillEntries * 4, 256);
}

void TraceBatch(const BatchRequest* table, uint32_t requests, uint32_t chunk, uint32_t geo_chunks, uint32_t unrecorded_chunks,
                uint32_t scene_chunks, void* scratch, hipStream_t s) {
    if (requests == 0 || geo_chunks + unrecorded_chunks + scene_chunks == 0) return;
    uint8_t* base = static_cast<uint8_t*>(scratch);
    uint32_t* head = reinterpret_cast<uint32_t*>(base);
    auto* spill = reinterpret_cast<uint32_t*>(base + kBatchTableOffset + BatchTableBytes(requests));
    const uint32_t mb = MinBusyLanesBatch();
    BatchInitKernel<<<1, 64, 0, s>>>(head);
    // The geometry kernel takes chunks [0, geo + unrecorded), the two-level
    // kernel [geo, geo + unrecorded + scene); a kernel with no chunk is not
    // launched.  Grids as a single query's of the chunks' lanes.
    const uint32_t n1 = geo_chunks + unrecorded_chunks, n2 = unrecorded_chunks + scene_chunks;
    if (n1) {
        const uint32_t b = std::min(BatchBlocks(), QueryBlocks((uint64_t)n1 * chunk, false));
        TraceBatchKernel<<<b, kBlock, 0, s>>>(table, requests, 0u, n1, head, spill, mb, chunk, DeviceErrorWords());
    }
    if (n2) {
        const uint32_t b = std::min(Batch2LBlocks(), QueryBlocks((uint64_t)n2 * chunk, true));
        TraceBatch2LKernel<<<b, kBlock, 0, s>>>(table, requests, geo_chunks, n2, head, spill, MinBusyLanes2L(), chunk,
                                                DeviceErrorWords());
    }
}

}  // namespace rr
