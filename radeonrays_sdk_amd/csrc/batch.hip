// batch.hip — batched BLAS build and refit for gfx950.
//
// Many small meshes (up to kBatchMaxTris triangles each) are built by three
// launches.  A persistent one in which each workgroup takes the next mesh
// from a queue (the host orders the table largest first) and runs the
// mesh-wide steps of the single-mesh builder with block-level
// synchronisation only — bounds (wave64 reductions), Morton codes and a
// stable LSD radix sort of the (code, index) pairs in four 8-bit passes;
// then Karras emission and the bottom-up refit, one thread per node or leaf
// of any mesh.  Every step computes exactly what the single-mesh kernels
// compute (min/max are exact, the sort is stable, emission and refit are the
// shared device functions), so each result equals BuildBlas of that mesh.
// Behavioural spec of each step: as build.hip (vlk/hlbvh_builder.cpp:159-351).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "batch.h"
#include "build_device.h"
#include "rr_device.h"
#include "sort.h"

namespace rr {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kChunkItems = 4;  // keys per thread per sort chunk
constexpr int kChunk = kBlock * kChunkItems;

// Exclusive scan over a workgroup of WAVES waves; contains two barriers.
template <int WAVES = kWaves>
__device__ __forceinline__ uint32_t BlockScan(uint32_t v, uint32_t* s_wave_tot) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_wave_tot[wave] = x;
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) add += (w < (int)wave) ? s_wave_tot[w] : 0u;
    __syncthreads();
    return x - v + add;
}

struct BatchShared {
    uint32_t mesh;
    uint32_t bounds[kWaves][6];
    uint32_t wave_tot[kWaves];
    uint32_t offs[kSortRadix];                 // running output base per digit
    uint32_t wave_hist[kWaves][kSortRadix];
};

// The mesh's buffers, addressed from `arg` (a kernel argument) so that
// accesses compile to global-memory instructions.
__device__ __forceinline__ MeshView ViewOfBatch(const BatchMesh& b, const void* arg) {
    return MeshView{RebaseOn<const uint8_t>(arg, b.vertices), RebaseOn<const void>(arg, b.indices),
                    b.vertex_stride, b.index16, b.vertex_count, b.tri_count};
}

// Stable LSD sort of n (key, value) pairs by the 8-bit digit at `shift`,
// from (ki, vi) to (ko, vo), by one workgroup.
__device__ void BlockSortPass(const uint32_t* __restrict__ ki, const uint32_t* __restrict__ vi, uint32_t* __restrict__ ko,
                              uint32_t* __restrict__ vo, uint32_t n, uint32_t shift, BatchShared& sh) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // Digit counts, then their exclusive scan: each digit's output base.
    for (uint32_t i = tid; i < kWaves * kSortRadix; i += kBlock) (&sh.wave_hist[0][0])[i] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < n; t += kBlock) atomicAdd(&sh.wave_hist[0][(ki[t] >> shift) & (kSortRadix - 1)], 1u);
    __syncthreads();
    const uint32_t total = sh.wave_hist[0][tid];
    const uint32_t base = BlockScan(total, sh.wave_tot);
    sh.offs[tid] = base;
    sh.wave_hist[0][tid] = 0;
    __syncthreads();
    // Chunks in order; waves in order within a chunk; lanes in order within
    // an item: keys keep their relative order within each digit.
    for (uint32_t c0 = 0; c0 < n; c0 += kChunk) {
        uint32_t k[kChunkItems], v[kChunkItems], rank[kChunkItems];
        bool ok[kChunkItems];
#pragma unroll
        for (int j = 0; j < kChunkItems; ++j) {
            const uint32_t idx = c0 + wave * (64 * kChunkItems) + j * 64 + lane;
            ok[j] = idx < n;
            k[j] = ok[j] ? ki[idx] : 0u;
            v[j] = ok[j] ? vi[idx] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kChunkItems; ++j) {
            const uint32_t d = (k[j] >> shift) & (kSortRadix - 1);
            const uint64_t valid = __ballot(ok[j]);
            uint32_t m_lo = (uint32_t)valid, m_hi = (uint32_t)(valid >> 32);
#pragma unroll
            for (int b = 0; b < kSortRadixBits; ++b) {
                const uint32_t neg = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);  // bit b of the digit
                const uint64_t bal = __ballot(neg != 0u);
                m_lo &= ~((uint32_t)bal ^ neg);
                m_hi &= ~((uint32_t)(bal >> 32) ^ neg);
            }
            const uint32_t before = __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
            const uint32_t pre = sh.wave_hist[wave][d];
            rank[j] = pre + before;
            __builtin_amdgcn_wave_barrier();
            if (ok[j] && before == 0) sh.wave_hist[wave][d] = pre + (uint32_t)(__popc(m_lo) + __popc(m_hi));
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // Per digit: exclusive offsets of the waves, and the chunk's count.
        uint32_t count = 0;
        {
            const uint32_t d = tid;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = sh.wave_hist[w][d];
                sh.wave_hist[w][d] = count;
                count += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kChunkItems; ++j) {
            if (!ok[j]) continue;
            const uint32_t d = (k[j] >> shift) & (kSortRadix - 1);
            const uint32_t pos = sh.offs[d] + sh.wave_hist[wave][d] + rank[j];
            ko[pos] = k[j];
            vo[pos] = v[j];
        }
        __syncthreads();
        sh.offs[tid] += count;
        for (int w = 0; w < kWaves; ++w) sh.wave_hist[w][tid] = 0;
        __syncthreads();
    }
}

__global__ void BatchQueueInitKernel(uint32_t* queue) {
    if (threadIdx.x == 0) queue[0] = 0;
    if (threadIdx.x == kBatchLdsQueueWord) queue[kBatchLdsQueueWord] = 0;
    if (threadIdx.x == kBatchLdsSmallQueueWord) queue[kBatchLdsSmallQueueWord] = 0;
}

__global__ __launch_bounds__(kBlock) void BatchBuildKernel(const BatchMesh* __restrict__ table, uint32_t meshes,
                                                           uint32_t* queue) {
    __shared__ BatchShared sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (;;) {
        if (tid == 0) sh.mesh = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t mi = sh.mesh;
        __syncthreads();
        if (mi >= meshes) break;
        const BatchMesh bm = table[mi];
        const MeshView m = ViewOfBatch(bm, queue);
        const uint32_t n = bm.tri_count;
        uint8_t* result = RebaseOn<uint8_t>(queue, bm.result);
        AccelHeader* hdr = reinterpret_cast<AccelHeader*>(result);
        if (tid == 0) {
            // As InitBuildKernel.
            const uint32_t serial = hdr->magic == kAccelMagic ? hdr->build_serial + 1 : 1;
            AccelHeader h = {};
            h.magic = kAccelMagic;
            h.kind = kAccelBlas;
            h.prim_count = n;
            h.node_count = NodeCount(n);
            h.nodes_offset = sizeof(AccelHeader);
            h.flags = bm.flags;
            h.build_serial = serial;
            *hdr = h;
        }
        if (n == 0) continue;

        // Bounds of all vertices (ordered-uint min/max: exact).
        uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
        for (uint32_t t = tid; t < n; t += kBlock) {
            float v[9];
            LoadTriangle(m, t, v);
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) {
                    const uint32_t o = FloatToOrdered(v[3 * c + a]);
                    mn[a] = min(mn[a], o);
                    mx[a] = max(mx[a], o);
                }
        }
        for (int a = 0; a < 3; ++a) {
            mn[a] = WaveMin(mn[a]);
            mx[a] = WaveMax(mx[a]);
        }
        if (lane == 0)
            for (int a = 0; a < 3; ++a) {
                sh.bounds[wave][a] = mn[a];
                sh.bounds[wave][3 + a] = mx[a];
            }
        __syncthreads();
        float lo[3], ext[3];
        for (int a = 0; a < 3; ++a) {
            uint32_t l = sh.bounds[0][a], h = sh.bounds[0][3 + a];
            for (int w = 1; w < kWaves; ++w) {
                l = min(l, sh.bounds[w][a]);
                h = max(h, sh.bounds[w][3 + a]);
            }
            lo[a] = OrderedToFloat(l);
            ext[a] = OrderedToFloat(h) - lo[a];
        }

        // Morton codes of the triangle-box centroids, values = triangle index.
        uint32_t* keys = RebaseOn<uint32_t>(queue, bm.work);
        uint32_t* vals = keys + AlignUp(n, 64);
        uint32_t* ktmp = vals + AlignUp(n, 64);
        uint32_t* vtmp = ktmp + AlignUp(n, 64);
        for (uint32_t t = tid; t < n; t += kBlock) {
            float c[3];
            TriangleCentroid(m, t, c);
            keys[t] = Morton30(UnitCoord(c[0], lo[0], ext[0]), UnitCoord(c[1], lo[1], ext[1]),
                               UnitCoord(c[2], lo[2], ext[2]));
            vals[t] = t;
        }
        __syncthreads();
        // Four passes: keys -> tmp -> keys -> tmp -> keys.
        for (int p = 0; p < SortPasses(30); ++p) {
            const bool even = (p & 1) == 0;
            BlockSortPass(even ? keys : ktmp, even ? vals : vtmp, even ? ktmp : keys, even ? vtmp : vals, n,
                          (uint32_t)(p * kSortRadixBits), sh);
        }
        __syncthreads();  // the hierarchy is emitted by BatchEmitKernel, one thread per node of all meshes
    }
}

// Meshes of at most kBatchLdsMaxTris triangles: one workgroup per mesh
// computes bounds and Morton codes, runs the four 8-bit passes of the
// stable sort with keys and values resident in LDS (128 KB for 16,384
// pairs), and emits the hierarchy from the LDS-resident codes.  Item j of lane l in wave w is position
// w * (64 * kLdsItems) + 64 j + l, and ranks follow that order (waves, then
// items, then lanes), so each pass is stable and the result equals the
// single-mesh sort of the same (code, index) pairs.  kBatchLdsMaxTris is a
// multiple of 64, so keys padded to 64 stay within LDS.
// BLOCK threads x ITEMS keys: 1024 x 16 for meshes of up to 16,384
// triangles (one workgroup per CU), 256 x 16 for meshes of up to 4,096
// (four per CU, so that one mesh's memory round trips overlap another's
// ranking).
template <int BLOCK, int ITEMS>
struct BatchLdsShared {
    static constexpr int kWavesPer = BLOCK / 64;
    uint32_t keys[BLOCK * ITEMS];
    uint32_t vals[BLOCK * ITEMS];
    uint32_t wave_hist[kWavesPer][kSortRadix];
    uint32_t digit_start[kSortRadix];
    uint32_t bounds[kWavesPer][6];
    uint32_t wave_tot[kWavesPer];
    uint32_t mesh;
};

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void BatchBuildLdsKernel(const BatchMesh* __restrict__ table, uint32_t first,
                                                             uint32_t meshes, uint32_t* queue, uint32_t queue_word) {
    constexpr int kLdsBlock = BLOCK;
    constexpr int kLdsWaves = BLOCK / 64;
    constexpr int kLdsItems = ITEMS;
    constexpr uint32_t kMaxTris = (uint32_t)(BLOCK * ITEMS);
    static_assert(BLOCK >= kSortRadix, "one thread per digit");
    __shared__ BatchLdsShared<BLOCK, ITEMS> sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t wbase = wave * (64 * kLdsItems);  // position of this wave's item 0, lane 0
    for (;;) {
        if (tid == 0) sh.mesh = first + atomicAdd(queue + queue_word, 1u);
        __syncthreads();
        const uint32_t mi = sh.mesh;
        __syncthreads();
        if (mi >= meshes) break;
        const BatchMesh bm = table[mi];
        const MeshView m = ViewOfBatch(bm, queue);
        uint8_t* result = RebaseOn<uint8_t>(queue, bm.result);
        AccelHeader* hdr = reinterpret_cast<AccelHeader*>(result);
        if (tid == 0) {
            // As InitBuildKernel.
            const uint32_t serial = hdr->magic == kAccelMagic ? hdr->build_serial + 1 : 1;
            AccelHeader h = {};
            h.magic = kAccelMagic;
            h.kind = kAccelBlas;
            h.prim_count = bm.tri_count;
            h.node_count = NodeCount(bm.tri_count);
            h.nodes_offset = sizeof(AccelHeader);
            h.flags = bm.flags;
            h.build_serial = serial;
            *hdr = h;
        }
        // The host routes larger meshes elsewhere; the clamp keeps a bad
        // table within LDS.
        const uint32_t n = min(bm.tri_count, kMaxTris);
        if (n == 0) continue;

        // Bounds of all vertices (ordered-uint min/max: exact).
        uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
        for (uint32_t t = tid; t < n; t += kLdsBlock) {
            float v[9];
            LoadTriangle(m, t, v);
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) {
                    const uint32_t o = FloatToOrdered(v[3 * c + a]);
                    mn[a] = min(mn[a], o);
                    mx[a] = max(mx[a], o);
                }
        }
        for (int a = 0; a < 3; ++a) {
            mn[a] = WaveMin(mn[a]);
            mx[a] = WaveMax(mx[a]);
        }
        if (lane == 0)
            for (int a = 0; a < 3; ++a) {
                sh.bounds[wave][a] = mn[a];
                sh.bounds[wave][3 + a] = mx[a];
            }
        __syncthreads();
        float lo[3], ext[3];
        for (int a = 0; a < 3; ++a) {
            uint32_t l = sh.bounds[0][a], h = sh.bounds[0][3 + a];
            for (int w = 1; w < kLdsWaves; ++w) {
                l = min(l, sh.bounds[w][a]);
                h = max(h, sh.bounds[w][3 + a]);
            }
            lo[a] = OrderedToFloat(l);
            ext[a] = OrderedToFloat(h) - lo[a];
        }
        // Morton codes of the triangle-box centroids into LDS, values =
        // triangle index (the mesh's second read of its triangles).  Keys are
        // padded to a multiple of 64 with all-ones keys: above every 30-bit
        // code, in the last digit of every pass, so a stable pass keeps them
        // after every real key, at positions [n, n64).
        const uint32_t n64 = (n + 63u) & ~63u;
        for (uint32_t t = tid; t < n64; t += kLdsBlock) {
            uint32_t key = 0xFFFFFFFFu;
            if (t < n) {
                float c[3];
                TriangleCentroid(m, t, c);
                key = Morton30(UnitCoord(c[0], lo[0], ext[0]), UnitCoord(c[1], lo[1], ext[1]),
                               UnitCoord(c[2], lo[2], ext[2]));
            }
            sh.keys[t] = key;
            sh.vals[t] = t;
        }

        for (int p = 0; p < SortPasses(30); ++p) {
            const uint32_t shift = (uint32_t)(p * kSortRadixBits);
            __syncthreads();  // the previous step's LDS writes are done
            // This wave's items are positions wbase + 64 j + lane for
            // j < items (whole 64-key chunks below n64).
            const uint32_t items = wbase < n64 ? min((n64 - wbase) / 64u, (uint32_t)kLdsItems) : 0u;
            uint32_t k[kLdsItems], v[kLdsItems], rank[kLdsItems];
#pragma unroll
            for (int j = 0; j < kLdsItems; ++j) {
                k[j] = (uint32_t)j < items ? sh.keys[wbase + 64u * j + lane] : 0xFFFFFFFFu;
                rank[j] = 0;
            }
            for (uint32_t i = tid; i < kLdsWaves * kSortRadix; i += kLdsBlock) (&sh.wave_hist[0][0])[i] = 0;
            __syncthreads();
            // Rank within the wave: lanes with the same digit are the AND
            // over the digit's bits of the bit's ballot or its complement.
#pragma unroll
            for (int j = 0; j < kLdsItems; ++j) {
                if ((uint32_t)j >= items) continue;  // uniform in the wave
                const uint32_t d = (k[j] >> shift) & (kSortRadix - 1);
                uint32_t m_lo = ~0u, m_hi = ~0u;
#pragma unroll
                for (int b = 0; b < kSortRadixBits; ++b) {
                    const uint32_t neg = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);  // bit b of the digit
                    const uint64_t bal = __ballot(neg != 0u);
                    m_lo &= ~((uint32_t)bal ^ neg);
                    m_hi &= ~((uint32_t)(bal >> 32) ^ neg);
                }
                const uint32_t before = __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
                const uint32_t pre = sh.wave_hist[wave][d];
                rank[j] = pre + before;
                __builtin_amdgcn_wave_barrier();
                if (before == 0) sh.wave_hist[wave][d] = pre + (uint32_t)(__popc(m_lo) + __popc(m_hi));
                __builtin_amdgcn_wave_barrier();
            }
            // Values only now, so they hold no registers through the ranking
            // (LDS is rewritten only after the barriers below).
#pragma unroll
            for (int j = 0; j < kLdsItems; ++j) v[j] = (uint32_t)j < items ? sh.vals[wbase + 64u * j + lane] : 0u;
            __syncthreads();
            // Per digit: exclusive offsets of the waves, then the digits' starts.
            uint32_t count = 0;
            if (tid < (uint32_t)kSortRadix) {
#pragma unroll
                for (int w = 0; w < kLdsWaves; ++w) {
                    const uint32_t c = sh.wave_hist[w][tid];
                    sh.wave_hist[w][tid] = count;
                    count += c;
                }
            }
            const uint32_t dstart = BlockScan<kLdsWaves>(count, sh.wave_tot);  // contains barriers
            if (tid < (uint32_t)kSortRadix) sh.digit_start[tid] = dstart;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kLdsItems; ++j) {
                if ((uint32_t)j >= items) continue;
                const uint32_t d = (k[j] >> shift) & (kSortRadix - 1);
                const uint32_t pos = sh.digit_start[d] + sh.wave_hist[wave][d] + rank[j];
                sh.keys[pos] = k[j];
                sh.vals[pos] = v[j];
            }
            __syncthreads();
        }
        // The hierarchy from the LDS-resident sorted codes (EmitNode, as
        // EmitHierarchyKernel); BatchRefitKernel refits it.
        BvhNode* nodes = reinterpret_cast<BvhNode*>(result + sizeof(AccelHeader));
        for (uint32_t i = tid; i < n; i += kLdsBlock) EmitNode(sh.keys, sh.vals, n, i, nodes);
        // The next mesh's first barrier orders these LDS reads before reuse.
    }
}

// Karras emission of every mesh, one thread per leaf index (node i and
// leaf i of its mesh), over the 64-leaf chunks of all meshes: each node's
// binary searches are a chain of dependent loads, so the whole GPU's
// threads run them rather than one workgroup per mesh.
__global__ __launch_bounds__(kBlock) void BatchEmitKernel(const BatchMesh* __restrict__ table, uint32_t meshes,
                                                          uint32_t total_chunks, uint32_t* scratch) {
    const uint32_t c = XcdContiguousBlock(blockIdx.x, gridDim.x) * (kBlock / kRefitChunk) + threadIdx.x / kRefitChunk;
    if (c >= total_chunks) return;
    uint32_t lo = 0, hi = meshes;  // last mesh with first_chunk <= c (empty meshes own no chunk)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (table[mid].first_chunk <= c) lo = mid;
        else hi = mid;
    }
    const BatchMesh bm = table[lo];
    const uint32_t i = (c - bm.first_chunk) * kRefitChunk + threadIdx.x % kRefitChunk;
    if (c < bm.first_chunk || i >= bm.tri_count) return;
    const uint32_t n = bm.tri_count;
    const uint32_t* keys = RebaseOn<const uint32_t>(scratch, bm.work);  // sorted codes after the four passes
    const uint32_t* vals = keys + AlignUp(n, 64);
    BvhNode* nodes = reinterpret_cast<BvhNode*>(RebaseOn<uint8_t>(scratch, bm.result) + sizeof(AccelHeader));
    EmitNode(keys, vals, n, i, nodes);
}

// Refit of every mesh in the table, one thread per leaf: thread t of block b
// takes leaf t % kRefitChunk of refit chunk b * (kBlock / kRefitChunk) +
// t / kRefitChunk, in the mesh that owns the chunk (binary search over
// first_chunk).  The chunks of one mesh within a block form one refit window
// (consecutive leaves); chunks of one mesh in different blocks meet in shared
// ancestors through ClimbRefit's agent-scope hand-off.
static_assert(kBlock % kRefitChunk == 0, "whole chunks per block");
__global__ __launch_bounds__(kBlock) void BatchRefitKernel(const BatchMesh* __restrict__ table, uint32_t meshes,
                                                           uint32_t total_chunks, uint32_t* scratch) {
    __shared__ RefitLds<kBlock> lds;
    constexpr uint32_t kChunksPerBlock = kBlock / kRefitChunk;
    // Consecutive chunks (one mesh's leaves) on one XCD, sharing its L2.
    const uint32_t cb = XcdContiguousBlock(blockIdx.x, gridDim.x) * kChunksPerBlock;
    const uint32_t c = cb + threadIdx.x / kRefitChunk;
    bool active = c < total_chunks;
    uint32_t lo_m = 0, hi_m = meshes;  // last mesh with first_chunk <= c (empty meshes own no chunk)
    while (active && hi_m - lo_m > 1) {
        const uint32_t mid = (lo_m + hi_m) / 2;
        if (table[mid].first_chunk <= c) lo_m = mid;
        else hi_m = mid;
    }
    BatchMesh bm{};
    if (active) bm = table[lo_m];
    active = active && c >= bm.first_chunk;
    const uint32_t i = (c - bm.first_chunk) * kRefitChunk + threadIdx.x % kRefitChunk;
    active = active && i < bm.tri_count;
    // This mesh's chunks within the block: [first, last).
    const uint32_t first = std::max(cb, bm.first_chunk);
    const uint32_t last = std::min(cb + kChunksPerBlock, bm.first_chunk + (bm.tri_count + kRefitChunk - 1) / kRefitChunk);
    BvhNode* nodes = reinterpret_cast<BvhNode*>(RebaseOn<uint8_t>(scratch, bm.result) + sizeof(AccelHeader));
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    uint32_t parent = kInvalid;
    if (active) parent = WriteLeaf(ViewOfBatch(bm, scratch), nodes, i, lo, hi);
    const uint32_t n = bm.tri_count;
    WindowRefit<kBlock>(nodes, 2 * n - 1, (first - bm.first_chunk) * kRefitChunk,
                        active ? (last - first) * kRefitChunk : 0, (first - cb) * kRefitChunk, active,
                        active && i + 1 < n ? i : kInvalid, n - 1 + i, parent, lo, hi, lds);
}

// Resident blocks of `kernel` (workgroups of `block` threads) on the current
// device, cached per device; `which` names the kernel's cache row.
uint32_t ResidentBlocks(const void* kernel, int block, int which) {
    static std::atomic<uint32_t> cache[3][64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    uint32_t b = cache[which][dev].load(std::memory_order_relaxed);
    if (!b) {
        int per_cu = 0, cus = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        b = (uint32_t)std::max(1, per_cu * std::max(1, cus));
        cache[which][dev].store(b, std::memory_order_relaxed);
    }
    return b;
}

}  // namespace

size_t BatchWorkBytes(uint32_t tris) {
    return std::max(AlignUp(size_t(AlignUp(tris, 64)) * 16, 256), RestructureScratchBytes(tris));
}

void BatchBuild(const BatchMesh* table, uint32_t meshes, uint32_t big, uint32_t small, uint32_t big_chunks,
                uint32_t total_chunks, void* scratch, hipStream_t s) {
    if (meshes == 0) return;
    big = std::min(big, meshes);
    uint32_t* queue = static_cast<uint32_t*>(scratch);
    BatchQueueInitKernel<<<1, 64, 0, s>>>(queue);
    if (big) {
        const uint32_t blocks =
            std::min(big, ResidentBlocks(reinterpret_cast<const void*>(&BatchBuildKernel), kBlock, 0));
        BatchBuildKernel<<<blocks, kBlock, 0, s>>>(table, big, queue);
        if (big_chunks) {
            const uint32_t eblocks = (uint32_t)(((uint64_t)big_chunks * kRefitChunk + kBlock - 1) / kBlock);
            BatchEmitKernel<<<eblocks, kBlock, 0, s>>>(table, big, big_chunks, queue);
        }
    }
    small = std::min(std::max(small, big), meshes);
    if (small > big) {
        const auto kernel = &BatchBuildLdsKernel<1024, 16>;
        const uint32_t blocks =
            std::min(small - big, ResidentBlocks(reinterpret_cast<const void*>(kernel), 1024, 1));
        kernel<<<blocks, 1024, 0, s>>>(table, big, small, queue, kBatchLdsQueueWord);
    }
    if (meshes > small) {
        const auto kernel = &BatchBuildLdsKernel<256, 16>;
        const uint32_t blocks =
            std::min(meshes - small, ResidentBlocks(reinterpret_cast<const void*>(kernel), 256, 2));
        kernel<<<blocks, 256, 0, s>>>(table, small, meshes, queue, kBatchLdsSmallQueueWord);
    }
    BatchRefit(table, meshes, total_chunks, scratch, s);
}

void BatchRefit(const BatchMesh* table, uint32_t meshes, uint32_t total_chunks, void* scratch, hipStream_t s) {
    if (meshes == 0 || total_chunks == 0) return;
    const uint32_t blocks = (uint32_t)(((uint64_t)total_chunks * kRefitChunk + kBlock - 1) / kBlock);
    BatchRefitKernel<<<blocks, kBlock, 0, s>>>(table, meshes, total_chunks, static_cast<uint32_t*>(scratch));
}

}  // namespace rr
