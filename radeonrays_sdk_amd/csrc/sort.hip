// sort.hip — onesweep LSD radix sort of (key, value) pairs for gfx950.
//
// Behavioural contract: a stable sort of uint32 keys carrying uint32
// values, as the reference's RadixSortKeyValue (vlk/radix_sort.cpp:157-338)
// provides to the LBVH builder.  The implementation is the onesweep scheme
// (Adinets & Merrill, "Onesweep: A Faster Least Significant Digit Radix Sort
// for GPUs", 2022) written for wave64:
//   * one histogram pass counts all digits of all passes at once;
//   * each pass kernel takes tiles in ticket order, ranks its keys with
//     64-bit ballots (stable within a tile), publishes per-digit tile counts
//     and resolves its global offsets by decoupled look-back, then scatters
//     through LDS so that each digit's run is written contiguously.
// Tickets make every tile's predecessors already running, so look-back only
// waits on work that is guaranteed to publish; the spin is bounded anyway.
#include <hip/hip_runtime.h>

#include "rr_device.h"
#include "sort.h"

namespace rr {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagInclusive = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t AtomicLoadAgent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void AtomicStoreAgent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t BlockExclusiveScan256(uint32_t v, uint32_t* s_wave_tot) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_wave_tot[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += s_wave_tot[w];
    __syncthreads();
    return x - v + add;
}

}  // namespace

__global__ __launch_bounds__(kBlock) void SortClearKernel(uint32_t* hist, uint32_t* tickets) {
    for (uint32_t i = threadIdx.x; i < kSortMaxPasses * kSortRadix; i += kBlock) hist[i] = 0;
    if (threadIdx.x < kSortMaxPasses) tickets[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void SortHistogramKernel(const uint32_t* __restrict__ keys, uint32_t n, int passes,
                                                              uint32_t* __restrict__ hist, uint32_t* __restrict__ status,
                                                              size_t status_words) {
    __shared__ uint32_t s_h[kSortMaxPasses][kSortRadix];
    for (uint32_t i = threadIdx.x; i < kSortMaxPasses * kSortRadix; i += kBlock) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * kBlock;
    // Every lane of a wave runs the same number of iterations so that the
    // ballots in AccumulateDigits see the whole wave.
    const uint32_t wave_base = blockIdx.x * kBlock + (threadIdx.x & ~63u);
    for (uint32_t b = wave_base; b < n; b += stride) {
        const uint32_t i = b + __lane_id();
        const bool valid = i < n;
        const uint32_t k = valid ? keys[i] : 0u;
        AccumulateDigits(k, valid, s_h, passes);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < (uint32_t)passes * kSortRadix; i += kBlock) {
        const uint32_t c = (&s_h[0][0])[i];
        if (c) atomicAdd(&hist[i], c);
    }
    for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < status_words; w += stride) status[w] = 0;
}

template <int ITEMS>
__global__ __launch_bounds__(kBlock) void SortPassKernel(const uint32_t* __restrict__ keys_in,
                                                         const uint32_t* __restrict__ vals_in,
                                                         uint32_t* __restrict__ keys_out,
                                                         uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
                                                         const uint32_t* __restrict__ hist, uint32_t* tickets,
                                                         uint32_t* status) {
    constexpr int kWaves = kBlock / 64;
    constexpr int kTile = kBlock * ITEMS;
    __shared__ uint32_t s_buf[kTile];
    __shared__ uint32_t s_wave_hist[kWaves][kSortRadix];
    __shared__ uint32_t s_digit_start[kSortRadix];
    __shared__ uint32_t s_global_off[kSortRadix];
    __shared__ uint32_t s_wave_tot[kWaves];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tickets, 1u);
    for (uint32_t i = tid; i < kWaves * kSortRadix; i += kBlock) (&s_wave_hist[0][0])[i] = 0;
    // Global exclusive digit offsets from the upfront histogram.
    const uint32_t hist_d = hist[tid];
    const uint32_t hist_scan = BlockExclusiveScan256(hist_d, s_wave_tot);  // contains __syncthreads
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * kTile + (size_t)wave * (64 * ITEMS);

    uint32_t k[ITEMS], v[ITEMS], rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const size_t idx = base + (size_t)j * 64 + lane;
        const bool valid = idx < n;
        k[j] = valid ? keys_in[idx] : 0xFFFFFFFFu;
        v[j] = valid ? vals_in[idx] : 0u;
    }
    const uint64_t lt = LaneMaskLt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> shift) & (kSortRadix - 1);
        uint64_t match = ~0ull;
#pragma unroll
        for (int b = 0; b < kSortRadixBits; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(match & lt);
        const uint32_t pre = s_wave_hist[wave][d];
        rank[j] = pre + before;
        __builtin_amdgcn_wave_barrier();
        if (before == 0) s_wave_hist[wave][d] = pre + (uint32_t)__popcll(match);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Per digit: exclusive offsets of each wave and the tile total.
    uint32_t count = 0;
    {
        const uint32_t d = tid;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = s_wave_hist[w][d];
            s_wave_hist[w][d] = count;
            count += c;
        }
    }
    // Decoupled look-back for digit `tid`.
    {
        const uint32_t d = tid;
        uint32_t excl = 0;
        if (tile == 0) {
            AtomicStoreAgent(&status[d], kFlagInclusive | count);
        } else {
            AtomicStoreAgent(&status[(size_t)tile * kSortRadix + d], kFlagAggregate | count);
            int64_t p = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (p >= 0) {
                const uint32_t s = AtomicLoadAgent(&status[(size_t)p * kSortRadix + d]);
                if ((s & ~kCountMask) == 0) {
                    if (++spins > kSpinLimit) break;
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += s & kCountMask;
                if ((s & ~kCountMask) == kFlagInclusive) break;
                --p;
            }
            AtomicStoreAgent(&status[(size_t)tile * kSortRadix + d], kFlagInclusive | (excl + count));
        }
        s_global_off[d] = hist_scan + excl;
    }
    const uint32_t dstart = BlockExclusiveScan256(count, s_wave_tot);  // contains __syncthreads
    s_digit_start[tid] = dstart;
    __syncthreads();

    // Keys: rank into LDS, then write each digit run contiguously.
    uint32_t pos[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> shift) & (kSortRadix - 1);
        pos[j] = s_digit_start[d] + s_wave_hist[wave][d] + rank[j];
        s_buf[pos[j]] = k[j];
    }
    __syncthreads();
    uint32_t gpos[ITEMS];  // n <= 2^30, so positions fit in 32 bits
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t i = (uint32_t)j * kBlock + tid;
        const uint32_t key = s_buf[i];
        const uint32_t d = (key >> shift) & (kSortRadix - 1);
        gpos[j] = s_global_off[d] + (i - s_digit_start[d]);
        if (gpos[j] < n) keys_out[gpos[j]] = key;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) s_buf[pos[j]] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t i = (uint32_t)j * kBlock + tid;
        if (gpos[j] < n) vals_out[gpos[j]] = s_buf[i];
    }
}

SortShape ChooseSortShape(uint32_t n) {
    if (n >= (1u << 22)) return {kBlock, 16};
    if (n >= (1u << 19)) return {kBlock, 8};
    return {kBlock, 4};
}

static size_t SortTiles(uint32_t n) {
    const SortShape s = ChooseSortShape(n);
    return n ? ((size_t)n + s.tile() - 1) / s.tile() : 1;
}

size_t SortScratchBytes(uint32_t n) {
    const size_t words = kSortMaxPasses * kSortRadix + 64 + (size_t)kSortMaxPasses * SortTiles(n) * kSortRadix;
    return AlignUp(words * 4, 256);
}

SortScratch CarveSortScratch(void* base, uint32_t n) {
    SortScratch s;
    uint32_t* p = static_cast<uint32_t*>(base);
    s.hist = p;
    s.tickets = p + kSortMaxPasses * kSortRadix;
    s.status = s.tickets + 64;
    s.status_words = (size_t)kSortMaxPasses * SortTiles(n) * kSortRadix;
    return s;
}

void LaunchSortClear(const SortScratch& sc, hipStream_t stream) {
    SortClearKernel<<<1, kBlock, 0, stream>>>(sc.hist, sc.tickets);
}

bool RadixSortPairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                    const SortScratch& sc, bool hist_ready, hipStream_t stream) {
    if (n <= 1) return true;
    const int passes = SortPasses(key_bits);
    if (!hist_ready) {
        LaunchSortClear(sc, stream);
        const uint32_t blocks = (uint32_t)std::min<size_t>(2048, (n + kBlock * 8 - 1) / (kBlock * 8));
        SortHistogramKernel<<<blocks, kBlock, 0, stream>>>(keys, n, passes, sc.hist, sc.status, sc.status_words);
    }
    const SortShape shape = ChooseSortShape(n);
    const size_t tiles = SortTiles(n);
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    for (int p = 0; p < passes; ++p) {
        uint32_t* st = sc.status + (size_t)p * tiles * kSortRadix;
        const uint32_t* h = sc.hist + p * kSortRadix;
        const uint32_t shift = (uint32_t)(p * kSortRadixBits);
        switch (shape.items) {
            case 16:
                SortPassKernel<16><<<(uint32_t)tiles, kBlock, 0, stream>>>(ki, vi, ko, vo, n, shift, h, sc.tickets + p, st);
                break;
            case 8:
                SortPassKernel<8><<<(uint32_t)tiles, kBlock, 0, stream>>>(ki, vi, ko, vo, n, shift, h, sc.tickets + p, st);
                break;
            default:
                SortPassKernel<4><<<(uint32_t)tiles, kBlock, 0, stream>>>(ki, vi, ko, vo, n, shift, h, sc.tickets + p, st);
                break;
        }
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return ki == keys;
}

}  // namespace rr
