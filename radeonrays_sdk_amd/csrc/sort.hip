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
// waits on work that is guaranteed to publish; the spin is bounded anyway,
// and a tile that gives up sets the look-back error word (device_errors.h):
// its offsets, and so the sort order, are then wrong, and the API reports it.
#include <hip/hip_runtime.h>

#include "device_errors.h"
#include "rr_device.h"
#include "sort.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>

namespace rr {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagInclusive = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
constexpr int kLookback = 4;

__device__ __forceinline__ uint32_t AtomicLoadAgent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void AtomicStoreAgent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a BLOCK-thread block.
template <int BLOCK>
__device__ __forceinline__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t* s_wave_tot) {
    constexpr int kWaves = BLOCK / 64;
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
#pragma unroll
    for (int w = 0; w < kWaves; ++w) add += (w < (int)wave) ? s_wave_tot[w] : 0u;
    __syncthreads();
    return x - v + add;
}

}  // namespace

// Clears the first `words` words of every histogram copy, and the tickets.
__global__ __launch_bounds__(kBlock) void SortClearKernel(uint32_t* hist, uint32_t* tickets, uint32_t words) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < kSortHistCopies * words; i += gridDim.x * kBlock)
        hist[(i / words) * kSortHistStride + i % words] = 0;
    if (blockIdx.x == 0 && threadIdx.x < kSortNarrowPasses) tickets[threadIdx.x] = 0;
}

// Digit counts of every pass in one read of the keys: each thread takes 16
// consecutive keys (four 16-byte loads) per step of a grid-stride loop.
constexpr int kHistKeysPerThread = 16;

// Digits of `key` for `passes` passes of BITS-bit digits, counted in LDS.
template <int BITS>
__device__ __forceinline__ void CountDigits(uint32_t key, uint32_t* s_h, int passes) {
    for (int p = 0; p < passes; ++p) atomicAdd(&s_h[(p << BITS) | ((key >> (p * BITS)) & ((1u << BITS) - 1))], 1u);
}

template <int BITS>
__global__ __launch_bounds__(kBlock) void SortHistogramKernel(const uint32_t* __restrict__ keys, uint32_t n, int passes,
                                                              uint32_t* __restrict__ hist, uint32_t* __restrict__ status,
                                                              size_t status_words) {
    constexpr int kWords = ((32 + BITS - 1) / BITS) << BITS;  // every pass of 32-bit keys
    static_assert(kWords <= kSortHistStride, "histogram copy too small");
    __shared__ uint32_t s_h[kWords];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kWords; i += kBlock) s_h[i] = 0;
    __syncthreads();
    constexpr uint32_t kStep = kBlock * kHistKeysPerThread;
    // Steps whose keys are all in range (none when keys is not 16-byte aligned).
    const uint32_t full = (reinterpret_cast<uintptr_t>(keys) & 15u) ? 0u : n / kStep;
    for (uint32_t step = blockIdx.x; step < full; step += gridDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(keys + (size_t)step * kStep);
        uint4 q[kHistKeysPerThread / 4];
#pragma unroll
        for (int j = 0; j < kHistKeysPerThread / 4; ++j) q[j] = src[j * kBlock + threadIdx.x];
#pragma unroll
        for (int j = 0; j < kHistKeysPerThread / 4; ++j) {
            CountDigits<BITS>(q[j].x, s_h, passes);
            CountDigits<BITS>(q[j].y, s_h, passes);
            CountDigits<BITS>(q[j].z, s_h, passes);
            CountDigits<BITS>(q[j].w, s_h, passes);
        }
    }
    // The rest, one key per thread.
    for (uint32_t i = full * kStep + blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        CountDigits<BITS>(keys[i], s_h, passes);
    __syncthreads();
    uint32_t* copy = HistCopyOfBlock(hist);
    for (uint32_t i = threadIdx.x; i < ((uint32_t)passes << BITS); i += kBlock) {
        const uint32_t c = s_h[i];
        if (c) atomicAdd(&copy[i], c);
    }
    for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < status_words; w += (size_t)gridDim.x * kBlock)
        status[w] = 0;
}

// One pass over one tile, in this order: load the keys; count the tile's
// digits (LDS atomics) and publish them at once, then resolve the tile's
// global offsets by look-back; only then rank the keys (ballots) and
// scatter.  Publishing before ranking lets successors' look-backs finish
// while this tile is still ranking.  Four waves per SIMD: with 16384-key
// tiles (1024 threads x 16 keys) one tile fills a CU's LDS and at most 128
// VGPRs keep all 16 waves resident.
//
// BITS is the digit width: 8, or with 1024-thread tiles 10 (three passes
// for 30-bit keys instead of four; digit runs get shorter, so tiles stay
// large) or 4 (the reference's radix: eight passes).
template <int BLOCK, int ITEMS, int BITS = kSortRadixBits>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void SortPassKernel(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift, const uint32_t* __restrict__ hist, uint32_t* tickets,
    uint32_t* status, uint32_t* next_status, uint32_t spin_limit, uint32_t* err) {
    constexpr int kRadix = 1 << BITS;
    static_assert(BLOCK >= kRadix && BLOCK % 64 == 0, "one thread per digit");
    constexpr int kWaves = BLOCK / 64;
    constexpr int kTile = BLOCK * ITEMS;
    __shared__ uint32_t s_buf[kTile];
    __shared__ uint32_t s_vbuf[kTile];
    __shared__ uint32_t s_wave_hist[kWaves][kRadix];
    __shared__ uint32_t s_count[kRadix];
    __shared__ uint32_t s_digit_start[kRadix];
    __shared__ uint32_t s_global_off[kRadix];
    __shared__ uint32_t s_wave_tot[kWaves];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool digit_thread = tid < (uint32_t)kRadix;  // thread d handles digit d
    if (tid == 0) s_tile = atomicAdd(tickets, 1u);
    for (uint32_t i = tid; i < kWaves * kRadix; i += BLOCK) (&s_wave_hist[0][0])[i] = 0;
    if (digit_thread) s_count[tid] = 0;
    // Global exclusive digit offsets from the upfront histogram (the sum of its copies).
    uint32_t hist_d = 0;
    if (digit_thread)
#pragma unroll
        for (int c = 0; c < kSortHistCopies; ++c) hist_d += hist[c * kSortHistStride + tid];
    const uint32_t hist_scan = BlockExclusiveScan<BLOCK>(hist_d, s_wave_tot);  // contains __syncthreads
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kTile + wave * (64 * ITEMS) + lane;  // n <= 2^30: 32-bit indices

    uint32_t k[ITEMS], v[ITEMS], rank[ITEMS];
    const bool full = (tile + 1) * kTile <= n;  // no bounds checks needed
    if (full) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) k[j] = keys_in[base + (uint32_t)j * 64];
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t idx = base + (uint32_t)j * 64;
            k[j] = idx < n ? keys_in[idx] : 0xFFFFFFFFu;
        }
    }
    // The tile's digit counts.  Padding keys (all ones) are counted in digit
    // 255 and are ranked last, so they never displace a real key.
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) atomicAdd(&s_count[(k[j] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();

    // Decoupled look-back for digit `tid`, kLookback predecessors at a time:
    // their words are loaded together, so a chain of aggregates costs one
    // memory round trip per kLookback tiles rather than per tile.
    if (digit_thread) {
        const uint32_t d = tid;
        const uint32_t count = s_count[d];
        uint32_t excl = 0;
        if (tile == 0) {
            AtomicStoreAgent(&status[d], kFlagInclusive | count);
        } else {
            AtomicStoreAgent(&status[(size_t)tile * kRadix + d], kFlagAggregate | count);
            int64_t p = (int64_t)tile - 1;
            uint32_t spins = 0;
            bool done = false;
            while (!done && p >= 0) {
                uint32_t w[kLookback];
#pragma unroll
                for (int q = 0; q < kLookback; ++q)
                    w[q] = (p - q >= 0) ? AtomicLoadAgent(&status[(size_t)(p - q) * kRadix + d]) : kFlagInclusive;
                int used = 0;
                bool blocked = false;
#pragma unroll
                for (int q = 0; q < kLookback; ++q) {
                    if (done || blocked) continue;
                    const uint32_t f = w[q] & ~kCountMask;
                    if (f == 0) {
                        blocked = true;  // not published yet: poll again from here
                        continue;
                    }
                    excl += w[q] & kCountMask;
                    used = q + 1;
                    done = f == kFlagInclusive;
                }
                p -= used;
                if (blocked) {
                    if (++spins > spin_limit) {
                        if (err) StG(err + kErrLookbackTimeout, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            AtomicStoreAgent(&status[(size_t)tile * kRadix + d], kFlagInclusive | (excl + count));
        }
        s_global_off[d] = hist_scan + excl;
        // The next pass finds this tile's look-back words cleared.
        if (next_status) next_status[(size_t)tile * kRadix + d] = 0;
    }

    // Rank within each wave: lanes holding the same digit are the AND over
    // the digit's bits of the bit's ballot or its complement (0 - bit
    // selects which), one three-input bit operation per half.
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> shift) & (kRadix - 1);
        const uint32_t pre = s_wave_hist[wave][d];
        uint32_t m_lo = ~0u, m_hi = ~0u;
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            // Bit b of the digit (not of the key: shift + b may pass bit 31).
            const uint32_t neg = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);  // 0 or ~0
            const uint64_t bal = __ballot(neg != 0u);
            m_lo &= ~((uint32_t)bal ^ neg);
            m_hi &= ~((uint32_t)(bal >> 32) ^ neg);
        }
        const uint32_t before = __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
        rank[j] = pre + before;
        __builtin_amdgcn_wave_barrier();
        if (before == 0) s_wave_hist[wave][d] = pre + (uint32_t)(__popc(m_lo) + __popc(m_hi));
        __builtin_amdgcn_wave_barrier();
    }
    // Values are loaded only now, so they do not hold registers through the ranking.
    if (full) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) v[j] = vals_in[base + (uint32_t)j * 64];
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t idx = base + (uint32_t)j * 64;
            v[j] = idx < n ? vals_in[idx] : 0u;
        }
    }
    __syncthreads();
    // Per digit: exclusive offsets of each wave within the tile.
    uint32_t count = 0;
    if (digit_thread) {
        const uint32_t d = tid;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = s_wave_hist[w][d];
            s_wave_hist[w][d] = count;
            count += c;
        }
    }
    const uint32_t dstart = BlockExclusiveScan<BLOCK>(count, s_wave_tot);  // contains __syncthreads
    if (digit_thread) s_digit_start[tid] = dstart;
    __syncthreads();

    // Keys and values: place by rank in LDS, then write each digit's run
    // contiguously.
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> shift) & (kRadix - 1);
        const uint32_t pos = s_digit_start[d] + s_wave_hist[wave][d] + rank[j];
        s_buf[pos] = k[j];
        s_vbuf[pos] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t i = (uint32_t)j * BLOCK + tid;
        const uint32_t key = s_buf[i];
        const uint32_t d = (key >> shift) & (kRadix - 1);
        const uint32_t gpos = s_global_off[d] + (i - s_digit_start[d]);  // n <= 2^30: fits in 32 bits
        if (gpos < n) {
            keys_out[gpos] = key;
            vals_out[gpos] = s_vbuf[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Hybrid sort of large inputs: one most-significant-digit pass, then every
// bucket sorted on its own by one workgroup.
//
// The MSD pass is the onesweep pass above with a digit of up to 10 bits
// taken from the top of the key: it moves each pair to its bucket, stably,
// through the same decoupled look-back.  A bucket of up to kBucketCapacity
// pairs is then loaded by one workgroup, sorted by the remaining low bits in
// LDS (stable LSD rounds of up to 8 bits, ranked with ballots) and written
// back contiguously; a larger bucket (skewed keys) is sorted by its
// workgroup in capacity-sized chunks through memory.  Two passes over the
// data instead of four, and the second reads and writes whole runs.

constexpr int kMsdBits = 10;
constexpr int kMsdRadix = 1 << kMsdBits;
constexpr int kMsdBlock = 1024;  // one thread per digit
constexpr int kMsdItems = 16;    // 16,384-key tiles
constexpr int kBucketBlock = 1024;
constexpr int kBucketItems = 20;
constexpr uint32_t kBucketCapacity = kBucketBlock * kBucketItems;
constexpr int kBucketRoundBits = 8;  // widest digit of one LDS round
static_assert(kMsdBlock == kMsdRadix, "the MSD pass gives each thread one digit");

namespace {

// Lanes of this wave whose digit d (of `bits` bits, at most MAXBITS) equals
// this lane's: an AND over the digit's bits of the bit's ballot or its
// complement.
template <int MAXBITS>
__device__ __forceinline__ void MatchDigit(uint32_t d, int bits, uint32_t& m_lo, uint32_t& m_hi) {
    m_lo = ~0u;
    m_hi = ~0u;
#pragma unroll
    for (int b = 0; b < MAXBITS; ++b) {
        if (b >= bits) continue;
        const uint32_t neg = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);  // 0 or ~0
        const uint64_t bal = __ballot(neg != 0u);
        m_lo &= ~((uint32_t)bal ^ neg);
        m_hi &= ~((uint32_t)(bal >> 32) ^ neg);
    }
}

}  // namespace

// Counts of the MSD digit (key >> shift) & mask; clears the MSD pass's
// look-back words.
__global__ __launch_bounds__(kBlock) void MsdHistogramKernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                             uint32_t shift, uint32_t mask, uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ status, size_t status_words) {
    __shared__ uint32_t s_h[kMsdRadix];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kMsdRadix; i += kBlock) s_h[i] = 0;
    __syncthreads();
    constexpr uint32_t kStep = kBlock * kHistKeysPerThread;
    const uint32_t full = (reinterpret_cast<uintptr_t>(keys) & 15u) ? 0u : n / kStep;
    for (uint32_t step = blockIdx.x; step < full; step += gridDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(keys + (size_t)step * kStep);
        uint4 q[kHistKeysPerThread / 4];
#pragma unroll
        for (int j = 0; j < kHistKeysPerThread / 4; ++j) q[j] = src[j * kBlock + threadIdx.x];
#pragma unroll
        for (int j = 0; j < kHistKeysPerThread / 4; ++j) {
            atomicAdd(&s_h[(q[j].x >> shift) & mask], 1u);
            atomicAdd(&s_h[(q[j].y >> shift) & mask], 1u);
            atomicAdd(&s_h[(q[j].z >> shift) & mask], 1u);
            atomicAdd(&s_h[(q[j].w >> shift) & mask], 1u);
        }
    }
    for (uint32_t i = full * kStep + blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        atomicAdd(&s_h[(keys[i] >> shift) & mask], 1u);
    __syncthreads();
    uint32_t* copy = HistCopyOfBlock(hist);
    for (uint32_t i = threadIdx.x; i <= mask; i += kBlock) {
        const uint32_t c = s_h[i];
        if (c) atomicAdd(&copy[i], c);
    }
    for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < status_words; w += (size_t)gridDim.x * kBlock)
        status[w] = 0;
}

// The MSD pass: as SortPassKernel with a digit of `bits` (<= 10) bits at
// `shift`, one thread per digit, and keys and values staged one after the
// other through one LDS buffer (the per-wave counts of 1,024 digits take
// the other half of the LDS).  Tile 0 also writes the buckets' start
// offsets (bucket_off[d], and n at bucket_off[1 << bits]) for the bucket
// kernel.
__global__ __launch_bounds__(kMsdBlock) __attribute__((amdgpu_waves_per_eu(4))) void MsdPassKernel(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift, int bits, const uint32_t* __restrict__ hist,
    uint32_t* tickets, uint32_t* status, uint32_t* __restrict__ bucket_off, uint32_t spin_limit, uint32_t* err) {
    constexpr int kWaves = kMsdBlock / 64;
    constexpr int kTile = kMsdBlock * kMsdItems;
    __shared__ uint32_t s_buf[kTile];
    __shared__ uint32_t s_wave_hist[kWaves][kMsdRadix];
    __shared__ uint32_t s_count[kMsdRadix];
    __shared__ uint32_t s_digit_start[kMsdRadix];
    __shared__ uint32_t s_global_off[kMsdRadix];
    __shared__ uint32_t s_wave_tot[kWaves];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    if (tid == 0) s_tile = atomicAdd(tickets, 1u);
    for (uint32_t i = tid; i < kWaves * kMsdRadix; i += kMsdBlock) (&s_wave_hist[0][0])[i] = 0;
    s_count[tid] = 0;
    uint32_t hist_d = 0;
#pragma unroll
    for (int c = 0; c < kSortHistCopies; ++c) hist_d += hist[c * kSortHistStride + tid];
    if (tid > mask) hist_d = 0;
    const uint32_t hist_scan = BlockExclusiveScan<kMsdBlock>(hist_d, s_wave_tot);  // contains __syncthreads
    const uint32_t tile = s_tile;
    if (tile == 0) {
        if (tid <= mask) bucket_off[tid] = hist_scan;
        if (tid == 0) bucket_off[mask + 1] = n;
    }
    const uint32_t base = tile * kTile + wave * (64 * kMsdItems) + lane;

    uint32_t k[kMsdItems], rank[kMsdItems];
    const bool full = (tile + 1) * kTile <= n;
    if (full) {
#pragma unroll
        for (int j = 0; j < kMsdItems; ++j) k[j] = keys_in[base + (uint32_t)j * 64];
    } else {
#pragma unroll
        for (int j = 0; j < kMsdItems; ++j) {
            const uint32_t idx = base + (uint32_t)j * 64;
            k[j] = idx < n ? keys_in[idx] : 0xFFFFFFFFu;  // padding: the last digit, ranked after every real key
        }
    }
#pragma unroll
    for (int j = 0; j < kMsdItems; ++j) atomicAdd(&s_count[(k[j] >> shift) & mask], 1u);
    __syncthreads();

    // Decoupled look-back for digit `tid` (digits above mask count zero).
    {
        const uint32_t count = s_count[tid];
        uint32_t excl = 0;
        if (tile == 0) {
            AtomicStoreAgent(&status[tid], kFlagInclusive | count);
        } else {
            AtomicStoreAgent(&status[(size_t)tile * kMsdRadix + tid], kFlagAggregate | count);
            int64_t p = (int64_t)tile - 1;
            uint32_t spins = 0;
            bool done = false;
            while (!done && p >= 0) {
                uint32_t w[kLookback];
#pragma unroll
                for (int q = 0; q < kLookback; ++q)
                    w[q] = (p - q >= 0) ? AtomicLoadAgent(&status[(size_t)(p - q) * kMsdRadix + tid]) : kFlagInclusive;
                int used = 0;
                bool blocked = false;
#pragma unroll
                for (int q = 0; q < kLookback; ++q) {
                    if (done || blocked) continue;
                    const uint32_t f = w[q] & ~kCountMask;
                    if (f == 0) {
                        blocked = true;
                        continue;
                    }
                    excl += w[q] & kCountMask;
                    used = q + 1;
                    done = f == kFlagInclusive;
                }
                p -= used;
                if (blocked) {
                    if (++spins > spin_limit) {
                        if (err) StG(err + kErrLookbackTimeout, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            AtomicStoreAgent(&status[(size_t)tile * kMsdRadix + tid], kFlagInclusive | (excl + count));
        }
        s_global_off[tid] = hist_scan + excl;
    }

    // Stable rank within each wave.
#pragma unroll
    for (int j = 0; j < kMsdItems; ++j) {
        const uint32_t d = (k[j] >> shift) & mask;
        const uint32_t pre = s_wave_hist[wave][d];
        uint32_t m_lo, m_hi;
        MatchDigit<kMsdBits>(d, bits, m_lo, m_hi);
        const uint32_t before = __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
        rank[j] = pre + before;
        __builtin_amdgcn_wave_barrier();
        if (before == 0) s_wave_hist[wave][d] = pre + (uint32_t)(__popc(m_lo) + __popc(m_hi));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    uint32_t count = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wave_hist[w][tid];
        s_wave_hist[w][tid] = count;
        count += c;
    }
    const uint32_t dstart = BlockExclusiveScan<kMsdBlock>(count, s_wave_tot);  // contains __syncthreads
    s_digit_start[tid] = dstart;
    __syncthreads();
    // Tile positions (the key's rank in the tile's stable digit order).
#pragma unroll
    for (int j = 0; j < kMsdItems; ++j) {
        const uint32_t d = (k[j] >> shift) & mask;
        rank[j] += s_digit_start[d] + s_wave_hist[wave][d];
        s_buf[rank[j]] = k[j];
    }
    __syncthreads();
    // Keys out, each digit's run contiguous; remember each slot's destination.
    uint32_t dst[kMsdItems];
#pragma unroll
    for (int j = 0; j < kMsdItems; ++j) {
        const uint32_t i = (uint32_t)j * kMsdBlock + tid;
        const uint32_t key = s_buf[i];
        const uint32_t d = (key >> shift) & mask;
        dst[j] = s_global_off[d] + (i - s_digit_start[d]);
        if (dst[j] < n) keys_out[dst[j]] = key;
    }
    // Values: loaded only now, through the same buffer.
    uint32_t v[kMsdItems];
    if (full) {
#pragma unroll
        for (int j = 0; j < kMsdItems; ++j) v[j] = vals_in[base + (uint32_t)j * 64];
    } else {
#pragma unroll
        for (int j = 0; j < kMsdItems; ++j) {
            const uint32_t idx = base + (uint32_t)j * 64;
            v[j] = idx < n ? vals_in[idx] : 0u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kMsdItems; ++j) s_buf[rank[j]] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kMsdItems; ++j) {
        const uint32_t i = (uint32_t)j * kMsdBlock + tid;
        if (dst[j] < n) vals_out[dst[j]] = s_buf[i];
    }
}

namespace {

// LDS of the bucket kernel.
struct BucketLds {
    uint32_t buf[kBucketCapacity];
    uint32_t wave_hist[kBucketBlock / 64][1 << kBucketRoundBits];
    uint32_t digit_start[1 << kBucketRoundBits];
    uint32_t run[1 << kBucketRoundBits];
    uint32_t wave_tot[kBucketBlock / 64];
    uint8_t slot_digit[kBucketCapacity];  // chunked path: digit of the key at each tile slot
};

// Tile positions are below kBucketCapacity < 2^16: two per register.
__device__ __forceinline__ uint32_t GetPos(const uint32_t (&pk)[kBucketItems / 2], int j) {
    return (pk[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
}
__device__ __forceinline__ void SetPos(uint32_t (&pk)[kBucketItems / 2], int j, uint32_t v) {
    if (j & 1) pk[j >> 1] = (pk[j >> 1] & 0xFFFFu) | (v << 16);
    else pk[j >> 1] = (pk[j >> 1] & 0xFFFF0000u) | v;
}

// Stable ranks of `it` keys per thread (item j of wave w, lane l at tile
// position w * 64 * it + j * 64 + l) by the digit (key >> sh) & ((1 << bits)
// - 1): position j becomes the key's position in the tile's stable order.
// Positions are packed two per register (GetPos / SetPos).  Returns the
// tile's count of digit threadIdx.x (0 for threads past the digits);
// lds.digit_start holds each digit's first position.
__device__ __forceinline__ uint32_t RankBucketTile(const uint32_t (&k)[kBucketItems], uint32_t (&pk)[kBucketItems / 2],
                                                  uint32_t it, uint32_t sh, int bits, BucketLds& lds) {
    constexpr int kWaves = kBucketBlock / 64;
    constexpr int kR = 1 << kBucketRoundBits;
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    for (uint32_t i = tid; i < kWaves * kR; i += kBucketBlock) (&lds.wave_hist[0][0])[i] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBucketItems; ++j) {
        if ((uint32_t)j >= it) continue;
        const uint32_t d = (k[j] >> sh) & mask;
        const uint32_t pre = lds.wave_hist[wave][d];
        uint32_t m_lo, m_hi;
        MatchDigit<kBucketRoundBits>(d, bits, m_lo, m_hi);
        const uint32_t before = __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
        SetPos(pk, j, pre + before);
        __builtin_amdgcn_wave_barrier();
        if (before == 0) lds.wave_hist[wave][d] = pre + (uint32_t)(__popc(m_lo) + __popc(m_hi));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    uint32_t count = 0;
    if (tid <= mask) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = lds.wave_hist[w][tid];
            lds.wave_hist[w][tid] = count;
            count += c;
        }
    }
    const uint32_t dstart = BlockExclusiveScan<kBucketBlock>(count, lds.wave_tot);  // contains __syncthreads
    if (tid <= mask) lds.digit_start[tid] = dstart;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBucketItems; ++j) {
        if ((uint32_t)j >= it) continue;
        const uint32_t d = (k[j] >> sh) & mask;
        SetPos(pk, j, GetPos(pk, j) + lds.digit_start[d] + lds.wave_hist[wave][d]);
    }
    return count;
}

// Digit width of round r of `rounds` over `bits` low bits (balanced).
__device__ __forceinline__ int RoundBits(uint32_t bits, int rounds, int r, uint32_t done) {
    const uint32_t rest = bits - done;
    return (int)((rest + (uint32_t)(rounds - r) - 1) / (uint32_t)(rounds - r));
}

}  // namespace

namespace {

// One stable LSD round over a bucket held in registers, ITEMS pairs per
// thread (item j of wave w, lane l at bucket position w * 64 * ITEMS + j * 64
// + l), by the digit (key >> sh) & mask of at most BITS bits: rank within
// each wave by ballots, offsets per digit across waves and digits, then keys
// and values reordered through LDS, after which each thread again holds the
// items at its positions, now in the round's order.  In the last round
// (out_k not null) the reordered keys and values go straight from LDS to
// out_k / out_v, `size` of each.
template <int ITEMS, int BITS>
__device__ __forceinline__ void BucketRound(uint32_t (&k)[ITEMS], uint32_t (&v)[ITEMS], uint32_t sh, uint32_t mask,
                                            BucketLds& lds, uint32_t* __restrict__ out_k, uint32_t* __restrict__ out_v,
                                            uint32_t size) {
    constexpr int kWaves = kBucketBlock / 64;
    constexpr int kR = 1 << BITS;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t wbase = wave * 64u * ITEMS + lane;
    for (uint32_t i = tid; i < kWaves * kR; i += kBucketBlock) lds.wave_hist[i / kR][i % kR] = 0;
    __syncthreads();
    uint32_t pos[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> sh) & mask;
        uint32_t m_lo = ~0u, m_hi = ~0u;
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const uint32_t neg = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);
            const uint64_t bal = __ballot(neg != 0u);
            m_lo &= ~((uint32_t)bal ^ neg);
            m_hi &= ~((uint32_t)(bal >> 32) ^ neg);
        }
        const uint32_t pre = lds.wave_hist[wave][d];
        const uint32_t before = __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
        pos[j] = pre + before;
        __builtin_amdgcn_wave_barrier();
        if (before == 0) lds.wave_hist[wave][d] = pre + (uint32_t)(__popc(m_lo) + __popc(m_hi));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    uint32_t count = 0;
    if (tid < (uint32_t)kR) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = lds.wave_hist[w][tid];
            lds.wave_hist[w][tid] = count;
            count += c;
        }
    }
    const uint32_t dstart = BlockExclusiveScan<kBucketBlock>(count, lds.wave_tot);  // contains __syncthreads
    if (tid < (uint32_t)kR) lds.digit_start[tid] = dstart;
    __syncthreads();
    // Each key's digit is computed again below rather than kept from the
    // ranking loop (with its LDS address, two more registers per key): an
    // empty asm statement hides from the compiler that the keys are unchanged.
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) asm volatile("" : "+v"(k[j]));
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> sh) & mask;
        pos[j] += lds.digit_start[d] + lds.wave_hist[wave][d];
        lds.buf[pos[j]] = k[j];
    }
    __syncthreads();
    // One base address per array, items at constant offsets from it.
    const uint32_t* kb = lds.buf + wbase;
    if (out_k) {
        for (uint32_t i = tid; i < size; i += kBucketBlock) out_k[i] = lds.buf[i];
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) k[j] = kb[j * 64];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) lds.buf[pos[j]] = v[j];
    __syncthreads();
    if (out_k) {
        for (uint32_t i = tid; i < size; i += kBucketBlock) out_v[i] = lds.buf[i];
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) v[j] = kb[j * 64];
    }
    __syncthreads();
}

// A bucket of at most 1024 * ITEMS pairs: load (padding with all-ones keys,
// which rank after every real key in every round), the LSD rounds, store.
template <int ITEMS>
__device__ __forceinline__ void SortBucketInLds(const uint32_t* __restrict__ keys_in,
                                                const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
                                                uint32_t* __restrict__ vals_out, uint32_t start, uint32_t size,
                                                uint32_t low_bits, BucketLds& lds) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t wbase = wave * 64u * ITEMS + lane;
    uint32_t k[ITEMS], v[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t p = wbase + (uint32_t)j * 64;
        const bool ok = p < size;
        k[j] = ok ? keys_in[start + p] : 0xFFFFFFFFu;
        v[j] = ok ? vals_in[start + p] : 0u;
    }
    // At least one round (a 0-bit one when there are no low bits: a copy).
    const int rounds = (int)max(1u, (low_bits + kBucketRoundBits - 1) / kBucketRoundBits);
    uint32_t sh = 0;
    for (int r = 0; r < rounds; ++r) {
        const int bits = RoundBits(low_bits, rounds, r, sh);
        const uint32_t mask = (1u << bits) - 1u;
        uint32_t* ok = r + 1 == rounds ? keys_out + start : nullptr;
        BucketRound<ITEMS, kBucketRoundBits>(k, v, sh, mask, lds, ok, vals_out + start, size);
        sh += (uint32_t)bits;
    }
}

}  // namespace

// One workgroup per MSD bucket of at most 1024 * ITEMS pairs: sort the
// bucket's pairs (read from keys_in / vals_in, where the MSD pass put them)
// by their low `low_bits` bits into keys_out / vals_out, in LDS.  Larger
// buckets are left to BucketSortChunkedKernel.
template <int ITEMS>
__global__ __launch_bounds__(kBucketBlock) void BucketSortLdsKernel(const uint32_t* __restrict__ keys_in,
                                                                    const uint32_t* __restrict__ vals_in,
                                                                    uint32_t* __restrict__ keys_out,
                                                                    uint32_t* __restrict__ vals_out,
                                                                    const uint32_t* __restrict__ bucket_off,
                                                                    uint32_t low_bits) {
    __shared__ BucketLds lds;
    const uint32_t start = bucket_off[blockIdx.x];
    const uint32_t size = bucket_off[blockIdx.x + 1] - start;
    if (size == 0 || size > kBucketBlock * ITEMS) return;
    SortBucketInLds<ITEMS>(keys_in, vals_in, keys_out, vals_out, start, size, low_bits, lds);
}

// One workgroup per MSD bucket larger than `cap` pairs (skewed keys).
__global__ __launch_bounds__(kBucketBlock) void BucketSortChunkedKernel(uint32_t* __restrict__ keys_in,
                                                                        uint32_t* __restrict__ vals_in,
                                                                        uint32_t* __restrict__ keys_out,
                                                                        uint32_t* __restrict__ vals_out,
                                                                        const uint32_t* __restrict__ bucket_off,
                                                                        uint32_t low_bits, uint32_t cap) {
    __shared__ BucketLds lds;
    const uint32_t start = bucket_off[blockIdx.x];
    const uint32_t size = bucket_off[blockIdx.x + 1] - start;
    if (size <= cap) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const int rounds = (int)((low_bits + kBucketRoundBits - 1) / kBucketRoundBits);
    uint32_t k[kBucketItems], v[kBucketItems], pk[kBucketItems / 2];

    // A bucket larger than the LDS: LSD rounds through memory, each over
    // the bucket in capacity-sized chunks (stable: chunks in order, ranks
    // within a chunk stable, each digit's output advancing chunk by chunk).
    // Rounds alternate between the two arrays; the last lands in keys_out.
    uint32_t *sk = keys_in + start, *sv = vals_in + start, *dk = keys_out + start, *dv = vals_out + start;
    const int passes = rounds + ((rounds & 1) ? 0 : 1);  // odd, so the data ends in keys_out
    uint32_t sh = 0;
    for (int r = 0; r < passes; ++r) {
        const int bits = r < rounds ? RoundBits(low_bits, rounds, r, sh) : 0;  // a 0-bit round copies
        const uint32_t mask = (1u << bits) - 1u;
        for (uint32_t i = tid; i < (1u << kBucketRoundBits); i += kBucketBlock) lds.run[i] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < size; i += kBucketBlock) atomicAdd(&lds.run[(sk[i] >> sh) & mask], 1u);
        __syncthreads();
        const uint32_t c = tid <= mask ? lds.run[tid] : 0u;
        const uint32_t excl = BlockExclusiveScan<kBucketBlock>(c, lds.wave_tot);  // contains __syncthreads
        if (tid <= mask) lds.run[tid] = excl;
        __syncthreads();
        for (uint32_t cbase = 0; cbase < size; cbase += kBucketCapacity) {
            const uint32_t csize = min(kBucketCapacity, size - cbase);
            const uint32_t it = (csize + kBucketBlock - 1) / kBucketBlock;
            const uint32_t wbase = wave * 64u * it + lane;
#pragma unroll
            for (int j = 0; j < kBucketItems; ++j) {
                if ((uint32_t)j >= it) continue;
                const uint32_t p = wbase + (uint32_t)j * 64;
                const bool ok = p < csize;
                k[j] = ok ? sk[cbase + p] : 0xFFFFFFFFu;
                v[j] = ok ? sv[cbase + p] : 0u;
            }
            const uint32_t cnt = RankBucketTile(k, pk, it, sh, bits, lds);
#pragma unroll
            for (int j = 0; j < kBucketItems; ++j) {
                if ((uint32_t)j >= it) continue;
                lds.buf[GetPos(pk, j)] = k[j];
            }
            __syncthreads();
            // Keys out by tile slot (slot i = tid + j * block; padding, ranked
            // last, falls at or past csize and is not written); each slot's
            // digit is kept for the values.
            for (uint32_t i = tid; i < csize; i += kBucketBlock) {
                const uint32_t key = lds.buf[i];
                const uint32_t d = (key >> sh) & mask;
                lds.slot_digit[i] = (uint8_t)d;
                dk[lds.run[d] + (i - lds.digit_start[d])] = key;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kBucketItems; ++j) {
                if ((uint32_t)j >= it) continue;
                lds.buf[GetPos(pk, j)] = v[j];
            }
            __syncthreads();
            for (uint32_t i = tid; i < csize; i += kBucketBlock) {
                const uint32_t d = lds.slot_digit[i];
                dv[lds.run[d] + (i - lds.digit_start[d])] = lds.buf[i];
            }
            __syncthreads();
            if (tid <= mask) lds.run[tid] += cnt;
            __syncthreads();
        }
        uint32_t* t = sk;
        sk = dk;
        dk = t;
        t = sv;
        sv = dv;
        dv = t;
        sh += (uint32_t)bits;
    }
}

// Tile shapes the pass kernel is instantiated for.
static bool KnownShape(int block, int items, int bits) {
    if (bits == kSortWideBits) return block == 1024 && (items == 8 || items == 4);
    if (bits == kSortNarrowBits) return block == 1024 && items == 16;
    return bits == 8 && ((block == 1024 && (items == 16 || items == 8)) || (block == 512 && (items == 16 || items == 8)) ||
                         (block == 256 && (items == 16 || items == 8 || items == 4)));
}

static std::atomic<uint32_t> g_spin_limit{kSortDefaultSpinLimit};
uint32_t SortSpinLimit() { return g_spin_limit.load(std::memory_order_relaxed); }
void SetSortSpinLimit(uint32_t polls) { g_spin_limit.store(polls, std::memory_order_relaxed); }

static std::atomic<int> g_digit_bits{8};
int SortDigitBits() { return g_digit_bits.load(std::memory_order_relaxed); }
bool SetSortDigitBits(int bits) {
    if (bits != 8 && bits != kSortWideBits && bits != kSortNarrowBits) return false;
    g_digit_bits.store(bits, std::memory_order_relaxed);
    return true;
}

// Inputs at least this large may take 10-bit digits (8192-key tiles) or
// 4-bit digits (16384-key tiles).
constexpr uint32_t kWideMinKeys = 1u << 20;

// The 8-bit shape of a sort of n keys, without tuning.
static SortShape DefaultShape(uint32_t n) {
    // Large inputs: 16384-key tiles (fewer tiles, so fewer look-back rounds,
    // and 64-key digit runs per tile on average, i.e. whole cache lines).
    if (n >= (1u << 22)) return {1024, 16, 8};
    // Mid-size inputs: 4096-key tiles (whole fast build of 1M triangles
    // 0.269 against 0.295 ms with 2048-key tiles; of 262,144 triangles 0.160
    // against 0.171 ms with 1024-key tiles; the same at 65,536).
    if (n >= (1u << 17)) return {512, 8, 8};
    return {kBlock, 4, 8};
}

// Shape of a sort of n keys.  wide_ok: digits other than 8 bits are allowed
// (not for callers that prepared an 8-bit histogram themselves).
static SortShape ChooseShape(uint32_t n, bool wide_ok) {
    // RR_SORT_SHAPE=<threads>x<keys per thread>[x<digit bits>] (one of
    // KnownShape) fixes the shape of sorts that allow it (tuning runs).
    static const SortShape forced = [] {
        const char* e = std::getenv("RR_SORT_SHAPE");
        int b = 0, i = 0, bits = 8;
        if (e && std::sscanf(e, "%dx%dx%d", &b, &i, &bits) >= 2 && KnownShape(b, i, bits)) return SortShape{b, i, bits};
        return SortShape{0, 0, 8};
    }();
    if (forced.block && (wide_ok || forced.bits == 8)) return forced;
    if (wide_ok && n >= kWideMinKeys && SortDigitBits() == kSortWideBits) return {1024, 8, kSortWideBits};
    if (wide_ok && n >= kWideMinKeys && SortDigitBits() == kSortNarrowBits) return {1024, 16, kSortNarrowBits};
    return DefaultShape(n);
}

SortShape ChooseSortShape(uint32_t n) { return ChooseShape(n, true); }

static size_t TilesOf(uint32_t n, const SortShape& s) { return n ? ((size_t)n + s.tile() - 1) / s.tile() : 1; }
static int PassesOf(int key_bits, const SortShape& s) { return (key_bits + s.bits - 1) / s.bits; }

// Look-back words of all passes for one shape (32-bit keys).
static size_t StatusWords(uint32_t n, const SortShape& s) {
    return (size_t)PassesOf(32, s) * TilesOf(n, s) * ((size_t)1 << s.bits);
}

static std::atomic<int> g_hybrid{0};
bool SortHybrid() { return g_hybrid.load(std::memory_order_relaxed) != 0; }
void SetSortHybrid(bool on) { g_hybrid.store(on ? 1 : 0, std::memory_order_relaxed); }

// MSD digit width of the hybrid sort for n keys of key_bits bits, or 0 when
// the LSD passes sort them.  The digit is as wide as gives buckets of about
// 4,096 keys on uniform keys, up to 10 bits (16,384 keys at 2^24 keys:
// well inside the 24,576 of one workgroup's LDS sort for such keys).
static int HybridBits(uint32_t n, int key_bits, bool hist_ready) {
    if (hist_ready || !SortHybrid() || n < kSortHybridMinKeys || n > kSortHybridMaxKeys || key_bits < 2) return 0;
    int lg = 0;
    while (lg < 31 && (1u << lg) < n) ++lg;
    return std::max(1, std::min({lg - 12, kMsdBits, key_bits}));
}

static size_t MsdTiles(uint32_t n) { return ((size_t)n + kMsdBlock * kMsdItems - 1) / (kMsdBlock * kMsdItems); }

// Words before the look-back words: histogram copies, tickets, bucket offsets.
constexpr size_t kSortHeadWords = (size_t)kSortHistCopies * kSortHistStride + 64 + (kMsdRadix + 1 + 63) / 64 * 64;

size_t SortScratchBytes(uint32_t n) {
    // Enough for any shape a sort of n keys may take, whatever the digit
    // width is set to when it runs.
    size_t status = std::max(StatusWords(n, ChooseShape(n, true)), StatusWords(n, ChooseShape(n, false)));
    if (n >= kWideMinKeys) {
        status = std::max(status, StatusWords(n, SortShape{1024, 8, kSortWideBits}));
        status = std::max(status, StatusWords(n, SortShape{1024, 16, kSortNarrowBits}));
    }
    if (n >= kSortHybridMinKeys && n <= kSortHybridMaxKeys) status = std::max(status, MsdTiles(n) * kMsdRadix);
    return AlignUp((kSortHeadWords + status) * 4, 256);
}

SortScratch CarveSortScratch(void* base, uint32_t n) {
    SortScratch s;
    uint32_t* p = static_cast<uint32_t*>(base);
    s.hist = p;
    s.tickets = p + kSortHistCopies * kSortHistStride;
    s.bucket_off = s.tickets + 64;
    s.status = p + kSortHeadWords;
    // Pass 0's look-back words for 8-bit digits (what callers that build the
    // histogram themselves clear); each pass clears the next pass's.
    const SortShape narrow = ChooseShape(n, false);
    s.status_words = TilesOf(n, narrow) * ((size_t)1 << narrow.bits);
    return s;
}

void LaunchSortClear(const SortScratch& sc, hipStream_t stream) {
    SortClearKernel<<<1, kBlock, 0, stream>>>(sc.hist, sc.tickets, kSortHistWords);
}

// The hybrid sort: MSD pass into the tmp arrays, then the bucket kernel
// back into keys/vals.
static void HybridSort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                       int key_bits, int bits, const SortScratch& sc, hipStream_t stream) {
    const uint32_t shift = (uint32_t)(key_bits - bits);
    const uint32_t mask = (1u << bits) - 1u;
    const size_t tiles = MsdTiles(n);
    SortClearKernel<<<(kSortHistCopies * kMsdRadix + kBlock - 1) / kBlock, kBlock, 0, stream>>>(sc.hist, sc.tickets,
                                                                                               kMsdRadix);
    const uint32_t blocks =
        (uint32_t)std::max<size_t>(1, std::min<size_t>(1024, (size_t)n / (kBlock * kHistKeysPerThread)));
    MsdHistogramKernel<<<blocks, kBlock, 0, stream>>>(keys, n, shift, mask, sc.hist, sc.status, tiles * kMsdRadix);
    MsdPassKernel<<<(uint32_t)tiles, kMsdBlock, 0, stream>>>(keys, vals, keys_tmp, vals_tmp, n, shift, bits, sc.hist,
                                                             sc.tickets, sc.status, sc.bucket_off, SortSpinLimit(),
                                                             DeviceErrorWords());
    // Buckets average n >> bits pairs; the LDS kernel takes those of up to
    // 1024 * items (items: that average plus a quarter, in steps of four),
    // the chunked kernel the rest.
    const uint32_t avg = (uint32_t)((n >> bits) + 1);
    const uint32_t items = std::min<uint32_t>(kBucketItems, (avg + avg / 4 + 4 * kBucketBlock - 1) / (4 * kBucketBlock) * 4);
    const uint32_t buckets = 1u << bits;
    switch (items) {
        case 4: BucketSortLdsKernel<4><<<buckets, kBucketBlock, 0, stream>>>(keys_tmp, vals_tmp, keys, vals, sc.bucket_off, shift); break;
        case 8: BucketSortLdsKernel<8><<<buckets, kBucketBlock, 0, stream>>>(keys_tmp, vals_tmp, keys, vals, sc.bucket_off, shift); break;
        case 12: BucketSortLdsKernel<12><<<buckets, kBucketBlock, 0, stream>>>(keys_tmp, vals_tmp, keys, vals, sc.bucket_off, shift); break;
        case 16: BucketSortLdsKernel<16><<<buckets, kBucketBlock, 0, stream>>>(keys_tmp, vals_tmp, keys, vals, sc.bucket_off, shift); break;
        default: BucketSortLdsKernel<20><<<buckets, kBucketBlock, 0, stream>>>(keys_tmp, vals_tmp, keys, vals, sc.bucket_off, shift); break;
    }
    BucketSortChunkedKernel<<<buckets, kBucketBlock, 0, stream>>>(keys_tmp, vals_tmp, keys, vals, sc.bucket_off, shift,
                                                                  kBucketBlock * std::max<uint32_t>(4, items));
}

bool RadixSortPairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                    const SortScratch& sc, bool hist_ready, hipStream_t stream, bool untuned) {
    if (n <= 1) return true;
    if (const int bits = untuned ? 0 : HybridBits(n, key_bits, hist_ready)) {
        HybridSort(keys, vals, keys_tmp, vals_tmp, n, key_bits, bits, sc, stream);
        return true;
    }
    const SortShape shape = untuned ? DefaultShape(n) : ChooseShape(n, !hist_ready);
    const int passes = PassesOf(key_bits, shape);
    const size_t tiles = TilesOf(n, shape);
    const size_t radix = (size_t)1 << shape.bits;
    if (!hist_ready) {
        const uint32_t words = (uint32_t)(passes * radix);
        SortClearKernel<<<(kSortHistCopies * words + kBlock - 1) / kBlock, kBlock, 0, stream>>>(sc.hist, sc.tickets,
                                                                                                words);
        const uint32_t blocks =
            (uint32_t)std::max<size_t>(1, std::min<size_t>(1024, (size_t)n / (kBlock * kHistKeysPerThread)));
        const size_t status0 = tiles * radix;
        if (shape.bits == kSortWideBits)
            SortHistogramKernel<kSortWideBits><<<blocks, kBlock, 0, stream>>>(keys, n, passes, sc.hist, sc.status, status0);
        else if (shape.bits == kSortNarrowBits)
            SortHistogramKernel<kSortNarrowBits><<<blocks, kBlock, 0, stream>>>(keys, n, passes, sc.hist, sc.status,
                                                                               status0);
        else
            SortHistogramKernel<8><<<blocks, kBlock, 0, stream>>>(keys, n, passes, sc.hist, sc.status, status0);
    }
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    const uint32_t spin = SortSpinLimit();
    uint32_t* err = DeviceErrorWords();
    for (int p = 0; p < passes; ++p) {
        uint32_t* st = sc.status + (size_t)p * tiles * radix;
        uint32_t* nst = p + 1 < passes ? st + tiles * radix : nullptr;
        const uint32_t* h = sc.hist + p * radix;
        const uint32_t shift = (uint32_t)(p * shape.bits);
#define RR_SORT_PASS(B, I, D) \
    SortPassKernel<B, I, D><<<(uint32_t)tiles, B, 0, stream>>>(ki, vi, ko, vo, n, shift, h, sc.tickets + p, st, nst, \
                                                              spin, err)
        if (shape.bits == kSortWideBits) {
            if (shape.items == 8) RR_SORT_PASS(1024, 8, kSortWideBits);
            else RR_SORT_PASS(1024, 4, kSortWideBits);
        } else if (shape.bits == kSortNarrowBits) {
            RR_SORT_PASS(1024, 16, kSortNarrowBits);
        } else if (shape.block == 1024) {
            if (shape.items == 16) RR_SORT_PASS(1024, 16, 8);
            else RR_SORT_PASS(1024, 8, 8);
        } else if (shape.block == 512) {
            if (shape.items == 16) RR_SORT_PASS(512, 16, 8);
            else RR_SORT_PASS(512, 8, 8);
        } else {
            if (shape.items == 16) RR_SORT_PASS(kBlock, 16, 8);
            else if (shape.items == 8) RR_SORT_PASS(kBlock, 8, 8);
            else RR_SORT_PASS(kBlock, 4, 8);
        }
#undef RR_SORT_PASS
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return ki == keys;
}

// ---------------------------------------------------------------------------
// Sort without allocation: the sort's scratch (histogram copies, tickets,
// look-back words) in a static device array instead of a caller's buffer,
// for callers that cannot allocate device memory.  The array is sized for
// kSortNoAllocMaxKeys pairs with the untuned 8-bit shapes: four passes of at
// most 1,024 tiles (16,384-key tiles from 2^22 keys, 4,096-key tiles below)
// of 256 look-back words.  Every GPU the library runs on holds its own copy
// until the process ends; all no-allocation sorts on one GPU share it, so
// such sorts must not overlap in time on one GPU (enqueue them on one
// stream, or order their streams).  The sort clears what must start at zero
// (histograms and tickets by SortClearKernel, look-back words by the
// histogram kernel), so a call does not depend on what the last one left.

constexpr size_t kSortNoAllocStatusWords = (size_t)kSortMaxPasses * 1024 * kSortRadix;
constexpr size_t kSortNoAllocWords = kSortHeadWords + kSortNoAllocStatusWords;

static __device__ __align__(256) uint32_t scratch[kSortNoAllocWords];

// Device address of this GPU's array, found at context creation (on the
// host the array's name only stands in for it and must not be passed to a
// kernel).
constexpr int kMaxDevices = 64;
static std::atomic<uint32_t*> g_noalloc_scratch[kMaxDevices];

bool InitSortNoAllocScratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return false;
    if (g_noalloc_scratch[dev].load(std::memory_order_acquire)) return true;
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p) {
        (void)hipGetLastError();
        return false;
    }
    g_noalloc_scratch[dev].store(static_cast<uint32_t*>(p), std::memory_order_release);
    return true;
}

size_t SortNoAllocScratchBytes() { return sizeof(uint32_t) * kSortNoAllocWords; }

int SortPairsNoAlloc(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                     hipStream_t stream) {
    if (n > kSortNoAllocMaxKeys) return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -2;
    uint32_t* base = g_noalloc_scratch[dev].load(std::memory_order_acquire);
    if (!base) return -2;
    if (n <= 1) return 1;
    SortScratch sc;
    sc.hist = base;
    sc.tickets = base + kSortHistCopies * kSortHistStride;
    sc.bucket_off = sc.tickets + 64;
    sc.status = base + kSortHeadWords;
    const SortShape shape = DefaultShape(n);
    sc.status_words = TilesOf(n, shape) * ((size_t)1 << shape.bits);
    if (kSortHeadWords + StatusWords(n, shape) > kSortNoAllocWords) return -1;  // not reached for n <= the maximum
    return RadixSortPairs(keys, vals, keys_tmp, vals_tmp, n, key_bits, sc, false, stream, true) ? 1 : 0;
}

}  // namespace rr
