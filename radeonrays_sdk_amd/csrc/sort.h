// sort.h — host interface of the onesweep LSD radix sort of (uint32 key,
// uint32 value) pairs.  Replaces the reference's 8-pass 4-bit sort
// (vlk/radix_sort.cpp:157-338 with radix_sort_group_histograms.comp,
// scan_exclusive_add*.comp and radix_sort_scatter_keys.comp): one upfront
// histogram of every digit, then one kernel per 8-bit digit that ranks a
// tile in LDS, finds its global offsets by decoupled look-back and scatters.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rr {

constexpr int kSortRadixBits = 8;
constexpr int kSortRadix = 1 << kSortRadixBits;
constexpr int kSortMaxPasses = 4;
// The global digit histogram is kept as kSortHistCopies copies: a block adds
// its counts into copy blockIdx.x % kSortHistCopies (one per XCD, as blocks
// are dispatched round-robin), which divides the same-address atomics of
// thousands of blocks by that factor; readers sum the copies.
constexpr int kSortHistCopies = 8;
constexpr int kSortHistWords = kSortMaxPasses * kSortRadix;  // 8-bit digits: words used in each copy
// Large sorts may use 10-bit digits (three passes for 30-bit keys, four for
// 32-bit keys) or the reference's 4-bit digits (eight passes for 32-bit
// keys); a copy is laid out for the largest (4 << 10 >= 4 << 8 >= 8 << 4).
constexpr int kSortWideBits = 10;
constexpr int kSortWidePasses = 4;
constexpr int kSortNarrowBits = 4;
constexpr int kSortNarrowPasses = 8;
constexpr int kSortHistStride = kSortWidePasses << kSortWideBits;  // words per copy
static_assert(kSortHistStride >= (kSortNarrowPasses << kSortNarrowBits), "histogram copy too small");
static_assert(kSortHistStride >= (kSortMaxPasses << kSortRadixBits), "histogram copy too small");

// Tile shape of one pass kernel.  Small inputs use small tiles so that the
// grid still spreads over all CUs.
struct SortShape {
    int block;   // threads per tile
    int items;   // keys per thread
    int bits;    // digit width: 8, or 10 or 4 (1024-thread tiles)
    int tile() const { return block * items; }
};
SortShape ChooseSortShape(uint32_t n);

// Digit width of sorts of at least 2^20 keys whose callers did not prepare
// the histogram (rrxCmdSortPairs): 8 (default), 10 or 4.  Process-wide.
int SortDigitBits();
bool SetSortDigitBits(int bits);

// Bound of a tile's look-back wait, in polls that find a predecessor's
// word unpublished (process-wide; default 2^22, far beyond any real wait).
// Past it the tile gives up and sets the look-back error word; 0 gives up
// at the first such poll (a test hook for the failure report).
constexpr uint32_t kSortDefaultSpinLimit = 1u << 22;
uint32_t SortSpinLimit();
void SetSortSpinLimit(uint32_t polls);

// With the hybrid switch on, sorts of kSortHybridMinKeys..kSortHybridMaxKeys
// pairs whose caller did not prepare the histogram take one
// most-significant-digit pass (up to 10 bits) and then sort each bucket in
// one workgroup's LDS, instead of the LSD passes (process-wide; off by
// default; the order is the same either way).
constexpr uint32_t kSortHybridMinKeys = 1u << 22;
constexpr uint32_t kSortHybridMaxKeys = 1u << 24;
bool SortHybrid();
void SetSortHybrid(bool on);

// Scratch needed to sort n pairs: histograms, tile tickets and look-back
// status words.  Everything in it is (re)initialised by the sort itself.
size_t SortScratchBytes(uint32_t n);

// Number of 8-bit passes for keys whose bits above key_bits are zero.
__host__ __device__ inline int SortPasses(int key_bits) { return (key_bits + kSortRadixBits - 1) / kSortRadixBits; }

struct SortScratch {
    uint32_t* hist;     // [kSortHistCopies][kSortMaxPasses][256] global digit counts
    uint32_t* tickets;  // [kSortNarrowPasses] tile tickets, one per pass
    uint32_t* bucket_off;  // [1025] hybrid sort: first position of each MSD bucket, then n
    uint32_t* status;   // [kSortMaxPasses][tiles][256] look-back words
    size_t status_words;  // words of pass 0's look-back region, which must be zero before a sort
};
SortScratch CarveSortScratch(void* base, uint32_t n);

// Sort keys/vals in place (ping-ponging through keys_tmp/vals_tmp) by the
// low key_bits bits; stable.  When hist_ready is true the caller has already
// cleared hist/tickets, accumulated the digit histograms into hist and
// cleared the first status_words look-back words (the build fuses that into
// its Morton kernel).
// Returns true if the result ended in keys/vals, false if in the tmp arrays.
// untuned: the default 8-bit LSD shapes whatever the tuning (the
// no-allocation sort, whose scratch is sized for them).
bool RadixSortPairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                    const SortScratch& sc, bool hist_ready, hipStream_t stream, bool untuned = false);

// Sort without allocation (rrxCmdSortPairsNoAlloc): the scratch is a static
// device array of SortNoAllocScratchBytes() bytes per GPU, shared by every
// such sort on that GPU (overlapping calls on one GPU race).  At most
// kSortNoAllocMaxKeys pairs.  InitSortNoAllocScratch finds the current
// device's array (context creation calls it with the device current).
// SortPairsNoAlloc returns 1 if the result is in keys/vals, 0 if in the tmp
// arrays, -1 if n is above the maximum, -2 if the array was not found.
constexpr uint32_t kSortNoAllocMaxKeys = 1u << 24;
bool InitSortNoAllocScratch();
size_t SortNoAllocScratchBytes();
int SortPairsNoAlloc(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                     hipStream_t stream);

// Launch helpers exposed for the build's fused histogram.
void LaunchSortClear(const SortScratch& sc, hipStream_t stream);

// This block's copy of the global histogram.
__device__ __forceinline__ uint32_t* HistCopyOfBlock(uint32_t* hist) {
    return hist + (blockIdx.x % kSortHistCopies) * kSortHistStride;
}

// Accumulate the 8-bit digits of `key` for `passes` passes into LDS
// histograms.  The top two digits of spatially ordered keys (Morton codes of
// a mesh in triangle order) are mostly equal across a wave, and same-address
// LDS atomics serialise, so for those a wave whose valid keys all share the
// digit adds once; lower digits are near-random and go straight to atomics.
// With CHECK_TOP false (keys of no particular order) every digit goes
// straight to the atomics.
template <bool CHECK_TOP = true>
__device__ __forceinline__ void AccumulateDigits(uint32_t key, bool valid, uint32_t (*s_h)[kSortRadix], int passes) {
    const uint64_t act = __ballot(valid);
    const uint32_t lane = __lane_id();
    for (int p = 0; p < passes; ++p) {
        const uint32_t d = (key >> (p * kSortRadixBits)) & (kSortRadix - 1);
        if (!CHECK_TOP || p + 2 < passes) {
            if (valid) atomicAdd(&s_h[p][d], 1u);
            continue;
        }
        const uint32_t d0 = __shfl(d, 0, 64);
        const bool same = __all(!valid || d == d0);
        if (same) {
            if (lane == 0 && valid) atomicAdd(&s_h[p][d0], (uint32_t)__popcll(act));
        } else if (valid) {
            atomicAdd(&s_h[p][d], 1u);
        }
    }
}

}  // namespace rr
