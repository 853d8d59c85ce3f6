/*
 * radeonrays_ext.h — extensions of the MI355X library with no counterpart
 * in the reference API.  They expose pieces of the hot path on their own
 * (for measurement and for callers that batch work across many
 * acceleration structures); none is needed to use the rr* API.
 */
#ifndef RADEONRAYS_EXT_H
#define RADEONRAYS_EXT_H

#include "radeonrays.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scratch bytes for rrxCmdSortPairs on n pairs. */
RR_API RRError rrxGetSortMemoryRequirements(RRContext context, uint32_t n, size_t* scratch_size);

/* Stable LSD radix sort of n (uint32 key, uint32 value) pairs by the low
 * key_bits bits of the key, in place in keys/values, using keys_tmp /
 * values_tmp (n uint32 each) and scratch.  The same onesweep sort the BVH
 * builder runs on Morton codes (reference counterpart: the internal
 * RadixSortKeyValue of vlk/radix_sort.cpp, which the public API does not
 * expose). */
RR_API RRError rrxCmdSortPairs(RRContext       context,
                               RRDevicePtr     keys,
                               RRDevicePtr     values,
                               RRDevicePtr     keys_tmp,
                               RRDevicePtr     values_tmp,
                               uint32_t        n,
                               uint32_t        key_bits,
                               RRDevicePtr     scratch,
                               RRCommandStream command_stream);

/* Closest- or any-hit query like rrCmdIntersect with FULL_HIT_T output that
 * also writes per-ray traversal counters to `stats` (4 uint32 per ray: inner
 * nodes visited, leaves visited, box tests, triangle tests), the figures
 * bvh_analyzer reports as TraversalStats (bvh_analyzer/intersection_primitives.h:70-100). */
RR_API RRError rrxCmdIntersectWithStats(RRContext        context,
                                        RRDevicePtr      scene_buffer,
                                        RRIntersectQuery query,
                                        RRDevicePtr      rays,
                                        uint32_t         ray_count,
                                        RRDevicePtr      hits,
                                        RRDevicePtr      scratch,
                                        RRDevicePtr      stats,
                                        RRCommandStream  command_stream);

#ifdef __cplusplus
}
#endif

#endif /* RADEONRAYS_EXT_H */
