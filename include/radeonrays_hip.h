/*
 * radeonrays_hip.h — HIP interop for the MI355X library.
 *
 * Plays the role that radeonrays_vlk.h (src/core/include/radeonrays_vlk.h:38-95)
 * and radeonrays_dx.h play for the Vulkan and DX12 backends: wrapping client
 * memory and queues, and allocating host-mappable buffers.  HIP handles are
 * passed as void* so that this header needs no HIP include; pass a
 * hipStream_t where a stream is asked for and a device address from hipMalloc
 * (or a torch tensor's data_ptr()) where memory is asked for.
 */
#ifndef RADEONRAYS_HIP_H
#define RADEONRAYS_HIP_H

#include "radeonrays.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Create a context on HIP device `device` (replaces rrCreateContextVk,
 * radeonrays_vlk.h:50).  If hip_stream is not NULL, internal command
 * streams submit onto it; otherwise the context creates its own stream. */
RR_API RRError rrCreateContextHip(uint32_t api_version, int device, void* hip_stream, RRContext* context);

/* Wrap a device address plus byte offset as an RRDevicePtr (replaces
 * rrGetDevicePtrFromVkBuffer, radeonrays_vlk.h:38).  The handle does not own
 * the memory; free the handle with rrReleaseDevicePtr. */
RR_API RRError rrGetDevicePtrFromHipPtr(RRContext context, void* device_address, size_t offset, RRDevicePtr* device_ptr);

/* Wrap a client hipStream_t as a command stream (replaces
 * rrGetCommandStreamFromVkCommandBuffer, radeonrays_vlk.h:64).  rrCmd* calls
 * on it enqueue immediately, in order, on that stream.  Free the wrapper
 * with rrReleaseExternalCommandStream. */
RR_API RRError rrGetCommandStreamFromHipStream(RRContext context, void* hip_stream, RRCommandStream* command_stream);

/* Allocate `size` bytes of device (HBM) memory owned by the returned handle
 * (radeonrays_vlk.h:77).  rrReleaseDevicePtr frees the memory too. */
RR_API RRError rrAllocateDeviceBuffer(RRContext context, size_t size, RRDevicePtr* device_ptr);

/* Map a buffer from rrAllocateDeviceBuffer for host access
 * (radeonrays_vlk.h:86).  The mapping is a host copy: mapping waits for the
 * device and copies the buffer to the host; unmapping copies it back. */
RR_API RRError rrMapDevicePtr(RRContext context, RRDevicePtr device_ptr, void** mapping_ptr);
RR_API RRError rrUnmapDevicePtr(RRContext context, RRDevicePtr device_ptr, void** mapping_ptr);

/* Device address (base + offset) behind an RRDevicePtr; NULL on bad input.
 * No counterpart in the reference: lets clients of rrAllocateDeviceBuffer
 * hand the memory to their own HIP kernels. */
RR_API void* rrGetHipAddress(RRDevicePtr device_ptr);

/* hipStream_t behind a command stream.  No counterpart in the reference. */
RR_API void* rrGetHipStream(RRCommandStream command_stream);

#ifdef __cplusplus
}
#endif

#endif /* RADEONRAYS_HIP_H */
