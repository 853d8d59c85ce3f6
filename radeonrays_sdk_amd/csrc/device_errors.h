// device_errors.h — device-side failure reporting.
//
// Kernels that cannot produce a correct result set a word here instead of
// returning wrong data silently: a traversal stack that overflows, a sort
// look-back that gives up, an acceleration-structure buffer whose header is
// not one this library wrote.  The words live in pinned, coherent host
// memory mapped into every device's address space, so a kernel stores 1
// into one with a plain vector store and the host reads it after the work
// has completed without a device copy.  The words are process-wide and
// sticky: rrWaitEvent reports any set word as RR_ERROR_INTERNAL and clears
// it; rrxGetDeviceErrors reads and clears them after waiting for the device.
#pragma once
#include <stdint.h>

namespace rr {

enum DeviceErrorKind : uint32_t {
    kErrStackOverflow = 0,    // a ray needed more than kStackCapacity stack entries
    kErrLookbackTimeout = 1,  // a sort tile's look-back gave up waiting for a predecessor
    kErrBadAccel = 2,         // a query named a buffer without a valid header
    kErrKinds = 3,
};

// The words (kErrKinds of them), or nullptr before InitDeviceErrors has run.
// Kernels receive this pointer as an argument and test it for null.
uint32_t* DeviceErrorWords();

// Allocate the words once per process (called when a context is created).
void InitDeviceErrors();

// Bit k set if word k was set; clears the words that were read as set.
uint32_t TakeDeviceErrors();

}  // namespace rr
