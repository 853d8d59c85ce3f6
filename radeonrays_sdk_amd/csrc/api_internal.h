// api_internal.h — what other translation units of the library need to know
// about a context beyond the public C ABI.
#pragma once
#include "../../include/radeonrays.h"

namespace rr {

// HIP device ordinal the context was created on.
int ContextDevice(RRContext context);

// True while `context` is a context the library created and has not destroyed.
bool IsLiveContext(RRContext context);

}  // namespace rr
