// api_internal.h — what other translation units of the library need to know
// about a context beyond the public C ABI.
#pragma once
#include <cstdint>
#include <cstring>

#include "../../include/radeonrays.h"

namespace rr {

// The bits of an enum value a caller passed.  A C caller may pass any
// integer where the ABI names an enum; reading such a value through the
// enum type is undefined in C++ when it lies outside the enumerators' range,
// so every check of caller-provided enums reads the representation instead.
template <typename E>
inline uint32_t Raw(const E& e) {
    static_assert(sizeof(E) == sizeof(uint32_t), "C enums are int-sized");
    uint32_t v;
    std::memcpy(&v, &e, sizeof v);
    return v;
}

// HIP device ordinal the context was created on.
int ContextDevice(RRContext context);

// True while `context` is a context the library created and has not destroyed.
bool IsLiveContext(RRContext context);

}  // namespace rr
