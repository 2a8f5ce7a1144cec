// Alignment rule for the 64-bit reduction keys of the material kernels.
// They are updated with 64-bit atomicMax, and a 64-bit atomic on an address
// that is not a multiple of 8 bytes faults the device.  Host-only and
// dependency-free, so a CPU test can compile it on its own.
#pragma once

#include <cstdint>

inline bool fd23_keys_aligned(const void *keys) {
    return keys != nullptr && (reinterpret_cast<std::uintptr_t>(keys) & 7u) == 0;
}
