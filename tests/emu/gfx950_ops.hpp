// Stand-in for mtgp_amd/csrc/gfx950_ops.hpp in the kernel emulation build:
// the same values computed in plain C++.
#pragma once
#include <stdint.h>
#include <string.h>

namespace mtgp {

inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; }

struct store_window_t {
    uint64_t base;
    uint32_t bytes;
};
inline store_window_t store_window(uint64_t base, uint32_t bytes) { return store_window_t{base, bytes}; }
inline void store_u32(store_window_t w, uint32_t off, uint32_t v) {
    if (off < w.bytes)
        memcpy(reinterpret_cast<char *>(w.base) + off, &v, 4);
}

}  // namespace mtgp
