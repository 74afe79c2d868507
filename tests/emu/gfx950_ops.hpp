// Stand-in for mtgp_amd/csrc/gfx950_ops.hpp in the kernel emulation build:
// the same values computed in plain C++.
#pragma once
#include <stdint.h>

namespace mtgp {

inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; }

}  // namespace mtgp
