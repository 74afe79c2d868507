// gfx950_ops.hpp -- single gfx950 instructions the compiler does not form
// from plain C in the kernels' recursion.
#pragma once
#include <stdint.h>

namespace mtgp {

// a ^ b ^ c in one v_bitop3_b32 (the compiler otherwise keeps two xors
// to reuse a partial result).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

}  // namespace mtgp
