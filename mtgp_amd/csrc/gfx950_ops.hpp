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

// A range-checked store window of `bytes` bytes at device address `base`
// (a raw buffer descriptor: the record count bounds the byte offset, so a
// lane's store beyond the window writes nothing).  Stores through it take a
// 32-bit per-lane offset and the window's scalar base: no 64-bit address
// arithmetic per step and no lane masks for a partial step.  Word 3 is the
// gfx9-family untyped 32-bit format.
typedef __amdgpu_buffer_rsrc_t store_window_t;
__device__ __forceinline__ store_window_t store_window(uint64_t base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void store_u32(store_window_t w, uint32_t off, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, w, (int)off, 0, 0);
}

}  // namespace mtgp
