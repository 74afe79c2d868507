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
// Output stores are device-scope (cache policy sc1): on gfx950 such a store
// writes through the XCD's L2 instead of leaving a dirty line there for a
// later write-back.  The power-bound generation launches ran 3% faster with
// it than with wave-scope stores (policy 0) and than with non-temporal ones.
constexpr int kOutPolicy = 16;
__device__ __forceinline__ void store_u32(store_window_t w, uint32_t off, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, w, (int)off, 0, kOutPolicy);
}
// An 8-byte store (lo at off, hi at off + 4); off and the window's size are
// multiples of 8, so a number is written whole or not at all.
__device__ __forceinline__ void store_u64(store_window_t w, uint32_t off, uint32_t lo, uint32_t hi) {
    typedef uint32_t v2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(v2{lo, hi}, w, (int)off, 0, kOutPolicy);
}

// Matrix-core operands of the fp4 (e2m1) 32 x 32 x 64 product: 32 values
// per lane in dwords 0..3 (value 8v + q in nibble q of dword v; dwords 4..7
// unused), fp32 accumulators.
typedef int mfma_v8i __attribute__((ext_vector_type(8)));
typedef float mfma_v16f __attribute__((ext_vector_type(16)));

// acc + A B with A 32 x 64 and B 64 x 32 in fp4, unit scales: lane L holds
// row L % 32 of A and column L % 32 of B, both for the 32 values of K group
// L / 32 (the same value index means the same k in both); acc[r] of lane L
// is row 8 (r / 4) + 4 (L / 32) + r % 4, column L % 32.
__device__ __forceinline__ mfma_v16f mfma_fp4_32x32x64(mfma_v8i a, mfma_v8i b, mfma_v16f acc) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc, 4, 4, 0, 127, 0, 127);
}

// v of lane (lane ^ D) within each group of 32 lanes (ds_swizzle, bit mode).
template <int D>
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (D << 10) | 0x1f);
}

// Pins a matrix-core accumulator to its AGPRs at this point, so that the
// copies a read-out makes come after it (one accumulator at a time).
__device__ __forceinline__ void keep_in_agprs(mfma_v16f &acc) { asm volatile("" : "+a"(acc)); }

// Bit L = p of lane L.
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

}  // namespace mtgp
