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
inline void store_u64(store_window_t w, uint32_t off, uint32_t lo, uint32_t hi) {
    if (off < w.bytes) {
        memcpy(reinterpret_cast<char *>(w.base) + off, &lo, 4);
        memcpy(reinterpret_cast<char *>(w.base) + off + 4, &hi, 4);
    }
}

typedef int mfma_v8i __attribute__((ext_vector_type(8)));
typedef float mfma_v16f __attribute__((ext_vector_type(16)));

// fp4 (e2m1): sign, two exponent bits (bias 1), one mantissa bit.
inline float emu_fp4(uint32_t code) {
    static const float mag[8] = {0.0f, 0.5f, 1.0f, 1.5f, 2.0f, 3.0f, 4.0f, 6.0f};
    return (code & 8) ? -mag[code & 7] : mag[code & 7];
}

// The 32 x 32 x 64 product over the workgroup's 64 lane threads, with the
// operand and result layout of the matrix-core instruction.
inline mfma_v16f mfma_fp4_32x32x64(mfma_v8i a, mfma_v8i b, mfma_v16f acc) {
    static uint32_t A[64][4], B[64][4];
    const int L = (int)threadIdx.x;
    for (int v = 0; v < 4; v++) {
        A[L][v] = (uint32_t)a[v];
        B[L][v] = (uint32_t)b[v];
    }
    __syncthreads();
    mfma_v16f r = acc;
    for (int i = 0; i < 16; i++) {
        const int m = 8 * (i / 4) + 4 * (L / 32) + i % 4, n = L % 32;
        float s = 0.0f;
        for (int g = 0; g < 2; g++)
            for (int t = 0; t < 32; t++)
                s += emu_fp4(A[m + 32 * g][t / 8] >> (4 * (t % 8))) * emu_fp4(B[n + 32 * g][t / 8] >> (4 * (t % 8)));
        r[i] += s;
    }
    __syncthreads();
    return r;
}

inline void keep_in_agprs(mfma_v16f &) {}

template <int D>
inline uint32_t lane_xor32(uint32_t v) {
    return emu_lane_value<uint32_t>(v, threadIdx.x ^ (uint32_t)D);
}

inline uint64_t wave_ballot(bool p) {
    const uint32_t L = threadIdx.x;
    g_emu_block.slots[L] = p ? 1u : 0u;
    __syncthreads();
    uint64_t m = 0;
    for (uint32_t i = 0; i < blockDim.x && i < 64; i++)
        m |= (uint64_t)(g_emu_block.slots[i] & 1u) << i;
    __syncthreads();
    return m;
}

}  // namespace mtgp
