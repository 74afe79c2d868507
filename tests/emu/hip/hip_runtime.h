// CPU stand-in for <hip/hip_runtime.h> used by the kernel emulation tests:
// a workgroup runs as blockDim.x host threads, __syncthreads() is a real
// barrier, cross-lane builtins go through a shared slot array, and
// workgroups of a launch run one after another.  Shared memory is a
// function-local static (one instance, shared by the workgroup's threads).
// Test infrastructure only: it is never part of libmtgp_amd.so.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <barrier>
#include <thread>
#include <vector>

#include "hip_runtime_api.h"

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __shared__ static
#define __launch_bounds__(...)

struct emu_idx {
    uint32_t x = 0, y = 0, z = 0;
};
inline thread_local emu_idx threadIdx, blockIdx;
inline emu_idx gridDim, blockDim;

struct emu_block {
    std::barrier<> *bar = nullptr;
    uint64_t slots[1024];
};
inline emu_block g_emu_block;

inline void __syncthreads() { g_emu_block.bar->arrive_and_wait(); }

template <class T>
inline T emu_lane_value(T v, uint32_t src) {
    g_emu_block.slots[threadIdx.x] = (uint64_t)v;
    __syncthreads();
    T r = (T)g_emu_block.slots[src];
    __syncthreads();
    return r;
}
#define __builtin_amdgcn_readfirstlane(v) emu_lane_value<uint32_t>((uint32_t)(v), 0u)
#define __builtin_amdgcn_readlane(v, src) emu_lane_value<uint32_t>((uint32_t)(v), (uint32_t)(src))
template <class T>
inline T __shfl(T v, int src) { return emu_lane_value<T>(v, (uint32_t)src); }

inline uint32_t atomicAdd(uint32_t *p, uint32_t v) { return std::atomic_ref<uint32_t>(*p).fetch_add(v); }
inline uint32_t atomicExch(uint32_t *p, uint32_t v) { return std::atomic_ref<uint32_t>(*p).exchange(v); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }

inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline double __hiloint2double(int hi, int lo) {
    uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    double d;
    memcpy(&d, &u, 8);
    return d;
}
inline int __double2hiint(double d) { uint64_t u; memcpy(&u, &d, 8); return (int)(uint32_t)(u >> 32); }
inline int __double2loint(double d) { uint64_t u; memcpy(&u, &d, 8); return (int)(uint32_t)u; }

struct uint4 { uint32_t x, y, z, w; };
struct uint2 { uint32_t x, y; };
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }

template <class... P, class... A>
inline void emu_launch(void (*kernel)(P...), dim3 grid, dim3 block, size_t, hipStream_t, A... args) {
    gridDim.x = grid.x;
    blockDim.x = block.x;
    for (uint32_t b = 0; b < grid.x; b++) {
        std::barrier<> bar((ptrdiff_t)block.x);
        g_emu_block.bar = &bar;
        std::vector<std::thread> lanes;
        for (uint32_t t = 0; t < block.x; t++)
            lanes.emplace_back([=] {
                threadIdx.x = t;
                blockIdx.x = b;
                kernel(args...);
            });
        for (auto &th : lanes)
            th.join();
    }
}
#define hipLaunchKernelGGL emu_launch
