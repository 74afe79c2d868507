// init_kernels.hip -- mtgp32/64_init_state and init_by_array on the GPU, one
// wave per generator, for generator sets too large to initialise on the
// host and upload.  Same windows as the host functions (mtgp32-fast.c:115-
// 149 and 151-235, mtgp64-fast.c:120-154 and 156-240): the window is built
// in LDS, lane 0 runs the (serial) recurrence, all lanes write it out.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "kernels.hpp"

namespace mtgp {
namespace {

#define GMEM __attribute__((address_space(1)))

constexpr int kWave = 64;
constexpr int kMaxWords = 1727;  // largest state (MTGP64-110503)

__device__ __forceinline__ uint32_t fold_byte(uint32_t v) {
    v += v >> 16;
    v += v >> 8;
    return v & 0xff;
}

// mtgp32/64 multipliers of the seed recurrence and the two key mixes.
struct Ops32 {
    typedef uint32_t W;
    static __device__ W step(W prev, int i) { return 1812433253u * (prev ^ (prev >> 30)) + (W)i; }
    static __device__ W mix1(W x) { return (x ^ (x >> 27)) * 1664525u; }
    static __device__ W mix2(W x) { return (x ^ (x >> 27)) * 1566083941u; }
    static __device__ W fill(uint64_t hidden) { return fold_byte((uint32_t)hidden) * 0x01010101u; }
};
struct Ops64 {
    typedef uint64_t W;
    static __device__ W step(W prev, int i) { return 6364136223846793005ull * (prev ^ (prev >> 62)) + (W)i; }
    static __device__ W mix1(W x) { return (x ^ (x >> 59)) * 2173292883993ull; }
    static __device__ W mix2(W x) { return (x ^ (x >> 59)) * 58885565329898161ull; }
    static __device__ W fill(uint64_t hidden) {
        return (uint64_t)fold_byte((uint32_t)(hidden >> 32)) * 0x0101010101010101ull;
    }
};

// The key mixing of init_by_array (two passes over the window with lag and
// mid offsets), run by one lane.
template <class O>
__device__ void mix_key(typename O::W *w, int n, typename O::W hidden, int length, const GMEM uint64_t *key) {
    typedef typename O::W W;
    const int lag = n >= 623 ? 11 : n >= 68 ? 7 : n >= 39 ? 5 : 3;
    const int mid = (n - lag) / 2;
    w[0] = hidden;
    const int rounds = (length + 1 > n ? length + 1 : n) - 1;
    W r = O::mix1(w[0] ^ w[mid] ^ w[n - 1]);
    w[mid] += r;
    r += (W)length;
    w[(mid + lag) % n] += r;
    w[0] = r;
    int i = 1;
    for (int j = 0; j < rounds; j++) {
        const int im = i + mid >= n ? i + mid - n : i + mid;
        const int il = im + lag >= n ? im + lag - n : im + lag;
        const int ip = i == 0 ? n - 1 : i - 1;
        r = O::mix1(w[i] ^ w[im] ^ w[ip]);
        w[im] += r;
        r += (j < length ? (W)key[j] : (W)0) + (W)i;
        w[il] += r;
        w[i] = r;
        i = i + 1 == n ? 0 : i + 1;
    }
    for (int j = 0; j < n; j++) {
        const int im = i + mid >= n ? i + mid - n : i + mid;
        const int il = im + lag >= n ? im + lag - n : im + lag;
        const int ip = i == 0 ? n - 1 : i - 1;
        r = O::mix2(w[i] + w[im] + w[ip]);
        w[im] ^= r;
        r -= (W)i;
        w[il] ^= r;
        w[i] = r;
        i = i + 1 == n ? 0 : i + 1;
    }
    if (w[n - 1] == 0)
        w[n - 1] = 0x4d544750u;
}

template <class O>
__device__ void init_one(typename O::W *w, const InitItem &I) {
    typedef typename O::W W;
    const int l = threadIdx.x, n = I.n;
    const W fill = O::fill(I.hidden);
    for (int j = l; j < n; j += kWave)
        w[j] = fill;
    __syncthreads();
    if (l == 0) {
        if (I.mode == 0) {
            w[0] = (W)I.seed;
            w[1] = (W)I.hidden;
            for (int i = 1; i < n; i++)
                w[i] ^= O::step(w[i - 1], i);
        } else {
            mix_key<O>(w, n, (W)I.hidden, I.key_len, reinterpret_cast<const GMEM uint64_t *>(I.seed));
        }
    }
    __syncthreads();
    GMEM W *out = reinterpret_cast<GMEM W *>(I.state);
    for (int j = l; j < n; j += kWave)
        out[j] = w[j];
}

__global__ __launch_bounds__(64) void init_kernel(const InitItem *__restrict__ items, int n_items) {
    __shared__ uint64_t buf[kMaxWords];
    if ((int)blockIdx.x >= n_items)
        return;
    const InitItem &I = items[blockIdx.x];
    if (I.bits == 32)
        init_one<Ops32>(reinterpret_cast<uint32_t *>(buf), I);
    else
        init_one<Ops64>(buf, I);
}

}  // namespace

hipError_t launch_init(const InitItem *items, int n_items, hipStream_t s) {
    if (n_items <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(init_kernel, dim3(n_items), dim3(kWave), 0, s, items, n_items);
    return hipGetLastError();
}

}  // namespace mtgp
