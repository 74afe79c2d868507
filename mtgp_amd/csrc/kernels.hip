// kernels.hip -- MTGP32 / MTGP64 generation and jump-ahead kernels (gfx950).
//
// Replaces the device samples (cuda-sample/mtgp32-cuda.cu:171-297,
// mtgp64-cuda.cu:208-392, openCL-sample/mtgp32.cl:167-399, and the device
// jump of openCL-sample/mtgp32-jump.cl:630-731) with one persistent kernel
// per word size that works through a list of Items (common.hpp).
//
// Design (one workgroup = one wave64 = one Item at a time; the two-wave
// cells of layout 2 are the exception, see mtgp32_items_kernel):
//  * The state window of N words lives in a linear LDS buffer.  A step
//    appends 256 words: lane l computes words 4l..4l+3, which are mutually
//    independent because every parameter set has N - pos >= 256.  The step
//    reads x1/x2 as one aligned 128-bit LDS load plus one word, y/t as five
//    words, looks the 4-bit indices up in 16-entry tables held in LDS, and
//    writes its outputs as one 128-bit global store per lane.
//  * Instead of a masked ring, the buffer is linear and the live window is
//    copied back to the start every K steps, so every LDS address of a step
//    is a per-lane base plus an immediate offset.
//  * Jump-ahead (optional per Item): with c the jump polynomial, the new
//    window is new[j] = XOR_{k : c_k = 1} w[k + j] over the generator's own
//    sequence w.  The kernel streams w through the same LDS buffer K steps
//    at a time and XOR-accumulates into registers (lane l owns j = l + 64m),
//    scanning c's set bits with scalar instructions.  This is the same
//    state as the reference's mtgp32_fast_jump (mtgp32-fast-jump.c:51-79).
//  * Work distribution: workgroup g starts with Item g, then takes further
//    Items from an atomic counter; the last workgroup to finish resets the
//    counter pair, so launches need no memset and are graph-capturable.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.hpp"
#include "kernels.hpp"
// Angle brackets: the kernel emulation build substitutes its own copy.
#include <gfx950_ops.hpp>

namespace mtgp {
namespace {

// Device-memory pointers carry address space 1 so that accesses through
// addresses taken from Items compile to global_* (not flat_*) instructions.
#define GMEM __attribute__((address_space(1)))
typedef GMEM uint32_t gu32;
typedef GMEM uint64_t gu64;
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef GMEM v4u gu4;
typedef GMEM v2u gu2;

constexpr int kWave = 64;
constexpr int kStep = 256;  // words per wave step (4 per lane)

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Orders the workgroup's LDS accesses across steps (a barrier of one wave,
// or of the two waves of a layout-2 cell).
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// The 16-entry tables (recursion tbl, tempering tmp, float tempering flt)
// are held in LDS, contiguous.  (Replicating each to 64 entries would make
// the lookup offset one SDWA shift, but entries 32 apart then share an LDS
// bank: measured 2-way conflicts and a slower step.)
constexpr int kTabRep = 16;
constexpr int kTabWords = 3 * kTabRep;

__device__ __forceinline__ uint32_t lookup(const uint32_t *t, uint32_t v) { return t[v & 15]; }

// Fills tab[kTabWords] from a parameter record: tbl, tmp and flt, each
// read through its own array (one table per group of 16 lanes).
__device__ __forceinline__ void fill_tables(uint32_t *tab, const DevParam &D) {
    const int l = threadIdx.x;
    if (l < kTabRep)
        tab[l] = D.tbl[l];
    else if (l < 2 * kTabRep)
        tab[l] = D.tmp[l - kTabRep];
    else if (l < kTabWords)
        tab[l] = D.flt[l - 2 * kTabRep];
}

// t ^= t >> 16; t ^= t >> 8 (mtgp32-fast.h:196-197); the low four bits
// index the tempering tables (the mask of the second shift keeps those bits
// and lets the compiler use a byte-select operand).
__device__ __forceinline__ uint32_t temper_mix(uint32_t t) {
    t ^= t >> 16;
    return t ^ ((t >> 8) & 0xffu);
}

// ----------------------------------------------------------------- MTGP32

struct P32 {
    int n, pos, sh1, sh2;
    uint32_t mask;
    int pm1;  // pos - 1: the tempering word's offset, kept in a scalar register
};

// mtgp32-fast.h:149-159
__device__ __forceinline__ uint32_t rec32(P32 p, const uint32_t *tbl, uint32_t x1, uint32_t x2,
                                          uint32_t y) {
    const uint32_t x = (x1 & p.mask) ^ x2;
    y = xor3(x, x << p.sh1, y >> p.sh2);
    return y ^ lookup(tbl, y);
}

// Output word for each kind (mtgp32-fast.h:192-346).  All float operations
// are exact: x - 1 and 2 - x for x in [1, 2) are representable.
template <int KIND>
__device__ __forceinline__ uint32_t out32(const uint32_t *tab, uint32_t r, uint32_t t) {
    const uint32_t ti = temper_mix(t);
    if (KIND == 0)
        return r ^ lookup(tab + kTabRep, ti);
    const uint32_t u = (r >> 9) ^ lookup(tab + 2 * kTabRep, ti);
    if (KIND == 1)
        return u;
    if (KIND == 2)
        return __float_as_uint(__uint_as_float(u) - 1.0f);
    if (KIND == 3)
        return __float_as_uint(2.0f - __uint_as_float(u));
    return __float_as_uint(__uint_as_float(u | 1u) - 1.0f);
}

// One 256-word step from window base b (b % 4 == 0).  With OUT, writes the
// m (<= 256) leading outputs to dst (16-byte aligned).
template <int KIND, bool OUT, bool FULL = false>
__device__ __forceinline__ void step32(uint32_t *buf, const uint32_t *tab, P32 p, int b,
                                       gu32 *dst, uint32_t m) {
    const int l = threadIdx.x;
    const uint32_t *xp = buf + b + 4 * l;
    const uint4 xa = *reinterpret_cast<const uint4 *>(xp);
    const uint32_t x4 = xp[4];
    const uint32_t *yp = xp + p.pos - 1;
    const uint32_t t0 = yp[0], t1 = yp[1], t2 = yp[2], t3 = yp[3], t4 = yp[4];
    const uint32_t r0 = rec32(p, tab, xa.x, xa.y, t1);
    const uint32_t r1 = rec32(p, tab, xa.y, xa.z, t2);
    const uint32_t r2 = rec32(p, tab, xa.z, xa.w, t3);
    const uint32_t r3 = rec32(p, tab, xa.w, x4, t4);
    uint32_t *wp = buf + b + p.n + 4 * l;
    wp[0] = r0;
    wp[1] = r1;
    wp[2] = r2;
    wp[3] = r3;
    if (OUT) {
        uint4 o;
        o.x = out32<KIND>(tab, r0, t0);
        o.y = out32<KIND>(tab, r1, t1);
        o.z = out32<KIND>(tab, r2, t2);
        o.w = out32<KIND>(tab, r3, t3);
        const uint32_t q = 4 * l;
        if (FULL || q + 4 <= m) {
            *reinterpret_cast<gu4 *>(dst + q) = v4u{o.x, o.y, o.z, o.w};
        } else if (q < m) {
            dst[q] = o.x;
            if (q + 1 < m)
                dst[q + 1] = o.y;
            if (q + 2 < m)
                dst[q + 2] = o.z;
        }
    }
}

// Lane-strided variant of step32: lane l computes words l + 64i (i < 4),
// so every LDS access of the step is lane-consecutive (no bank conflicts)
// and every store instruction writes 256 contiguous bytes.  No alignment
// requirement on b or dst.  With W waves per workgroup and WPL words per
// lane the step is 64*W*WPL words (thread l computes l + 64Wi, i < WPL),
// which needs N - pos >= 64*W*WPL.
template <int KIND, bool OUT, bool FULL = false, int W = 1, int WPL = 4>
__device__ __forceinline__ void step32s(uint32_t *buf, const uint32_t *tab, P32 p, int b,
                                        gu32 *dst, uint32_t m) {
    constexpr int T = kWave * W;
    const int l = threadIdx.x;
    const uint32_t *xp = buf + b + l;
    const uint32_t *yp = xp + p.pm1;
    uint32_t r[WPL], t[WPL];
#pragma unroll
    for (int i = 0; i < WPL; i++) {
        t[i] = yp[T * i];
        r[i] = rec32(p, tab, xp[T * i], xp[T * i + 1], yp[T * i + 1]);
    }
    uint32_t *wp = buf + b + p.n + l;
#pragma unroll
    for (int i = 0; i < WPL; i++)
        wp[T * i] = r[i];
    if (OUT) {
        uint32_t o[WPL];
#pragma unroll
        for (int i = 0; i < WPL; i++)
            o[i] = out32<KIND>(tab, r[i], t[i]);
        // the m leading outputs: lanes past them fall outside the window
        const store_window_t w = store_window((uint64_t)(uintptr_t)dst, FULL ? 4u * T * WPL : 4u * m);
#pragma unroll
        for (int i = 0; i < WPL; i++)
            store_u32(w, 4u * (uint32_t)(l + T * i), o[i]);
    }
}

// Copies the window [b, b + n) to [to, to + n) through registers (the
// ranges may overlap); NJ words per thread of a W-wave workgroup.
template <int NJ, int W = 1>
__device__ __forceinline__ void relocate32(uint32_t *buf, int b, int n, int to = 0) {
    constexpr int T = kWave * W;
    const int l = threadIdx.x;
    uint32_t v[NJ];
#pragma unroll
    for (int m = 0; m < NJ; m++)
        v[m] = (l + T * m < n) ? buf[b + l + T * m] : 0u;
    wave_sync();
#pragma unroll
    for (int m = 0; m < NJ; m++)
        if (l + T * m < n)
            buf[to + l + T * m] = v[m];
    wave_sync();
}

// Folds the chunk of positions [a, a + 256*steps) (clipped to `total`) of
// the jump polynomial into acc: acc[m] ^= w[k + l + 64m] for every set
// coefficient k, with buf[i] = w[a + i].  Positions are visited by residue
// rho = k mod 64: for a fixed rho every set coefficient k = a + 64q + rho
// needs the same lane column buf[rho + l + 64r] (r = q + m), which each lane
// loads once into registers; the XORs then read registers only.  This
// halves the LDS traffic of visiting the coefficients one by one.
template <int NJ, int K>
__device__ __forceinline__ void fold32(const uint32_t *buf, const gu32 *__restrict__ poly, int a, int total,
                                       int steps, uint32_t (&acc)[NJ]) {
    constexpr int QMAX = 4 * K;  // 64-position groups per chunk
    const int l = threadIdx.x;
    const int end = a + steps * kStep < total ? a + steps * kStep : total;
    const int nw = (end - a) / 32;  // coefficient words in this chunk
    uint32_t cw[2 * QMAX];
#pragma unroll
    for (int i = 0; i < 2 * QMAX; i++)
        cw[i] = i < nw ? uniform(poly[a / 32 + i]) : 0u;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        for (int r0 = 0; r0 < 32; r0++) {
            const int rho = 32 * half + r0;
            uint32_t any = 0;
#pragma unroll
            for (int q = 0; q < QMAX; q++)
                any |= cw[2 * q + half] >> r0 & 1u;
            if (!any)
                continue;
            uint32_t R[QMAX + NJ];
            const uint32_t *col = buf + rho + l;
#pragma unroll
            for (int r = 0; r < QMAX + NJ; r++)
                R[r] = col[kWave * r];
#pragma unroll
            for (int q = 0; q < QMAX; q++) {
                if (cw[2 * q + half] >> r0 & 1u) {
#pragma unroll
                    for (int m = 0; m < NJ; m++)
                        acc[m] ^= R[q + m];
                }
            }
        }
    }
}

// Window after jumping: acc[m] = new[l + 64m] (see file header).
template <int NJ, int K, int NBUF>
__device__ void jump32(uint32_t *buf, const uint32_t *tab, P32 p, const gu32 *__restrict__ in,
                       const gu32 *__restrict__ poly, int poly_words, uint32_t (&acc)[NJ]) {
    const int l = threadIdx.x;
#pragma unroll
    for (int m = 0; m < NJ; m++)
        acc[m] = 0;
    for (int j = l; j < p.n; j += kWave)
        buf[j] = in[j];
    wave_sync();
    const int total = poly_words * 32;
    for (int a = 0; a < total;) {
        const int left = total - a;
        const int steps = left >= K * kStep ? K : (left + kStep - 1) / kStep;
        for (int s = 0; s < steps; s++) {
            step32<0, false>(buf, tab, p, s * kStep, nullptr, 0);
            wave_sync();
        }
        // buf now holds w[a .. a + n + 256*steps); fold positions
        // [a, a + 256*steps) with set coefficients into acc.
        fold32<NJ, K>(buf, poly, a, total, steps, acc);
        a += steps * kStep;
        if (a < total)
            relocate32<NJ>(buf, steps * kStep, p.n);
    }
    wave_sync();
}

// Numbers before the first aligned store of an output at `out`: to 16 bytes
// (layout 0) or to a 128-byte line (layout 1).
template <int LAYOUT>
__device__ __forceinline__ uint32_t head32(uint64_t out, uint64_t count) {
    const uint32_t h = LAYOUT == 0 ? ((16u - (uint32_t)(out & 15)) & 15) >> 2
                                   : ((128u - (uint32_t)(out & 127)) & 127) >> 2;
    return (uint64_t)h > count ? (uint32_t)count : h;
}

// Window base for a segment with head h: layout 0 needs (b + h) % 4 == 0;
// both need room for one more step.  Moves the window when either fails.
template <int NJ, int NBUF, int LAYOUT, int W = 1, int WPL = 4>
__device__ __forceinline__ int place32(uint32_t *buf, P32 p, int b, uint32_t h) {
    const int to = LAYOUT == 0 ? (int)((4u - h) & 3u) : 0;
    const bool misaligned = LAYOUT == 0 && (((uint32_t)b + h) & 3u) != 0;
    if (misaligned || b + p.n + kWave * W * WPL > NBUF) {
        relocate32<NJ, W>(buf, b, p.n, to);
        return to;
    }
    return b;
}

// Steps the window over `count` numbers without output; returns the new base.
template <int NJ, int NBUF, int W = 1, int WPL = 4>
__device__ __forceinline__ int advance32(uint32_t *buf, const uint32_t *tab, P32 p, int b, uint64_t count) {
    constexpr int S = kWave * W * WPL;
    for (uint64_t s = count / S; s > 0; s--) {
        if (b + p.n + S > NBUF) {
            relocate32<NJ, W>(buf, b, p.n);
            b = 0;
        }
        step32s<0, false, false, W, WPL>(buf, tab, p, b, nullptr, 0);
        wave_sync();
        b += S;
    }
    const uint32_t m = (uint32_t)(count % S);
    if (m) {
        if (b + p.n + S > NBUF) {
            relocate32<NJ, W>(buf, b, p.n);
            b = 0;
        }
        step32s<0, false, false, W, WPL>(buf, tab, p, b, nullptr, 0);
        wave_sync();
        b += (int)m;
    }
    return b;
}

// Writes `count` numbers of KIND to out from the window at base b (placed
// by place32 for head h); returns the window base after the last number.
template <int KIND, int NJ, int NBUF, int LAYOUT, int W = 1, int WPL = 4>
__device__ __forceinline__ int generate32(uint32_t *buf, const uint32_t *tab, P32 p, int b, uint32_t h,
                          gu32 *__restrict__ out, uint64_t count) {
    static_assert(LAYOUT == 1 || (W == 1 && WPL == 4), "multi-wave cells are lane-strided");
    constexpr int S = kWave * W * WPL;
    const int l = threadIdx.x;
    gu32 *dst = out;
    uint64_t rest = count;
    if (h > 0) {
        if (LAYOUT == 0) {
            // h < 4 words, one per lane, so that the main loop's stores are
            // 16-byte aligned (the window base becomes b + h == 0 mod 4).
            if ((uint32_t)l < h) {
                const int q = b + l;
                const uint32_t r = rec32(p, tab, buf[q], buf[q + 1], buf[q + p.pos]);
                buf[q + p.n] = r;
                out[l] = out32<KIND>(tab, r, buf[q + p.pos - 1]);
            }
        } else {
            // h < 32 numbers so that every later 256-byte store starts on a
            // 128-byte line (no partially written lines).
            step32s<KIND, true, false, W, WPL>(buf, tab, p, b, dst, h);
        }
        wave_sync();
        b += (int)h;
        dst += h;
        rest -= (uint64_t)h;
    }
    // Whole steps, then one partial step.  (Written without a select on the
    // 64-bit remainder.)
    for (uint64_t s = rest / S; s > 0; s--) {
        if (b + p.n + S > NBUF) {
            relocate32<NJ, W>(buf, b, p.n);
            b = 0;
        }
        if (LAYOUT == 0)
            step32<KIND, true, true>(buf, tab, p, b, dst, kStep);
        else
            step32s<KIND, true, true, W, WPL>(buf, tab, p, b, dst, S);
        wave_sync();
        b += S;
        dst += S;
    }
    const uint32_t m = (uint32_t)(rest % S);
    if (m) {
        if (b + p.n + S > NBUF) {
            relocate32<NJ, W>(buf, b, p.n);
            b = 0;
        }
        if (LAYOUT == 0)
            step32<KIND, true>(buf, tab, p, b, dst, m);
        else
            step32s<KIND, true, false, W, WPL>(buf, tab, p, b, dst, m);
        wave_sync();
        b += (int)m;
    }
    return b;
}

// One segment (Segment, common.hpp) from window base b; returns the new base.
template <int NJ, int NBUF, int LAYOUT, int W = 1, int WPL = 4>
__device__ __forceinline__ int segment32_body(uint32_t *buf, const uint32_t *tab, P32 p, int b, uint64_t out,
                                              uint64_t count, int kind) {
    if (count == 0)
        return b;
    if (out == 0)
        return advance32<NJ, NBUF, W, WPL>(buf, tab, p, b, count);
    const uint32_t h = head32<LAYOUT>(out, count);
    b = place32<NJ, NBUF, LAYOUT, W, WPL>(buf, p, b, h);
    gu32 *dst = reinterpret_cast<gu32 *>(out);
    switch (kind) {
    case 0: return generate32<0, NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, h, dst, count);
    case 1: return generate32<1, NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, h, dst, count);
    case 2: return generate32<2, NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, h, dst, count);
    case 3: return generate32<3, NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, h, dst, count);
    default: return generate32<4, NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, h, dst, count);
    }
}

// The small state classes (NJ <= 6) inline the segment into the kernel, so
// that LDS offsets fold into the instructions; the large ones call an
// out-of-line copy (fully inlined, their kernels spill registers).
template <int NJ, int NBUF, int LAYOUT, int W, int WPL>
__device__ __noinline__ int segment32_call(uint32_t *buf, const uint32_t *tab, P32 p, int b, uint64_t out,
                                           uint64_t count, int kind) {
    return segment32_body<NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, out, count, kind);
}
template <int NJ, int NBUF, int LAYOUT, int W = 1, int WPL = 4>
__device__ __forceinline__ int segment32(uint32_t *buf, const uint32_t *tab, P32 p, int b, uint64_t out,
                                         uint64_t count, int kind) {
    if constexpr (NJ <= 6)
        return segment32_body<NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, out, count, kind);
    else
        return segment32_call<NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, out, count, kind);
}

// W waves per workgroup, WPL words per lane and step (W > 1: lane-strided
// only, every Item's set has N - pos >= 64*W*WPL, and no Item carries a
// fused jump polynomial; the planner guarantees all three).  The fused
// jump steps 256 words at a time, so single-wave kernels keep WPL = 4.
template <int NMAX, int K, int LAYOUT, int W = 1, int WPL = 4>
__global__ __launch_bounds__(64 * W) void mtgp32_items_kernel(const Item *__restrict__ items, int n_items,
                                                              const DevParam *__restrict__ params,
                                                              uint32_t *__restrict__ queue) {
    static_assert(WPL == 4 || (W > 1 && WPL == 2), "words per lane");
    constexpr int T = kWave * W;
    constexpr int NJ = (NMAX + T - 1) / T;
    constexpr int NBUF = K * T * WPL + T * NJ + T;
    __shared__ __attribute__((aligned(16))) uint32_t buf[NBUF];
    __shared__ uint32_t tab[kTabWords];
    __shared__ uint32_t next_item;
    const int l = threadIdx.x;
    uint32_t it = blockIdx.x;
    while (it < (uint32_t)n_items) {
        const Item &I = items[it];
        const DevParam &D = params[I.param];
        P32 p;
        p.n = D.n;
        p.pos = D.pos;
        p.pm1 = D.pos - 1;
        p.sh1 = D.sh1;
        p.sh2 = D.sh2;
        p.mask = D.mask_lo;
        fill_tables(tab, D);
        gu32 *state_out = reinterpret_cast<gu32 *>(I.state_out);
        const gu32 *state_in = reinterpret_cast<const gu32 *>(I.state_in);
        wave_sync();
        bool fused = false;
        if constexpr (W == 1) {
            if (I.poly) {
                fused = true;
                uint32_t acc[NJ];
                jump32<NJ, K, NBUF>(buf, tab, p, state_in, reinterpret_cast<const gu32 *>(I.poly),
                                    I.poly_words, acc);
#pragma unroll
                for (int m = 0; m < NJ; m++)
                    if (l + kWave * m < p.n)
                        buf[l + kWave * m] = acc[m];
            }
        }
        if (!fused) {
            const int parts = I.parts > 1 ? I.parts : 1;
            const uint64_t pstride = I.part_bytes / 4;
            for (int j = l; j < p.n; j += T) {
                uint32_t v = state_in[j];
                for (int q = 1; q < parts; q++)
                    v ^= state_in[(uint64_t)q * pstride + (uint64_t)j];
                buf[j] = v;
            }
        }
        wave_sync();
        int b = 0;
        if (I.nsegs > 0) {
            const GMEM Segment *segs = reinterpret_cast<const GMEM Segment *>(I.segs);
            for (int si = 0; si < I.nsegs; si++)
                b = segment32<NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, segs[si].out, segs[si].count,
                                                        segs[si].kind);
        } else {
            b = segment32<NJ, NBUF, LAYOUT, W, WPL>(buf, tab, p, b, I.out, I.count, I.kind);
        }
        if (state_out)
            for (int j = l; j < p.n; j += T)
                state_out[j] = buf[b + j];
        wave_sync();
        if constexpr (W == 1) {
            uint32_t next = 0;
            if (l == 0)
                next = atomicAdd(&queue[0], 1u);
            it = gridDim.x + uniform(__shfl(next, 0));
        } else {
            // One claim per workgroup, handed to every wave through LDS.
            if (l == 0)
                next_item = atomicAdd(&queue[0], 1u);
            wave_sync();
            it = gridDim.x + uniform(next_item);
            wave_sync();
        }
    }
    if (l == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// ----------------------------------------------------------------- MTGP64
// Words are kept as separate high/low 32-bit planes in LDS.

struct P64 {
    int n, pos, sh1, sh2;
    uint32_t mask_hi, mask_lo;
};

// mtgp64-fast.h:148-169: (rh, rl) = rec(x1, x2, y); only the high half of
// a table entry is non-zero.
__device__ __forceinline__ void rec64(P64 p, const uint32_t *tbl, uint32_t x1h, uint32_t x1l,
                                      uint32_t x2h, uint32_t x2l, uint32_t yh, uint32_t yl, uint32_t &rh,
                                      uint32_t &rl) {
    const uint32_t xh = (x1h & p.mask_hi) ^ x2h;
    const uint32_t xl = (x1l & p.mask_lo) ^ x2l;
    // halves shifted separately, then crossed (xh ^= xh << sh1, ...)
    const uint32_t nh = xor3(xl, xl << p.sh1, yh >> p.sh2);
    const uint32_t nl = xor3(xh, xh << p.sh1, yl >> p.sh2);
    rh = nh ^ lookup(tbl, nl);
    rl = nl;
}

// mtgp64-fast.h:202-356.  The double tables' low halves are zero.
template <int KIND>
__device__ __forceinline__ void out64(const uint32_t *tab, uint32_t rh, uint32_t rl, uint32_t tl, uint32_t &oh,
                                      uint32_t &ol) {
    const uint32_t ti = temper_mix(tl);
    if (KIND == 0) {
        oh = rh ^ lookup(tab + kTabRep, ti);
        ol = rl;
        return;
    }
    uint32_t lo = (rl >> 12) | (rh << 20);
    const uint32_t hi = (rh >> 12) ^ lookup(tab + 2 * kTabRep, ti);
    if (KIND == 4)
        lo |= 1u;
    if (KIND == 1) {
        oh = hi;
        ol = lo;
        return;
    }
    const double d = __hiloint2double((int)hi, (int)lo);
    double v;
    if (KIND == 3)
        v = 2.0 - d;
    else
        v = d - 1.0;
    oh = (uint32_t)__double2hiint(v);
    ol = (uint32_t)__double2loint(v);
}

template <int KIND, bool OUT, bool FULL = false>
__device__ __forceinline__ void step64(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b,
                                       gu32 *dst, uint32_t m) {
    const int l = threadIdx.x;
    const int q = b + 4 * l;
    const uint4 xh = *reinterpret_cast<const uint4 *>(bh + q);
    const uint4 xl = *reinterpret_cast<const uint4 *>(bl + q);
    const uint32_t x4h = bh[q + 4], x4l = bl[q + 4];
    const int yq = q + p.pos - 1;
    const uint32_t th0 = bh[yq], th1 = bh[yq + 1], th2 = bh[yq + 2], th3 = bh[yq + 3], th4 = bh[yq + 4];
    const uint32_t tl0 = bl[yq], tl1 = bl[yq + 1], tl2 = bl[yq + 2], tl3 = bl[yq + 3], tl4 = bl[yq + 4];
    uint32_t rh0, rl0, rh1, rl1, rh2, rl2, rh3, rl3;
    rec64(p, tab, xh.x, xl.x, xh.y, xl.y, th1, tl1, rh0, rl0);
    rec64(p, tab, xh.y, xl.y, xh.z, xl.z, th2, tl2, rh1, rl1);
    rec64(p, tab, xh.z, xl.z, xh.w, xl.w, th3, tl3, rh2, rl2);
    rec64(p, tab, xh.w, xl.w, x4h, x4l, th4, tl4, rh3, rl3);
    (void)th0;
    const int w = q + p.n;
    bh[w] = rh0;
    bh[w + 1] = rh1;
    bh[w + 2] = rh2;
    bh[w + 3] = rh3;
    bl[w] = rl0;
    bl[w + 1] = rl1;
    bl[w + 2] = rl2;
    bl[w + 3] = rl3;
    if (OUT) {
        uint4 o0, o1;
        out64<KIND>(tab, rh0, rl0, tl0, o0.y, o0.x);
        out64<KIND>(tab, rh1, rl1, tl1, o0.w, o0.z);
        out64<KIND>(tab, rh2, rl2, tl2, o1.y, o1.x);
        out64<KIND>(tab, rh3, rl3, tl3, o1.w, o1.z);
        const uint32_t e = 4 * l;  // first output element of this lane
        gu4 *d4 = reinterpret_cast<gu4 *>(dst + 2 * e);
        if (FULL || e + 4 <= m) {
            d4[0] = v4u{o0.x, o0.y, o0.z, o0.w};
            d4[1] = v4u{o1.x, o1.y, o1.z, o1.w};
        } else if (e < m) {
            gu2 *d2 = reinterpret_cast<gu2 *>(dst + 2 * e);
            d2[0] = v2u{o0.x, o0.y};
            if (e + 1 < m)
                d2[1] = v2u{o0.z, o0.w};
            if (e + 2 < m)
                d2[2] = v2u{o1.x, o1.y};
        }
    }
}

// Lane-strided variant of step64 (see step32s); outputs are 8-byte stores,
// 512 contiguous bytes per instruction.  With WV waves per workgroup the
// step is 64 * WV * WPL words (thread l computes l + 64WVi).
template <int KIND, bool OUT, bool FULL = false, int WPL = 4, int WV = 1>
__device__ __forceinline__ void step64s(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b,
                                        gu32 *dst, uint32_t m) {
    constexpr int T = kWave * WV;
    const int l = threadIdx.x;
    const int q = b + l;
    const int yq = q + p.pos - 1;
    uint32_t rh[WPL], rl[WPL], tl[WPL];
#pragma unroll
    for (int i = 0; i < WPL; i++) {
        const int x = q + T * i, y = yq + T * i;
        tl[i] = bl[y];
        rec64(p, tab, bh[x], bl[x], bh[x + 1], bl[x + 1], bh[y + 1], bl[y + 1], rh[i], rl[i]);
    }
    const int w = q + p.n;
#pragma unroll
    for (int i = 0; i < WPL; i++) {
        bh[w + T * i] = rh[i];
        bl[w + T * i] = rl[i];
    }
    if (OUT) {
        uint32_t oh[WPL], ol[WPL];
#pragma unroll
        for (int i = 0; i < WPL; i++)
            out64<KIND>(tab, rh[i], rl[i], tl[i], oh[i], ol[i]);
        // the m leading outputs: lanes past them fall outside the window
        const store_window_t sw = store_window((uint64_t)(uintptr_t)dst, FULL ? 8u * T * WPL : 8u * m);
#pragma unroll
        for (int i = 0; i < WPL; i++)
            store_u64(sw, 8u * (uint32_t)(l + T * i), ol[i], oh[i]);
    }
}

template <int NJ, int WV = 1>
__device__ __forceinline__ void relocate64(uint32_t *bh, uint32_t *bl, int b, int n, int to = 0) {
    constexpr int T = kWave * WV;
    const int l = threadIdx.x;
    uint32_t vh[NJ], vl[NJ];
#pragma unroll
    for (int m = 0; m < NJ; m++) {
        const bool ok = l + T * m < n;
        vh[m] = ok ? bh[b + l + T * m] : 0u;
        vl[m] = ok ? bl[b + l + T * m] : 0u;
    }
    wave_sync();
#pragma unroll
    for (int m = 0; m < NJ; m++)
        if (l + T * m < n) {
            bh[to + l + T * m] = vh[m];
            bl[to + l + T * m] = vl[m];
        }
    wave_sync();
}

// fold32 for the two 32-bit planes of MTGP64 words.
template <int NJ, int K>
__device__ __forceinline__ void fold64(const uint32_t *bh, const uint32_t *bl, const gu32 *__restrict__ poly, int a,
                                       int total, int steps, uint32_t (&ah)[NJ], uint32_t (&al)[NJ]) {
    constexpr int QMAX = 4 * K;
    const int l = threadIdx.x;
    const int end = a + steps * kStep < total ? a + steps * kStep : total;
    const int nw = (end - a) / 32;  // coefficient words in this chunk
    uint32_t cw[2 * QMAX];
#pragma unroll
    for (int i = 0; i < 2 * QMAX; i++)
        cw[i] = i < nw ? uniform(poly[a / 32 + i]) : 0u;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        for (int r0 = 0; r0 < 32; r0++) {
            const int rho = 32 * half + r0;
            uint32_t any = 0;
#pragma unroll
            for (int q = 0; q < QMAX; q++)
                any |= cw[2 * q + half] >> r0 & 1u;
            if (!any)
                continue;
            uint32_t RH[QMAX + NJ], RL[QMAX + NJ];
#pragma unroll
            for (int r = 0; r < QMAX + NJ; r++) {
                RH[r] = bh[rho + l + kWave * r];
                RL[r] = bl[rho + l + kWave * r];
            }
#pragma unroll
            for (int q = 0; q < QMAX; q++) {
                if (cw[2 * q + half] >> r0 & 1u) {
#pragma unroll
                    for (int m = 0; m < NJ; m++) {
                        ah[m] ^= RH[q + m];
                        al[m] ^= RL[q + m];
                    }
                }
            }
        }
    }
}

template <int NJ, int K, int NBUF>
__device__ void jump64(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p,
                       const gu64 *__restrict__ in, const gu32 *__restrict__ poly, int poly_words,
                       uint32_t (&ah)[NJ], uint32_t (&al)[NJ]) {
    const int l = threadIdx.x;
#pragma unroll
    for (int m = 0; m < NJ; m++)
        ah[m] = al[m] = 0;
    for (int j = l; j < p.n; j += kWave) {
        const uint64_t v = in[j];
        bh[j] = (uint32_t)(v >> 32);
        bl[j] = (uint32_t)v;
    }
    wave_sync();
    const int total = poly_words * 32;
    for (int a = 0; a < total;) {
        const int left = total - a;
        const int steps = left >= K * kStep ? K : (left + kStep - 1) / kStep;
        for (int s = 0; s < steps; s++) {
            step64<0, false>(bh, bl, tab, p, s * kStep, nullptr, 0);
            wave_sync();
        }
        fold64<NJ, K>(bh, bl, poly, a, total, steps, ah, al);
        a += steps * kStep;
        if (a < total)
            relocate64<NJ>(bh, bl, steps * kStep, p.n);
    }
    wave_sync();
}

// As head32 for 8-byte outputs.
template <int LAYOUT>
__device__ __forceinline__ uint32_t head64(uint64_t out, uint64_t count) {
    const uint32_t h = LAYOUT == 0 ? ((out & 15) ? 1u : 0u) : ((128u - (uint32_t)(out & 127)) & 127) >> 3;
    return (uint64_t)h > count ? (uint32_t)count : h;
}

// WPL (words per lane and step): 4 = 256-word steps; 2 = 128-word steps for
// states with N - pos < 256 (the MTGP64 dynamic-creator sets, N = 176),
// lane-strided layout only.
template <int NJ, int NBUF, int LAYOUT, int WPL = 4, int WV = 1>
__device__ __forceinline__ int place64(uint32_t *bh, uint32_t *bl, P64 p, int b, uint32_t h) {
    const int to = LAYOUT == 0 ? (int)((4u - h) & 3u) : 0;
    const bool misaligned = LAYOUT == 0 && (((uint32_t)b + h) & 3u) != 0;
    if (misaligned || b + p.n + kWave * WV * WPL > NBUF) {
        relocate64<NJ, WV>(bh, bl, b, p.n, to);
        return to;
    }
    return b;
}

template <int NJ, int NBUF, int WPL = 4, int WV = 1>
__device__ __forceinline__ int advance64(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b, uint64_t count) {
    constexpr int W = kWave * WV * WPL;
    for (uint64_t s = count / W; s > 0; s--) {
        if (b + p.n + W > NBUF) {
            relocate64<NJ, WV>(bh, bl, b, p.n);
            b = 0;
        }
        step64s<0, false, false, WPL, WV>(bh, bl, tab, p, b, nullptr, 0);
        wave_sync();
        b += W;
    }
    const uint32_t m = (uint32_t)(count % W);
    if (m) {
        if (b + p.n + W > NBUF) {
            relocate64<NJ, WV>(bh, bl, b, p.n);
            b = 0;
        }
        step64s<0, false, false, WPL, WV>(bh, bl, tab, p, b, nullptr, 0);
        wave_sync();
        b += (int)m;
    }
    return b;
}

template <int KIND, int NJ, int NBUF, int LAYOUT, int WPL = 4, int WV = 1>
__device__ __forceinline__ int generate64(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b, uint32_t h,
                          gu32 *__restrict__ out, uint64_t count) {
    static_assert(LAYOUT == 1 || (WPL == 4 && WV == 1), "128-word steps and multi-wave cells are lane-strided");
    constexpr int W = kWave * WV * WPL;
    const int l = threadIdx.x;
    gu32 *dst = out;
    uint64_t rest = count;
    if (h > 0) {
        if (LAYOUT == 0) {
            if ((uint32_t)l < h) {
                const int q = b + l;
                uint32_t rh, rl, oh, ol;
                rec64(p, tab, bh[q], bl[q], bh[q + 1], bl[q + 1], bh[q + p.pos], bl[q + p.pos], rh, rl);
                bh[q + p.n] = rh;
                bl[q + p.n] = rl;
                out64<KIND>(tab, rh, rl, bl[q + p.pos - 1], oh, ol);
                reinterpret_cast<gu2 *>(out)[l] = v2u{ol, oh};
            }
        } else {
            step64s<KIND, true, false, WPL, WV>(bh, bl, tab, p, b, dst, h);
        }
        wave_sync();
        b += (int)h;
        dst += 2 * h;
        rest -= (uint64_t)h;
    }
    for (uint64_t s = rest / W; s > 0; s--) {
        if (b + p.n + W > NBUF) {
            relocate64<NJ, WV>(bh, bl, b, p.n);
            b = 0;
        }
        if (LAYOUT == 0)
            step64<KIND, true, true>(bh, bl, tab, p, b, dst, kStep);
        else
            step64s<KIND, true, true, WPL, WV>(bh, bl, tab, p, b, dst, W);
        wave_sync();
        b += W;
        dst += 2 * W;
    }
    const uint32_t m = (uint32_t)(rest % W);
    if (m) {
        if (b + p.n + W > NBUF) {
            relocate64<NJ, WV>(bh, bl, b, p.n);
            b = 0;
        }
        if (LAYOUT == 0)
            step64<KIND, true>(bh, bl, tab, p, b, dst, m);
        else
            step64s<KIND, true, false, WPL, WV>(bh, bl, tab, p, b, dst, m);
        wave_sync();
        b += (int)m;
    }
    return b;
}

template <int NJ, int NBUF, int LAYOUT, int WPL, int WV = 1>
__device__ __forceinline__ int segment64_body(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b,
                                              uint64_t out, uint64_t count, int kind) {
    if (count == 0)
        return b;
    if (out == 0)
        return advance64<NJ, NBUF, WPL, WV>(bh, bl, tab, p, b, count);
    const uint32_t h = head64<LAYOUT>(out, count);
    b = place64<NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, p, b, h);
    gu32 *dst = reinterpret_cast<gu32 *>(out);
    switch (kind) {
    case 0: return generate64<0, NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, h, dst, count);
    case 1: return generate64<1, NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, h, dst, count);
    case 2: return generate64<2, NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, h, dst, count);
    case 3: return generate64<3, NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, h, dst, count);
    default: return generate64<4, NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, h, dst, count);
    }
}

// Inlined for the 2-words-per-lane class, an out-of-line call otherwise:
// inlined, the 4-words-per-lane kernels need about 180 VGPRs (2 waves per
// SIMD) against about 100 with the call.
template <int NJ, int NBUF, int LAYOUT, int WPL, int WV>
__device__ __noinline__ int segment64_call(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b,
                                           uint64_t out, uint64_t count, int kind) {
    return segment64_body<NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, out, count, kind);
}
template <int NJ, int NBUF, int LAYOUT, int WPL = 4, int WV = 1>
__device__ __forceinline__ int segment64(uint32_t *bh, uint32_t *bl, const uint32_t *tab, P64 p, int b,
                                         uint64_t out, uint64_t count, int kind) {
    if constexpr (WPL == 2)
        return segment64_body<NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, out, count, kind);
    else
        return segment64_call<NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, out, count, kind);
}

// WV waves per workgroup: as for mtgp32_items_kernel.
template <int NMAX, int K, int LAYOUT, int WPL = 4, int WV = 1>
__global__ __launch_bounds__(64 * WV) void mtgp64_items_kernel(const Item *__restrict__ items, int n_items,
                                                               const DevParam *__restrict__ params,
                                                               uint32_t *__restrict__ queue) {
    constexpr int T = kWave * WV;
    constexpr int NJ = (NMAX + T - 1) / T;
    constexpr int NBUF = K * T * WPL + T * NJ + T;
    __shared__ __attribute__((aligned(16))) uint32_t bh[NBUF];
    __shared__ __attribute__((aligned(16))) uint32_t bl[NBUF];
    __shared__ uint32_t tab[kTabWords];
    __shared__ uint32_t next_item;
    const int l = threadIdx.x;
    uint32_t it = blockIdx.x;
    while (it < (uint32_t)n_items) {
        const Item &I = items[it];
        const DevParam &D = params[I.param];
        P64 p;
        p.n = D.n;
        p.pos = D.pos;
        p.sh1 = D.sh1;
        p.sh2 = D.sh2;
        p.mask_hi = D.mask_hi;
        p.mask_lo = D.mask_lo;
        fill_tables(tab, D);
        gu64 *state_out = reinterpret_cast<gu64 *>(I.state_out);
        const gu64 *state_in = reinterpret_cast<const gu64 *>(I.state_in);
        wave_sync();
        // (128-word steps and multi-wave cells: the host gives such items
        // batched start states, so the fused jump, which steps 256 words at a
        // time on one wave, is compiled out.)
        if (WPL == 4 && WV == 1 && I.poly) {
            uint32_t ah[NJ], al[NJ];
            jump64<NJ, K, NBUF>(bh, bl, tab, p, state_in, reinterpret_cast<const gu32 *>(I.poly),
                                I.poly_words, ah, al);
#pragma unroll
            for (int m = 0; m < NJ; m++)
                if (l + kWave * m < p.n) {
                    bh[l + kWave * m] = ah[m];
                    bl[l + kWave * m] = al[m];
                }
        } else {
            const int parts = I.parts > 1 ? I.parts : 1;
            const uint64_t pstride = I.part_bytes / 8;
            for (int j = l; j < p.n; j += T) {
                uint64_t v = state_in[j];
                for (int q = 1; q < parts; q++)
                    v ^= state_in[(uint64_t)q * pstride + (uint64_t)j];
                bh[j] = (uint32_t)(v >> 32);
                bl[j] = (uint32_t)v;
            }
        }
        wave_sync();
        int b = 0;
        if (I.nsegs > 0) {
            const GMEM Segment *segs = reinterpret_cast<const GMEM Segment *>(I.segs);
            for (int si = 0; si < I.nsegs; si++)
                b = segment64<NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, segs[si].out, segs[si].count,
                                                         segs[si].kind);
        } else {
            b = segment64<NJ, NBUF, LAYOUT, WPL, WV>(bh, bl, tab, p, b, I.out, I.count, I.kind);
        }
        if (state_out)
            for (int j = l; j < p.n; j += T)
                state_out[j] = ((uint64_t)bh[b + j] << 32) | bl[b + j];
        wave_sync();
        if constexpr (WV == 1) {
            uint32_t next = 0;
            if (l == 0)
                next = atomicAdd(&queue[0], 1u);
            it = gridDim.x + uniform(__shfl(next, 0));
        } else {
            // One claim per workgroup, handed to every wave through LDS.
            if (l == 0)
                next_item = atomicAdd(&queue[0], 1u);
            wave_sync();
            it = gridDim.x + uniform(next_item);
            wave_sync();
        }
    }
    if (l == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// ------------------------------------------------------ batched jump kernels
// One workgroup takes a JumpItem: the base window's sequence is generated
// chunk by chunk in LDS (kJumpSteps steps per chunk) and every chunk is
// folded into kJumpBatch accumulators at once.  For each residue rho the
// lane column buf[rho + l + 64r] is loaded into registers once and shared
// by all jumps of the batch, so LDS traffic per jump drops by the batch
// size; the host supplies each jump's coefficients as per-residue masks.

typedef GMEM uint64_t gu64c;

// Lane `src`'s 64-bit value as a wave-uniform value.
__device__ __forceinline__ uint64_t lane_mask64(uint64_t v, int src) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// acc[m] ^= XOR of R[q + m] over the set bits q of a wave-uniform mask, one
// uniform branch per bit.  (Measured against taking the bits in pairs with
// v_xor3, and against branch-free acc ^ (R & s) with scalar masks: both were
// 1.8x and 2.6x slower on the batched jump launch.)
template <int QMAX, int NJ>
__device__ __forceinline__ void fold_column(uint32_t mask, const uint32_t (&R)[QMAX + NJ], uint32_t (&acc)[NJ]) {
    if (!mask)
        return;
#pragma unroll
    for (int q = 0; q < QMAX; q++)
        if (mask >> q & 1u) {
#pragma unroll
            for (int m = 0; m < NJ; m++)
                acc[m] ^= R[q + m];
        }
}

// fold_column for the two 32-bit planes of MTGP64 words.
template <int QMAX, int NJ>
__device__ __forceinline__ void fold_column2(uint32_t mask, const uint32_t (&RH)[QMAX + NJ],
                                             const uint32_t (&RL)[QMAX + NJ], uint32_t (&ah)[NJ], uint32_t (&al)[NJ]) {
    if (!mask)
        return;
#pragma unroll
    for (int q = 0; q < QMAX; q++)
        if (mask >> q & 1u) {
#pragma unroll
            for (int m = 0; m < NJ; m++) {
                ah[m] ^= RH[q + m];
                al[m] ^= RL[q + m];
            }
        }
}

template <int NMAX>
__global__ __launch_bounds__(64) void mtgp32_jump_kernel(const JumpItem *__restrict__ items, int n_items,
                                                         const DevParam *__restrict__ params,
                                                         uint32_t *__restrict__ queue) {
    constexpr int K = kJumpSteps;
    constexpr int NJ = (NMAX + kWave - 1) / kWave;
    constexpr int QMAX = 4 * K;
    constexpr int B = kJumpBatch;
    constexpr int NBUF = K * kStep + kWave * NJ + kWave;
    static_assert(B * 16 == 64 && QMAX == 16, "one 64-bit mask word per residue");
    __shared__ __attribute__((aligned(16))) uint32_t buf[NBUF];
    __shared__ uint32_t tab[kTabWords];
    const int l = threadIdx.x;
    uint32_t it = blockIdx.x;
    while (it < (uint32_t)n_items) {
        const JumpItem &I = items[it];
        const DevParam &D = params[I.param];
        P32 p;
        p.n = D.n;
        p.pos = D.pos;
        p.pm1 = D.pos - 1;
        p.sh1 = D.sh1;
        p.sh2 = D.sh2;
        p.mask = D.mask_lo;
        fill_tables(tab, D);
        const gu32 *in = reinterpret_cast<const gu32 *>(I.state_in);
        const gu64c *mk = reinterpret_cast<const gu64c *>(I.masks);
        const int in_parts = I.in_parts;
        const uint64_t in_stride = I.in_part_bytes / 4;
        wave_sync();
        for (int j = l; j < p.n; j += kWave) {
            uint32_t v = in[j];
            for (int q = 1; q < in_parts; q++)
                v ^= in[(uint64_t)q * in_stride + (uint64_t)j];
            buf[j] = v;
        }
        wave_sync();
        uint32_t acc[B][NJ];
#pragma unroll
        for (int b = 0; b < B; b++)
#pragma unroll
            for (int m = 0; m < NJ; m++)
                acc[b][m] = 0;
        for (int c = 0; c < I.c1; c++) {
            for (int s = 0; s < K; s++) {
                step32s<0, false>(buf, tab, p, s * kStep, nullptr, 0);
                wave_sync();
            }
            // Lane rho holds residue rho's masks; readlane hands them out.
            const uint64_t lane_masks = c >= I.c0 ? mk[(uint64_t)c * kWave + l] : 0ull;
            for (int rho = 0; c >= I.c0 && rho < kWave; rho++) {
                const uint64_t mm = lane_mask64(lane_masks, rho);
                if (!mm)
                    continue;
                uint32_t R[QMAX + NJ];
                const uint32_t *col = buf + rho + l;
#pragma unroll
                for (int r = 0; r < QMAX + NJ; r++)
                    R[r] = col[kWave * r];
#pragma unroll
                for (int b = 0; b < B; b++)
                    fold_column<QMAX, NJ>((uint32_t)(mm >> (16 * b)) & 0xffffu, R, acc[b]);
            }
            relocate32<NJ>(buf, K * kStep, p.n);
        }
#pragma unroll
        for (int b = 0; b < B; b++) {
            if (b < I.nb) {
                gu32 *o = reinterpret_cast<gu32 *>(I.states_out + (uint64_t)b * (uint64_t)I.state_bytes);
#pragma unroll
                for (int m = 0; m < NJ; m++)
                    if (l + kWave * m < p.n)
                        o[l + kWave * m] = acc[b][m];
            }
        }
        wave_sync();
        uint32_t next = 0;
        if (l == 0)
            next = atomicAdd(&queue[0], 1u);
        it = gridDim.x + uniform(__shfl(next, 0));
    }
    if (l == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

template <int NMAX, int WPL = 4>
__global__ __launch_bounds__(64) void mtgp64_jump_kernel(const JumpItem *__restrict__ items, int n_items,
                                                         const DevParam *__restrict__ params,
                                                         uint32_t *__restrict__ queue) {
    constexpr int K = kJumpChunk / (kWave * WPL);  // steps per chunk
    constexpr int NJ = (NMAX + kWave - 1) / kWave;
    constexpr int QMAX = kJumpChunk / kWave;
    constexpr int B = kJumpBatch;
    constexpr int NBUF = kJumpChunk + kWave * NJ + kWave;
    __shared__ __attribute__((aligned(16))) uint32_t bh[NBUF];
    __shared__ __attribute__((aligned(16))) uint32_t bl[NBUF];
    __shared__ uint32_t tab[kTabWords];
    const int l = threadIdx.x;
    uint32_t it = blockIdx.x;
    while (it < (uint32_t)n_items) {
        const JumpItem &I = items[it];
        const DevParam &D = params[I.param];
        P64 p;
        p.n = D.n;
        p.pos = D.pos;
        p.sh1 = D.sh1;
        p.sh2 = D.sh2;
        p.mask_hi = D.mask_hi;
        p.mask_lo = D.mask_lo;
        fill_tables(tab, D);
        const gu64 *in = reinterpret_cast<const gu64 *>(I.state_in);
        const gu64c *mk = reinterpret_cast<const gu64c *>(I.masks);
        const int in_parts = I.in_parts;
        const uint64_t in_stride = I.in_part_bytes / 8;
        wave_sync();
        for (int j = l; j < p.n; j += kWave) {
            uint64_t v = in[j];
            for (int q = 1; q < in_parts; q++)
                v ^= in[(uint64_t)q * in_stride + (uint64_t)j];
            bh[j] = (uint32_t)(v >> 32);
            bl[j] = (uint32_t)v;
        }
        wave_sync();
        uint32_t ah[B][NJ], al[B][NJ];
#pragma unroll
        for (int b = 0; b < B; b++)
#pragma unroll
            for (int m = 0; m < NJ; m++)
                ah[b][m] = al[b][m] = 0;
        for (int c = 0; c < I.c1; c++) {
            for (int s = 0; s < K; s++) {
                step64s<0, false, false, WPL>(bh, bl, tab, p, s * kWave * WPL, nullptr, 0);
                wave_sync();
            }
            // Lane rho holds residue rho's masks; readlane hands them out.
            const uint64_t lane_masks = c >= I.c0 ? mk[(uint64_t)c * kWave + l] : 0ull;
            for (int rho = 0; c >= I.c0 && rho < kWave; rho++) {
                const uint64_t mm = lane_mask64(lane_masks, rho);
                if (!mm)
                    continue;
                uint32_t RH[QMAX + NJ], RL[QMAX + NJ];
#pragma unroll
                for (int r = 0; r < QMAX + NJ; r++) {
                    RH[r] = bh[rho + l + kWave * r];
                    RL[r] = bl[rho + l + kWave * r];
                }
#pragma unroll
                for (int b = 0; b < B; b++)
                    fold_column2<QMAX, NJ>((uint32_t)(mm >> (16 * b)) & 0xffffu, RH, RL, ah[b], al[b]);
            }
            relocate64<NJ>(bh, bl, kJumpChunk, p.n);
        }
#pragma unroll
        for (int b = 0; b < B; b++) {
            if (b < I.nb) {
                gu64 *o = reinterpret_cast<gu64 *>(I.states_out + (uint64_t)b * (uint64_t)I.state_bytes);
#pragma unroll
                for (int m = 0; m < NJ; m++)
                    if (l + kWave * m < p.n)
                        o[l + kWave * m] = ((uint64_t)ah[b][m] << 32) | al[b][m];
            }
        }
        wave_sync();
        uint32_t next = 0;
        if (l == 0)
            next = atomicAdd(&queue[0], 1u);
        it = gridDim.x + uniform(__shfl(next, 0));
    }
    if (l == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// Steps between window relocations (compile-time; -DMTGP_RELOC_STEPS=n).
#ifndef MTGP_RELOC_STEPS
#define MTGP_RELOC_STEPS 3
#endif
constexpr int kK = MTGP_RELOC_STEPS;

// Largest state of the 64-bit narrow class (128-word steps).
constexpr int kNarrowN = 176;

template <int B, int NMAX, int LAYOUT, int WPL = 4>
struct KernelOf;
// Layout 2: lane-strided two-wave cells, 512-word steps (256-word steps of
// two words per lane for the N <= 351/363 classes, whose N - pos can be as
// small as 258).  Layout 3: four-wave cells, 1024-word steps (512-word
// steps of two words per lane for the N <= 726/696 classes).
template <int NMAX, int LAYOUT>
struct KernelOf<32, NMAX, LAYOUT, 4> {
    static constexpr int waves = LAYOUT == 3 ? 4 : LAYOUT == 2 ? 2 : 1;
    static constexpr int step_layout = LAYOUT >= 2 ? 1 : LAYOUT;
    static constexpr int wpl = (LAYOUT == 2 && NMAX <= 351) || (LAYOUT == 3 && NMAX <= 726) ? 2 : 4;
    static constexpr auto fn = mtgp32_items_kernel<NMAX, kK, step_layout, waves, wpl>;
    static std::string name() {
        return "mtgp32_items_kernel<" + std::to_string(NMAX) + "," + std::to_string(kK) + "," +
               std::to_string(step_layout) + "," + std::to_string(waves) + "," + std::to_string(wpl) + ">";
    }
};
template <int NMAX, int LAYOUT, int WPL>
struct KernelOf<64, NMAX, LAYOUT, WPL> {
    static constexpr int waves = LAYOUT == 3 ? 4 : LAYOUT == 2 ? 2 : 1;
    static constexpr int step_layout = LAYOUT >= 2 ? 1 : LAYOUT;
    static constexpr int wpl = (LAYOUT == 2 && NMAX <= 363) || (LAYOUT == 3 && NMAX <= 696) ? 2 : WPL;
    static constexpr auto fn = mtgp64_items_kernel<NMAX, kK, step_layout, wpl, waves>;
    static std::string name() {
        return "mtgp64_items_kernel<" + std::to_string(NMAX) + "," + std::to_string(kK) + "," +
               std::to_string(step_layout) + "," + std::to_string(wpl) + "," + std::to_string(waves) + ">";
    }
};

template <class KO>
hipError_t launch_one(const Item *items, int n_items, const DevParam *params, uint32_t *queue, int grid,
                      size_t dyn_lds, hipStream_t s) {
    hipLaunchKernelGGL(KO::fn, dim3(grid), dim3(kWave * KO::waves), dyn_lds, s, items, n_items, params, queue);
    return hipGetLastError();
}

// Calls f.template operator()<KernelOf<...>>() for (bits, cls, layout).
// Class 3 (64-bit only) is the narrow class: 128-word steps, layout 1.
template <class F>
auto with_kernel(int bits, int cls, int layout, F &&f) {
    if (!layout_supported(bits, cls, layout))
        layout = 1;
    if (bits == 64 && cls == 3)
        return f.template operator()<KernelOf<64, kNarrowN, 1, 2>>();
#define MTGP_CASE(B, N, L) \
    if (bits == B && N_OF(B, cls) == N && layout == L) return f.template operator()<KernelOf<B, N, L>>();
#define N_OF(B, c) ((B) == 32 ? ((c) == 0 ? 351 : (c) == 1 ? 726 : 1391) : ((c) == 0 ? 363 : (c) == 1 ? 696 : 1727))
    MTGP_CASE(32, 351, 0) MTGP_CASE(32, 351, 1) MTGP_CASE(32, 726, 0) MTGP_CASE(32, 726, 1)
    MTGP_CASE(32, 1391, 0) MTGP_CASE(32, 1391, 1) MTGP_CASE(32, 351, 2) MTGP_CASE(32, 726, 2)
    MTGP_CASE(32, 1391, 2) MTGP_CASE(32, 726, 3) MTGP_CASE(32, 1391, 3) MTGP_CASE(64, 363, 0) MTGP_CASE(64, 363, 1)
    MTGP_CASE(64, 363, 2)
    MTGP_CASE(64, 696, 0) MTGP_CASE(64, 696, 1) MTGP_CASE(64, 1727, 0) MTGP_CASE(64, 1727, 1)
    MTGP_CASE(64, 696, 2) MTGP_CASE(64, 1727, 2) MTGP_CASE(64, 696, 3) MTGP_CASE(64, 1727, 3)
#undef MTGP_CASE
#undef N_OF
    return f.template operator()<KernelOf<64, 1727, 1>>();
}

}  // namespace

// Class c of each word size: N <= 351/726/1391 (32-bit), 363/696/1727 (64).
int kernel_class(int bits, int n) {
    if (bits == 32)
        return n <= 351 ? 0 : n <= 726 ? 1 : n <= 1391 ? 2 : -1;
    return n <= 363 ? 0 : n <= 696 ? 1 : n <= 1727 ? 2 : -1;
}

int kernel_class(int bits, int n, int pos) {
    if (n - pos >= 256)
        return kernel_class(bits, n);
    if (bits == 64 && n <= kNarrowN && n - pos >= 128)
        return 3;
    return -1;
}

const void *kernel_fn(int bits, int cls, int layout) {
    return with_kernel(bits, cls, layout, []<class KO>() { return reinterpret_cast<const void *>(KO::fn); });
}

int kernel_threads(int bits, int cls, int layout) {
    return with_kernel(bits, cls, layout, []<class KO>() { return kWave * KO::waves; });
}

std::string kernel_name(int bits, int cls, int layout) {
    return with_kernel(bits, cls, layout, []<class KO>() { return KO::name(); });
}

bool layout_supported(int bits, int cls, int layout) {
    if (layout < 0 || layout >= kLayouts)
        return false;
    if (bits == 64 && cls == 3)
        return layout == 1;
    if (layout == 2)
        return cls <= 2;
    if (layout == 3)
        return cls == 1 || cls == 2;
    return true;
}

int layout_step_words(int bits, int cls, int layout) {
    if (bits == 64 && cls == 3)
        return 128;
    if (layout == 3)
        return cls == 1 ? 512 : 1024;
    if (layout == 2)
        return cls == 0 ? 256 : 512;
    return 256;
}

hipError_t launch_items(int bits, int cls, int layout, const Item *items, int n_items, const DevParam *params,
                        uint32_t *queue, int grid, size_t dyn_lds, hipStream_t s) {
    return with_kernel(bits, cls, layout,
                       [&]<class KO>() { return launch_one<KO>(items, n_items, params, queue, grid, dyn_lds, s); });
}

namespace {
template <class F>
auto with_jump_kernel(int bits, int cls, F &&f) {
    if (bits == 32)
        return cls == 0 ? f(mtgp32_jump_kernel<351>) : cls == 1 ? f(mtgp32_jump_kernel<726>) : f(mtgp32_jump_kernel<1391>);
    return cls == 0   ? f(mtgp64_jump_kernel<363>)
           : cls == 1 ? f(mtgp64_jump_kernel<696>)
           : cls == 2 ? f(mtgp64_jump_kernel<1727>)
                      : f(mtgp64_jump_kernel<kNarrowN, 2>);
}
}  // namespace

const void *jump_kernel_fn(int bits, int cls) {
    return with_jump_kernel(bits, cls, [](auto fn) { return reinterpret_cast<const void *>(fn); });
}

hipError_t launch_jumps(int bits, int cls, const JumpItem *items, int n_items, const DevParam *params,
                        uint32_t *queue, int grid, size_t dyn_lds, hipStream_t s) {
    return with_jump_kernel(bits, cls, [&](auto fn) {
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kWave), dyn_lds, s, items, n_items, params, queue);
        return hipGetLastError();
    });
}

namespace {
// Plain streaming copy (16 bytes per lane per iteration, grid-stride): the
// reference point for hardware byte counters.
__global__ __launch_bounds__(256) void copy16_kernel(uint64_t src_addr, uint64_t dst_addr, uint64_t n16) {
    const gu4 *src = reinterpret_cast<const gu4 *>(src_addr);
    gu4 *dst = reinterpret_cast<gu4 *>(dst_addr);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}
}  // namespace

namespace {
// Plain streaming stores: each single-wave workgroup writes its own
// contiguous rows of 256 words, one 256-byte dword store per instruction
// (the store pattern of the generation kernels, without the generator).
// Its stores are wave-scope: without a generator beside them they are
// faster than the generation kernels' device-scope stores (5.64-5.65
// against 5.43-5.48 TB/s over 16 GiB), so this is the write ceiling.
__global__ __launch_bounds__(64) void fill_rows_kernel(uint64_t dst_addr, uint64_t n_words, uint64_t per_wave,
                                                       uint32_t value) {
    gu32 *dst = reinterpret_cast<gu32 *>(dst_addr);
    const uint64_t begin = (uint64_t)blockIdx.x * per_wave;
    const uint64_t end = begin + per_wave < n_words ? begin + per_wave : n_words;
    for (uint64_t k = begin + threadIdx.x; k < end; k += kWave)
        dst[k] = value;
}
}  // namespace

hipError_t launch_fill(void *dst, uint64_t n_words, uint32_t value, int waves, hipStream_t s) {
    if (n_words == 0)
        return hipSuccess;
    const uint64_t per = ((n_words + (uint64_t)waves - 1) / (uint64_t)waves + 255) & ~(uint64_t)255;
    const int grid = (int)((n_words + per - 1) / per);
    hipLaunchKernelGGL(fill_rows_kernel, dim3(grid), dim3(kWave), 0, s, (uint64_t)(uintptr_t)dst, n_words, per,
                       value);
    return hipGetLastError();
}

hipError_t launch_copy16(const void *src, void *dst, uint64_t n16, int grid, hipStream_t s) {
    hipLaunchKernelGGL(copy16_kernel, dim3(grid), dim3(256), 0, s, (uint64_t)(uintptr_t)src,
                       (uint64_t)(uintptr_t)dst, n16);
    return hipGetLastError();
}

namespace {
// One workgroup per JumpItem: the window at states_out (state_bytes) is the
// XOR of the in_parts partial windows at state_in, in_part_bytes apart.
__global__ __launch_bounds__(64) void combine_kernel(const JumpItem *__restrict__ items) {
    const JumpItem &I = items[blockIdx.x];
    const gu32 *in = reinterpret_cast<const gu32 *>(I.state_in);
    gu32 *out = reinterpret_cast<gu32 *>(I.states_out);
    const int words = I.state_bytes / 4;
    const int parts = I.in_parts > 1 ? I.in_parts : 1;
    const uint64_t stride = I.in_part_bytes / 4;
    for (int j = threadIdx.x; j < words; j += kWave) {
        uint32_t v = in[j];
        for (int q = 1; q < parts; q++)
            v ^= in[(uint64_t)q * stride + (uint64_t)j];
        out[j] = v;
    }
}
}  // namespace

hipError_t launch_combine(const JumpItem *items, int n_items, hipStream_t s) {
    if (n_items <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(combine_kernel, dim3(n_items), dim3(kWave), 0, s, items);
    return hipGetLastError();
}

// ------------------------------------------------- jumps on the matrix cores
// The nj jumps of one base window are one GF(2) product: bit p of jump b's
// new word j is the parity of sum_k c_b[k] * bit_p(w[k + j]) over the
// base's sequence w (the fold of mtgp32_jump_kernel, one bit at a time).
// With the coefficient bits and the sequence bits as fp4 values 0 and 1.0,
// one 32 x 32 x 64 matrix-core product adds 65536 such terms exactly (the
// counts stay below 2^24 in fp32), and the parity of each accumulator is
// the jumped bit:
//  * rows: 32 jumps (coefficients k in groups of 64);
//  * columns: the 32 bits of one window word j, bit p read from the bit
//    plane of the sequence, so a column's 64 values are the plane's bits
//    k + j for the product step's 64 k.
// The planes kernel writes each group's planes first: the sequence in 32 x
// 32 bit blocks transposed in place (planes[32 x + p] holds bit p of
// w[32 x + t] in bit t), so a plane's window of 32 + kMjWords bits at any
// offset is a funnel shift of three words.
namespace {

constexpr uint32_t kFp4Ones = 0x22222222u;  // fp4 1.0 (0b0010) in every nibble

// 32 coefficient bits (bit t = k offset t) as an fp4 operand: value 8v + q
// is 1.0 when bit 4q + v is set.  B operands use the same order.
__device__ __forceinline__ mfma_v8i fp4_of_bits(uint32_t c) {
    mfma_v8i a = {0, 0, 0, 0, 0, 0, 0, 0};
    a[0] = (int)((c << 1) & kFp4Ones);
    a[1] = (int)(c & kFp4Ones);
    a[2] = (int)((c >> 1) & kFp4Ones);
    a[3] = (int)((c >> 2) & kFp4Ones);
    return a;
}

// Low word of the 64-bit (hi:lo) shifted right by s < 32 (v_alignbit_b32).
__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
}

// One round of the 32 x 32 bit transpose: lanes l and l ^ D swap the
// off-diagonal D x D blocks of their 2D x 2D block (lo masks the low D bits
// of every 2D).
template <int D>
__device__ __forceinline__ uint32_t transpose_round(uint32_t v, uint32_t l, uint32_t lo) {
    const uint32_t p = lane_xor32<D>(v);
    return (l & D) ? ((v & ~lo) | ((p >> D) & lo)) : ((v & lo) | ((p << D) & ~lo));
}

// Each group of 32 lanes holds a 32 x 32 bit matrix, row t in lane t;
// returns column l of it to lane l (bit t = row t's bit l).
__device__ __forceinline__ uint32_t transpose32(uint32_t v) {
    const uint32_t l = threadIdx.x;
    v = transpose_round<16>(v, l, 0x0000ffffu);
    v = transpose_round<8>(v, l, 0x00ff00ffu);
    v = transpose_round<4>(v, l, 0x0f0f0f0fu);
    v = transpose_round<2>(v, l, 0x33333333u);
    v = transpose_round<1>(v, l, 0x55555555u);
    return v;
}

// One workgroup of kPlaneWaves waves per MJumpGroup: the base window (n
// words), then the recursion 256 words at a time (one word per thread) in a
// linear LDS buffer, relocated as in the generation kernels; every complete
// row of 64 positions is transposed (one row per wave) and stored, so
// nothing is dropped by a relocation (n >= 64).
constexpr int kPlaneWaves = 4;
__global__ __launch_bounds__(64 * kPlaneWaves) void mtgp32_planes_kernel(const MJumpGroup *__restrict__ groups,
                                                                        int n_groups,
                                                                        const DevParam *__restrict__ params) {
    constexpr int W = kPlaneWaves;
    constexpr int NMAX = 1391;
    constexpr int K = 4;
    constexpr int NJ = (NMAX + kWave * W - 1) / (kWave * W);
    constexpr int NBUF = K * kStep + kWave * W * NJ + kWave;
    __shared__ __attribute__((aligned(16))) uint32_t buf[NBUF];
    __shared__ uint32_t tab[kTabWords];
    const int t = threadIdx.x, l = t & (kWave - 1), wv = t / kWave;
    for (int gi = blockIdx.x; gi < n_groups; gi += gridDim.x) {
        const MJumpGroup &G = groups[gi];
        const DevParam &D = params[G.param];
        P32 p;
        p.n = D.n;
        p.pos = D.pos;
        p.pm1 = D.pos - 1;
        p.sh1 = D.sh1;
        p.sh2 = D.sh2;
        p.mask = D.mask_lo;
        fill_tables(tab, D);
        const gu32 *in = reinterpret_cast<const gu32 *>(G.state_in);
        const int in_parts = G.in_parts;
        const uint64_t in_stride = G.in_part_bytes / 4;
        for (int j = t; j < p.n; j += kWave * W) {
            uint32_t v = in[j];
            for (int q = 1; q < in_parts; q++)
                v ^= in[(uint64_t)q * in_stride + (uint64_t)j];
            buf[j] = v;
        }
        wave_sync();
        gu32 *T = reinterpret_cast<gu32 *>(G.planes);
        const int total = 32 * G.pw;
        int off = 0, len = p.n, done = 0;  // buf[i] = w[off + i] for i < len
        for (;;) {
            // complete rows, one per wave
            const int avail = off + len < total ? off + len : total;
            while (done + kWave <= avail) {
                const int row = done + kWave * wv;
                // (every wave transposes; a wave without a row of its own
                // repeats the first one and does not store it)
                const uint32_t v = transpose32(buf[(row + kWave <= avail ? row : done) - off + l]);
                if (row + kWave <= avail)
                    T[row + l] = v;
                done += kWave * ((avail - done) / kWave < W ? (avail - done) / kWave : W);
            }
            if (done >= total)
                break;
            wave_sync();
            if (len + kStep > NBUF) {
                relocate32<NJ, W>(buf, len - p.n, p.n);
                off += len - p.n;
                len = p.n;
            }
            step32s<0, false, false, W, 1>(buf, tab, p, len - p.n, nullptr, 0);
            wave_sync();
            len += kStep;
        }
        wave_sync();
    }
}

// One wave per MJumpItem: accumulators for RT x 32 jumps (the item's jumps
// fill RT row tiles) and JW words, product steps of 64 coefficients; then
// each accumulator's parity, gathered by ballots into window words.
template <int RT, int JW>
__global__ __launch_bounds__(64) void mtgp32_jump_mfma_kernel(const MJumpGroup *__restrict__ groups,
                                                              const MJumpItem *__restrict__ items) {
    const MJumpItem I = items[blockIdx.x];
    const MJumpGroup &G = groups[I.group];
    const int L = threadIdx.x, col = L & 31, kg = L >> 5;
    const int nj = G.nj, njp = G.njp, r0 = I.r0, j0 = I.j0;
    if (r0 < 0 || r0 >= nj || j0 < 0 || j0 >= G.n)
        return;  // not an item of this group (the host makes none)
    const gu32 *T = reinterpret_cast<const gu32 *>(G.planes);
    const gu32 *C = reinterpret_cast<const gu32 *>(G.coef);
    const uint32_t sh = (uint32_t)j0 & 31u;
    mfma_v16f acc[RT][JW];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < JW; c++)
#pragma unroll
            for (int i = 0; i < 16; i++)
                acc[r][c][i] = 0.0f;
    const int nks = G.kpad >> 6;
    // Operands of step ks: coefficient words cw[r] and plane words d0..d2;
    // the next step's are loaded before this step's products.
    const gu32 *cp = C + (uint64_t)kg * (uint64_t)njp + (uint64_t)(r0 + col);
    const gu32 *tp = T + (uint64_t)(kg + (j0 >> 5)) * 32 + (uint64_t)col;
    const uint64_t cstep = 2 * (uint64_t)njp;
    uint32_t cw[RT], d0, d1, d2;
#pragma unroll
    for (int r = 0; r < RT; r++)
        cw[r] = cp[32 * r];
    d0 = tp[0];
    d1 = tp[32];
    d2 = tp[64];
    for (int ks = 0; ks < nks; ks++) {
        mfma_v8i A[RT];
#pragma unroll
        for (int r = 0; r < RT; r++)
            A[r] = fp4_of_bits(cw[r]);
        // This lane's 32 positions start at k = 64 ks + 32 kg; the column of
        // word j0 + jj reads plane `col` from bit k + j0 + jj on.
        const uint32_t wlo = funnel(d1, d0, sh), whi = funnel(d2, d1, sh);
        if (ks + 1 < nks) {
            cp += cstep;
            tp += 64;
#pragma unroll
            for (int r = 0; r < RT; r++)
                cw[r] = cp[32 * r];
            d0 = tp[0];
            d1 = tp[32];
            d2 = tp[64];
        }
        // Value 8v + q of word jj's operand is bit 4q + v + jj of the window
        // (wlo, whi): nibble q of U[u] = (window >> u) masked to bit 1, with
        // u = v + jj - 1 (U[-1] = (window << 1) masked); U[4a + r] is P_r =
        // U[r] over the whole 64-bit window, shifted right by 4a.
        uint32_t plo[4], phi[4], U[JW + 4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            plo[r] = funnel(whi, wlo, (uint32_t)r) & kFp4Ones;
            phi[r] = (whi >> r) & kFp4Ones;
        }
        U[0] = (wlo << 1) & kFp4Ones;
#pragma unroll
        for (int u = 0; u < JW + 3; u++)
            U[u + 1] = (u >> 2) == 0 ? plo[u & 3] : funnel(phi[u & 3], plo[u & 3], 4u * (uint32_t)(u >> 2));
#pragma unroll
        for (int jj = 0; jj < JW; jj++) {
            const mfma_v8i B = {(int)U[jj], (int)U[jj + 1], (int)U[jj + 2], (int)U[jj + 3], 0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < RT; r++)
                acc[r][jj] = mfma_fp4_32x32x64(A[r], B, acc[r][jj]);
        }
    }
    // acc[r][jj][i] of lane L: jump r0 + 32 r + 8 (i / 4) + 4 (L / 32) + i % 4,
    // bit L % 32 of word j0 + jj.  A ballot of the parities gives that word
    // for two jumps (lanes 0..31 and 32..63); lanes 0 and 1 store them.
#pragma unroll
    for (int r = 0; r < RT; r++) {
#pragma unroll
        for (int jj = 0; jj < JW; jj++) {
            // read out one accumulator at a time
            keep_in_agprs(acc[r][jj]);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint64_t m = wave_ballot(((uint32_t)acc[r][jj][i] & 1u) != 0);
                const int jump = r0 + 32 * r + 8 * (i / 4) + 4 * L + i % 4, word = j0 + jj;
                if (L < 2 && jump < nj && word < G.n) {
                    gu32 *o = reinterpret_cast<gu32 *>(G.out + (uint64_t)jump * G.out_stride);
                    o[word] = (uint32_t)(m >> (32 * L));
                }
            }
        }
    }
}

}  // namespace

hipError_t launch_planes32(const MJumpGroup *groups, int n_groups, const DevParam *params, hipStream_t s) {
    if (n_groups <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(mtgp32_planes_kernel, dim3(n_groups), dim3(kWave * kPlaneWaves), 0, s, groups, n_groups,
                       params);
    return hipGetLastError();
}

hipError_t launch_jump_mfma32(const MJumpGroup *groups, const MJumpItem *items, int n_items, int row_tiles,
                              hipStream_t s) {
    if (n_items <= 0)
        return hipSuccess;
    static_assert(kMjRowTiles == 2, "one kernel per row-tile count");
    if (row_tiles < 1 || row_tiles > kMjRowTiles)
        return hipErrorLaunchFailure;
    auto *fn = row_tiles == 1 ? mtgp32_jump_mfma_kernel<1, kMjWords> : mtgp32_jump_mfma_kernel<2, kMjWords / 2>;
    hipLaunchKernelGGL(fn, dim3(n_items), dim3(kWave), 0, s, groups, items);
    return hipGetLastError();
}

}  // namespace mtgp
