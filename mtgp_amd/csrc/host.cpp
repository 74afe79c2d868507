// host.cpp -- the reference's host-side C API: parameter tables,
// initialisation, free and id strings, for MTGP32 and MTGP64.
//
// Replaces mtgp32-fast.c:36-360 and mtgp64-fast.c:36-365 (and the tables
// of mtgp32-param-fast.c / mtgp64-param-fast.c).  The parameter values are
// the reference's (they define the generators); they are kept in the
// compact basis form of param_sets.inc and expanded at compile time.
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <mutex>
#include <string>

#include "mtgp32-fast.h"
#include "mtgp64-fast.h"
#include "common.hpp"

namespace {

constexpr int hexval(char c) {
    return c >= '0' && c <= '9' ? c - '0' : c >= 'a' && c <= 'f' ? c - 'a' + 10 : 0;
}

// mtgp32-fast.h:47-58: tbl/tmp_tbl are the GF(2) span of four basis
// vectors; flt_tmp_tbl and mask follow from tmp_tbl and mexp.
constexpr mtgp32_params_fast_t make32(int mexp, int pos, int sh1, int sh2,
                                      uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3,
                                      uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3,
                                      const char *sha) {
    mtgp32_params_fast_t p{};
    const uint32_t a[4] = {a0, a1, a2, a3}, t[4] = {t0, t1, t2, t3};
    p.mexp = mexp;
    p.pos = pos;
    p.sh1 = sh1;
    p.sh2 = sh2;
    for (int j = 0; j < 16; j++) {
        uint32_t x = 0, y = 0;
        for (int b = 0; b < 4; b++)
            if (j >> b & 1) {
                x ^= a[b];
                y ^= t[b];
            }
        p.tbl[j] = x;
        p.tmp_tbl[j] = y;
        p.flt_tmp_tbl[j] = (y >> 9) | 0x3f800000u;
    }
    const int n = mexp / 32 + 1;
    p.mask = (uint32_t)(0xffffffffull << (32 * n - mexp));
    for (int i = 0; i < 20; i++)
        p.poly_sha1[i] = (unsigned char)(hexval(sha[2 * i]) * 16 + hexval(sha[2 * i + 1]));
    return p;
}

// mtgp64-fast.h:46-57; basis values are the high halves.
constexpr mtgp64_params_fast_t make64(int mexp, int pos, int sh1, int sh2,
                                      uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3,
                                      uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3,
                                      const char *sha) {
    mtgp64_params_fast_t p{};
    const uint32_t a[4] = {a0, a1, a2, a3}, t[4] = {t0, t1, t2, t3};
    p.mexp = mexp;
    p.pos = pos;
    p.sh1 = sh1;
    p.sh2 = sh2;
    for (int j = 0; j < 16; j++) {
        uint64_t x = 0, y = 0;
        for (int b = 0; b < 4; b++)
            if (j >> b & 1) {
                x ^= (uint64_t)a[b] << 32;
                y ^= (uint64_t)t[b] << 32;
            }
        p.tbl[j] = x;
        p.tmp_tbl[j] = y;
        p.dbl_tmp_tbl[j] = (y >> 12) | 0x3ff0000000000000ull;
    }
    const int n = mexp / 64 + 1;
    p.mask = ~0ull << (64 * n - mexp);
    for (int i = 0; i < 20; i++)
        p.poly_sha1[i] = (unsigned char)(hexval(sha[2 * i]) * 16 + hexval(sha[2 * i + 1]));
    return p;
}

#define MTGP32_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha) \
    make32(m, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha),
#define MTGP64_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha)
constexpr mtgp32_params_fast_t all32[] = {
#include "param_sets.inc"
};
#undef MTGP32_SET
#undef MTGP64_SET
#define MTGP32_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha)
#define MTGP64_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha) \
    make64(m, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha),
constexpr mtgp64_params_fast_t all64[] = {
#include "param_sets.inc"
};
#undef MTGP32_SET
#undef MTGP64_SET
static_assert(sizeof(all32) / sizeof(all32[0]) == 384, "mtgp32 sets");
static_assert(sizeof(all64) / sizeof(all64[0]) == 384, "mtgp64 sets");
static_assert(all32[0].mexp == 11213 && all32[128].mexp == 23209 && all32[256].mexp == 44497, "order");
static_assert(all64[0].mexp == 23209 && all64[128].mexp == 44497 && all64[256].mexp == 110503, "order");

// Dynamic-creator sets (cuda-sample/mtgp32dc-param-11213.c,
// mtgp64dc-param-11213.c), same record form.
#define MTGP32DC_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha) \
    make32(m, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha),
#define MTGP64DC_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha)
constexpr mtgp32_params_fast_t dc32[] = {
#include "dc_param_sets.inc"
};
#undef MTGP32DC_SET
#undef MTGP64DC_SET
#define MTGP32DC_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha)
#define MTGP64DC_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha) \
    make64(m, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha),
constexpr mtgp64_params_fast_t dc64[] = {
#include "dc_param_sets.inc"
};
#undef MTGP32DC_SET
#undef MTGP64DC_SET
static_assert(sizeof(dc32) / sizeof(dc32[0]) == 200 && sizeof(dc64) / sizeof(dc64[0]) == 200, "dc sets");

}  // namespace

#define R4(T, b) T[(b)], T[(b) + 1], T[(b) + 2], T[(b) + 3]
#define R16(T, b) R4(T, b), R4(T, (b) + 4), R4(T, (b) + 8), R4(T, (b) + 12)
#define R64(T, b) R16(T, b), R16(T, (b) + 16), R16(T, (b) + 32), R16(T, (b) + 48)
#define R128(T, b) R64(T, b), R64(T, (b) + 64)
#define R200(T) R128(T, 0), R64(T, 128), R4(T, 192), R4(T, 196)

extern "C" {
mtgp32_params_fast_t mtgp32_params_fast_11213[128] = {R128(all32, 0)};
mtgp32_params_fast_t mtgp32_params_fast_23209[128] = {R128(all32, 128)};
mtgp32_params_fast_t mtgp32_params_fast_44497[128] = {R128(all32, 256)};
mtgp64_params_fast_t mtgp64_params_fast_23209[128] = {R128(all64, 0)};
mtgp64_params_fast_t mtgp64_params_fast_44497[128] = {R128(all64, 128)};
mtgp64_params_fast_t mtgp64_params_fast_110503[128] = {R128(all64, 256)};
mtgp32_params_fast_t mtgp32dc_params_fast_11213[200] = {R200(dc32)};
mtgp64_params_fast_t mtgp64dc_params_fast_11213[200] = {R200(dc64)};
}

namespace mtgp {

int table_of(int bits, int mexp) {
    if (bits == 32)
        return mexp == 11213 ? 0 : mexp == 23209 ? 1 : mexp == 44497 ? 2 : -1;
    if (bits == 64)
        return mexp == 23209 ? 3 : mexp == 44497 ? 4 : mexp == 110503 ? 5 : -1;
    return -1;
}

static DevParam expand32(const mtgp32_params_fast_t &p) {
    DevParam d{};
    d.bits = 32;
    d.mexp = p.mexp;
    d.n = p.mexp / 32 + 1;
    d.pos = p.pos;
    d.sh1 = p.sh1;
    d.sh2 = p.sh2;
    d.mask_lo = p.mask;
    d.mask_hi = 0;
    for (int j = 0; j < 16; j++) {
        d.tbl[j] = p.tbl[j];
        d.tmp[j] = p.tmp_tbl[j];
        d.flt[j] = p.flt_tmp_tbl[j];
    }
    return d;
}

static DevParam expand64(const mtgp64_params_fast_t &p) {
    DevParam d{};
    d.bits = 64;
    d.mexp = p.mexp;
    d.n = p.mexp / 64 + 1;
    d.pos = p.pos;
    d.sh1 = p.sh1;
    d.sh2 = p.sh2;
    d.mask_lo = (uint32_t)p.mask;
    d.mask_hi = (uint32_t)(p.mask >> 32);
    for (int j = 0; j < 16; j++) {
        d.tbl[j] = (uint32_t)(p.tbl[j] >> 32);
        d.tmp[j] = (uint32_t)(p.tmp_tbl[j] >> 32);
        d.flt[j] = (uint32_t)(p.dbl_tmp_tbl[j] >> 32);
    }
    return d;
}

const DevParam &host_param(int gid) {
    static DevParam table[kNumSets];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int k = 0; k < 384; k++) {
            table[k] = expand32(all32[k]);
            table[384 + k] = expand64(all64[k]);
        }
    });
    return table[gid];
}

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
const char *last_error() { return g_error.c_str(); }

}  // namespace mtgp

// ------------------------------------------------------------------ init
namespace {

// Least power of two >= size (mtgp32-fast.c:36-58 computes the same).
int ring_size(int size) {
    int x = 1;
    while (x < size)
        x <<= 1;
    return x;
}

template <class Gen, class Status, class Word, class Params>
int alloc_status(Gen *g, const Params *para, int bits) {
    const int size = para->mexp / bits + 1;
    const int large = ring_size(size);
    Status *st = (Status *)malloc(sizeof(Status) + sizeof(Word) * (size_t)large);
    if (!st)
        return -1;
    memset(st->array, 0, sizeof(Word) * (size_t)large);
    g->params = *para;
    g->status = st;
    st->size = size;
    st->large_size = large;
    st->large_mask = large - 1;
    st->idx = size - 1;
    return 0;
}

// Fill byte of the initial state: the bytes of the hidden seed folded
// (mtgp32-fast.c:126-130; mtgp64 folds the high half, mtgp64-fast.c:131-134).
inline uint32_t fold_byte(uint32_t v) {
    v += v >> 16;
    v += v >> 8;
    return v & 0xff;
}

// Key mixing shared by init_by_array and init_by_str (mtgp32-fast.c:151-235,
// :250-334; mtgp64-fast.c:156-240, :255-339).  Word is uint32_t/uint64_t,
// KeyAt(j) returns the value the reference adds as "array[j] + i" without i.
template <class Word, class Mix1, class Mix2, class KeyAt>
void mix_key(Word *w, int n, Word hidden, int length, Mix1 f1, Mix2 f2, KeyAt key) {
    const int lag = n >= 623 ? 11 : n >= 68 ? 7 : n >= 39 ? 5 : 3;
    const int mid = (n - lag) / 2;
    w[0] = hidden;
    int rounds = (length + 1 > n ? length + 1 : n) - 1;
    Word r = f1(w[0] ^ w[mid] ^ w[n - 1]);
    w[mid] += r;
    r += (Word)length;
    w[(mid + lag) % n] += r;
    w[0] = r;
    int i = 1, j = 0;
    for (; j < rounds; j++) {
        r = f1(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += (j < length ? key(j) : (Word)0) + (Word)i;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (j = 0; j < n; j++) {
        r = f2(w[i] + w[(i + mid) % n] + w[(i + n - 1) % n]);
        w[(i + mid) % n] ^= r;
        r -= (Word)i;
        w[(i + mid + lag) % n] ^= r;
        w[i] = r;
        i = (i + 1) % n;
    }
    if (w[n - 1] == 0)
        w[n - 1] = 0x4d544750u;
}

inline uint32_t m32a(uint32_t x) { return (x ^ (x >> 27)) * 1664525u; }
inline uint32_t m32b(uint32_t x) { return (x ^ (x >> 27)) * 1566083941u; }
inline uint64_t m64a(uint64_t x) { return (x ^ (x >> 59)) * 2173292883993ull; }
inline uint64_t m64b(uint64_t x) { return (x ^ (x >> 59)) * 58885565329898161ull; }

}  // namespace

extern "C" {

void mtgp32_init_state(uint32_t array[], const mtgp32_params_fast_t *para, uint32_t seed) {
    const int n = para->mexp / 32 + 1;
    const uint32_t hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    memset(array, (int)fold_byte(hidden), sizeof(uint32_t) * (size_t)n);
    array[0] = seed;
    array[1] = hidden;
    for (int i = 1; i < n; i++)
        array[i] ^= 1812433253u * (array[i - 1] ^ (array[i - 1] >> 30)) + (uint32_t)i;
}

int mtgp32_init(mtgp32_fast_t *g, const mtgp32_params_fast_t *para, uint32_t seed) {
    if (alloc_status<mtgp32_fast_t, mtgp32_status_fast_t, uint32_t>(g, para, 32))
        return -1;
    mtgp32_init_state(g->status->array, para, seed);
    return 0;
}

int mtgp32_init_by_array(mtgp32_fast_t *g, const mtgp32_params_fast_t *para,
                         uint32_t *array, int length) {
    if (alloc_status<mtgp32_fast_t, mtgp32_status_fast_t, uint32_t>(g, para, 32))
        return -1;
    const int n = g->status->size;
    const uint32_t hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    uint32_t *w = g->status->array;
    memset(w, (int)fold_byte(hidden), sizeof(uint32_t) * (size_t)n);
    mix_key<uint32_t>(w, n, hidden, length, m32a, m32b,
                      [&](int j) { return array[j]; });
    return 0;
}

int mtgp32_init_by_str(mtgp32_fast_t *g, const mtgp32_params_fast_t *para, char *str) {
    if (alloc_status<mtgp32_fast_t, mtgp32_status_fast_t, uint32_t>(g, para, 32))
        return -1;
    const int n = g->status->size;
    const uint32_t hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    uint32_t *w = g->status->array;
    memset(w, (int)fold_byte(hidden), sizeof(uint32_t) * (size_t)n);
    mix_key<uint32_t>(w, n, hidden, (int)strlen(str), m32a, m32b,
                      [&](int j) { return (uint32_t)(int32_t)(signed char)str[j]; });
    return 0;
}

void mtgp32_free(mtgp32_fast_t *g) {
    free(g->status);
    g->status = nullptr;
}

void mtgp32_print_idstring(const mtgp32_fast_t *g, FILE *fp) {
    fprintf(fp, "mtgp32:%d:", g->params.mexp);
    for (int i = 0; i < 20; i++)
        fprintf(fp, "%02x", (unsigned)g->params.poly_sha1[i]);
    fputc('\n', fp);
}

void mtgp64_init_state(uint64_t array[], const mtgp64_params_fast_t *para, uint64_t seed) {
    const int n = para->mexp / 64 + 1;
    const uint64_t hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    memset(array, (int)fold_byte((uint32_t)(hidden >> 32)), sizeof(uint64_t) * (size_t)n);
    array[0] = seed;
    array[1] = hidden;
    for (int i = 1; i < n; i++)
        array[i] ^= 6364136223846793005ull * (array[i - 1] ^ (array[i - 1] >> 62)) + (uint64_t)i;
}

int mtgp64_init(mtgp64_fast_t *g, const mtgp64_params_fast_t *para, uint64_t seed) {
    if (alloc_status<mtgp64_fast_t, mtgp64_status_fast_t, uint64_t>(g, para, 64))
        return -1;
    mtgp64_init_state(g->status->array, para, seed);
    return 0;
}

int mtgp64_init_by_array(mtgp64_fast_t *g, const mtgp64_params_fast_t *para,
                         uint64_t *array, int length) {
    if (alloc_status<mtgp64_fast_t, mtgp64_status_fast_t, uint64_t>(g, para, 64))
        return -1;
    const int n = g->status->size;
    const uint64_t hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    uint64_t *w = g->status->array;
    memset(w, (int)fold_byte((uint32_t)(hidden >> 32)), sizeof(uint64_t) * (size_t)n);
    mix_key<uint64_t>(w, n, hidden, length, m64a, m64b,
                      [&](int j) { return array[j]; });
    return 0;
}

int mtgp64_init_by_str(mtgp64_fast_t *g, const mtgp64_params_fast_t *para, char *str) {
    if (alloc_status<mtgp64_fast_t, mtgp64_status_fast_t, uint64_t>(g, para, 64))
        return -1;
    const int n = g->status->size;
    const uint64_t hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    uint64_t *w = g->status->array;
    memset(w, (int)fold_byte((uint32_t)(hidden >> 32)), sizeof(uint64_t) * (size_t)n);
    mix_key<uint64_t>(w, n, hidden, (int)strlen(str), m64a, m64b,
                      [&](int j) { return (uint64_t)(int64_t)(signed char)str[j]; });
    return 0;
}

void mtgp64_free(mtgp64_fast_t *g) {
    free(g->status);
    g->status = nullptr;
}

void mtgp64_print_idstring(const mtgp64_fast_t *g, FILE *fp) {
    fprintf(fp, "mtgp64:%d:", g->params.mexp);
    for (int i = 0; i < 20; i++)
        fprintf(fp, "%02x", (unsigned)g->params.poly_sha1[i]);
    fputc('\n', fp);
}

}  // extern "C"
