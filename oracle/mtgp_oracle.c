/*
 * mtgp_oracle.c -- CPU restatement of MTGP, used ONLY as a test checker.
 *
 * This file is test infrastructure.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only to check or time
 * against; the product library (mtgp_amd/) never links or calls it.
 *
 * It restates, in plain C and in the simplest form that is still fast
 * enough for tests, the algorithms of the reference library:
 *   - parameter expansion        mtgp32-fast.h:47-58, mtgp64-fast.h:46-57
 *   - recursion / tempering      mtgp32-fast.h:149-239, mtgp64-fast.h:148-249
 *   - genrand family             mtgp32-fast.h:252-346, mtgp64-fast.h:262-356
 *   - init / init_by_array / str mtgp32-fast.c:119-334, mtgp64-fast.c:124-339
 *   - jump apply                 mtgp32-fast-jump.c:29-79 (and 64-bit twin)
 *   - GF(2)[x] jump polynomials  mtgp-calc-jump.hpp:34-162 (NTL PowerXMod,
 *                                polytostring / stringtopoly), and the
 *                                characteristic polynomial of
 *                                mtgp32-calc-poly.cpp:36-45 (NTL MinPolySeq)
 *                                by textbook Berlekamp-Massey.
 *
 * State convention: a generator state is its canonical window w[0..N-1],
 * oldest word first (the reference ring read as
 * array[(idx - N + 1 + j) & large_mask], mtgp32-fast.h:165-183).  Output k
 * is temper(w[N+k], w[k+pos-1]) with w[N+k] = rec(w[k], w[k+1], w[k+pos]).
 *
 * Pinning (what checks each part; see tests/):
 *   - params: compact tables expand to the reference tables (golden file
 *     written from the reference build) and to poly_sha1 via BM + SHA-1;
 *   - generation / init: bit-compare with the reference library built from
 *     its own sources (oracle/_ref) and with committed golden vectors;
 *   - jump: compared with the reference mtgp32_fast_jump, with the shipped
 *     3^162 jump string and jump table, and with stepping (test-jump32.cpp);
 *   - polynomials: x^(3^162) mod CHAR32_11213 against the shipped string,
 *     table[k+1] == table[k]^4 mod P, BM against poly_sha1.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

typedef struct {
    int bits;        /* 32 or 64 */
    int mexp;
    int index;       /* index of the set in its table */
    int n;           /* state words: mexp / bits + 1 */
    int pos, sh1, sh2;
    uint64_t mask;
    uint64_t tbl[16], tmp[16], flt[16];
    char sha1[41];
} orc_params_t;

typedef struct {
    int bits, mexp, index, pos, sh1, sh2;
    uint32_t b[8];
    const char *sha1;
} orc_set_t;

#define MTGP32_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha) \
    {32, m, i, p, s1, s2, {a0, a1, a2, a3, t0, t1, t2, t3}, sha},
#define MTGP64_SET(m, i, p, s1, s2, a0, a1, a2, a3, t0, t1, t2, t3, sha) \
    {64, m, i, p, s1, s2, {a0, a1, a2, a3, t0, t1, t2, t3}, sha},
static const orc_set_t orc_sets[] = {
#include "../mtgp_amd/csrc/param_sets.inc"
};
#undef MTGP32_SET
#undef MTGP64_SET

/* Expand one compact set to the full record (mtgp32-fast.h:47-58): tbl and
 * tmp_tbl are the GF(2) span of four basis vectors, flt_tmp_tbl is
 * (tmp >> 9) | 0x3f800000 (64-bit: (tmp >> 12) | 0x3ff0...), mask keeps the
 * top mexp - (N-1)*bits bits of word 0. */
int orc_params_get(int bits, int mexp, int index, orc_params_t *p)
{
    size_t k;
    for (k = 0; k < sizeof(orc_sets) / sizeof(orc_sets[0]); k++) {
        const orc_set_t *s = &orc_sets[k];
        int j, b;
        if (s->bits != bits || s->mexp != mexp || s->index != index)
            continue;
        memset(p, 0, sizeof(*p));
        p->bits = bits;
        p->mexp = mexp;
        p->index = index;
        p->n = mexp / bits + 1;
        p->pos = s->pos;
        p->sh1 = s->sh1;
        p->sh2 = s->sh2;
        for (j = 0; j < 16; j++) {
            for (b = 0; b < 4; b++) {
                if (j >> b & 1) {
                    uint64_t t = s->b[b], m = s->b[4 + b];
                    if (bits == 64) {
                        t <<= 32;
                        m <<= 32;
                    }
                    p->tbl[j] ^= t;
                    p->tmp[j] ^= m;
                }
            }
            if (bits == 32)
                p->flt[j] = (p->tmp[j] >> 9) | 0x3f800000u;
            else
                p->flt[j] = (p->tmp[j] >> 12) | 0x3ff0000000000000ull;
        }
        if (bits == 32)
            p->mask = (0xffffffffull << (32 * p->n - mexp)) & 0xffffffffull;
        else
            p->mask = ~0ull << (64 * p->n - mexp);
        memcpy(p->sha1, s->sha1, 40);
        p->sha1[40] = 0;
        return 0;
    }
    return -1;
}

/* ---------------------------------------------------------------- 32-bit */

/* mtgp32-fast.h:149-159 */
static uint32_t rec32(const orc_params_t *p, uint32_t x1, uint32_t x2, uint32_t y)
{
    uint32_t x = (x1 & (uint32_t)p->mask) ^ x2;
    x ^= x << p->sh1;
    y = x ^ (y >> p->sh2);
    return y ^ (uint32_t)p->tbl[y & 15];
}

/* mtgp32-fast.h:192-198 (index part shared by temper / temper_float) */
static int temper_index(uint64_t t)
{
    t ^= t >> 16;
    t ^= t >> 8;
    return (int)(t & 15);
}

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static double u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

/* Output kinds: 0 integer, 1 [1,2), 2 [0,1), 3 (0,1], 4 (0,1)
 * (mtgp32-fast.h:252-346, mtgp64-fast.h:262-356). */
static void emit32(const orc_params_t *p, int kind, uint32_t r, uint32_t t, void *out, uint64_t k)
{
    int ti = temper_index(t);
    uint32_t u;
    switch (kind) {
    case 0:
        ((uint32_t *)out)[k] = r ^ (uint32_t)p->tmp[ti];
        return;
    case 1:
        ((uint32_t *)out)[k] = (r >> 9) ^ (uint32_t)p->flt[ti];
        return;
    case 2:
        ((float *)out)[k] = u2f((r >> 9) ^ (uint32_t)p->flt[ti]) - 1.0f;
        return;
    case 3:
        ((float *)out)[k] = 2.0f - u2f((r >> 9) ^ (uint32_t)p->flt[ti]);
        return;
    default:
        u = ((r >> 9) ^ (uint32_t)p->flt[ti]) | 1u;
        ((float *)out)[k] = u2f(u) - 1.0f;
        return;
    }
}

#define ORC_CHUNK 4096

/* Continue the stream of window `w` (N words) by `count` outputs of `kind`,
 * writing them to `out` (may be NULL to only advance) and updating w. */
int orc32_generate(const orc_params_t *p, uint32_t *w, int kind, uint64_t count, void *out)
{
    int n = p->n, pos = p->pos;
    uint32_t *b = (uint32_t *)malloc(sizeof(uint32_t) * (n + ORC_CHUNK));
    uint64_t done = 0;
    if (!b)
        return -1;
    memcpy(b, w, sizeof(uint32_t) * n);
    while (done < count) {
        uint64_t c = count - done < ORC_CHUNK ? count - done : ORC_CHUNK;
        uint64_t k;
        for (k = 0; k < c; k++) {
            b[n + k] = rec32(p, b[k], b[k + 1], b[k + pos]);
            if (out)
                emit32(p, kind, b[n + k], b[k + pos - 1], out, done + k);
        }
        memmove(b, b + c, sizeof(uint32_t) * n);
        done += c;
    }
    memcpy(w, b, sizeof(uint32_t) * n);
    free(b);
    return 0;
}

/* mtgp32-fast.c:119-137 */
void orc32_init_state(const orc_params_t *p, uint32_t seed, uint32_t *w)
{
    int i, n = p->n;
    uint32_t hidden = (uint32_t)p->tbl[4] ^ ((uint32_t)p->tbl[8] << 16);
    uint32_t fill = hidden;
    fill += fill >> 16;
    fill += fill >> 8;
    memset(w, (int)(fill & 0xff), sizeof(uint32_t) * n);
    w[0] = seed;
    w[1] = hidden;
    for (i = 1; i < n; i++)
        w[i] ^= 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
}

static uint32_t mix1_32(uint32_t x) { return (x ^ (x >> 27)) * 1664525u; }
static uint32_t mix2_32(uint32_t x) { return (x ^ (x >> 27)) * 1566083941u; }

/* The shared body of mtgp32_init_by_array (mtgp32-fast.c:151-235) and
 * mtgp32_init_by_str (:250-334).  key(j) gives the j-th key element already
 * converted the way the reference adds it ("r += array[j] + i" with a
 * uint32_t array or a signed char string). */
static void init_mix32(const orc_params_t *p, uint32_t *w, int length,
                       const uint32_t *key32, const char *keystr)
{
    int n = p->n, lag, mid, i, j, count;
    uint32_t hidden, fill, r;
    lag = n >= 623 ? 11 : n >= 68 ? 7 : n >= 39 ? 5 : 3;
    mid = (n - lag) / 2;
    hidden = (uint32_t)p->tbl[4] ^ ((uint32_t)p->tbl[8] << 16);
    fill = hidden;
    fill += fill >> 16;
    fill += fill >> 8;
    memset(w, (int)(fill & 0xff), sizeof(uint32_t) * n);
    w[0] = hidden;
    count = length + 1 > n ? length + 1 : n;
    r = mix1_32(w[0] ^ w[mid] ^ w[n - 1]);
    w[mid] += r;
    r += (uint32_t)length;
    w[(mid + lag) % n] += r;
    w[0] = r;
    count--;
    for (i = 1, j = 0; j < count && j < length; j++) {
        uint32_t kv = key32 ? key32[j] + (uint32_t)i
                            : (uint32_t)((int)(signed char)keystr[j] + i);
        r = mix1_32(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += kv;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (; j < count; j++) {
        r = mix1_32(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += (uint32_t)i;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (j = 0; j < n; j++) {
        r = mix2_32(w[i] + w[(i + mid) % n] + w[(i + n - 1) % n]);
        w[(i + mid) % n] ^= r;
        r -= (uint32_t)i;
        w[(i + mid + lag) % n] ^= r;
        w[i] = r;
        i = (i + 1) % n;
    }
    if (w[n - 1] == 0)
        w[n - 1] = 0x4d544750u;
}

void orc32_init_by_array(const orc_params_t *p, const uint32_t *key, int len, uint32_t *w)
{
    init_mix32(p, w, len, key, NULL);
}

void orc32_init_by_str(const orc_params_t *p, const char *s, uint32_t *w)
{
    init_mix32(p, w, (int)strlen(s), NULL, s);
}

/* ---------------------------------------------------------------- 64-bit */

/* mtgp64-fast.h:148-169 */
static uint64_t rec64(const orc_params_t *p, uint64_t x1, uint64_t x2, uint64_t y)
{
    uint64_t x = (x1 & p->mask) ^ x2;
    uint32_t xh = (uint32_t)(x >> 32), xl = (uint32_t)x;
    uint32_t yh = (uint32_t)(y >> 32), yl = (uint32_t)y;
    xh ^= xh << p->sh1;
    xl ^= xl << p->sh1;
    yh = xl ^ (yh >> p->sh2);
    yl = xh ^ (yl >> p->sh2);
    return (((uint64_t)yh << 32) | yl) ^ p->tbl[yl & 15];
}

static void emit64(const orc_params_t *p, int kind, uint64_t r, uint64_t t, void *out, uint64_t k)
{
    int ti = temper_index(t);
    switch (kind) {
    case 0:
        ((uint64_t *)out)[k] = r ^ p->tmp[ti];
        return;
    case 1:
        ((uint64_t *)out)[k] = (r >> 12) ^ p->flt[ti];
        return;
    case 2:
        ((double *)out)[k] = u2d((r >> 12) ^ p->flt[ti]) - 1.0;
        return;
    case 3:
        ((double *)out)[k] = 2.0 - u2d((r >> 12) ^ p->flt[ti]);
        return;
    default:
        ((double *)out)[k] = u2d(((r >> 12) | 1) ^ p->flt[ti]) - 1.0;
        return;
    }
}

int orc64_generate(const orc_params_t *p, uint64_t *w, int kind, uint64_t count, void *out)
{
    int n = p->n, pos = p->pos;
    uint64_t *b = (uint64_t *)malloc(sizeof(uint64_t) * (n + ORC_CHUNK));
    uint64_t done = 0;
    if (!b)
        return -1;
    memcpy(b, w, sizeof(uint64_t) * n);
    while (done < count) {
        uint64_t c = count - done < ORC_CHUNK ? count - done : ORC_CHUNK;
        uint64_t k;
        for (k = 0; k < c; k++) {
            b[n + k] = rec64(p, b[k], b[k + 1], b[k + pos]);
            if (out)
                emit64(p, kind, b[n + k], b[k + pos - 1], out, done + k);
        }
        memmove(b, b + c, sizeof(uint64_t) * n);
        done += c;
    }
    memcpy(w, b, sizeof(uint64_t) * n);
    free(b);
    return 0;
}

/* mtgp64-fast.c:124-142 */
void orc64_init_state(const orc_params_t *p, uint64_t seed, uint64_t *w)
{
    int i, n = p->n;
    uint64_t hidden = p->tbl[4] ^ (p->tbl[8] << 16);
    uint64_t fill = hidden >> 32;
    fill += fill >> 16;
    fill += fill >> 8;
    memset(w, (int)(fill & 0xff), sizeof(uint64_t) * n);
    w[0] = seed;
    w[1] = hidden;
    for (i = 1; i < n; i++)
        w[i] ^= 6364136223846793005ull * (w[i - 1] ^ (w[i - 1] >> 62)) + (uint64_t)i;
}

static uint64_t mix1_64(uint64_t x) { return (x ^ (x >> 59)) * 2173292883993ull; }
static uint64_t mix2_64(uint64_t x) { return (x ^ (x >> 59)) * 58885565329898161ull; }

/* mtgp64-fast.c:156-240 and :255-339 */
static void init_mix64(const orc_params_t *p, uint64_t *w, int length,
                       const uint64_t *key64, const char *keystr)
{
    int n = p->n, lag, mid, i, j, count;
    uint64_t hidden, fill, r;
    lag = n >= 623 ? 11 : n >= 68 ? 7 : n >= 39 ? 5 : 3;
    mid = (n - lag) / 2;
    hidden = p->tbl[4] ^ (p->tbl[8] << 16);
    fill = hidden >> 32;
    fill += fill >> 16;
    fill += fill >> 8;
    memset(w, (int)(fill & 0xff), sizeof(uint64_t) * n);
    w[0] = hidden;
    count = length + 1 > n ? length + 1 : n;
    r = mix1_64(w[0] ^ w[mid] ^ w[n - 1]);
    w[mid] += r;
    r += (uint64_t)length;
    w[(mid + lag) % n] += r;
    w[0] = r;
    count--;
    for (i = 1, j = 0; j < count && j < length; j++) {
        uint64_t kv = key64 ? key64[j] + (uint64_t)i
                            : (uint64_t)(int64_t)((int)(signed char)keystr[j] + i);
        r = mix1_64(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += kv;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (; j < count; j++) {
        r = mix1_64(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += (uint64_t)i;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (j = 0; j < n; j++) {
        r = mix2_64(w[i] + w[(i + mid) % n] + w[(i + n - 1) % n]);
        w[(i + mid) % n] ^= r;
        r -= (uint64_t)i;
        w[(i + mid + lag) % n] ^= r;
        w[i] = r;
        i = (i + 1) % n;
    }
    if (w[n - 1] == 0)
        w[n - 1] = 0x4d544750u;
}

void orc64_init_by_array(const orc_params_t *p, const uint64_t *key, int len, uint64_t *w)
{
    init_mix64(p, w, len, key, NULL);
}

void orc64_init_by_str(const orc_params_t *p, const char *s, uint64_t *w)
{
    init_mix64(p, w, (int)strlen(s), NULL, s);
}

/* ------------------------------------------------------------------ jump */

static int hexval(int c)
{
    if (!isxdigit(c))
        return -1;
    c = tolower(c);
    return c >= 'a' ? c - 'a' + 10 : c - '0';
}

/* mtgp32-fast-jump.c:51-79: s <- sum_k c_k T^k s, where the coefficients
 * c_k come from the hex string (nibble i holds c_{4i}..c_{4i+3}, lowest
 * bit first).  Written out on the canonical window this is
 * new[j] = XOR over k with c_k = 1 of w[k + j]. Returns -1 on a non-hex
 * character (the reference asserts). */
int orc32_jump(const orc_params_t *p, uint32_t *w, const char *hex)
{
    int n = p->n, j;
    size_t len = strlen(hex), k, steps = 4 * len;
    uint32_t *seq, *acc;
    for (k = 0; k < len; k++)
        if (hexval((unsigned char)hex[k]) < 0)
            return -1;
    seq = (uint32_t *)malloc(sizeof(uint32_t) * (steps + n));
    acc = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
    if (!seq || !acc) {
        free(seq);
        free(acc);
        return -2;
    }
    memcpy(seq, w, sizeof(uint32_t) * n);
    for (k = 0; k < steps; k++)
        seq[n + k] = rec32(p, seq[k], seq[k + 1], seq[k + p->pos]);
    for (k = 0; k < steps; k++) {
        if (hexval((unsigned char)hex[k / 4]) >> (k % 4) & 1)
            for (j = 0; j < n; j++)
                acc[j] ^= seq[k + j];
    }
    memcpy(w, acc, sizeof(uint32_t) * n);
    free(seq);
    free(acc);
    return 0;
}

int orc64_jump(const orc_params_t *p, uint64_t *w, const char *hex)
{
    int n = p->n, j;
    size_t len = strlen(hex), k, steps = 4 * len;
    uint64_t *seq, *acc;
    for (k = 0; k < len; k++)
        if (hexval((unsigned char)hex[k]) < 0)
            return -1;
    seq = (uint64_t *)malloc(sizeof(uint64_t) * (steps + n));
    acc = (uint64_t *)calloc((size_t)n, sizeof(uint64_t));
    if (!seq || !acc) {
        free(seq);
        free(acc);
        return -2;
    }
    memcpy(seq, w, sizeof(uint64_t) * n);
    for (k = 0; k < steps; k++)
        seq[n + k] = rec64(p, seq[k], seq[k + 1], seq[k + p->pos]);
    for (k = 0; k < steps; k++) {
        if (hexval((unsigned char)hex[k / 4]) >> (k % 4) & 1)
            for (j = 0; j < n; j++)
                acc[j] ^= seq[k + j];
    }
    memcpy(w, acc, sizeof(uint64_t) * n);
    free(seq);
    free(acc);
    return 0;
}

/* ------------------------------------------------------------ GF(2)[x] */
/* Polynomials are little-endian bit vectors in uint64_t words: bit i of the
 * vector is the coefficient of x^i.  Everything here is schoolbook. */

static int bit_get(const uint64_t *a, long i) { return (int)(a[i >> 6] >> (i & 63) & 1); }
static void bit_flip(uint64_t *a, long i) { a[i >> 6] ^= 1ull << (i & 63); }

/* Degree of a polynomial with `words` words, -1 for zero. */
long orc_poly_deg(const uint64_t *a, long words)
{
    long i;
    for (i = words - 1; i >= 0; i--)
        if (a[i])
            return i * 64 + 63 - __builtin_clzll(a[i]);
    return -1;
}

/* out = a * b mod m, all of `words` words; deg m = d, deg a,b < d. */
void orc_poly_mulmod(const uint64_t *a, const uint64_t *b, const uint64_t *m,
                     long words, uint64_t *out)
{
    long d = orc_poly_deg(m, words), i, k;
    uint64_t *r = (uint64_t *)calloc((size_t)words, 8);
    /* Horner over the bits of b from the top: r = r*x mod m, r ^= a. */
    for (i = d - 1; i >= 0; i--) {
        int top = bit_get(r, d - 1);
        for (k = words - 1; k > 0; k--)
            r[k] = (r[k] << 1) | (r[k - 1] >> 63);
        r[0] <<= 1;
        if (top) {
            for (k = 0; k < words; k++)
                r[k] ^= m[k];   /* clears bit d too */
        }
        if (bit_get(b, i))
            for (k = 0; k < words; k++)
                r[k] ^= a[k];
    }
    memcpy(out, r, (size_t)words * 8);
    free(r);
}

/* out = x^e mod m; the exponent is a little-endian array of ewords words. */
void orc_poly_xpow_mod(const uint64_t *e, long ewords, const uint64_t *m,
                       long words, uint64_t *out)
{
    long ebits = ewords * 64, i;
    uint64_t *r = (uint64_t *)calloc((size_t)words, 8);
    uint64_t *x = (uint64_t *)calloc((size_t)words, 8);
    r[0] = 1;
    x[0] = 2;   /* the moduli used here have degree >= 2, so x mod m = x */
    while (ebits > 0 && !bit_get(e, ebits - 1))
        ebits--;
    for (i = ebits - 1; i >= 0; i--) {
        orc_poly_mulmod(r, r, m, words, r);
        if (bit_get(e, i))
            orc_poly_mulmod(r, x, m, words, r);
    }
    memcpy(out, r, (size_t)words * 8);
    free(r);
    free(x);
}

/* Berlekamp-Massey over GF(2): minimal polynomial of the bit sequence s of
 * length len (the connection polynomial reversed, so that the result is the
 * characteristic polynomial with the NTL MinPolySeq convention used in
 * mtgp32-calc-poly.cpp:36-45).  out must hold `words` words.  Returns the
 * degree L. */
long orc_bm(const uint8_t *s, long len, uint64_t *out, long words)
{
    long n, L = 0, m = 1, i;
    uint64_t *C = (uint64_t *)calloc((size_t)words, 8);
    uint64_t *B = (uint64_t *)calloc((size_t)words, 8);
    uint64_t *T = (uint64_t *)calloc((size_t)words, 8);
    C[0] = 1;
    B[0] = 1;
    for (n = 0; n < len; n++) {
        int d = s[n] & 1;
        for (i = 1; i <= L; i++)
            if (bit_get(C, i))
                d ^= s[n - i] & 1;
        if (!d) {
            m++;
            continue;
        }
        memcpy(T, C, (size_t)words * 8);
        for (i = 0; i + m < words * 64; i++)
            if (bit_get(B, i))
                bit_flip(C, i + m);
        if (2 * L <= n) {
            L = n + 1 - L;
            memcpy(B, T, (size_t)words * 8);
            m = 1;
        } else {
            m++;
        }
    }
    /* characteristic polynomial = x^L C(1/x) */
    memset(out, 0, (size_t)words * 8);
    for (i = 0; i <= L; i++)
        if (bit_get(C, i))
            bit_flip(out, L - i);
    free(C);
    free(B);
    free(T);
    return L;
}

/* mtgp-calc-jump.hpp:34-58 (polytostring): hex digits, nibble i holding
 * coefficients 4i..4i+3 lowest bit first, up to the degree's nibble. */
void orc_poly_to_hex(const uint64_t *a, long words, char *out)
{
    long d = orc_poly_deg(a, words), i;
    static const char digits[] = "0123456789abcdef";
    for (i = 0; i <= d / 4 && d >= 0; i++)
        out[i] = digits[(a[(4 * i) >> 6] >> ((4 * i) & 63)) & 15];
    out[d >= 0 ? d / 4 + 1 : 0] = 0;
}

/* mtgp-calc-jump.hpp:93-126 (stringtopoly); returns -1 on a bad digit. */
int orc_hex_to_poly(const char *hex, uint64_t *out, long words)
{
    long i, len = (long)strlen(hex);
    memset(out, 0, (size_t)words * 8);
    for (i = 0; i < len; i++) {
        int v = hexval((unsigned char)hex[i]), b;
        if (v < 0)
            return -1;
        for (b = 0; b < 4; b++)
            if (v >> b & 1) {
                if (4 * i + b >= words * 64)
                    return -2;
                bit_flip(out, 4 * i + b);
            }
    }
    return 0;
}
