/*
 * ref_shim.c -- flat C entry points over the reference MTGP library, so that
 * tests can drive the reference code itself through ctypes.
 *
 * Test infrastructure only.  oracle/Makefile compiles this file together
 * with the reference's own sources, where they lie under the reference tree,
 * into oracle/_ref/*.so.  Exactly one of the variants below is selected:
 *   SHIM_FAST32  mtgp32-fast.{h,c}, mtgp32-param-fast.c, mtgp32-fast-jump.c
 *   SHIM_FAST64  mtgp64-fast.{h,c}, mtgp64-param-fast.c, mtgp64-fast-jump.c
 *   SHIM_REF32   mtgp32-ref.{h,c}, mtgp32-param-ref.c
 *   SHIM_REF64   mtgp64-ref.{h,c}, mtgp64-param-ref.c
 * With SHIM_DC as well (fast variants), mexp 11213 selects the 200-set
 * dynamic-creator table (cuda-sample/mtgp*dc-param-11213.c) instead.
 * The genrand family is `inline static` in the fast headers, so a bulk
 * "fill" loop here is the only way to call it from outside C.
 *
 * Output kinds: 0 integer, 1 [1,2), 2 [0,1), 3 (0,1], 4 (0,1).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#if defined(SHIM_FAST32) || defined(SHIM_REF32)
typedef uint32_t word_t;
#define W 32
#else
typedef uint64_t word_t;
#define W 64
#endif

#if defined(SHIM_FAST32)
#include "mtgp32-fast.h"
#include "mtgp32-fast-jump.h"
typedef mtgp32_fast_t gen_t;
typedef mtgp32_params_fast_t par_t;
#define TBL(m) (m == 11213 ? mtgp32_params_fast_11213 : m == 23209 ? mtgp32_params_fast_23209 : m == 44497 ? mtgp32_params_fast_44497 : NULL)
#define P(x) mtgp32_##x
#elif defined(SHIM_FAST64)
#include "mtgp64-fast.h"
#include "mtgp64-fast-jump.h"
typedef mtgp64_fast_t gen_t;
typedef mtgp64_params_fast_t par_t;
#define TBL(m) (m == 23209 ? mtgp64_params_fast_23209 : m == 44497 ? mtgp64_params_fast_44497 : m == 110503 ? mtgp64_params_fast_110503 : NULL)
#define P(x) mtgp64_##x
#elif defined(SHIM_REF32)
#include "mtgp32-ref.h"
typedef mtgp32_ref_t gen_t;
typedef mtgp32_params_ref_t par_t;
#define TBL(m) (m == 11213 ? mtgp32_params_ref_11213 : m == 23209 ? mtgp32_params_ref_23209 : m == 44497 ? mtgp32_params_ref_44497 : NULL)
#define P(x) mtgp32_##x
#elif defined(SHIM_REF64)
#include "mtgp64-ref.h"
typedef mtgp64_ref_t gen_t;
typedef mtgp64_params_ref_t par_t;
#define TBL(m) (m == 23209 ? mtgp64_params_ref_23209 : m == 44497 ? mtgp64_params_ref_44497 : m == 110503 ? mtgp64_params_ref_110503 : NULL)
#define P(x) mtgp64_##x
#else
#error "select a shim variant"
#endif

#if defined(SHIM_DC)
#if W == 32
extern par_t mtgp32dc_params_fast_11213[];
#define DCTBL mtgp32dc_params_fast_11213
#else
extern par_t mtgp64dc_params_fast_11213[];
#define DCTBL mtgp64dc_params_fast_11213
#endif
static const par_t *lookup(int mexp, int index)
{
    if (mexp != 11213 || index < 0 || index > 199)
        return NULL;
    return DCTBL + index;
}
#else
static const par_t *lookup(int mexp, int index)
{
    const par_t *t = TBL(mexp);
    if (!t || index < 0 || index > 127)
        return NULL;
    return t + index;
}
#endif

void *rs_new_seed(int mexp, int index, uint64_t seed)
{
    const par_t *p = lookup(mexp, index);
    gen_t *g = (gen_t *)malloc(sizeof(gen_t));
    if (!p || !g || P(init)(g, p, (word_t)seed)) {
        free(g);
        return NULL;
    }
    return g;
}

void *rs_new_array(int mexp, int index, const word_t *key, int len)
{
    const par_t *p = lookup(mexp, index);
    gen_t *g = (gen_t *)malloc(sizeof(gen_t));
    if (!p || !g || P(init_by_array)(g, p, (word_t *)key, len)) {
        free(g);
        return NULL;
    }
    return g;
}

void *rs_new_str(int mexp, int index, const char *s)
{
    const par_t *p = lookup(mexp, index);
    gen_t *g = (gen_t *)malloc(sizeof(gen_t));
    if (!p || !g || P(init_by_str)(g, p, (char *)s)) {
        free(g);
        return NULL;
    }
    return g;
}

void rs_free(void *h)
{
    gen_t *g = (gen_t *)h;
    if (g) {
        P(free)(g);
        free(g);
    }
}

int rs_size(void *h) { return ((gen_t *)h)->status->size; }

void rs_fill(void *h, int kind, uint64_t n, void *out)
{
    gen_t *g = (gen_t *)h;
    uint64_t i;
#if W == 32
    float *f = (float *)out;
    uint32_t *u = (uint32_t *)out;
    for (i = 0; i < n; i++) {
        switch (kind) {
        case 0: u[i] = P(genrand_uint32)(g); break;
        case 1: f[i] = P(genrand_close1_open2)(g); break;
        case 2: f[i] = P(genrand_close_open)(g); break;
        case 3: f[i] = P(genrand_open_close)(g); break;
        default: f[i] = P(genrand_open_open)(g); break;
        }
    }
#else
    double *f = (double *)out;
    uint64_t *u = (uint64_t *)out;
    for (i = 0; i < n; i++) {
        switch (kind) {
        case 0: u[i] = P(genrand_uint64)(g); break;
        case 1: f[i] = P(genrand_close1_open2)(g); break;
        case 2: f[i] = P(genrand_close_open)(g); break;
        case 3: f[i] = P(genrand_open_close)(g); break;
        default: f[i] = P(genrand_open_open)(g); break;
        }
    }
#endif
}

/* Canonical window, oldest word first. */
void rs_window(void *h, word_t *out)
{
    gen_t *g = (gen_t *)h;
    int n = g->status->size, j;
#if defined(SHIM_FAST32) || defined(SHIM_FAST64)
    for (j = 0; j < n; j++)
        out[j] = g->status->array[(g->status->idx - n + 1 + j + g->status->large_size)
                                  & g->status->large_mask];
#else
    for (j = 0; j < n; j++)
        out[j] = g->status->array[(g->status->idx + 1 + j) % n];
#endif
}

#if defined(SHIM_FAST32) || defined(SHIM_FAST64)
/* Make the canonical window equal to w (idx = N - 1, as after init). */
void rs_set_window(void *h, const word_t *w)
{
    gen_t *g = (gen_t *)h;
    int n = g->status->size;
    memcpy(g->status->array, w, sizeof(word_t) * n);
    g->status->idx = n - 1;
}

void rs_jump(void *h, const char *hex)
{
    P(fast_jump)((gen_t *)h, hex);
}

void rs_init_state(int mexp, int index, uint64_t seed, word_t *out)
{
    P(init_state)(out, lookup(mexp, index), (word_t)seed);
}

/* Full parameter record, widened to 64-bit words:
 * [mexp,pos,sh1,sh2, tbl[16], tmp_tbl[16], flt/dbl_tmp_tbl[16], mask,
 *  sha1 bytes[20]] = 73 values. */
int rs_params(int mexp, int index, uint64_t *out)
{
    const par_t *p = lookup(mexp, index);
    int j;
    if (!p)
        return -1;
    out[0] = (uint64_t)p->mexp;
    out[1] = (uint64_t)p->pos;
    out[2] = (uint64_t)p->sh1;
    out[3] = (uint64_t)p->sh2;
    for (j = 0; j < 16; j++) {
        out[4 + j] = p->tbl[j];
        out[20 + j] = p->tmp_tbl[j];
#if W == 32
        out[36 + j] = p->flt_tmp_tbl[j];
#else
        out[36 + j] = p->dbl_tmp_tbl[j];
#endif
    }
    out[52] = p->mask;
    for (j = 0; j < 20; j++)
        out[53 + j] = p->poly_sha1[j];
    return 0;
}
#endif

/* The text mtgp*_print_idstring writes. */
int rs_idstring(void *h, char *buf, int len)
{
    char *mem = NULL;
    size_t sz = 0;
    FILE *fp = open_memstream(&mem, &sz);
    if (!fp)
        return -1;
    P(print_idstring)((gen_t *)h, fp);
    fclose(fp);
    snprintf(buf, (size_t)len, "%s", mem);
    free(mem);
    return (int)sz;
}
