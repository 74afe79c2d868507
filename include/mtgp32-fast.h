/*
 * mtgp32-fast.h -- drop-in MTGP32 interface of the MI355X MTGP library.
 *
 * Replaces the reference header mtgp32-fast.h: the same struct layouts
 * (mtgp32-fast.h:47-58, :75-81, :87-90), the same exported functions
 * (:108-119) and the same per-number inline generators (:121-346), so code
 * written against the reference compiles and links unchanged against
 * libmtgp_amd.so.  Bulk generation on the GPU is declared in mtgp_hip.h.
 *
 * State layout: `status->array` is a ring of `large_size` words (the least
 * power of two >= size); `idx` is the newest word.  The live window is the
 * `size` words ending at idx, oldest first:
 *     window[j] = array[(idx - size + 1 + j) & large_mask].
 */
#ifndef MTGP32_FAST_H
#define MTGP32_FAST_H

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#if defined(__cplusplus)
extern "C" {
#endif

/* One MTGP32 parameter set (reference mtgp32-fast.h:47-58). */
typedef struct MTGP32_PARAMS_FAST_T {
    int mexp;                   /* Mersenne exponent */
    int pos;                    /* pick-up position */
    int sh1;                    /* left shift of the recursion */
    int sh2;                    /* right shift of the recursion */
    uint32_t tbl[16];           /* recursion matrix, indexed by 4 bits */
    uint32_t tmp_tbl[16];       /* tempering matrix */
    uint32_t flt_tmp_tbl[16];   /* (tmp_tbl >> 9) | 0x3f800000 */
    uint32_t mask;              /* keeps mexp bits of the state */
    unsigned char poly_sha1[21];/* SHA-1 of the characteristic polynomial */
} mtgp32_params_fast_t;

/* Ring-buffer state (reference mtgp32-fast.h:75-81). */
typedef struct MTGP32_STATUS_FAST_T {
    int idx;
    int size;
    int large_size;
    int large_mask;
    uint32_t array[];
} mtgp32_status_fast_t;

/* Generator = parameters + state (reference mtgp32-fast.h:87-90). */
typedef struct MTGP32_FAST_T {
    mtgp32_params_fast_t params;
    mtgp32_status_fast_t *status;
} mtgp32_fast_t;

/* 128 parameter sets per Mersenne exponent (mtgp32-fast.h:96-106). */
extern mtgp32_params_fast_t mtgp32_params_fast_11213[128];
extern mtgp32_params_fast_t mtgp32_params_fast_23209[128];
extern mtgp32_params_fast_t mtgp32_params_fast_44497[128];

/* mtgp32-fast.c:94-104.  Allocates status; 0 on success, -1 on failure. */
int mtgp32_init(mtgp32_fast_t *mtgp32, const mtgp32_params_fast_t *para,
                uint32_t seed);
/* mtgp32-fast.c:119-137.  Fills the first mexp/32+1 words of array. */
void mtgp32_init_state(uint32_t array[], const mtgp32_params_fast_t *para,
                       uint32_t seed);
/* mtgp32-fast.c:151-235. */
int mtgp32_init_by_array(mtgp32_fast_t *mtgp32,
                         const mtgp32_params_fast_t *para,
                         uint32_t *array, int length);
/* mtgp32-fast.c:250-334 (characters are taken as signed char). */
int mtgp32_init_by_str(mtgp32_fast_t *mtgp32,
                       const mtgp32_params_fast_t *para, char *str);
/* mtgp32-fast.c:342-344. */
void mtgp32_free(mtgp32_fast_t *mtgp32);
/* mtgp32-fast.c:353-360: "mtgp32:<mexp>:<sha1 hex>\n". */
void mtgp32_print_idstring(const mtgp32_fast_t *mtgp32, FILE *fp);

/* ---- per-number generation on the host (reference mtgp32-fast.h:149-346) */

/* One step of the recursion: the word `size` positions after x1. */
inline static void mtgp32_do_recursion(uint32_t *r, uint32_t x1, uint32_t x2,
                                       uint32_t y, int sh1, int sh2,
                                       uint32_t mask, uint32_t tbl[16])
{
    uint32_t x = (x1 & mask) ^ x2;
    x ^= x << sh1;
    y = x ^ (y >> sh2);
    *r = y ^ tbl[y & 15];
}

/* Appends one word to the ring. */
inline static void mtgp32_next_state(mtgp32_fast_t *mtgp32)
{
    mtgp32_status_fast_t *st = mtgp32->status;
    uint32_t m = (uint32_t)st->large_mask;
    int i = (st->idx + 1) & (int)m;
    int back = st->large_size - st->size;   /* i - size, kept non-negative */
    st->idx = i;
    mtgp32_do_recursion(&st->array[i],
                        st->array[(uint32_t)(i + back) & m],
                        st->array[(uint32_t)(i + back + 1) & m],
                        st->array[(uint32_t)(i + back + mtgp32->params.pos) & m],
                        mtgp32->params.sh1, mtgp32->params.sh2,
                        mtgp32->params.mask, mtgp32->params.tbl);
}

/* Index into the tempering tables taken from the helper word t. */
inline static int mtgp32_temper_index(uint32_t t)
{
    t ^= t >> 16;
    t ^= t >> 8;
    return (int)(t & 15);
}

inline static uint32_t mtgp32_temper(const uint32_t tmp_tbl[16], uint32_t r,
                                     uint32_t t)
{
    return r ^ tmp_tbl[mtgp32_temper_index(t)];
}

inline static float mtgp32_bits_to_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

/* Tempered value as a float in [1, 2). */
inline static float mtgp32_temper_float(const uint32_t flt_tmp_tbl[16],
                                        uint32_t r, uint32_t t)
{
    return mtgp32_bits_to_float((r >> 9) ^ flt_tmp_tbl[mtgp32_temper_index(t)]);
}

/* Tempered value as a float in (1, 2): the lowest mantissa bit forced on. */
inline static float mtgp32_temper_float_open(const uint32_t flt_tmp_tbl[16],
                                             uint32_t r, uint32_t t)
{
    return mtgp32_bits_to_float(((r >> 9) ^ flt_tmp_tbl[mtgp32_temper_index(t)]) | 1u);
}

/* Helper word of the newest output: array[idx + pos - 1 - size]. */
inline static uint32_t mtgp32_temper_helper(const mtgp32_fast_t *mtgp32)
{
    const mtgp32_status_fast_t *st = mtgp32->status;
    return st->array[(uint32_t)(st->idx + mtgp32->params.pos - 1 - st->size
                                + st->large_size) & (uint32_t)st->large_mask];
}

inline static uint32_t mtgp32_genrand_uint32(mtgp32_fast_t *mtgp32)
{
    mtgp32_next_state(mtgp32);
    return mtgp32_temper(mtgp32->params.tmp_tbl,
                         mtgp32->status->array[mtgp32->status->idx],
                         mtgp32_temper_helper(mtgp32));
}

/* [1, 2) */
inline static float mtgp32_genrand_close1_open2(mtgp32_fast_t *mtgp32)
{
    mtgp32_next_state(mtgp32);
    return mtgp32_temper_float(mtgp32->params.flt_tmp_tbl,
                               mtgp32->status->array[mtgp32->status->idx],
                               mtgp32_temper_helper(mtgp32));
}

/* [0, 1) */
inline static float mtgp32_genrand_close_open(mtgp32_fast_t *mtgp32)
{
    return mtgp32_genrand_close1_open2(mtgp32) - 1.0F;
}

/* (0, 1] */
inline static float mtgp32_genrand_open_close(mtgp32_fast_t *mtgp32)
{
    return 2.0F - mtgp32_genrand_close1_open2(mtgp32);
}

/* (0, 1) */
inline static float mtgp32_genrand_open_open(mtgp32_fast_t *mtgp32)
{
    mtgp32_next_state(mtgp32);
    return mtgp32_temper_float_open(mtgp32->params.flt_tmp_tbl,
                                    mtgp32->status->array[mtgp32->status->idx],
                                    mtgp32_temper_helper(mtgp32)) - 1.0F;
}

#if defined(__cplusplus)
}
#endif
#endif /* MTGP32_FAST_H */
