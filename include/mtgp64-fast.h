/*
 * mtgp64-fast.h -- drop-in MTGP64 interface of the MI355X MTGP library.
 *
 * Replaces the reference header mtgp64-fast.h: the same struct layouts
 * (mtgp64-fast.h:46-57, :74-80, :86-89), exported functions (:107-118) and
 * per-number inline generators (:148-356).  Bulk generation on the GPU is
 * declared in mtgp_hip.h.  Ring convention as in mtgp32-fast.h.
 */
#ifndef MTGP64_FAST_H
#define MTGP64_FAST_H

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#if defined(__cplusplus)
extern "C" {
#endif

/* One MTGP64 parameter set (reference mtgp64-fast.h:46-57). */
typedef struct MTGP64_PARAMS_FAST_T {
    int mexp;
    int pos;
    int sh1;
    int sh2;
    uint64_t tbl[16];           /* recursion matrix */
    uint64_t tmp_tbl[16];       /* tempering matrix */
    uint64_t dbl_tmp_tbl[16];   /* (tmp_tbl >> 12) | 0x3ff0000000000000 */
    uint64_t mask;
    unsigned char poly_sha1[21];
} mtgp64_params_fast_t;

typedef struct MTGP64_STATUS_FAST_T {
    int idx;
    int size;
    int large_size;
    int large_mask;
    uint64_t array[];
} mtgp64_status_fast_t;

typedef struct MTGP64_FAST_T {
    mtgp64_params_fast_t params;
    mtgp64_status_fast_t *status;
} mtgp64_fast_t;

/* mtgp64-fast.h:95-105 */
extern mtgp64_params_fast_t mtgp64_params_fast_23209[128];
extern mtgp64_params_fast_t mtgp64_params_fast_44497[128];
extern mtgp64_params_fast_t mtgp64_params_fast_110503[128];

/* mtgp64-fast.c:99-109, :124-142, :156-240, :255-339, :347-349, :358-365 */
int mtgp64_init(mtgp64_fast_t *mtgp64, const mtgp64_params_fast_t *para,
                uint64_t seed);
void mtgp64_init_state(uint64_t array[], const mtgp64_params_fast_t *para,
                       uint64_t seed);
int mtgp64_init_by_array(mtgp64_fast_t *mtgp64,
                         const mtgp64_params_fast_t *para,
                         uint64_t *array, int length);
int mtgp64_init_by_str(mtgp64_fast_t *mtgp64,
                       const mtgp64_params_fast_t *para, char *str);
void mtgp64_free(mtgp64_fast_t *mtgp64);
void mtgp64_print_idstring(const mtgp64_fast_t *mtgp64, FILE *fp);

/* Recursion on two 32-bit halves with the halves crossed over
 * (reference mtgp64-fast.h:148-169). */
inline static void mtgp64_do_recursion(uint64_t *r, uint64_t x1, uint64_t x2,
                                       uint64_t y, int sh1, int sh2,
                                       uint64_t mask, uint64_t tbl[16])
{
    uint64_t x = (x1 & mask) ^ x2;
    uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    uint32_t nh, nl;
    hi ^= hi << sh1;
    lo ^= lo << sh1;
    nh = lo ^ ((uint32_t)(y >> 32) >> sh2);
    nl = hi ^ ((uint32_t)y >> sh2);
    *r = (((uint64_t)nh << 32) | nl) ^ tbl[nl & 15];
}

inline static void mtgp64_next_state(mtgp64_fast_t *mtgp64)
{
    mtgp64_status_fast_t *st = mtgp64->status;
    uint32_t m = (uint32_t)st->large_mask;
    int i = (st->idx + 1) & (int)m;
    int back = st->large_size - st->size;
    st->idx = i;
    mtgp64_do_recursion(&st->array[i],
                        st->array[(uint32_t)(i + back) & m],
                        st->array[(uint32_t)(i + back + 1) & m],
                        st->array[(uint32_t)(i + back + mtgp64->params.pos) & m],
                        mtgp64->params.sh1, mtgp64->params.sh2,
                        mtgp64->params.mask, mtgp64->params.tbl);
}

inline static int mtgp64_temper_index(uint64_t t)
{
    t ^= t >> 16;
    t ^= t >> 8;
    return (int)(t & 15);
}

inline static uint64_t mtgp64_temper(const uint64_t tmp_tbl[16], uint64_t r,
                                     uint64_t t)
{
    return r ^ tmp_tbl[mtgp64_temper_index(t)];
}

inline static double mtgp64_bits_to_double(uint64_t u)
{
    double d;
    memcpy(&d, &u, sizeof d);
    return d;
}

/* [1, 2) */
inline static double mtgp64_temper_double(const uint64_t dbl_tmp_tbl[16],
                                          uint64_t r, uint64_t t)
{
    return mtgp64_bits_to_double((r >> 12) ^ dbl_tmp_tbl[mtgp64_temper_index(t)]);
}

/* (1, 2) */
inline static double mtgp64_temper_double_open(const uint64_t dbl_tmp_tbl[16],
                                               uint64_t r, uint64_t t)
{
    return mtgp64_bits_to_double(((r >> 12) | 1) ^ dbl_tmp_tbl[mtgp64_temper_index(t)]);
}

inline static uint64_t mtgp64_temper_helper(const mtgp64_fast_t *mtgp64)
{
    const mtgp64_status_fast_t *st = mtgp64->status;
    return st->array[(uint32_t)(st->idx + mtgp64->params.pos - 1 - st->size
                                + st->large_size) & (uint32_t)st->large_mask];
}

inline static uint64_t mtgp64_genrand_uint64(mtgp64_fast_t *mtgp64)
{
    mtgp64_next_state(mtgp64);
    return mtgp64_temper(mtgp64->params.tmp_tbl,
                         mtgp64->status->array[mtgp64->status->idx],
                         mtgp64_temper_helper(mtgp64));
}

inline static double mtgp64_genrand_close1_open2(mtgp64_fast_t *mtgp64)
{
    mtgp64_next_state(mtgp64);
    return mtgp64_temper_double(mtgp64->params.dbl_tmp_tbl,
                                mtgp64->status->array[mtgp64->status->idx],
                                mtgp64_temper_helper(mtgp64));
}

inline static double mtgp64_genrand_close_open(mtgp64_fast_t *mtgp64)
{
    return mtgp64_genrand_close1_open2(mtgp64) - 1.0;
}

inline static double mtgp64_genrand_open_close(mtgp64_fast_t *mtgp64)
{
    return 2.0 - mtgp64_genrand_close1_open2(mtgp64);
}

inline static double mtgp64_genrand_open_open(mtgp64_fast_t *mtgp64)
{
    mtgp64_next_state(mtgp64);
    return mtgp64_temper_double_open(mtgp64->params.dbl_tmp_tbl,
                                     mtgp64->status->array[mtgp64->status->idx],
                                     mtgp64_temper_helper(mtgp64)) - 1.0;
}

#if defined(__cplusplus)
}
#endif
#endif /* MTGP64_FAST_H */
