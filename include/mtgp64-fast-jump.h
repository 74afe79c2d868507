/*
 * mtgp64-fast-jump.h -- jump-ahead for MTGP64 (drop-in for the reference
 * mtgp64-fast-jump.h:24-36).
 *
 * mtgp64_fast_jump replaces the state s by sum_k c_k T^k s, where the c_k
 * are the coefficients of a jump polynomial given as a hex string (nibble i
 * holds c_{4i}..c_{4i+3}, lowest bit first; the format written by
 * polytostring, mtgp-calc-jump.hpp:34-58).  With c = x^s mod P (P the
 * characteristic polynomial) this advances the stream by s numbers.  In this
 * library the jump runs on the GPU (libmtgp_amd.so); a string with a
 * non-hex character leaves the state unchanged and sets an error readable
 * with mtgp_last_error().  mtgp_hip.h computes jump strings.
 */
#ifndef MTGP64_FAST_JUMP_H
#define MTGP64_FAST_JUMP_H

#include "mtgp64-fast.h"

#if defined(__cplusplus)
extern "C" {
#endif

void mtgp64_fast_jump(mtgp64_fast_t *mtgp64, const char *jump_string);

/* Copies parameters and the whole ring (reference mtgp64-fast-jump.h:26-36);
 * dest->status must already hold large_size words. */
inline static void mtgp64_copy(mtgp64_fast_t *dest, mtgp64_fast_t *src)
{
    dest->params = src->params;
    dest->status->idx = src->status->idx;
    dest->status->size = src->status->size;
    dest->status->large_size = src->status->large_size;
    dest->status->large_mask = src->status->large_mask;
    memcpy(dest->status->array, src->status->array,
           sizeof(uint64_t) * (size_t)src->status->large_size);
}

#if defined(__cplusplus)
}
#endif
#endif /* MTGP64_FAST_JUMP_H */
