/*
 * mtgp_hip.h -- GPU entry points of the MI355X MTGP library (libmtgp_amd.so).
 *
 * These are additions to the reference surface (mtgp32-fast.h,
 * mtgp64-fast.h, mtgp*-fast-jump.h, which this library also exports): the
 * reference computes every number on the host through inline genrand_*
 * calls (mtgp32-fast.h:252-346) and its device code exists only as samples
 * (cuda-sample/mtgp32-cuda.cu:171-297: one block per generator, state in
 * LDS, state written back for the next call).  Here the same streams are
 * produced by gfx950 kernels, bit-exact with the host functions.
 *
 * Conventions
 *  - Functions return 0 (or a non-negative count) on success and -1 on
 *    failure; mtgp_last_error() then describes the failure (thread-local).
 *  - A "window" is a generator state as N words oldest first, N =
 *    mexp/32+1 (uint32_t) or mexp/64+1 (uint64_t): the reference ring read
 *    as array[(idx - N + 1 + j) & large_mask].
 *  - Output kinds (MTGP_KIND_*) select genrand_uint32/uint64,
 *    close1_open2, close_open, open_close or open_open; 4-byte outputs for
 *    mtgp32, 8-byte outputs for mtgp64.
 *  - `out` pointers of the device entry points are device memory; the
 *    mtgp*_fill entry points also accept host memory.
 *  - Streams are hipStream_t passed as void* (NULL = the null stream).
 */
#ifndef MTGP_HIP_H
#define MTGP_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "mtgp32-fast.h"
#include "mtgp64-fast.h"

#if defined(__cplusplus)
extern "C" {
#endif

enum {
    MTGP_KIND_INT = 0,          /* genrand_uint32 / genrand_uint64 */
    MTGP_KIND_CLOSE1_OPEN2 = 1, /* [1, 2) */
    MTGP_KIND_CLOSE_OPEN = 2,   /* [0, 1) */
    MTGP_KIND_OPEN_CLOSE = 3,   /* (0, 1] */
    MTGP_KIND_OPEN_OPEN = 4     /* (0, 1) */
};

/* ------------------------------------------- dynamic-creator parameters
 * The 200-set MEXP-11213 tables of the CUDA sample
 * (cuda-sample/mtgp32dc-param-11213.c, mtgp64dc-param-11213.c), usable
 * with the init family, the inline genrand family, mtgp*_fast_jump, the
 * bulk fills and generator sets (mtgp_genset_create_params).  MTGP64-DC
 * sets have N = 176 and N - pos < 256; they run on kernels that step 128
 * words at a time. */
extern mtgp32_params_fast_t mtgp32dc_params_fast_11213[200];
extern mtgp64_params_fast_t mtgp64dc_params_fast_11213[200];

/* ------------------------------------------------------------ plumbing */
const char *mtgp_last_error(void);
/* Makes `device` current for this library (default: HIP's current device). */
int mtgp_set_device(int device);
/* Compute units and HBM bytes of the current device. */
int mtgp_device_info(int *compute_units, size_t *hbm_bytes);
/* The planner's tuning knobs (defaults measured on MI355X; each also set
 * by an environment variable at start-up): "segment_split" (MTGP_SEGMENT_
 * SPLIT, share of each stream in first-round cells of the two-size grid),
 * "small_rounds" (MTGP_SMALL_ROUNDS, small cells per slot of that grid),
 * "segment_rounds" (MTGP_SEGMENT_ROUNDS), "two_size_rho" (MTGP_TWO_SIZE_RHO,
 * largest jump cost / one-round cell size that keeps the two-size grid),
 * "jump_floor_ms" (MTGP_JUMP_FLOOR_MS: beyond it, one round of 4, 8 or 12
 * cells per CU chosen by a step-time model whose jump launch costs this
 * much plus the jumps' own time), "jump_waves_per_cu"
 * (MTGP_JUMP_WAVES), "jump_max_parts" (MTGP_JUMP_MAX_PARTS, most partial
 * jumps a jump is split into), "jump_mfma_min" (MTGP_JUMP_MFMA_MIN: an
 * MTGP32 generator's start-state jumps of one plan run on the matrix cores
 * when there are at least this many, default 256; 0 = never), "overlap"
 * (MTGP_OVERLAP, default two-phase launch).
 * They apply to plans created afterwards; set them while no other thread
 * is creating a plan.  -1 for an unknown name or a value out of range. */
int mtgp_tuning_set(const char *name, double value);
int mtgp_tuning_get(const char *name, double *value);
void *mtgp_dev_malloc(size_t bytes);
int mtgp_dev_free(void *ptr);
int mtgp_memcpy_htod(void *dst, const void *src, size_t bytes);
int mtgp_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (enqueued on `stream`,
 * NULL = default stream).  Reference point for hardware byte counters. */
int mtgp_device_copy(void *dst, const void *src, size_t bytes, void *stream);
/* Fills `bytes` (a multiple of 4) at dst with a 32-bit value by a plain
 * streaming-store kernel (kernel nodes only, so it can be captured in a
 * graph); also the reference point for the HBM write ceiling. */
int mtgp_device_fill(void *dst, uint32_t value, size_t bytes, void *stream);
int mtgp_synchronize(void);
/* Streams and events for callers without their own HIP code (timing). */
void *mtgp_stream_create(void);
int mtgp_stream_destroy(void *stream);
int mtgp_stream_synchronize(void *stream);
void *mtgp_event_create(void);
int mtgp_event_destroy(void *event);
int mtgp_event_record(void *event, void *stream);
int mtgp_event_elapsed_ms(void *start, void *stop, float *ms);

/* ------------------------------------ bulk generation of host generators
 * Continue the stream of a generator created by mtgp32_init* (or
 * mtgp64_init*) on the GPU: writes n numbers of `kind` to out (device or
 * host memory) and leaves `mtgp` exactly where n calls of the matching
 * genrand_* would have left it (same window; the ring is rewritten with
 * idx = size - 1).  Long requests are split across the GPU by jump-ahead.
 * A call of the same shape as the previous one on the device (parameter
 * set, kind, n, out; no tuning change since) relaunches the plan built
 * for it.  Calls on one device serialise.
 * Replaces a loop of mtgp32_genrand_* (mtgp32-fast.h:252-346). */
int mtgp32_fill(mtgp32_fast_t *mtgp32, int kind, void *out, size_t n);
int mtgp64_fill(mtgp64_fast_t *mtgp64, int kind, void *out, size_t n);

/* ------------------------------------------------- jump polynomials
 * NTL-free replacements for calc_characteristic (mtgp32-calc-poly.cpp:
 * 36-45) and calc_jump (mtgp-calc-jump.hpp:135-162).  Results are hex
 * strings in the polytostring format accepted by mtgp*_fast_jump; the
 * return value is the string length (buffer too small: -1).  The
 * characteristic polynomial is checked against params->poly_sha1.
 * `steps` is a little-endian array of nwords 64-bit words. */
int mtgp32_charpoly_hex(const mtgp32_params_fast_t *params, char *buf, size_t cap);
int mtgp64_charpoly_hex(const mtgp64_params_fast_t *params, char *buf, size_t cap);
int mtgp32_jump_hex(const mtgp32_params_fast_t *params, const uint64_t *steps, int nwords,
                    char *buf, size_t cap);
int mtgp64_jump_hex(const mtgp64_params_fast_t *params, const uint64_t *steps, int nwords,
                    char *buf, size_t cap);

/* ------------------------------------------------ no-allocation jump
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same jump and the same resulting state as
 * mtgp32_fast_jump, but the window, coefficient masks, partial windows of
 * the split jump, work items and queue counters the GPU needs live in one
 * static device array of the library
 * (about 133 KiB per GPU, held until the process ends) instead of per-call
 * hipMalloc; the call makes no device allocation or free.  The string may
 * hold at most 16384 hex characters (x^s mod P of every MTGP32 table fits);
 * a longer one, a non-hex character or a set that fits no device kernel
 * returns -1 with mtgp_last_error() set and the state unchanged.  The
 * parameter record is used as given (its poly_sha1 is not checked).
 * Synchronous on the null stream of the current device.  All calls on one
 * GPU share that array: calls that overlap on one GPU (several host threads)
 * race and must be serialised by the caller. */
int mtgp32_fast_jump_noalloc(mtgp32_fast_t *mtgp32, const char *jump_string);

/* -------------------------------------------- device generator sets
 * A set of generators whose states stay in device memory.  Generator g
 * uses parameter set index[g] of the table for (bits[g], mexp[g]). */
typedef struct mtgp_genset mtgp_genset_t;

mtgp_genset_t *mtgp_genset_create(int count, const int *bits, const int *mexp, const int *index);
/* The same from parameter records: params[g] points to an
 * mtgp32_params_fast_t (bits[g] == 32) or mtgp64_params_fast_t (64), e.g. a
 * dynamic-creator set; a record that is not a table set must carry the
 * poly_sha1 of its own characteristic polynomial (it is checked). */
mtgp_genset_t *mtgp_genset_create_params(int count, const int *bits, const void *const *params);
void mtgp_genset_destroy(mtgp_genset_t *set);
int mtgp_genset_count(const mtgp_genset_t *set);
/* Words in generator g's window (N). */
int mtgp_genset_state_words(const mtgp_genset_t *set, int g);
/* mtgp32_init / mtgp64_init with seeds[g] (low 32 bits for mtgp32). */
int mtgp_genset_init_seed(mtgp_genset_t *set, const uint64_t *seeds);
/* mtgp*_init_by_array: generator g uses keys[offsets[g] .. +lengths[g]]
 * (low 32 bits of each key word for mtgp32). */
int mtgp_genset_init_array(mtgp_genset_t *set, const uint64_t *keys, const int *offsets,
                           const int *lengths);
/* The same two, computed on the GPU (one wave per generator, window built
 * in LDS and written to the set's current slot): the same windows, without
 * host work or uploads for large sets.  Synchronous on `stream`. */
int mtgp_genset_init_seed_device(mtgp_genset_t *set, const uint64_t *seeds, void *stream);
int mtgp_genset_init_array_device(mtgp_genset_t *set, const uint64_t *keys, const int *offsets,
                                  const int *lengths, void *stream);
/* Host <-> device copy of one window (uint32_t or uint64_t words). */
int mtgp_genset_set_state(mtgp_genset_t *set, int g, const void *window);
int mtgp_genset_get_state(mtgp_genset_t *set, int g, void *window);
/* Replace generator g's state by its jump by a polynomial in hex (as
 * mtgp*_fast_jump does), on the device. */
int mtgp_genset_jump(mtgp_genset_t *set, int g, const char *jump_hex);
/* Make generator g of `set` the substream `i` of a base state: the base
 * window jumped i times by `jump_hex` (host-composed into one polynomial).
 * Used for the reference's substream scheme (mtgp32-jump.cl:697-731). */
int mtgp_genset_substreams(mtgp_genset_t *set, const void *base_window, const char *jump_hex);

/* One request: `count` numbers of `kind` from generator `gen`, after first
 * skipping `skip` numbers, written to device memory `out`.  Requests on the
 * same generator are consecutive parts of its stream, in array order. */
typedef struct mtgp_request {
    int32_t gen;
    int32_t kind;
    uint64_t count;
    uint64_t skip;
    void *out;
} mtgp_request_t;

/* Batched generation.  A plan fixes the requests and cuts each generator's
 * stream positions on a grid of `segment` numbers (0 = chosen from the size
 * of the batch) into cells: one work item per cell, which starts from a
 * device jump-ahead (or the generator's state) and runs the pieces of the
 * consecutive requests that fall into the cell, stepping over short skips.
 * The work list lives in device memory.  mtgp_plan_launch enqueues, per
 * (word size, state size class), a jump launch for the cells' start states
 * and a persistent generation launch on `stream`; it can be called again
 * (the same numbers are produced again, from the same states) and captured
 * in a hipGraph.  mtgp_plan_commit makes the states after the plan the
 * generators' current states; it fails (-1, nothing changes) if the plan
 * was not launched since it was built or rebased, or if a generator's state
 * has changed since then (another plan's commit, mtgp_genset_jump,
 * set_state, init or substreams), since committing would rewind it.  A
 * second commit of the same launch does nothing. */
typedef struct mtgp_plan mtgp_plan_t;
mtgp_plan_t *mtgp_plan_create(mtgp_genset_t *set, const mtgp_request_t *reqs, int nreq,
                              uint64_t segment);
int mtgp_plan_launch(mtgp_plan_t *plan, void *stream);
int mtgp_plan_commit(mtgp_plan_t *plan);
/* Re-targets a plan at its generators' current states (the same requests,
 * outputs, cells and jump polynomials): after a commit, the next launch
 * continues every stream with the next numbers, the way the reference's
 * device samples continue from the state they wrote back
 * (cuda-sample/mtgp32-cuda.cu:151-161).  Host work is patching the work
 * items' state addresses and one small asynchronous upload on `stream`,
 * which must be the stream the plan launches on (or ordered after its last
 * launch); no polynomial work.  A captured graph stays valid. */
int mtgp_plan_rebase(mtgp_plan_t *plan, void *stream);
void mtgp_plan_destroy(mtgp_plan_t *plan);
/* Tuning knobs (the autotuner's): step layout (0 = four consecutive
 * numbers per lane, 128-bit stores; 1 = lane-strided, conflict-free LDS;
 * 2 = layout 1 with two waves per cell: 512-word steps, or 256-word
 * steps of two words per lane for MTGP32-11213 and MTGP64-23209; for
 * launches whose sets all have N - pos of at least that step and whose
 * start states come from the batched jump launch; 3 = four waves per cell,
 * 1024-word steps for MTGP32-44497 and MTGP64-110503, 512-word steps of two
 * words per lane for MTGP32-23209 and MTGP64-44497 -- launches that do not
 * qualify take the next narrower layout) and resident workgroups per CU; -1 leaves a knob
 * unchanged. */
int mtgp_plan_configure(mtgp_plan_t *plan, int layout, int blocks_per_cu);
/* Two-phase launch order (off by default; MTGP_OVERLAP=1 changes the
 * default): on the automatic two-size grid, mtgp_plan_launch runs the large
 * cells' jumps, then their generation on `stream` while the small cells'
 * jumps and generation run on a side stream of the plan, which `stream`
 * waits for at the end.  Returns 1 if the plan now launches in two phases,
 * 0 if it launches serially (no small cells, or off), -1 on error. */
int mtgp_plan_set_overlap(mtgp_plan_t *plan, int on);
/* Enqueues one part of a plan, in serial order on `stream`: 0 = the
 * start-state jump launches, 1 = the generation launches, one per (word
 * size, class) over every cell (part 1 reads what part 0 wrote: order
 * them). */
int mtgp_plan_launch_part(mtgp_plan_t *plan, int part, void *stream);
/* Records the plan's launches (part 0 then part 1) in a hipGraph on
 * `stream` (not the null stream) and instantiates it; the graph contains
 * kernel nodes only (the work-queue counters reset themselves).
 * mtgp_plan_launch_graph replays it on any stream.  Configuring the plan
 * drops the graph. */
int mtgp_plan_capture(mtgp_plan_t *plan, void *stream);
int mtgp_plan_launch_graph(mtgp_plan_t *plan, void *stream);
/* Autotuner: runs the start-state jumps once, then times the generation
 * launches `reps` times for each step layout and resident workgroups per CU
 * (4..32, capped by occupancy), configures the fastest and makes it the
 * default for later plans of the same word size and state-size class.
 * Writes the choice and its time per generation pass.  The outputs are
 * written (identically) by every trial. */
int mtgp_plan_autotune(mtgp_plan_t *plan, void *stream, int reps, int *layout, int *blocks_per_cu,
                       float *best_ms);
/* Work items, kernel launches of one mtgp_plan_launch (jump and
 * generation), and host planning time of a plan. */
int mtgp_plan_stats(const mtgp_plan_t *plan, int *items, int *launches, double *plan_ms);
/* Start-state jumps of one mtgp_plan_launch: all of them, those run on the
 * matrix cores, and the generator groups those form. */
int mtgp_plan_jump_stats(const mtgp_plan_t *plan, int *jumps, int *matrix_jumps, int *matrix_groups);
/* Generation launch k of mtgp_plan_launch_part(plan, 1, ...), in order:
 * word size, state-size class, step layout, workgroups, work items and the
 * kernel's template name as profilers print it (written to kernel[cap] when
 * kernel is not NULL).  Returns the number of such launches (k < 0: only
 * that), or -1. */
int mtgp_plan_launch_info(const mtgp_plan_t *plan, int k, int *bits, int *cls, int *layout, int *grid,
                          int *items, char *kernel, size_t cap);
/* plan + launch + commit + synchronize.  Requests equal to the previous
 * call's on this set (and no tuning change since) rebase the plan already
 * built instead of planning again. */
int mtgp_genset_generate(mtgp_genset_t *set, const mtgp_request_t *reqs, int nreq, void *stream);

/* ---------------------------------------------------------------- serving
 * A request server over a generator set (no counterpart in the reference,
 * whose samples launch one kernel per call).  Requests are submitted at any
 * time from any thread; a scheduler thread admits them in arrival order
 * while their output fits the output-memory budget (one device arena),
 * coalesces everything admitted into batches -- one plan per state-size
 * class -- and launches each batch as soon as it is planned, on one of
 * several streams.  A batch waits (device-side event) only for earlier
 * in-flight batches that advance one of its generators, so new work
 * overlaps work already on the GPU.  Per-generator order is the submission
 * order: every result is exactly what the same requests made one at a time
 * would produce.  While a server runs it owns the set's states.
 * flags: MTGP_SERVER_PLAIN = one request per launch on a single stream, in
 * arrival order (the baseline); MTGP_SERVER_FUSED_JUMPS, below. */
typedef struct mtgp_server mtgp_server_t;
typedef struct mtgp_server_config {
    size_t arena_bytes;   /* output-memory budget (0: 4 GiB) */
    uint64_t cell;        /* numbers per work item of a large request (power of two; 0: 2^17) */
    int cached_cells;     /* jump polynomials kept per parameter set: x^(k cell), k <= this (0: 512) */
    int streams;          /* device streams (0: 4) */
    int max_inflight;     /* batches on the device at once (0: 4) */
    int max_batch;        /* requests per batch at most (0: 256) */
    int flags;
} mtgp_server_config_t;
#define MTGP_SERVER_PLAIN 1
/* Work items jump to their start state themselves (one launch per batch)
 * instead of a batched jump launch first. */
#define MTGP_SERVER_FUSED_JUMPS 2
typedef struct mtgp_completion {
    int64_t ticket;
    void *out;          /* device address of the output, in the arena */
    uint64_t bytes;
    double submit_us;   /* server clock (microseconds since creation) */
    double launch_us;
    double done_us;     /* when the scheduler saw the batch finish */
} mtgp_completion_t;
typedef struct mtgp_server_stats {
    int64_t submitted, completed, batches, launches;
    double numbers;           /* numbers produced */
    double plan_us;           /* host planning time, total */
    int64_t max_batch_requests;
} mtgp_server_stats_t;
mtgp_server_t *mtgp_server_create(mtgp_genset_t *set, const mtgp_server_config_t *config);
/* Drains, stops the scheduler and frees the arena (outputs included). */
int mtgp_server_destroy(mtgp_server_t *server);
/* Enqueues `count` numbers of `kind` from generator `gen` after skipping
 * `skip` numbers; returns a ticket >= 0, or -1 (mtgp_last_error()). */
int64_t mtgp_server_submit(mtgp_server_t *server, int gen, int kind, uint64_t count, uint64_t skip);
/* Up to `max` completions, oldest first; waits up to wait_ms for one
 * (0: no wait).  Returns how many were written. */
int mtgp_server_poll(mtgp_server_t *server, mtgp_completion_t *out, int max, int wait_ms);
/* Gives a completed request's output space back to the budget. */
int mtgp_server_release(mtgp_server_t *server, int64_t ticket);
/* Waits until every request submitted so far has completed. */
int mtgp_server_drain(mtgp_server_t *server);
int mtgp_server_stats(mtgp_server_t *server, mtgp_server_stats_t *stats);
/* Replays a trace: entry i is submitted `t_us[i]` microseconds after the
 * start, from this thread; a second thread polls the completions and
 * releases them.  Writes each request's latency (completion - submission,
 * microseconds), optionally its queueing delay (launch - submission), and
 * with `digests` != NULL a 64-bit digest of each output (copied to the
 * host: slower, for checking).  Returns 0 or -1. */
int mtgp_server_replay(mtgp_server_t *server, int n, const double *t_us, const int32_t *gen,
                       const int32_t *kind, const uint64_t *count, const uint64_t *skip,
                       double *latency_us, double *queue_us, uint64_t *digests, double *wall_us);
/* The 64-bit digest mtgp_server_replay computes (of any host bytes). */
uint64_t mtgp_digest64(const void *data, size_t bytes);

#if defined(__cplusplus)
}
#endif
#endif /* MTGP_HIP_H */
