// common.hpp -- definitions shared by the host library and the HIP kernels.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <string>

namespace mtgp {

// One piece of a work item's stream: `count` numbers of `kind` written to
// `out`, or, with out == 0, `count` numbers stepped over without output.
struct alignas(16) Segment {
    uint64_t out;
    uint64_t count;
    int32_t kind;
    int32_t pad0;
    uint64_t pad1;
};
static_assert(sizeof(Segment) == 32, "Segment layout");

// One unit of device work: continue (or first jump, then continue) the
// stream of one generator for `count` numbers of one output kind, or, with
// nsegs > 0, through the segments segs[0, nsegs) in order (consecutive
// requests of one generator run by one wave with no jump between them).
//
//  - state_in  : canonical window (N words, oldest first; 64-bit words for
//                MTGP64) the stream continues from.
//  - poly      : optional jump polynomial, 32 coefficients per uint32_t word
//                (bit b of word i = coefficient of x^(32i+b)).  When set the
//                window is first replaced by sum_k c_k T^k window.
//  - out       : output address (device memory), 4 or 8 bytes per number.
//  - state_out : optional; receives the window after the last number.
struct alignas(16) Item {
    uint64_t out;
    uint64_t count;
    uint64_t state_in;
    uint64_t state_out;
    uint64_t poly;
    int32_t poly_words;  // number of uint32_t coefficient words
    int32_t param;       // index into the device parameter table
    int32_t kind;        // 0 int, 1 [1,2), 2 [0,1), 3 (0,1], 4 (0,1)
    int32_t parts;       // >1: the start window is the XOR of `parts`
                         // partial windows, part_bytes apart, at state_in
    uint64_t part_bytes;
    uint64_t segs;       // const Segment * (used when nsegs > 0)
    int32_t nsegs;
    int32_t pad0;
};
static_assert(sizeof(Item) == 80, "Item layout");

// Batched jump: nb (<= kJumpBatch) jumps of one base window, each written
// to states_out + b * state_bytes.  The coefficients come pre-arranged by
// chunk and residue: the 64-bit word masks[c * 64 + rho] holds, in bit
// 16 b + q, the coefficient of x^(c * kJumpChunk + 64 q + rho) of jump b.
constexpr int kJumpBatch = 4;
constexpr int kJumpSteps = 4;                  // 256-word steps per chunk
constexpr int kJumpChunk = kJumpSteps * 256;   // positions per chunk
// Chunks [c0, c1) are folded; the sequence is generated from chunk 0.  A
// jump split into parts writes partial windows whose XOR is the jumped
// window (Item.parts).  The base window may itself be such a split result:
// with in_parts > 1 it is the XOR of in_parts windows in_part_bytes apart
// at state_in.
struct alignas(16) JumpItem {
    uint64_t state_in;
    uint64_t masks;       // const uint64_t *: [chunk][64 residues]
    uint64_t states_out;
    int32_t nb;
    int32_t c0, c1;
    int32_t param;
    int32_t state_bytes;
    int32_t in_parts;     // 0 or 1: one window at state_in
    uint64_t in_part_bytes;
    uint64_t pad;
};
static_assert(sizeof(JumpItem) == 64, "JumpItem layout");

// Jumps on the matrix cores: nj jumps of one MTGP32 base window, as one
// GF(2) product of their coefficient bits with the bit planes of the base's
// sequence (kernels.hip, mtgp32_jump_mfma_kernel).
//  - planes : 32 * pw words the planes kernel fills (positions [0, 32 pw)).
//  - coef   : coefficient words, coef[w * njp + b] = word w of jump b
//             (bit t = coefficient of x^(32 w + t)), kpad / 32 words per
//             jump, jumps zero-padded to njp (a multiple of 32).
//  - out    : jump b's window (n words) at out + b * out_stride bytes.
struct alignas(16) MJumpGroup {
    uint64_t state_in;       // base window; with in_parts > 1 the XOR of
    uint64_t in_part_bytes;  // in_parts windows this many bytes apart
    uint64_t planes;
    uint64_t coef;
    uint64_t out;
    uint64_t out_stride;
    int32_t nj, njp, kpad, pw;
    int32_t param, in_parts, n, pad;
};
static_assert(sizeof(MJumpGroup) == 80, "MJumpGroup layout");

// One wave of the matrix-core jump: jumps [r0, r0 + 32 row_tiles) of a
// group (as far as it has them) and window words [j0, j0 + mj_words).
struct alignas(16) MJumpItem {
    int32_t group, r0, j0, pad;
};
static_assert(sizeof(MJumpItem) == 16, "MJumpItem layout");
constexpr int kMjRowTiles = 2;  // row tiles of 32 jumps per wave, at most
constexpr int kMjWords = 8;     // window words per wave with one row tile
// Words per item of a launch with `row_tiles` row tiles (the same
// accumulator count).
inline int mj_words(int row_tiles) { return kMjWords / row_tiles; }
// Coefficients per jump, rounded up to the 64 of one product step.
inline int mj_kpad(int mexp) { return (mexp + 63) / 64 * 64; }
// Plane words per bit (32 positions each) a group needs: every window the
// products read ends below position 32 pw; even, so the planes kernel's
// 64-position rows tile it.
inline int mj_pw(int kpad, int n) {
    const int pw = kpad / 32 + (n + kMjWords + 31) / 32 + 3;
    return (pw + 1) & ~1;
}

// Device initialisation of one generator's window (mtgp*_init_state or
// mtgp*_init_by_array computed on the GPU).
struct alignas(16) InitItem {
    uint64_t state;    // window out: N words (uint32_t or uint64_t)
    uint64_t seed;     // mode 0: the seed; mode 1: address of key_len uint64_t keys
    uint64_t hidden;   // tbl[4] ^ (tbl[8] << 16) of the parameter set
    int32_t n, bits, key_len, mode;
};
static_assert(sizeof(InitItem) == 48, "InitItem layout");

// Expanded parameter record as the kernels read it.  For MTGP64 the
// tables hold the high 32 bits of each entry (all low halves are zero).
struct alignas(16) DevParam {
    int32_t bits, mexp, n, pos, sh1, sh2;
    uint32_t mask_lo, mask_hi;
    uint32_t tbl[16];
    uint32_t tmp[16];
    uint32_t flt[16];
};
static_assert(sizeof(DevParam) == 224, "DevParam layout");

// Global parameter-set ids: 6 tables x 128 sets, in the order
// mtgp32 11213, 23209, 44497, mtgp64 23209, 44497, 110503.
constexpr int kSetsPerTable = 128;
constexpr int kNumTables = 6;
constexpr int kNumSets = kSetsPerTable * kNumTables;

int table_of(int bits, int mexp);                 // -1 if unknown
inline int global_set_id(int bits, int mexp, int index) {
    int t = table_of(bits, mexp);
    return (t < 0 || index < 0 || index >= kSetsPerTable) ? -1 : t * kSetsPerTable + index;
}
const DevParam &host_param(int gid);              // expanded record
inline int state_words(int bits, int mexp) { return mexp / bits + 1; }

// Thread-local error text (mtgp_last_error).
void set_error(const std::string &msg);
const char *last_error();

}  // namespace mtgp
