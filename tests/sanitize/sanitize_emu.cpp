// The kernels (mtgp_amd/csrc/kernels.hip, init_kernels.hip) compiled as
// host C++ against the stand-in HIP headers of tests/emu, run under
// AddressSanitizer and UndefinedBehaviorSanitizer: generation items of every
// step layout and output kind with segment chains, heads, tails and gaps,
// batched jumps split into parts (and read back as split inputs), and the
// init kernel, each checked bit for bit against the inline genrand family
// of include/mtgp*-fast.h and the host init family.  Test infrastructure:
// built and run by tests/test_sanitizers.py; exit status 0 and a last line
// "SANITIZE-EMU-OK" when every check passed and no sanitizer reported.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "common.hpp"
#include "gf2x.hpp"
#include "kernels.hpp"
#include "mtgp32-fast.h"
#include "mtgp64-fast.h"

extern "C" int emu_run_items(int bits, int cls, int layout, const mtgp::Item *items, int n_items, int grid);
extern "C" int emu_run_jumps(int bits, int cls, const mtgp::JumpItem *items, int n_items, int grid);
extern "C" int emu_run_init(const mtgp::InitItem *items, int n_items);

using namespace mtgp;

static int failures = 0;
#define CHECK(c, ...)                       \
    do {                                    \
        if (!(c)) {                         \
            fprintf(stderr, __VA_ARGS__);   \
            fprintf(stderr, "\n");          \
            failures++;                     \
        }                                   \
    } while (0)

// Host generator of either word size, through the reference-shaped API.
template <int B>
struct Host;
template <>
struct Host<32> {
    using W = uint32_t;
    using P = mtgp32_params_fast_t;
    mtgp32_fast_t g;
    Host(const P &p, uint32_t seed) { mtgp32_init(&g, &p, seed); }
    ~Host() { mtgp32_free(&g); }
    W next(int kind) {
        float f;
        switch (kind) {
        case 0: return mtgp32_genrand_uint32(&g);
        case 1: f = mtgp32_genrand_close1_open2(&g); break;
        case 2: f = mtgp32_genrand_close_open(&g); break;
        case 3: f = mtgp32_genrand_open_close(&g); break;
        default: f = mtgp32_genrand_open_open(&g); break;
        }
        W u;
        memcpy(&u, &f, 4);
        return u;
    }
};
template <>
struct Host<64> {
    using W = uint64_t;
    using P = mtgp64_params_fast_t;
    mtgp64_fast_t g;
    Host(const P &p, uint64_t seed) { mtgp64_init(&g, &p, seed); }
    ~Host() { mtgp64_free(&g); }
    W next(int kind) {
        double f;
        switch (kind) {
        case 0: return mtgp64_genrand_uint64(&g);
        case 1: f = mtgp64_genrand_close1_open2(&g); break;
        case 2: f = mtgp64_genrand_close_open(&g); break;
        case 3: f = mtgp64_genrand_open_close(&g); break;
        default: f = mtgp64_genrand_open_open(&g); break;
        }
        W u;
        memcpy(&u, &f, 8);
        return u;
    }
};

template <class G>
static std::vector<typename G::W> window(const G &h) {
    const auto *st = h.g.status;
    std::vector<typename G::W> w((size_t)st->size);
    for (int j = 0; j < st->size; j++)
        w[(size_t)j] = st->array[(st->idx - st->size + 1 + j + st->large_size) & st->large_mask];
    return w;
}

// One cell: [count a of kind ka to an unaligned output][gap numbers stepped
// over][count b of kind kb], from a seeded window, in `layout`.
template <int B>
static void check_items(const typename Host<B>::P &p, int set_id, int layout, int ka, int kb, uint64_t a, uint64_t gap,
                        uint64_t b, int head) {
    using W = typename Host<B>::W;
    Host<B> h(p, 1000 + (uint32_t)set_id);
    const auto w0 = window(h);
    const int N = (int)w0.size();
    const int cls = kernel_class(B, N, p.pos);
    if (cls < 0 || !layout_supported(B, cls, layout))
        return;
    std::vector<W> out(a + b + 2 * (size_t)head + 8, 0), sout((size_t)N, 0);
    std::vector<Segment> segs(3);
    segs[0] = Segment{(uint64_t)(uintptr_t)(out.data() + head), a, ka, 0, 0};
    segs[1] = Segment{0, gap, 0, 0, 0};
    segs[2] = Segment{(uint64_t)(uintptr_t)(out.data() + head + a + 1), b, kb, 0, 0};
    Item it{};
    it.state_in = (uint64_t)(uintptr_t)w0.data();
    it.state_out = (uint64_t)(uintptr_t)sout.data();
    it.param = set_id;
    it.segs = (uint64_t)(uintptr_t)segs.data();
    it.nsegs = 3;
    CHECK(emu_run_items(B, cls, layout, &it, 1, 1) == 0, "MTGP%d set %d layout %d: run", B, set_id, layout);
    bool ok = true;
    for (uint64_t k = 0; k < a; k++)
        ok &= out[(size_t)head + k] == h.next(ka);
    for (uint64_t k = 0; k < gap; k++)
        (void)h.next(0);
    for (uint64_t k = 0; k < b; k++)
        ok &= out[(size_t)head + a + 1 + k] == h.next(kb);
    ok &= out[(size_t)head + a] == 0 && out[(size_t)head - 1] == 0 && out[(size_t)head + a + 1 + b] == 0;
    CHECK(ok, "MTGP%d set %d layout %d kinds %d/%d: outputs", B, set_id, layout, ka, kb);
    CHECK(sout == window(h), "MTGP%d set %d layout %d: final window", B, set_id, layout);
}

// x^s mod P of a set's characteristic polynomial (Berlekamp-Massey on the
// LSBs after init(0)) as 32-coefficient words.
template <int B>
static std::vector<uint32_t> jump_words(const typename Host<B>::P &p, uint64_t s) {
    Host<B> h(p, 0);
    std::vector<uint8_t> bits((size_t)(2 * p.mexp));
    for (auto &v : bits)
        v = (uint8_t)(h.next(0) & 1);
    gf2::Modulus M(gf2::berlekamp_massey(bits));
    return gf2::to_words32(M.xpow(s), (p.mexp + 31) / 32);
}

// Per-residue masks of one jump (JumpItem.masks layout, bits 0..15).
static int append_jump_masks_host(const std::vector<uint32_t> &w, std::vector<uint64_t> &out) {
    const size_t ncoeff = 32 * w.size();
    const int chunks = (int)((ncoeff + kJumpChunk - 1) / kJumpChunk);
    out.assign((size_t)chunks * 64, 0);
    for (size_t k = 0; k < ncoeff; k++)
        if (w[k / 32] >> (k % 32) & 1)
            out[(k / kJumpChunk) * 64 + k % 64] |= 1ull << ((k % kJumpChunk) / 64);
    return chunks;
}

// A batched jump by s of a seeded window, split into 3 parts; then a second
// jump by s2 whose base is read as the XOR of those parts (in_parts); both
// against stepping (word 0 under the mask: its other bits lie outside the
// state).
template <int B>
static void check_jumps(const typename Host<B>::P &p, int set_id, uint64_t s, uint64_t s2) {
    using W = typename Host<B>::W;
    Host<B> h(p, 77);
    const auto w0 = window(h);
    const int N = (int)w0.size();
    const int cls = kernel_class(B, N, p.pos);
    std::vector<uint64_t> masks, masks2;
    const auto j1 = jump_words<B>(p, s), j2 = jump_words<B>(p, s2);
    const int c1 = append_jump_masks_host(j1, masks), c2 = append_jump_masks_host(j2, masks2);
    const int P = 3;
    std::vector<W> parts((size_t)P * N, 0), out((size_t)N, 0);
    std::vector<JumpItem> items((size_t)P);
    for (int q = 0; q < P; q++) {
        JumpItem &ji = items[(size_t)q];
        ji.state_in = (uint64_t)(uintptr_t)w0.data();
        ji.masks = (uint64_t)(uintptr_t)masks.data();
        ji.states_out = (uint64_t)(uintptr_t)(parts.data() + (size_t)q * N);
        ji.nb = 1;
        ji.c0 = c1 * q / P;
        ji.c1 = c1 * (q + 1) / P;
        ji.param = set_id;
        ji.state_bytes = N * (int)sizeof(W);
    }
    CHECK(emu_run_jumps(B, cls, items.data(), P, 2) == 0, "MTGP%d set %d: jump run", B, set_id);
    JumpItem j{};
    j.state_in = (uint64_t)(uintptr_t)parts.data();
    j.in_parts = P;
    j.in_part_bytes = (uint64_t)N * sizeof(W);
    j.masks = (uint64_t)(uintptr_t)masks2.data();
    j.states_out = (uint64_t)(uintptr_t)out.data();
    j.nb = 1;
    j.c0 = 0;
    j.c1 = c2;
    j.param = set_id;
    j.state_bytes = N * (int)sizeof(W);
    CHECK(emu_run_jumps(B, cls, &j, 1, 1) == 0, "MTGP%d set %d: split-input jump run", B, set_id);
    for (uint64_t k = 0; k < s; k++)
        (void)h.next(0);
    auto w1 = window(h);
    std::vector<W> x((size_t)N, 0);
    for (int q = 0; q < P; q++)
        for (int i = 0; i < N; i++)
            x[(size_t)i] ^= parts[(size_t)q * N + i];
    const W mask = (W)p.mask;
    x[0] &= mask;
    w1[0] &= mask;
    CHECK(x == w1, "MTGP%d set %d: jump by %llu", B, set_id, (unsigned long long)s);
    for (uint64_t k = 0; k < s2; k++)
        (void)h.next(0);
    auto w2 = window(h);
    out[0] &= mask;
    w2[0] &= mask;
    CHECK(out == w2, "MTGP%d set %d: jump from split input", B, set_id);
}

template <int B>
static void check_init(const typename Host<B>::P &p, int set_id) {
    using W = typename Host<B>::W;
    const int N = p.mexp / B + 1;
    std::vector<W> w1((size_t)N), w2((size_t)N), ref((size_t)N);
    std::vector<uint64_t> keys((size_t)N + 5);
    for (size_t i = 0; i < keys.size(); i++)
        keys[i] = 0x9E3779B97F4A7C15ull * (i + 3);
    InitItem it[2] = {};
    it[0].state = (uint64_t)(uintptr_t)w1.data();
    it[0].seed = 4321;
    it[0].hidden = (uint64_t)(p.tbl[4] ^ (p.tbl[8] << 16));
    it[0].n = N;
    it[0].bits = B;
    it[0].mode = 0;
    it[1] = it[0];
    it[1].state = (uint64_t)(uintptr_t)w2.data();
    it[1].seed = (uint64_t)(uintptr_t)keys.data();
    it[1].key_len = N + 5;
    it[1].mode = 1;
    CHECK(emu_run_init(it, 2) == 0, "init run");
    if constexpr (B == 32) {
        mtgp32_init_state(ref.data(), &p, 4321);
        CHECK(w1 == ref, "MTGP32 set %d: device init_state", set_id);
        std::vector<uint32_t> k32(keys.begin(), keys.end());
        mtgp32_fast_t g;
        mtgp32_init_by_array(&g, &p, k32.data(), N + 5);
        std::vector<uint32_t> wa((size_t)N);
        for (int j = 0; j < N; j++)
            wa[(size_t)j] = g.status->array[(g.status->idx - N + 1 + j + g.status->large_size) & g.status->large_mask];
        CHECK(w2 == wa, "MTGP32 set %d: device init_by_array", set_id);
        mtgp32_free(&g);
    } else {
        mtgp64_init_state(ref.data(), &p, 4321);
        CHECK(w1 == ref, "MTGP64 set %d: device init_state", set_id);
        mtgp64_fast_t g;
        mtgp64_init_by_array(&g, &p, keys.data(), N + 5);
        std::vector<uint64_t> wa((size_t)N);
        for (int j = 0; j < N; j++)
            wa[(size_t)j] = g.status->array[(g.status->idx - N + 1 + j + g.status->large_size) & g.status->large_mask];
        CHECK(w2 == wa, "MTGP64 set %d: device init_by_array", set_id);
        mtgp64_free(&g);
    }
}

int main() {
    // table sets: global ids 128 t + i (tables 11213, 23209, 44497 / 23209, 44497, 110503)
    struct S32 {
        const mtgp32_params_fast_t *p;
        int id;
    } s32[] = {{&mtgp32_params_fast_11213[9], 9}, {&mtgp32_params_fast_23209[3], 128 + 3},
               {&mtgp32_params_fast_44497[1], 256 + 1}};
    struct S64 {
        const mtgp64_params_fast_t *p;
        int id;
    } s64[] = {{&mtgp64_params_fast_23209[2], 384 + 2}, {&mtgp64_params_fast_44497[4], 512 + 4},
               {&mtgp64_params_fast_110503[0], 640}};
    int kinds = 0;
    for (const auto &s : s32)
        for (int layout = 0; layout < kLayouts; layout++) {
            check_items<32>(*s.p, s.id, layout, kinds % 5, (kinds + 2) % 5, 3001, 517, 1300 + 64 * layout, 3);
            kinds++;
        }
    for (const auto &s : s64)
        for (int layout = 0; layout < kLayouts; layout++) {
            check_items<64>(*s.p, s.id, layout, kinds % 5, (kinds + 3) % 5, 2100, 300, 999, 1);
            kinds++;
        }
    check_jumps<32>(*s32[0].p, s32[0].id, 5000, 1234);
    check_jumps<64>(*s64[0].p, s64[0].id, 3000, 777);
    check_init<32>(*s32[1].p, s32[1].id);
    check_init<64>(*s64[2].p, s64[2].id);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("SANITIZE-EMU-OK\n");
    return 0;
}
