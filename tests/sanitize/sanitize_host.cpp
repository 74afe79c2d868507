// Host library objects (host.cpp: parameter tables and the init family;
// gf2x.cpp: GF(2)[x], Berlekamp-Massey, hex codecs, SHA-1) and the inline
// genrand family of include/mtgp*-fast.h, run under AddressSanitizer and
// UndefinedBehaviorSanitizer.  Test infrastructure: built and run by
// tests/test_sanitizers.py; exit status 0 and a last line "SANITIZE-HOST-OK"
// when every check passed and no sanitizer reported.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "gf2x.hpp"
#include "mtgp32-fast.h"
#include "mtgp64-fast.h"
#include "mtgp_hip.h"

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

template <class Fast>
static std::vector<uint64_t> window_of(const Fast &g) {
    const auto *st = g.status;
    std::vector<uint64_t> w((size_t)st->size);
    for (int j = 0; j < st->size; j++)
        w[(size_t)j] = st->array[(st->idx - st->size + 1 + j + st->large_size) & st->large_mask];
    return w;
}

// Characteristic polynomial from 2 MEXP output LSBs after init(0) (the
// reference's calc_characteristic pattern) and its SHA-1 against the
// record's poly_sha1.
template <class Fast, class Params, class Gen>
static void check_charpoly(const Params &p, int (*init)(Fast *, const Params *, decltype(p.mask)), Gen gen,
                           void (*release)(Fast *), const char *what) {
    Fast g;
    CHECK(init(&g, &p, 0) == 0, "%s: init", what);
    std::vector<uint8_t> bits((size_t)(2 * p.mexp));
    for (auto &b : bits)
        b = (uint8_t)(gen(&g) & 1);
    release(&g);
    const gf2::Poly P = gf2::berlekamp_massey(bits);
    const std::string c = gf2::coeff_string(P);
    unsigned char d[20];
    gf2::sha1(c.data(), c.size(), d);
    CHECK(memcmp(d, p.poly_sha1, 20) == 0, "%s: poly_sha1 differs", what);
    // x^(3^162) mod P, through the hex codecs and back
    gf2::Modulus M(P);
    std::vector<uint64_t> e(5, 0);
    // 3^162 as little-endian 64-bit words
    std::vector<uint32_t> big{1};
    for (int i = 0; i < 162; i++) {
        uint64_t carry = 0;
        for (auto &v : big) {
            const uint64_t t = (uint64_t)v * 3 + carry;
            v = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry)
            big.push_back((uint32_t)carry);
    }
    for (size_t i = 0; i < big.size(); i++)
        e[i / 2] |= (uint64_t)big[i] << (32 * (i % 2));
    const gf2::Poly J = M.xpow(e);
    const std::string hex = gf2::to_hex(J);
    gf2::Poly back;
    CHECK(gf2::from_hex(hex.c_str(), back), "%s: from_hex", what);
    CHECK(gf2::to_hex(M.reduce_any(back)) == hex, "%s: hex round trip", what);
    CHECK(gf2::to_hex(M.mulmod(J, M.xpow(1))) == gf2::to_hex(M.mulx(J)), "%s: mulx", what);
}

template <class Fast, class Params, class Word>
static void check_init(const Params &p, int bits, const char *what) {
    const int N = p.mexp / bits + 1;
    // init_by_array with key lengths 0 .. 2N + 3 and init_by_str with signed bytes
    for (int len : {0, 1, 4, N - 1, N, N + 3, 2 * N + 3}) {
        std::vector<Word> key((size_t)len);
        for (int i = 0; i < len; i++)
            key[(size_t)i] = (Word)(0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
        Fast g;
        int rc;
        if constexpr (sizeof(Word) == 4)
            rc = mtgp32_init_by_array(&g, &p, key.data(), len);
        else
            rc = mtgp64_init_by_array(&g, &p, key.data(), len);
        CHECK(rc == 0, "%s: init_by_array %d", what, len);
        const auto w = window_of(g);
        bool nonzero = false;
        for (uint64_t v : w)
            nonzero |= v != 0;
        CHECK(nonzero, "%s: zero state", what);
        if constexpr (sizeof(Word) == 4)
            mtgp32_free(&g);
        else
            mtgp64_free(&g);
    }
    char s[] = "MTGP\xff\x80\x7f!";
    Fast g;
    int rc;
    if constexpr (sizeof(Word) == 4)
        rc = mtgp32_init_by_str(&g, &p, s);
    else
        rc = mtgp64_init_by_str(&g, &p, s);
    CHECK(rc == 0, "%s: init_by_str", what);
    // the whole float family through the inline headers
    for (int k = 0; k < 3 * N + 7; k++) {
        if constexpr (sizeof(Word) == 4) {
            (void)mtgp32_genrand_uint32(&g);
            const float f = mtgp32_genrand_close1_open2(&g) + mtgp32_genrand_close_open(&g) +
                            mtgp32_genrand_open_close(&g) + mtgp32_genrand_open_open(&g);
            CHECK(f > 1.0f && f < 5.0f, "%s: float range", what);
        } else {
            (void)mtgp64_genrand_uint64(&g);
            const double f = mtgp64_genrand_close1_open2(&g) + mtgp64_genrand_close_open(&g) +
                             mtgp64_genrand_open_close(&g) + mtgp64_genrand_open_open(&g);
            CHECK(f > 1.0 && f < 5.0, "%s: double range", what);
        }
    }
    if constexpr (sizeof(Word) == 4)
        mtgp32_free(&g);
    else
        mtgp64_free(&g);
}

int main() {
    const int m32[3] = {11213, 23209, 44497}, m64[3] = {23209, 44497, 110503};
    mtgp32_params_fast_t *t32[3] = {mtgp32_params_fast_11213, mtgp32_params_fast_23209, mtgp32_params_fast_44497};
    mtgp64_params_fast_t *t64[3] = {mtgp64_params_fast_23209, mtgp64_params_fast_44497, mtgp64_params_fast_110503};
    char what[64];
    for (int t = 0; t < 3; t++)
        for (int i : {0, 77, 127}) {
            snprintf(what, sizeof what, "MTGP32-%d #%d", m32[t], i);
            check_init<mtgp32_fast_t, mtgp32_params_fast_t, uint32_t>(t32[t][i], 32, what);
            snprintf(what, sizeof what, "MTGP64-%d #%d", m64[t], i);
            check_init<mtgp64_fast_t, mtgp64_params_fast_t, uint64_t>(t64[t][i], 64, what);
        }
    for (int i : {0, 199}) {
        snprintf(what, sizeof what, "MTGP32-DC #%d", i);
        check_init<mtgp32_fast_t, mtgp32_params_fast_t, uint32_t>(mtgp32dc_params_fast_11213[i], 32, what);
        snprintf(what, sizeof what, "MTGP64-DC #%d", i);
        check_init<mtgp64_fast_t, mtgp64_params_fast_t, uint64_t>(mtgp64dc_params_fast_11213[i], 64, what);
    }
    check_charpoly<mtgp32_fast_t>(t32[0][0], mtgp32_init, [](mtgp32_fast_t *g) { return (uint64_t)mtgp32_genrand_uint32(g); },
                                  mtgp32_free, "charpoly MTGP32-11213 #0");
    check_charpoly<mtgp64_fast_t>(t64[0][5], mtgp64_init, [](mtgp64_fast_t *g) { return mtgp64_genrand_uint64(g); },
                                  mtgp64_free, "charpoly MTGP64-23209 #5");
    check_charpoly<mtgp64_fast_t>(mtgp64dc_params_fast_11213[3], mtgp64_init,
                                  [](mtgp64_fast_t *g) { return mtgp64_genrand_uint64(g); }, mtgp64_free,
                                  "charpoly MTGP64-DC #3");
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("SANITIZE-HOST-OK\n");
    return 0;
}
