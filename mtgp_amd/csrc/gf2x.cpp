// gf2x.cpp -- GF(2)[x] arithmetic for jump-ahead (see gf2x.hpp).
#include "gf2x.hpp"

#include <string.h>
#include <algorithm>
#include <immintrin.h>

namespace mtgp {
namespace gf2 {

namespace {

// ---------------------------------------------------- carry-less 64 x 64
struct U128 {
    uint64_t lo, hi;
};

U128 clmul_portable(uint64_t a, uint64_t b) {
    U128 r{0, 0};
    for (int i = 0; i < 64; i++)
        if (b >> i & 1) {
            r.lo ^= a << i;
            if (i)
                r.hi ^= a >> (64 - i);
        }
    return r;
}

__attribute__((target("pclmul,sse4.1"))) void mul_school_pclmul(const uint64_t *a, int na,
                                                              const uint64_t *b, int nb,
                                                              uint64_t *r) {
    for (int i = 0; i < na; i++) {
        if (!a[i])
            continue;
        const __m128i ai = _mm_set_epi64x(0, (long long)a[i]);
        for (int j = 0; j < nb; j++) {
            const __m128i p = _mm_clmulepi64_si128(ai, _mm_set_epi64x(0, (long long)b[j]), 0x00);
            r[i + j] ^= (uint64_t)_mm_cvtsi128_si64(p);
            r[i + j + 1] ^= (uint64_t)_mm_extract_epi64(p, 1);
        }
    }
}

void mul_school_portable(const uint64_t *a, int na, const uint64_t *b, int nb, uint64_t *r) {
    for (int i = 0; i < na; i++) {
        if (!a[i])
            continue;
        for (int j = 0; j < nb; j++) {
            U128 p = clmul_portable(a[i], b[j]);
            r[i + j] ^= p.lo;
            r[i + j + 1] ^= p.hi;
        }
    }
}

bool have_pclmul() {
    static const bool ok = __builtin_cpu_supports("pclmul") && __builtin_cpu_supports("sse4.1");
    return ok;
}

// r (na + nb words) ^= a * b, schoolbook.
void mul_school(const uint64_t *a, int na, const uint64_t *b, int nb, uint64_t *r) {
    if (have_pclmul())
        mul_school_pclmul(a, na, b, nb, r);
    else
        mul_school_portable(a, na, b, nb, r);
}

// r (2n words, zeroed by the caller) = a * b for n-word a, b (Karatsuba).
void mul_kara(const uint64_t *a, const uint64_t *b, int n, uint64_t *r) {
    if (n <= 24) {
        mul_school(a, n, b, n, r);
        return;
    }
    const int h = n - n / 2, t = n / 2;  // low half h words, high half t <= h
    // z0 = a0 b0 -> r[0, 2h); z2 = a1 b1 -> r[2h, 2h + 2t)
    mul_kara(a, b, h, r);
    std::vector<uint64_t> a1(h, 0), b1(h, 0);
    std::copy(a + h, a + n, a1.begin());
    std::copy(b + h, b + n, b1.begin());
    std::vector<uint64_t> z2(2 * h, 0);
    mul_kara(a1.data(), b1.data(), h, z2.data());
    for (int i = 0; i < 2 * t; i++)
        r[2 * h + i] ^= z2[i];
    // z1 = (a0 + a1)(b0 + b1) + z0 + z2, added at word offset h
    std::vector<uint64_t> sa(h), sb(h), z1(2 * h, 0);
    for (int i = 0; i < h; i++) {
        sa[i] = a[i] ^ a1[i];
        sb[i] = b[i] ^ b1[i];
    }
    mul_kara(sa.data(), sb.data(), h, z1.data());
    for (int i = 0; i < 2 * h; i++)
        z1[i] ^= r[i] ^ z2[i];
    for (int i = 0; i < 2 * h; i++)
        r[h + i] ^= z1[i];
}

inline uint64_t spread32(uint64_t x) {
    x &= 0xffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

int used_words(const Poly &a) {
    int n = (int)a.size();
    while (n > 0 && !a[n - 1])
        n--;
    return n;
}

inline int get_bit(const Poly &a, long i) {
    long w = i >> 6;
    return w < (long)a.size() ? (int)(a[w] >> (i & 63) & 1) : 0;
}

// bits [from, from + count) of a, as a vector of ceil(count/64) words.
Poly extract(const Poly &a, long from, long count) {
    const long nw = (count + 63) / 64;
    Poly r(nw, 0);
    const long w0 = from >> 6;
    const int sh = (int)(from & 63);
    for (long i = 0; i < nw; i++) {
        uint64_t lo = (w0 + i) < (long)a.size() ? a[w0 + i] : 0;
        uint64_t hi = (w0 + i + 1) < (long)a.size() ? a[w0 + i + 1] : 0;
        r[i] = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
    }
    if (count & 63)
        r[nw - 1] &= (1ull << (count & 63)) - 1;
    return r;
}

// a ^= b << s (bit shift), growing a as needed.
void xor_shifted(Poly &a, const Poly &b, long s) {
    const long ws = s >> 6;
    const int sh = (int)(s & 63);
    const long need = ws + (long)b.size() + 1;
    if ((long)a.size() < need)
        a.resize(need, 0);
    for (long i = 0; i < (long)b.size(); i++) {
        if (!b[i])
            continue;
        a[ws + i] ^= b[i] << sh;
        if (sh)
            a[ws + i + 1] ^= b[i] >> (64 - sh);
    }
}

}  // namespace

long degree(const Poly &a) {
    for (long i = (long)a.size() - 1; i >= 0; i--)
        if (a[i])
            return i * 64 + 63 - __builtin_clzll(a[i]);
    return -1;
}

Poly multiply(const Poly &a, const Poly &b) {
    const int na = used_words(a), nb = used_words(b);
    if (!na || !nb)
        return Poly(1, 0);
    const int n = std::max(na, nb);
    if (std::min(na, nb) * 2 < n || n <= 24) {
        Poly r(na + nb + 1, 0);
        mul_school(a.data(), na, b.data(), nb, r.data());
        return r;
    }
    Poly pa(n, 0), pb(n, 0), r(2 * n, 0);
    std::copy(a.begin(), a.begin() + na, pa.begin());
    std::copy(b.begin(), b.begin() + nb, pb.begin());
    mul_kara(pa.data(), pb.data(), n, r.data());
    return r;
}

Poly square(const Poly &a) {
    const int n = used_words(a);
    Poly r(2 * n + 1, 0);
    for (int i = 0; i < n; i++) {
        r[2 * i] = spread32(a[i]);
        r[2 * i + 1] = spread32(a[i] >> 32);
    }
    return r;
}

Modulus::Modulus(const Poly &m) : m_(m) {
    d_ = degree(m);
    w_ = (int)(d_ / 64 + 1);
    m_.resize(w_, 0);
    // mu = floor(x^(2d) / m) by long division: the window R holds the d+1
    // numerator bits [k-d, k]; the numerator x^(2d) has no other set bit.
    Poly R(w_, 0), mu((d_ + 1) / 64 + 1, 0);
    R[d_ >> 6] |= 1ull << (d_ & 63);
    for (long k = 2 * d_; k >= d_; k--) {
        if (R[d_ >> 6] >> (d_ & 63) & 1) {
            for (int i = 0; i < w_; i++)
                R[i] ^= m_[i];
            mu[(k - d_) >> 6] |= 1ull << ((k - d_) & 63);
        }
        for (int i = w_ - 1; i > 0; i--)
            R[i] = (R[i] << 1) | (R[i - 1] >> 63);
        R[0] <<= 1;
    }
    mu_ = mu;
}

Poly Modulus::reduce(const Poly &c) const {
    // Barrett: q = floor(floor(c / x^d) * mu / x^d); c mod m = c + q m.
    Poly hi = extract(c, d_, d_ + 1);
    Poly q = extract(multiply(hi, mu_), d_, d_ + 1);
    Poly qm = multiply(q, m_);
    Poly r(w_, 0);
    for (int i = 0; i < w_; i++) {
        uint64_t cv = i < (int)c.size() ? c[i] : 0;
        uint64_t qv = i < (int)qm.size() ? qm[i] : 0;
        r[i] = cv ^ qv;
    }
    if (d_ & 63)
        r[w_ - 1] &= (1ull << (d_ & 63)) - 1;
    else
        r[w_ - 1] = 0;
    return r;
}

Poly Modulus::reduce_any(const Poly &c) const {
    // Fold the top d bits down repeatedly: c = hi * x^k + lo with
    // deg(hi * x^k) < 2d handled by reduce().
    Poly r = c;
    long dc = degree(r);
    while (dc >= 2 * d_) {
        // split off bits [dc - 2d + 1, dc]: reduce that window, put it back
        const long lo = dc - 2 * d_ + 1;
        Poly top = extract(r, lo, 2 * d_);
        Poly red = reduce(top);  // deg < d
        // r = (r without the window) + red * x^lo
        for (long i = lo; i <= dc; i++)
            if (get_bit(r, i))
                r[i >> 6] ^= 1ull << (i & 63);
        xor_shifted(r, red, lo);
        dc = degree(r);
    }
    return reduce(r);
}

Poly Modulus::mulmod(const Poly &a, const Poly &b) const { return reduce(multiply(a, b)); }
Poly Modulus::sqrmod(const Poly &a) const { return reduce(square(a)); }

Poly Modulus::mulx(const Poly &a) const {
    Poly r(w_, 0);
    for (int i = 0; i < w_; i++) {
        uint64_t v = i < (int)a.size() ? a[i] : 0;
        uint64_t below = (i > 0 && i - 1 < (int)a.size()) ? a[i - 1] : 0;
        r[i] = (v << 1) | (below >> 63);
    }
    if (r[d_ >> 6] >> (d_ & 63) & 1)
        for (int i = 0; i < w_; i++)
            r[i] ^= m_[i];
    return r;
}

Poly Modulus::xpow(const std::vector<uint64_t> &e) const {
    long top = -1;
    for (long i = (long)e.size() * 64 - 1; i >= 0; i--)
        if (e[i >> 6] >> (i & 63) & 1) {
            top = i;
            break;
        }
    Poly r(w_, 0);
    r[0] = 1;
    for (long i = top; i >= 0; i--) {
        r = sqrmod(r);
        if (e[i >> 6] >> (i & 63) & 1)
            r = mulx(r);
    }
    return r;
}

Poly berlekamp_massey(const std::vector<uint8_t> &bits) {
    const long N = (long)bits.size();
    const long nw = N / 64 + 3;
    // rev[k] = s[N-1-k], so that s[n-i] = rev[off + i] with off = N-1-n.
    Poly rev(nw + 2, 0), C(nw, 0), B(nw, 0), T;
    for (long k = 0; k < N; k++)
        if (bits[N - 1 - k] & 1)
            rev[k >> 6] |= 1ull << (k & 63);
    C[0] = B[0] = 1;
    long L = 0, m = 1;
    for (long n = 0; n < N; n++) {
        const long off = N - 1 - n;
        const long cw = L / 64 + 1;
        const long w0 = off >> 6;
        const int sh = (int)(off & 63);
        uint64_t acc = 0;
        for (long q = 0; q < cw; q++) {
            uint64_t lo = rev[w0 + q], hi = (w0 + q + 1) < (long)rev.size() ? rev[w0 + q + 1] : 0;
            uint64_t win = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
            acc ^= C[q] & win;
        }
        if (!(__builtin_popcountll(acc) & 1)) {
            m++;
            continue;
        }
        if (2 * L <= n) {
            T = C;
            xor_shifted(C, B, m);
            L = n + 1 - L;
            B = T;
            m = 1;
        } else {
            xor_shifted(C, B, m);
            m++;
        }
    }
    Poly p(L / 64 + 1, 0);
    for (long i = 0; i <= L; i++)
        if (get_bit(C, i))
            p[(L - i) >> 6] |= 1ull << ((L - i) & 63);
    return p;
}

std::string to_hex(const Poly &a) {
    static const char digits[] = "0123456789abcdef";
    const long d = degree(a);
    std::string s;
    if (d < 0)
        return s;
    s.resize((size_t)(d / 4 + 1));
    for (long i = 0; i <= d / 4; i++)
        s[(size_t)i] = digits[(a[(4 * i) >> 6] >> ((4 * i) & 63)) & 15];
    return s;
}

bool from_hex(const char *hex, Poly &out) {
    const size_t len = strlen(hex);
    out.assign(len * 4 / 64 + 1, 0);
    for (size_t i = 0; i < len; i++) {
        int c = hex[i], v;
        if (c >= '0' && c <= '9')
            v = c - '0';
        else if (c >= 'a' && c <= 'f')
            v = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F')
            v = c - 'A' + 10;
        else
            return false;
        out[(4 * i) >> 6] |= (uint64_t)v << ((4 * i) & 63);
    }
    return true;
}

std::vector<uint32_t> to_words32(const Poly &a, int nwords) {
    std::vector<uint32_t> w((size_t)nwords, 0);
    for (int i = 0; i < nwords; i++) {
        const long k = i / 2;
        if (k < (long)a.size())
            w[(size_t)i] = (uint32_t)(a[k] >> (32 * (i & 1)));
    }
    return w;
}

std::string coeff_string(const Poly &a) {
    const long d = degree(a);
    std::string s;
    for (long i = 0; i <= d; i++)
        s.push_back(get_bit(a, i) ? '1' : '0');
    return s;
}

// ------------------------------------------------------------------ SHA-1
namespace {
inline uint32_t rol(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

void sha1_block(uint32_t h[5], const unsigned char *p) {
    uint32_t w[80];
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 |
               (uint32_t)p[4 * i + 2] << 8 | (uint32_t)p[4 * i + 3];
    for (int i = 16; i < 80; i++)
        w[i] = rol(w[i - 3] ^ w[i - 8] ^ w[i - 14] ^ w[i - 16], 1);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
    for (int i = 0; i < 80; i++) {
        uint32_t f, k;
        if (i < 20) {
            f = (b & c) | (~b & d);
            k = 0x5a827999u;
        } else if (i < 40) {
            f = b ^ c ^ d;
            k = 0x6ed9eba1u;
        } else if (i < 60) {
            f = (b & c) | (b & d) | (c & d);
            k = 0x8f1bbcdcu;
        } else {
            f = b ^ c ^ d;
            k = 0xca62c1d6u;
        }
        uint32_t t = rol(a, 5) + f + e + k + w[i];
        e = d;
        d = c;
        c = rol(b, 30);
        b = a;
        a = t;
    }
    h[0] += a;
    h[1] += b;
    h[2] += c;
    h[3] += d;
    h[4] += e;
}
}  // namespace

void sha1(const void *data, size_t len, unsigned char out[20]) {
    uint32_t h[5] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u, 0xc3d2e1f0u};
    const unsigned char *p = (const unsigned char *)data;
    size_t i = 0;
    for (; i + 64 <= len; i += 64)
        sha1_block(h, p + i);
    unsigned char tail[128];
    size_t rem = len - i;
    memset(tail, 0, sizeof tail);
    memcpy(tail, p + i, rem);
    tail[rem] = 0x80;
    const size_t tl = rem + 9 <= 64 ? 64 : 128;
    const uint64_t bitlen = (uint64_t)len * 8;
    for (int k = 0; k < 8; k++)
        tail[tl - 1 - k] = (unsigned char)(bitlen >> (8 * k));
    sha1_block(h, tail);
    if (tl == 128)
        sha1_block(h, tail + 64);
    for (int k = 0; k < 5; k++) {
        out[4 * k] = (unsigned char)(h[k] >> 24);
        out[4 * k + 1] = (unsigned char)(h[k] >> 16);
        out[4 * k + 2] = (unsigned char)(h[k] >> 8);
        out[4 * k + 3] = (unsigned char)h[k];
    }
}

}  // namespace gf2
}  // namespace mtgp
