// gf2x.hpp -- polynomials over GF(2) for MTGP jump-ahead, without NTL.
//
// The reference computes characteristic and jump polynomials with NTL
// (MinPolySeq in mtgp32-calc-poly.cpp:36-45, PowerXMod in
// mtgp-calc-jump.hpp:135-162).  This is the library's own replacement:
// carry-less multiplication (PCLMULQDQ when the host CPU has it), Barrett
// reduction, square-and-multiply powers of x, word-packed Berlekamp-Massey,
// and the hex / word-array formats of mtgp-calc-jump.hpp:34-126.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

namespace mtgp {
namespace gf2 {

// Little-endian bit vector: bit i of the vector is the coefficient of x^i.
using Poly = std::vector<uint64_t>;

long degree(const Poly &a);  // -1 for the zero polynomial
Poly multiply(const Poly &a, const Poly &b);
Poly square(const Poly &a);

// Arithmetic modulo a fixed polynomial m of degree d >= 2.
class Modulus {
  public:
    explicit Modulus(const Poly &m);
    long deg() const { return d_; }
    int words() const { return w_; }            // words of a reduced value
    const Poly &poly() const { return m_; }
    Poly reduce(const Poly &c) const;           // any c with deg c < 2d
    Poly reduce_any(const Poly &c) const;       // any degree
    Poly mulmod(const Poly &a, const Poly &b) const;
    Poly sqrmod(const Poly &a) const;
    Poly mulx(const Poly &a) const;             // a * x mod m
    // x^e mod m; e is a little-endian array of 64-bit words.
    Poly xpow(const std::vector<uint64_t> &e) const;
    Poly xpow(uint64_t e) const { return xpow(std::vector<uint64_t>{e}); }

  private:
    Poly m_, mu_;  // mu = floor(x^(2d) / m)
    long d_;
    int w_;
};

// Minimal polynomial of a bit sequence (bits[i] = s_i), Berlekamp-Massey.
// Returned with NTL's MinPolySeq convention: monic, sum_i p_i s_{n+i} = 0.
Poly berlekamp_massey(const std::vector<uint8_t> &bits);

// polytostring / stringtopoly (mtgp-calc-jump.hpp:34-58, :93-126).
std::string to_hex(const Poly &a);
bool from_hex(const char *hex, Poly &out);
// polytoarray (mtgp-calc-jump.hpp:67-91): bit i of word i/32.
std::vector<uint32_t> to_words32(const Poly &a, int nwords);
// The '0'/'1' coefficient string, degree 0 first (what poly_sha1 hashes).
std::string coeff_string(const Poly &a);

// SHA-1 (FIPS 180-4) of a byte string.
void sha1(const void *data, size_t len, unsigned char out[20]);

}  // namespace gf2
}  // namespace mtgp
