/*
 * xojump.cpp -- jump-ahead for xoshiro256+ (srcs/random.hpp:69-199) by any
 * 128-bit distance.
 *
 * The xoshiro256 transition M is linear over GF(2); advancing by d steps is
 * q(M) with q(x) = x^d mod p(x), p the characteristic polynomial of M.  p is
 * obtained once by Berlekamp-Massey on a state bit; applying q costs 256
 * steps (the same scheme as the generator's published fixed jumps, which are
 * x^(2^128) and x^(2^192) mod p).
 */
#include <cstring>
#include <mutex>
#include <vector>

#include "ftr_host.hpp"

namespace ftrh {

static inline void xo_next(uint64_t* s) {
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = (s[3] << 45) | (s[3] >> 19);
}

void xo_seed(uint64_t seed, uint64_t* st) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    st[i] = z ^ (z >> 31);
  }
}

static Poly g_p;
static std::once_flag g_once;

static void charpoly_init() {
  const int NS = 768;
  std::vector<uint8_t> s(NS), C(NS + 1, 0), B(NS + 1, 0), T;
  uint64_t st[4];
  xo_seed(0x243f6a8885a308d3ULL, st);
  for (int t = 0; t < NS; t++) {
    s[t] = (uint8_t)(st[0] & 1);
    xo_next(st);
  }
  C[0] = B[0] = 1;
  int L = 0, m = 1;
  for (int n = 0; n < NS; n++) {
    uint8_t d = s[n];
    for (int i = 1; i <= L; i++) d ^= (uint8_t)(C[i] & s[n - i]);
    if (!d) {
      m++;
      continue;
    }
    T = C;
    for (int i = 0; i + m <= NS; i++) C[i + m] ^= B[i];
    if (2 * L <= n) {
      L = n + 1 - L;
      B = T;
      m = 1;
    } else {
      m++;
    }
  }
  if (L != 256) throw Error(-3, "xoshiro256 characteristic polynomial: unexpected degree");
  memset(&g_p, 0, sizeof g_p);
  for (int i = 1; i <= L; i++)
    if (C[i]) {
      const int j = L - i;
      g_p.w[j >> 6] |= UINT64_C(1) << (j & 63);
    }
}

static const Poly& charpoly() {
  std::call_once(g_once, charpoly_init);
  return g_p;
}

void xo_charpoly(uint64_t* w5) {
  const Poly& p = charpoly();
  for (int i = 0; i < 4; i++) w5[i] = p.w[i];
  w5[4] = 1;
}

static Poly mulmod(const Poly& a, const Poly& b) {
  const Poly& p = charpoly();
  uint64_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 256; i++) {
    if (!((a.w[i >> 6] >> (i & 63)) & 1)) continue;
    const int ws = i >> 6, bs = i & 63;
    for (int k = 0; k < 4; k++) {
      r[k + ws] ^= b.w[k] << bs;
      if (bs) r[k + ws + 1] ^= b.w[k] >> (64 - bs);
    }
  }
  for (int k = 511; k >= 256; k--) {
    if (!((r[k >> 6] >> (k & 63)) & 1)) continue;
    r[k >> 6] ^= UINT64_C(1) << (k & 63);
    const int sh = k - 256, ws = sh >> 6, bs = sh & 63;
    for (int q = 0; q < 4; q++) {
      r[q + ws] ^= p.w[q] << bs;
      if (bs && q + ws + 1 < 8) r[q + ws + 1] ^= p.w[q] >> (64 - bs);
    }
  }
  Poly o;
  for (int k = 0; k < 4; k++) o.w[k] = r[k];
  return o;
}

Poly xo_poly_square(const Poly& a) { return mulmod(a, a); }
Poly xo_poly_mul(const Poly& a, const Poly& b) { return mulmod(a, b); }

Poly xo_jump_poly(uint64_t lo, uint64_t hi) {
  Poly res, base;
  memset(&res, 0, sizeof res);
  memset(&base, 0, sizeof base);
  res.w[0] = 1;
  base.w[0] = 2;
  int top = 0;
  for (int b = 0; b < 128; b++)
    if (((b < 64 ? lo : hi) >> (b & 63)) & 1) top = b;
  for (int b = 0; b <= top; b++) {
    if (((b < 64 ? lo : hi) >> (b & 63)) & 1) res = mulmod(res, base);
    if (b < top) base = mulmod(base, base);
  }
  return res;
}

void xo_apply(const Poly& q, uint64_t* st) {
  uint64_t acc[4] = {0, 0, 0, 0}, s[4];
  memcpy(s, st, 32);
  for (int i = 0; i < 256; i++) {
    if ((q.w[i >> 6] >> (i & 63)) & 1)
      for (int k = 0; k < 4; k++) acc[k] ^= s[k];
    xo_next(s);
  }
  memcpy(st, acc, 32);
}

}  // namespace ftrh
