/*
 * ftr_math.hpp -- FP64 elementary functions of the integration kernels:
 * log, sin/cos, atan2 and a reciprocal, cheaper than the library's (which
 * spend 80-100 instructions on double-double accuracy), each within two ulp.
 *
 * Compiled as HIP for the device (hipRTC, embedded ahead of ftr_device.hpp)
 * and as plain C++ by the CPU accuracy test (tests/test_math.py), which runs
 * every function against a 50-digit reference.  The only primitives that
 * differ between the two are below: on the host the exact quotient stands in
 * for v_rcp_f64 (whose error the two Newton steps remove either way), and
 * frexp and the table's address space take their host forms.
 */
#pragma once

#if defined(__HIP__) || defined(__HIPCC_RTC__)
#define FTR_MATH_FN __device__ inline
#define FTR_MATH_TABLE __constant__
namespace ftrm {
FTR_MATH_FN double rcp_seed(double d) { return __builtin_amdgcn_rcp(d); }
FTR_MATH_FN double frexp_mant(double y) { return __builtin_amdgcn_frexp_mant(y); }
FTR_MATH_FN int frexp_exp(double y) { return __builtin_amdgcn_frexp_exp(y); }
}  // namespace ftrm
#else
#include <cmath>
#include <cstdint>
#define FTR_MATH_FN inline
#define FTR_MATH_TABLE static const
namespace ftrm {
inline double rcp_seed(double d) { return 1.0 / d; }
inline double frexp_mant(double y) {
  int e;
  return std::frexp(y, &e);
}
inline int frexp_exp(double y) {
  int e;
  std::frexp(y, &e);
  return e;
}
}  // namespace ftrm
#endif

namespace ftrm {

/* high 64 bits of the 128-bit product */
FTR_MATH_FN uint64_t umulhi(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

/* 1/d from a reciprocal seed and two Newton steps: within an ulp of the IEEE
 * quotient for finite nonzero normal d (five instructions on the device
 * instead of the eleven of the IEEE division sequence).  Zero or infinite d
 * give NaN where the division gives an infinity or zero; the callers pass
 * positive weights and pivots, and a sample where one is zero or infinite
 * is non-finite, hence dropped, either way. */
FTR_MATH_FN double rcp_nr(double d) {
  double r = rcp_seed(d);
  r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
  return r;
}

/* Natural log of a normal y > 0, without special cases (the samplers call
 * it on 1 - U, which lies in [2^-53, 1]).  The reduction is fdlibm's
 * (e_log.c): y = m 2^k with m in [sqrt(1/2), sqrt(2)), f = m - 1,
 * s = f / (2 + f), and
 *   log(1 + f) = f - (f^2/2 - s (f^2/2 + R)),  R = 2 atanh(s)/s - 2 = z P(z),
 * z = s^2.  P has degree 6, fitted by a weighted Remez exchange on
 * z in [0, 0.02944] (error of R below 2.5e-18; log_poly_fit.py).  Error
 * below 0.82 ulp, in about 35 device instructions (the library's: ~80). */
FTR_MATH_FN double log_pos(double y) {
  double m = frexp_mant(y);
  int k = frexp_exp(y);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m += m;
    k -= 1;
  }
  const double f = m - 1.0, d = 2.0 + f;
  const double s = f * rcp_nr(d), z = s * s;
  double P = 0x1.2f0563862e5fdp-3;
  P = __builtin_fma(z, P, 0x1.39a1bab73fcacp-3);
  P = __builtin_fma(z, P, 0x1.74663ee86e841p-3);
  P = __builtin_fma(z, P, 0x1.c71c52095d16fp-3);
  P = __builtin_fma(z, P, 0x1.24924941f124fp-2);
  P = __builtin_fma(z, P, 0x1.999999997fdb7p-2);
  P = __builtin_fma(z, P, 0x1.5555555555592p-1);
  const double hfsq = 0.5 * f * f, dk = (double)k;
  /* log 2 = 0x1.62e42p-1 (k times it is exact) + 0x1.fdf473de6af28p-22 */
  const double t = __builtin_fma(s, __builtin_fma(z, P, hfsq), dk * 0x1.fdf473de6af28p-22);
  return __builtin_fma(dk, 0x1.62e42p-1, f - (hfsq - t));
}

/* Natural log of any double without branches: log_pos (subnormals scaled
 * by 2^54 first); zero gives -inf, negative or NaN x NaN, +inf +inf, as the
 * library's log. */
FTR_MATH_FN double log_any(double x) {
  const bool sub = x < 0x1p-1022;
  double r = log_pos(sub ? x * 0x1p54 : x) - (sub ? 54.0 * 0x1.62e42fefa39efp-1 : 0.0);
  if (x == 0.0) r = -__builtin_inf();
  if (!(x >= 0.0)) r = __builtin_nan("");
  if (x == __builtin_inf()) r = x;
  return r;
}

/* 2/pi in binary, most significant bit first, after 64 zero bits (so that
 * a window may start up to 64 bits before the binary point). */
FTR_MATH_TABLE uint64_t two_over_pi_bits[20] = {
    0x0000000000000000ull, 0xa2f9836e4e441529ull, 0xfc2757d1f534ddc0ull, 0xdb6295993c439041ull,
    0xfe5163abdebbc561ull, 0xb7246e3a424dd2e0ull, 0x06492eea09d1921cull, 0xfe1deb1cb129a73eull,
    0xe88235f52ebb4484ull, 0xe99c7026b45f7e41ull, 0x3991d639835339f4ull, 0x9c845f8bbdf9283bull,
    0x1ff897ffde05980full, 0xef2f118b5a0a6d1full, 0x6d367ecf27cb09b7ull, 0x4f463f669e5fea2dull,
    0x7527bac7ebe5f17bull, 0x3d0739f78a5292eaull, 0x6bfb5fb11f8d5d08ull, 0x56033046fc7b6babull};

/* Payne-Hanek reduction of a finite |x| > 2^20 (integer arithmetic): with
 * |x| = M 2^E (M the 53-bit significand) and 2/pi = sum b_i 2^-i, the bits
 * with i <= E - 3 only add multiples of 8 to |x| 2/pi, so the 192-bit window
 * S of b_(E-2) .. b_(E+189) gives |x| 2/pi = M S 2^-189 (mod 8) with an
 * error below 2^-137.  The bits 61.. of the 245-bit product M S come from
 * three 64x64 multiplies.  Returns r = (|x| 2/pi - k) pi/2 with k the
 * nearest integer, |r| <= pi/4, and k mod 4 in *q, both for x itself (sign
 * applied). */
FTR_MATH_FN double reduce_pio2_large(double x, int* q) {
  const uint64_t bits = __builtin_bit_cast(uint64_t, x);
  const int eb = (int)((bits >> 52) & 0x7ff);
  const uint64_t M = (bits & 0xfffffffffffffull) | 0x10000000000000ull;
  /* b_i sits at table bit i + 63; the window starts at i = E - 2 = eb - 1077 */
  int t = eb - 1014;
  t = t < 0 ? 0 : (t > 1033 ? 1033 : t);
  const int wi = t >> 6, o = t & 63;
  const uint64_t a0 = two_over_pi_bits[wi], a1 = two_over_pi_bits[wi + 1], a2 = two_over_pi_bits[wi + 2],
                 a3 = two_over_pi_bits[wi + 3];
  const uint64_t s0 = o ? (a0 << o) | (a1 >> (64 - o)) : a0;
  const uint64_t s1 = o ? (a1 << o) | (a2 >> (64 - o)) : a1;
  const uint64_t s2 = o ? (a2 << o) | (a3 >> (64 - o)) : a2;
  /* M S = (...) 2^192 + w2 2^128 + w1 2^64 + w0; bits 192.. are multiples of 8 */
  const uint64_t p1h = umulhi(M, s1), p1l = M * s1;
  const uint64_t p2h = umulhi(M, s2), w0 = M * s2;
  const uint64_t w1 = p1l + p2h;
  const uint64_t w2 = M * s0 + p1h + (w1 < p1l ? 1 : 0);
  /* integer part (mod 8): bits 189..191; fraction: F (bits 125..188), then G */
  const uint64_t F = (w2 << 3) | (w1 >> 61), G = (w1 << 3) | (w0 >> 61);
  int k = (int)((w2 >> 61) & 7) + (int)(F >> 63);
  /* f = F 2^-64 - [F >= 2^63] + G 2^-128, as a rounded double: the top 53
   * bits of the signed F exactly, the rest added once */
  const long long Fi = (long long)F;
  const long long Fh = Fi & ~0x7ffLL;
  const double f = (double)Fh + ((double)(Fi - Fh) + (double)G * 0x1p-64);
  /* r = f 2^-64 pi/2 with pi/2 as a double plus its rounding error */
  double r = __builtin_fma(f, 0x1.921fb54442d18p+0, f * 0x1.1a62633145c07p-54) * 0x1p-64;
  if (bits >> 63) {
    r = -r;
    k = -k;
  }
  *q = k & 3;
  return r;
}

/* sin and cos of x.  |x| <= 2^20: Cody-Waite reduction by pi/2 in three
 * parts (the first two with 33 bits, so that k times them is exact); larger
 * x: reduce_pio2_large (a branch that finite deformation angles of
 * ordinary samples never take).  Then r + r^3 S(r^2) and
 * 1 - r^2/2 + r^4 C(r^2) on |r| <= pi/4 with S, C of degree 5 fitted by
 * weighted Remez exchanges (errors 2^-58, 2^-64), and the quadrant's swap
 * and signs.  Error at most 2 ulp; NaN and infinite x give NaN, as the
 * library does. */
FTR_MATH_FN void sincos_fast(double x, double* sp, double* cp) {
  double r;
  int q;
  if (__builtin_fabs(x) <= 0x1p20) {
    /* + 0.0 turns a -0 from rint into +0, so that for k = 0 the three
     * FMAs add -0 to x and return x itself, a zero's sign included */
    const double k = __builtin_rint(x * 0x1.45f306dc9c883p-1) + 0.0;
    r = __builtin_fma(-k, 0x1.921fb54400000p+0, x);
    r = __builtin_fma(-k, 0x1.0b4611a600000p-34, r);
    r = __builtin_fma(-k, 0x1.3198a2e037073p-69, r);
    q = (int)k;
  } else {
    r = reduce_pio2_large(x, &q);
    if (!(__builtin_fabs(x) <= 0x1.fffffffffffffp+1023)) r = x - x; /* inf, NaN -> NaN */
  }
  const double z = r * r;
  double S = 0x1.5d932f7c7cd31p-33;
  S = __builtin_fma(z, S, -0x1.ae5e66d592248p-26);
  S = __builtin_fma(z, S, 0x1.71de357875383p-19);
  S = __builtin_fma(z, S, -0x1.a01a019c126c0p-13);
  S = __builtin_fma(z, S, 0x1.111111110f881p-7);
  S = __builtin_fma(z, S, -0x1.5555555555549p-3);
  /* sin r has the sign of r on [-pi/4, pi/4]; copysign also keeps -0 */
  const double sn = __builtin_copysign(__builtin_fma(r * z, S, r), r);
  double C = -0x1.8faecf6d0da9cp-37;
  C = __builtin_fma(z, C, 0x1.1ee9ec482516ep-29);
  C = __builtin_fma(z, C, -0x1.27e4f80a76fa7p-22);
  C = __builtin_fma(z, C, 0x1.a01a019cb26fdp-16);
  C = __builtin_fma(z, C, -0x1.6c16c16c15184p-10);
  C = __builtin_fma(z, C, 0x1.555555555554cp-5);
  /* 1 - z/2 + z^2 C, with the rounding error of 1 - z/2 added back */
  const double hz = 0.5 * z, w = 1.0 - hz;
  const double cs = w + (((1.0 - w) - hz) + z * z * C);
  const double s = (q & 1) ? cs : sn, c = (q & 1) ? sn : cs;
  *sp = (q & 2) ? -s : s;
  *cp = ((q + 1) & 2) ? -c : c;
}

/* atan2(y, x) for finite (y, x); atan2(+-0, +-0) as the library.  With
 * a = |y|, b = |x|: atan(a/b) = base + atan(u), |u| <= tan(pi/8), where
 * (u, base) is (a/b, 0), (-b/a, pi/2) or ((a-b)/(a+b), pi/4), so one
 * division; atan(u) = u + u^3 A(u^2) with A of degree 10 fitted by a
 * weighted Remez exchange (error 2^-59.5); then the quadrant.  Error at
 * most 2 ulp, in about 35 device instructions (the library's: ~90). */
FTR_MATH_FN double atan2_fast(double y, double x) {
  const double a = __builtin_fabs(y), b = __builtin_fabs(x);
  const bool lo = a <= 0x1.a827999fcef32p-2 * b, hi = a >= 0x1.3504f333f9de6p+1 * b;
  const double num = lo ? a : (hi ? -b : a - b);
  const double den = lo ? b : (hi ? a : a + b);
  const double u = num * rcp_nr(den), z = u * u;
  double P = -0x1.25a380c6fe3f9p-6;
  P = __builtin_fma(z, P, 0x1.37f008c0cef30p-5);
  P = __builtin_fma(z, P, -0x1.9cdd3954e051fp-5);
  P = __builtin_fma(z, P, 0x1.df11a2fcdfa6bp-5);
  P = __builtin_fma(z, P, -0x1.10ebe199b75e7p-4);
  P = __builtin_fma(z, P, 0x1.3b111cbb98e55p-4);
  P = __builtin_fma(z, P, -0x1.745cf90fd5e80p-4);
  P = __builtin_fma(z, P, 0x1.c71c70e7caca7p-4);
  P = __builtin_fma(z, P, -0x1.24924922b183ap-3);
  P = __builtin_fma(z, P, 0x1.99999999958fap-3);
  P = __builtin_fma(z, P, -0x1.555555555553dp-2);
  /* bases pi/2 and pi/4 as a double plus its rounding error */
  const double bh = lo ? 0.0 : (hi ? 0x1.921fb54442d18p+0 : 0x1.921fb54442d18p-1);
  const double bl = lo ? 0.0 : (hi ? 0x1.1a62633145c07p-54 : 0x1.1a62633145c07p-55);
  double t = bh + (u + __builtin_fma(u * z, P, bl));
  if (a == 0.0 && b == 0.0) t = 0.0;
  if (__builtin_signbit(x)) t = (0x1.921fb54442d18p+1 - t) + 0x1.1a62633145c07p-53;
  return __builtin_signbit(y) ? -t : t;
}

}  // namespace ftrm
