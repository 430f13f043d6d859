/*
 * ftr_device.hpp -- HIP device code for the tropical sampling hot path
 * (gfx950, FP64).  One sample per lane; every graph-dependent index is a
 * compile-time constant of the graph spec G, so all per-sample matrices live
 * in registers.
 *
 * The file is included by the ahead-of-time kernel units and is also embedded
 * verbatim in the runtime library for hipRTC specialisation of new graphs.
 *
 * Graph spec G (generated per graph):
 *   static constexpr int E, V, NEXT;             edges, vertices, |ext|
 *   static constexpr int hi[E], lo[E];           edge endpoints, hi > lo
 *   static constexpr int ord[V], pos[V];         elimination order: internal
 *                                                vertices, then the NEXT
 *                                                vertices with nonzero P rows
 *
 * Reference semantics followed (michibo/feyntrop):
 *   sampler   srcs/tropical_sampling.hpp:468-564
 *   integrand srcs/feynman_integral.hpp:166-362, srcs/laplacian.hpp:14-215
 *   RNG       srcs/random.hpp:16-140
 * Differences in *method* (same mathematics, different rounding):
 *   - the deleted vertex v* (feynman_integral.hpp:166-194) is kept in place and
 *     its row/column are replaced by the identity, so the V x V Laplacian has
 *     static structure; det and inverse of the reduced matrix are unchanged;
 *   - the real Laplacian (SPD) is factorised by LDL^T in static order without
 *     pivoting instead of Eigen's diagonally pivoted LDLT;
 *   - the complex Laplacian La(1/z) and the deformation Jacobian are factorised
 *     as complex-symmetric LDL^T without pivoting: det(I + i c H X) =
 *     det(X^-1 + i c H) * prod x, whose real part diag(1/x) is positive
 *     definite (Eigen: FullPivLU / PartialPivLU);
 *   - tr(La^-1 P) is taken over the external vertices only (P is zero on the
 *     others): internal vertices are eliminated first and the solves run on
 *     the trailing external block.
 */
#pragma once

#if !defined(__HIPCC_RTC__) && !defined(FTR_HOST_BUILD)
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif
#include "ftr_abi.hpp"
#include "ftr_math.hpp"


namespace ftr {

/* ----------------------------------------------------------------- RNG -- */
struct Xoshiro {
  uint64_t s0, s1, s2, s3;
  __host__ __device__ inline uint64_t next() {
    const uint64_t res = s0 + s3;
    const uint64_t t = s1 << 17;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = (s3 << 45) | (s3 >> 19);
    return res;
  }
  /* random.hpp:16-27 */
  __host__ __device__ inline double uniform() {
    return (double)(next() >> 11) * (1.0 / 9007199254740992.0);
  }
};

/* state <- q(M) state for a jump polynomial q (4 words, bit i = x^i). */
__device__ inline void apply_poly(Xoshiro& g, const uint64_t* q) {
  Xoshiro acc{0, 0, 0, 0};
  for (int w = 0; w < 4; w++) {
    const uint64_t bits = q[w];
    for (int b = 0; b < 64; b++) {
      if ((bits >> b) & 1) { acc.s0 ^= g.s0; acc.s1 ^= g.s1; acc.s2 ^= g.s2; acc.s3 ^= g.s3; }
      g.next();
    }
  }
  g = acc;
}

/* ---------------------------------------------------------- elementary -- */
using ftrm::atan2_fast;
using ftrm::log_any;
using ftrm::log_pos;
using ftrm::rcp_nr;
using ftrm::sincos_fast;

/* ------------------------------------------------------------- complex -- */
struct cplx {
  double re, im;
  __host__ __device__ constexpr cplx() : re(0.0), im(0.0) {}
  __host__ __device__ constexpr cplx(double r) : re(r), im(0.0) {}
  __host__ __device__ constexpr cplx(double r, double i) : re(r), im(i) {}
};
__host__ __device__ inline cplx cmk(double a, double b) { return cplx(a, b); }
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx operator-(cplx a) { return cplx{-a.re, -a.im}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
  return cplx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ inline cplx operator*(double s, cplx a) { return cplx{s * a.re, s * a.im}; }
/* y - a b: four FMAs for complex operands, one for real ones */
__host__ __device__ inline cplx fms(cplx y, cplx a, cplx b) {
  return cplx{fma(-a.re, b.re, fma(a.im, b.im, y.re)), fma(-a.re, b.im, fma(-a.im, b.re, y.im))};
}
__host__ __device__ inline double fms(double y, double a, double b) { return fma(-a, b, y); }
__host__ __device__ inline cplx crecip(cplx a) {
  const double inv = rcp_nr(a.re * a.re + a.im * a.im);
  return cplx{a.re * inv, -a.im * inv};
}
/* principal log: (log|z|, arg z) */
__host__ __device__ inline cplx clog_(cplx z) {
  return cplx{log(hypot(z.re, z.im)), atan2(z.im, z.re)};
}
/* The same without branches from log_pos and atan2_fast: z is scaled by
 * 2^-e, e the exponent of max(|re|, |im|), so that |z 2^-e|^2 lies in
 * [1/4, 2), and log|z| = log(|z 2^-e|^2)/2 + e log 2.  z = 0 gives -inf,
 * infinite or NaN parts give inf or NaN, as clog_ does. */
__host__ __device__ inline cplx clog_fast(cplx z) {
  const double mx = fmax(fabs(z.re), fabs(z.im));
  const int e = ftrm::frexp_exp(mx);
  const double zr = __builtin_ldexp(z.re, -e), zi = __builtin_ldexp(z.im, -e);
  double lr = fma(0.5, log_pos(fma(zr, zr, zi * zi)), (double)e * 0x1.62e42fefa39efp-1);
  if (mx == 0.0) lr = -__builtin_inf();
  if (!__builtin_isfinite(mx)) lr = mx;
  return cplx{lr, atan2_fast(z.im, z.re)};
}
/* libstdc++ pow(complex<double>, double): real fast path for positive reals,
 * else polar(exp(y Re log z), y Im log z)  (complex:1028-1039) */
__host__ __device__ inline cplx cpow_real(cplx z, double y) {
  if (z.im == 0.0 && z.re > 0.0) return cplx{pow(z.re, y), 0.0};
  const cplx t = clog_fast(z);
  const double rho = exp(y * t.re);
  double s, c;
  sincos_fast(y * t.im, &s, &c);
  return cplx{rho * c, rho * s};
}
__host__ __device__ inline bool cfinite(cplx z) { return __builtin_isfinite(z.re) && __builtin_isfinite(z.im); }

/* Two lanes per sample: lanes 2m and 2m+1 of a wave evaluate one sample.
 * pair_bcast<SRC> gives every lane of a pair the value held by its lane SRC
 * (0 or 1): a DPP quad permutation, [0,0,2,2] or [1,1,3,3], on both 32-bit
 * halves -- two VALU moves, no LDS.  With LPS = 1 it is the identity. */
template <int LPS, int SRC>
__host__ __device__ inline double pair_bcast(double v) {
#ifdef FTR_HOST_BUILD
  return v;
#else
  if constexpr (LPS == 1) {
    return v;
  } else {
    static_assert(LPS == 2 && (SRC == 0 || SRC == 1), "pairs of lanes only");
    constexpr int ctrl = SRC == 0 ? 0xA0 : 0xF5;
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, ctrl, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), ctrl, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
  }
#endif
}
template <int LPS, int SRC>
__host__ __device__ inline cplx pair_bcast(cplx z) {
  return cplx{pair_bcast<LPS, SRC>(z.re), pair_bcast<LPS, SRC>(z.im)};
}
/* the same with the source lane as a value that folds to a constant once
 * the caller's loop is unrolled */
template <int LPS, class T>
__host__ __device__ inline T pair_from(T v, int src) {
  if constexpr (LPS == 1) return v;
  else return src ? pair_bcast<LPS, 1>(v) : pair_bcast<LPS, 0>(v);
}

/* z^n for an integer n by repeated squaring (n uniform across the wave);
 * the principal branch of pow(z, n) coincides with it for integer n. */
__host__ __device__ inline cplx cipow(cplx z, int n) {
  cplx r = cmk(1.0, 0.0), b = z;
  unsigned m = n < 0 ? (unsigned)(-n) : (unsigned)n;
  while (m) {
    if (m & 1u) r = r * b;
    m >>= 1;
    if (m) b = b * b;
  }
  return n < 0 ? crecip(r) : r;
}
__host__ __device__ inline double dipow(double x, int n) {
  double r = 1.0, b = x;
  unsigned m = n < 0 ? (unsigned)(-n) : (unsigned)n;
  while (m) {
    if (m & 1u) r *= b;
    m >>= 1;
    if (m) b *= b;
  }
  return n < 0 ? 1.0 / r : r;
}
/* x^(-D/2), x^(-W) with the integer fast path when it applies */
template <class PT>
__host__ __device__ inline double pow_mD(const PT& p, double x) {
  return p.use_iD != 0.0 ? dipow(x, (int)p.iD) : pow(x, p.mhalfD);
}
template <class PT>
__host__ __device__ inline double pow_mW(const PT& p, double x) {
  return p.use_iW != 0.0 ? dipow(x, (int)p.iW) : pow(x, p.mW);
}
template <class PT>
__host__ __device__ inline cplx cpow_mD(const PT& p, cplx z) {
  return p.use_iD != 0.0 ? cipow(z, (int)p.iD) : cpow_real(z, p.mhalfD);
}
template <class PT>
__host__ __device__ inline cplx cpow_mW(const PT& p, cplx z) {
  return p.use_iW != 0.0 ? cipow(z, (int)p.iW) : cpow_real(z, p.mW);
}

/* lower-triangular packed index for i >= j */
#define FTR_TRI(i, j) ((i) * ((i) + 1) / 2 + (j))

/* ------------------------------------------------------------- sampler -- */
/* tropical_sampling.hpp:503-564, table mode 4 (folded rows).  X[e] filled
 * in edge order; returns log psi_tr and log G. */
template <class G, bool PF, bool UNR, class TV>
__device__ inline void sample_folded(Xoshiro& rng, const TV& tv, double (&X)[G::E], double& lpsi_tr, double& lG) {
  constexpr int E = G::E;
  /* Folded rows: the state's row holds the running sums T_q of its cut
   * weights, so J(g) = T_{E-1} (the reference's J: the same sum in the
   * same order) and the cut is the first q with T_q > R = U J(g) -- the
   * number of q with T_q <= R, since the sums are nondecreasing and rise
   * only at edges of g; when no sum exceeds R the last edge of g is
   * taken.  The row of the state a cut reaches also carries the 1/r and
   * the flags of the cut's own step, so a step makes one read (fewer HBM
   * bytes than a row and a record when the table is far larger than the
   * L2).  PF: that whole row is requested in the cut's step, so that the
   * sums the next step scans are in flight during the radial draw;
   * otherwise they are read at the start of the next step.  Same cuts,
   * same arithmetic. */
  uint32_t g = (1u << E) - 1u;
  const double* rg = tv.row(g);
  const uint32_t b0 = tv.mmL(rg);
  bool MM = b0 & 1;
  int oL = (int)(b0 >> 1);
  double lGv = 0.0, lF = 0.0, lX = 0.0, XX = 1.0;
  double rv[E];
  if constexpr (PF) {
#pragma unroll
    for (int q = 0; q < E; q++) rv[q] = rg[q];
  }
  auto step = [&](int j) {
    if constexpr (!PF) {
      const double* r = tv.row(g);
#pragma unroll
      for (int q = 0; q < E; q++) rv[q] = r[q];
    }
    const double R = rng.uniform() * rv[E - 1];
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < E; q++) cnt += (rv[q] <= R) ? 1 : 0;
    const int e = (cnt < E) ? cnt : 31 - __builtin_clz(g);
    const uint32_t c = g ^ (1u << e);
    const double* rc = tv.row(c);
    const double ir = tv.ir(rc);
    const uint32_t b = tv.mmL(rc);
    if constexpr (PF) {
      if (j < E - 1) {
#pragma unroll
        for (int q = 0; q < E; q++) rv[q] = rc[q];
      }
    }
    const bool mm = b & 1;
    const int L = (int)(b >> 1);
    g = c;
    if (MM && !mm) lGv = lX;
    MM = mm;
    tv.xs[e * G::BLK] = XX;
    if (L < oL) {
      lF += lX;
      oL = L;
    }
    if (j < E - 1) {
      const double U = rng.uniform();
      lX += ir * log_pos(1. - U);
      XX = exp(lX);
    }
  };
  if constexpr (UNR) {
#pragma unroll
    for (int j = 0; j < E; j++) step(j);
  } else {
    for (int j = 0; j < E; j++) step(j);
  }
#pragma unroll
  for (int k = 0; k < E; k++) X[k] = tv.xs[k * G::BLK];
  lpsi_tr = lF;
  lG = lGv;
}

/* tropical_sampling.hpp:503-564, table modes 0-3.  Table access goes
 * through the TableView (LDS or global).  X[e] filled in edge order; returns
 * log psi_tr and log G.
 * JSUM (row placement): R = U J(g) takes J(g) as the last running sum of the
 * state's row instead of from its record -- the reference's J is that sum,
 * in the same order, so the cut is the same; measured faster on the real
 * path (cubic5 -7 % generic, -12 % specialised; K4 wheel -1 %) and slower
 * on the deformed one (double box generic +2 %). */
template <class G, bool JSUM = false, bool PF = false, bool UNR = false, class TV>
__device__ inline void sample_tables(Xoshiro& rng, const TV& tv, double (&X)[G::E], double& lpsi_tr,
                                    double& lG) {
  /* The radius product x = prod (1-U)^(1/r) is carried as its logarithm, so
   * each radial draw costs one log and one exp (the reference: one pow); the
   * samples agree with the reference's to ~1e-15 relative.  Returns
   * log psi_tr = log F and log G, with xi_tr = G F. */
  constexpr int E = G::E;
  if constexpr (TV::TM == 3 && PF && !JSUM) {
    /* PF: the row of the state a cut reaches is requested in the cut's own
     * step, after its record, so that it is in flight during the radial
     * draw instead of being requested at the next step.  Same cuts, same
     * arithmetic.  Measured on the deformed path: generic kernels faster
     * (triple ladder 25.2 -> 24.6 ms per 1e8, double box 14.2 -> 13.9),
     * kernels specialised to a kinematic point slower (double box 11.8 ->
     * 12.7): the spec generator enables it for the former only. */
    uint32_t g = (1u << E) - 1u;
    double JJr = tv.jr(g);
    uint32_t b0 = tv.mmL(g);
    bool MM = b0 & 1;
    int oL = (int)(b0 >> 1);
    double lGv = 0.0, lF = 0.0, lX = 0.0, XX = 1.0;
    double rv[E];
    {
      const double* row = tv.row(g);
#pragma unroll
      for (int q = 0; q < E; q++) rv[q] = row[q];
    }
    auto step = [&](int j) {
      const double R = rng.uniform() * JJr;
      double T = 0.0;
      int cnt = 0;
#pragma unroll
      for (int q = 0; q < E; q++) {
        T += rv[q];
        cnt += (T <= R) ? 1 : 0;
      }
      const int e = (cnt < E) ? cnt : 31 - __builtin_clz(g);
      const uint32_t c = g ^ (1u << e);
      JJr = tv.jr(c);
      const double ir = tv.ir(c);
      const uint32_t b = tv.mmL(c);
      if (j < E - 1) {
        const double* row = tv.row(c);
#pragma unroll
        for (int q = 0; q < E; q++) rv[q] = row[q];
      }
      const bool mm = b & 1;
      const int L = (int)(b >> 1);
      g = c;
      if (MM && !mm) lGv = lX;
      MM = mm;
      tv.xs[e * G::BLK] = XX;
      if (L < oL) {
        lF += lX;
        oL = L;
      }
      if (j < E - 1) {
        const double U = rng.uniform();
        lX += ir * log_pos(1. - U);
        XX = exp(lX);
      }
    };
    /* UNR: the steps unrolled, so that the sampler is straight-line code
     * that the scheduler can interleave with other work (integrate_body's
     * pipelined loop) */
    if constexpr (UNR) {
#pragma unroll
      for (int j = 0; j < E; j++) step(j);
    } else {
      for (int j = 0; j < E; j++) step(j);
    }
#pragma unroll
    for (int k = 0; k < E; k++) X[k] = tv.xs[k * G::BLK];
    lpsi_tr = lF;
    lG = lGv;
    return;
  }
  uint32_t g = (E == 32) ? 0xffffffffu : ((1u << E) - 1u);
  double JJr = tv.jr(g);
  uint32_t b0 = tv.mmL(g);
  bool MM = b0 & 1;
  int oL = (int)(b0 >> 1);
  double lGv = 0.0, lF = 0.0, lX = 0.0, XX = 1.0;
#pragma unroll
  for (int e = 0; e < E; e++) X[e] = 0.0;
  auto step = [&](int j) {
    int e = 0;
    if constexpr (TV::TM == 3 && JSUM) {
      const double* row = tv.row(g);
      double rv[E];
#pragma unroll
      for (int q = 0; q < E; q++) rv[q] = row[q];
      double T = 0.0;
#pragma unroll
      for (int q = 0; q < E; q++) {
        T += rv[q];
        rv[q] = T;
      }
      const double R = rng.uniform() * T;
      int cnt = 0;
#pragma unroll
      for (int q = 0; q < E; q++) cnt += (rv[q] <= R) ? 1 : 0;
      e = (cnt < E) ? cnt : 31 - __builtin_clz(g);
    } else if constexpr (TV::TM == 3) {
      const double R = rng.uniform() * JJr;
      /* the state's row: all E weights loaded together, absent edges 0 */
      const double* row = tv.row(g);
      double rv[E];
#pragma unroll
      for (int q = 0; q < E; q++) rv[q] = row[q];
      /* The running sums T_q are nondecreasing (weights >= 0, absent
       * edges 0), so the first q with T_q > R is the number of q with
       * T_q <= R; that q carries a positive weight, so it is in g.  When
       * no sum exceeds R the last edge of g is taken.  (Rows of
       * precomputed sums were measured 1.5 % slower on the double box.) */
      double T = 0.0;
      int cnt = 0;
#pragma unroll
      for (int q = 0; q < E; q++) {
        T += rv[q];
        cnt += (T <= R) ? 1 : 0;
      }
      e = (cnt < E) ? cnt : 31 - __builtin_clz(g);
    } else {
      const double R = rng.uniform() * JJr;
      double T = 0.0;
      uint32_t m = g;
      while (m) {
        const int q = __builtin_ctz(m);
        m &= m - 1;
        T += tv.w(g ^ (1u << q));
        e = q;
        if (T > R) break;
      }
    }
    const uint32_t c = g ^ (1u << e);
    if constexpr (!(TV::TM == 3 && JSUM)) JJr = tv.jr(c);
    const double ir = tv.ir(c);
    const uint32_t b = tv.mmL(c);
    const bool mm = b & 1;
    const int L = (int)(b >> 1);
    g = c;
    if (MM && !mm) lGv = lX;
    MM = mm;
    if constexpr (TV::TM == 3) {
      tv.xs[e * G::BLK] = XX;
    } else {
#pragma unroll
      for (int k = 0; k < E; k++) X[k] = (k == e) ? XX : X[k];
    }
    if (L < oL) {
      lF += lX;
      oL = L;
    }
    if (j < E - 1) {
      const double U = rng.uniform();
      lX += ir * log_pos(1. - U);
      XX = exp(lX);
    }
  };
  if constexpr (UNR) {
#pragma unroll
    for (int j = 0; j < E; j++) step(j);
  } else {
    for (int j = 0; j < E; j++) step(j);
  }
  if constexpr (TV::TM == 3) {
#pragma unroll
    for (int k = 0; k < E; k++) X[k] = tv.xs[k * G::BLK];
  }
  lpsi_tr = lF;
  lG = lGv;
}

/* The sampler: table mode 4 (folded rows) or the others. */
template <class G, bool JSUM = false, bool PF = false, bool UNR = false, class TV>
__device__ inline void sample_point(Xoshiro& rng, const TV& tv, double (&X)[G::E], double& lpsi_tr,
                                    double& lG) {
  if constexpr (TV::TM == 4)
    sample_folded<G, PF, UNR>(rng, tv, X, lpsi_tr, lG);
  else
    sample_tables<G, JSUM, PF, UNR>(rng, tv, X, lpsi_tr, lG);
}

/* ----------------------------------------------------------- integrand -- */
/* Matrices are stored lower-triangular-packed in elimination order: vertex
 * G::ord[p] sits at position p (internal vertices first, the vertices that
 * carry external momentum last; G::pos is the inverse).  Eliminating the
 * internal vertices first leaves the Schur complement on the external block,
 * so tr(La^-1 P) needs solves on that block only and the factors of the
 * internal columns die right after their elimination step. */
#define FTR_SYM(A, i, j) (((i) >= (j)) ? A[FTR_TRI(i, j)] : A[FTR_TRI(j, i)])

/* In-place right-looking LDL^T of a packed symmetric matrix (real or
 * complex-symmetric); returns prod0 times the product of the pivots, and
 * their inverses in dinv. */
template <int V>
__host__ __device__ inline double ldlt_inplace(double (&M)[V * (V + 1) / 2], double (&dinv)[V],
                                               double prod = 1.0) {
#pragma unroll
  for (int k = 0; k < V; k++) {
    const double d = M[FTR_TRI(k, k)];
    prod *= d;
    const double inv = rcp_nr(d);
    dinv[k] = inv;
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      const double aik = M[FTR_TRI(i, k)];
#pragma unroll
      for (int j = k + 1; j <= i; j++) M[FTR_TRI(i, j)] -= aik * (M[FTR_TRI(j, k)] * inv);
    }
#pragma unroll
    for (int i = k + 1; i < V; i++) M[FTR_TRI(i, k)] *= inv;
  }
  return prod;
}

template <int V>
__host__ __device__ inline cplx ldlt_inplace(cplx (&M)[V * (V + 1) / 2], cplx (&dinv)[V], double prod0 = 1.0) {
  cplx prod = cmk(prod0, 0.0);
#pragma unroll
  for (int k = 0; k < V; k++) {
    const cplx d = M[FTR_TRI(k, k)];
    prod = prod * d;
    const cplx inv = crecip(d);
    dinv[k] = inv;
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      const cplx lik = M[FTR_TRI(i, k)] * inv;
#pragma unroll
      for (int j = k + 1; j <= i; j++) M[FTR_TRI(i, j)] = fms(M[FTR_TRI(i, j)], lik, M[FTR_TRI(j, k)]);
    }
#pragma unroll
    for (int i = k + 1; i < V; i++) M[FTR_TRI(i, k)] = M[FTR_TRI(i, k)] * inv;
  }
  return prod;
}

/* tr(A^-1 P) from the factors of A (packed L below the diagonal, dinv).  P's
 * block on the external vertices is sum_r eta_r q_r q_r^T (PR terms,
 * host-computed), so tr(A^-1 P) = sum_r eta_r q_r^T A^-1 q_r, and with
 * A = L D L^T, q^T A^-1 q = sum_i y_i^2 / d_i for y = L^-1 q.  y vanishes on
 * the internal positions (they come first and q is zero there), so each term
 * is one forward substitution on the trailing external block: no back
 * substitution, PR <= NEXT - 1 terms instead of NEXT solves.  Complex
 * symmetric A: the same with y_i^2 (no conjugate).  For A = La + alpha
 * e_v e_v^T (the Laplacian with v*'s diagonal raised, see integrand) A^-1 is
 * the reduced inverse plus 1 1^T / alpha, whose trace against P vanishes
 * (q_r is orthogonal to 1), so this is tr(La_red^-1 P_red). */
template <class G, class T, class PT>
__host__ __device__ inline T trace_inv_P(const T (&M)[G::V * (G::V + 1) / 2], const T (&dinv)[G::V],
                                         const PT& p) {
  constexpr int V = G::V, NX = G::NEXT, N0 = G::V - G::NEXT;
  T tr = T{};
#pragma unroll
  for (int r = 0; r < G::PR; r++) {
    T y[NX > 0 ? NX : 1];
#pragma unroll
    for (int a = 0; a < NX; a++) y[a] = T{p.Pq[r * FTR_MAX_V + a]};
#pragma unroll
    for (int a = 0; a < NX; a++)
#pragma unroll
      for (int b = 0; b < a; b++) y[a] = fms(y[a], M[FTR_TRI(N0 + a, N0 + b)], y[b]);
    T s = T{};
#pragma unroll
    for (int a = 0; a < NX; a++) s = s + (y[a] * y[a]) * dinv[N0 + a];
    tr = tr + p.Peta[r] * s;
  }
  (void)V;
  return tr;
}

/* log of the projective rescale factor (feynman_integral.hpp:201-207):
 * scale = (psi_tr^(-D/2) (psi_tr/xi_tr)^W)^(1/(D L/2 + W)), from
 * log psi_tr and log G (xi_tr = G psi_tr). */
template <class PT>
__host__ __device__ inline double log_scale(const PT& p, double lpsi, double lG) {
  const double t = p.mhalfD * lpsi - (p.W != 0.0 ? p.W * lG : 0.0);
  return p.inv_dlw * t;
}

/* Determinant of the deformation Jacobian's complex-symmetric matrix
 * B = X^-1 + i c H (c = def_ref, H in dd, packed lower, entry stride DS:
 * 1 in registers, samples per block in LDS), multiplied into jdet: left-looking
 * (Crout) LDL^T, column j needing rows >= j of the finished columns only,
 * so a factor row dies as the sweep passes it and only the pivots are kept.
 *
 * LPS = 1: one lane holds every row.  LPS = 2: the pair's lane h holds the
 * rows i = 2r + h (local row r); for column j the lane owning row j forms
 * w_k = d_k L[j][k] and the pivot d_j, and pair_bcast hands them to its
 * partner; each lane then updates its own rows below j.  The arithmetic of
 * every entry, and its order, is that of one lane, so both give the same
 * bits; the pair holds half of the live factor rows each.  A lane's local
 * row whose index is not below j in that lane (the pair's boundary row) is
 * updated too and never read back. */
template <class G, int LPS, int DS, class PT, class DDT>
__host__ __device__ inline __attribute__((always_inline)) void crout_ldlt(const PT& p, const double (&Xi)[G::E], DDT dd, cplx& jdet) {
  constexpr int E = G::E;
  constexpr int RL = (E + LPS - 1) / LPS;
#ifdef FTR_HOST_BUILD
  const int h = 0;
#else
  const int h = LPS == 1 ? 0 : (int)(threadIdx.x & (LPS - 1));
#endif
  /* packed factor rows: local row r (row r LPS + h) holds its entries
   * k < (r + 1) LPS - 1 from offset LPS r (r + 1) / 2 - r */
  cplx L[LPS * RL * (RL + 1) / 2 - RL + 1];
  cplx d[E];
#pragma unroll
  for (int j = 0; j < E; j++) {
    const int oj = j % LPS, rj = j / LPS; /* owner lane and local row of row j */
    cplx w[E];
#pragma unroll
    for (int k = 0; k < j; k++) w[k] = pair_from<LPS>(d[k] * L[LPS * rj * (rj + 1) / 2 - rj + k], oj);
    cplx dj = cmk(Xi[j], p.def_ref * dd[FTR_TRI(j, j) * DS]);
#pragma unroll
    for (int k = 0; k < j; k++) dj = fms(dj, L[LPS * rj * (rj + 1) / 2 - rj + k], w[k]);
    dj = pair_from<LPS>(dj, oj);
    d[j] = dj;
    jdet = jdet * dj;
    const cplx inv = crecip(dj);
#pragma unroll
    for (int r = 0; r < RL; r++) {
      if ((r + 1) * LPS - 1 > j) {
        /* this lane's row i = r LPS + h (with LPS = 1: row r); the pair's
         * last local row lacks its second row when E is odd */
        if (r * LPS + LPS - 1 < E || r * LPS + h < E) {
          cplx v = cmk(0.0, p.def_ref * dd[(FTR_TRI(r * LPS, j) + h * (r * LPS) + h * (h + 1) / 2) * DS]);
#pragma unroll
          for (int k = 0; k < j; k++) v = fms(v, L[LPS * r * (r + 1) / 2 - r + k], w[k]);
          L[LPS * r * (r + 1) / 2 - r + j] = v * inv;
        }
      }
    }
  }
}

/* The integrand of one sample: out[k] (complex, imag 0 on the real path).
 * Returns false when dropped (non-finite R0, feynman_integral.hpp:227-234,
 * :254-261, :329-347).  PT: the problem's parameters, ftr::DevParams read
 * from memory or a type with the same members as compile-time constants
 * (kernels specialised to one kinematic point).  DDL: the deformation's
 * second derivatives in LDS at dd_slots; LPS: lanes per sample (2: the
 * lane pair shares the Jacobian's factorisation, crout_ldlt). */
template <class G, int K, int MODE, class PT = DevParams, bool DDL = false, int LPS = 1>
__host__ __device__ inline bool integrand(const PT& p, double (&X)[G::E], double lscale,
                                          cplx (&out)[K], double* dd_slots = nullptr) {
  constexpr int E = G::E, V = G::V, NT = V * (V + 1) / 2;
  /* vertex weights and v* (first minimum) -- feynman_integral.hpp:166-194 */
  double XV[V];
#pragma unroll
  for (int v = 0; v < V; v++) XV[v] = 0.0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    XV[G::hi[e]] += X[e];
    XV[G::lo[e]] += X[e];
  }
  int vs = 0;
  double mv = XV[0];
#pragma unroll
  for (int v = 1; v < V; v++)
    if (mv > XV[v]) {
      mv = XV[v];
      vs = v;
    }
  /* projective rescale -- :201-207 */
  const double scale = exp(lscale);
  double Xi[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    X[e] *= scale;
    Xi[e] = rcp_nr(X[e]);
  }
  double xprod = 1.0;
#pragma unroll
  for (int e = 0; e < E; e++) xprod *= X[e];
  /* Laplacian with weights 1/x (laplacian.hpp:14-67) in elimination order.
   * The reference deletes v*'s row and column; here v*'s diagonal is raised
   * by alpha instead: for a connected graph's Laplacian La (null vector 1),
   *   det(La + alpha e_v e_v^T) = alpha det(La_red),
   *   (La + alpha e_v e_v^T)^-1 = La_red^-1 (zero row/col at v) + 1 1^T / alpha,
   * and the 1 1^T term drops out of every quantity used below (edge
   * differences u_e^T G u_f with u_e^T 1 = 0, and contractions with P, whose
   * rows sum to zero).  alpha = 2^40 sum(1/x), a power of two far above
   * every entry: v*'s pivot is ~alpha, its couplings reach the other pivots
   * at 2^-40 relative, so the elimination is numerically the deletion, and
   * dividing alpha out (the product starts at 1/alpha) is exact.  One select
   * per diagonal entry replaces one per matrix entry. */
  double alpha;
  {
    double sxi = 0.0;
#pragma unroll
    for (int e = 0; e < E; e++) sxi += Xi[e];
    alpha = __builtin_ldexp(1.0, ftrm::frexp_exp(sxi) + 40);
  }
  const double ialpha = 1.0 / alpha;
  double M[NT];
#pragma unroll
  for (int q = 0; q < NT; q++) M[q] = 0.0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int a = G::pos[G::lo[e]], b = G::pos[G::hi[e]];
    M[FTR_TRI(a, a)] += Xi[e];
    M[FTR_TRI(b, b)] += Xi[e];
    M[(a > b) ? FTR_TRI(a, b) : FTR_TRI(b, a)] -= Xi[e];
  }
#pragma unroll
  for (int i = 0; i < V; i++) M[FTR_TRI(i, i)] += (G::ord[i] == vs) ? alpha : 0.0;
  double dinv[V];
  const double psi = ldlt_inplace<V>(M, dinv, ialpha) * xprod;

  if constexpr (MODE != 2) {
    double R = p.IGtr * pow_mD(p, psi);
    if (!__builtin_isfinite(R)) return false;
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = cmk(0.0, 0.0);
    if constexpr (MODE == 1) {
      /* W = 0 and one order: no second Symanzik needed (:237-241) */
      out[0] = cmk(R, 0.0);
      return true;
    }
    const double tr = trace_inv_P<G, double, PT>(M, dinv, p);
    double dot = 0.0;
#pragma unroll
    for (int e = 0; e < E; e++) dot += p.m2[e] * X[e];
    const double pxi = -tr + dot;
    R *= pow_mW(p, pxi);
    if (!__builtin_isfinite(R)) return false;
    out[0] = cmk(R, 0.0);
    if (K > 1) {
      const double ef = log_any(psi) - p.Lf * log_any(pxi);
#pragma unroll
      for (int k = 1; k < K; k++) {
        R *= ef / (double)k;
        out[k] = cmk(R, 0.0);
      }
    }
    return true;
  } else {
    /* ---- analytic continuation (feynman_integral.hpp:278-362) ---- */
    /* Gm = (La + alpha e_v e_v^T)^-1 = L^-T D^-1 L^-1 (positions): the
     * reduced inverse plus 1 1^T / alpha, which neither the edge differences
     * nor LPGPLT = Gm P Gm (P 1 = 0) see */
    const double (&dz)[V] = dinv;
    double Gm[NT];
    {
      double Li[NT]; /* L^-1, unit lower (diagonal unused) */
#pragma unroll
      for (int i = 0; i < V; i++)
#pragma unroll
        for (int j = 0; j < i; j++) {
          double s = -M[FTR_TRI(i, j)];
#pragma unroll
          for (int q = j + 1; q < i; q++) s -= M[FTR_TRI(i, q)] * Li[FTR_TRI(q, j)];
          Li[FTR_TRI(i, j)] = s;
        }
      /* Gm[i][j] = sum_{k >= i} L^-1[k][i] (L^-1[k][j] d_k^-1): the
       * products in parentheses once per (k, j), then one FMA per term */
      double Wk[NT];
#pragma unroll
      for (int k = 0; k < V; k++) {
        Wk[FTR_TRI(k, k)] = dz[k];
#pragma unroll
        for (int j = 0; j < k; j++) Wk[FTR_TRI(k, j)] = Li[FTR_TRI(k, j)] * dz[k];
      }
#pragma unroll
      for (int i = 0; i < V; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
          double s = Wk[FTR_TRI(i, j)];
#pragma unroll
          for (int k = i + 1; k < V; k++) s = fma(Li[FTR_TRI(k, i)], Wk[FTR_TRI(k, j)], s);
          Gm[FTR_TRI(i, j)] = s;
        }
    }
    /* LPGPLT = Gm P Gm is needed only through its edge differences
     * u_a^T LPGPLT u_b (u_a = e_{k_a} - e_{l_a}); with P's external block
     * sum_r eta_r q_r q_r^T these are sum_r eta_r dq[r][a] dq[r][b],
     * dq[r][a] = (Gm q_r)[k_a] - (Gm q_r)[l_a]: no V x V product and no
     * cancellation of large LPGPLT entries in the differences */
    constexpr int PRK = G::PR > 0 ? G::PR : 1;
    double dq[PRK][E];
    {
      constexpr int N0 = V - G::NEXT;
#pragma unroll
      for (int r = 0; r < G::PR; r++) {
        double u[V];
#pragma unroll
        for (int w = 0; w < V; w++) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < G::NEXT; a++) s += FTR_SYM(Gm, w, N0 + a) * p.Pq[r * FTR_MAX_V + a];
          u[w] = s;
        }
#pragma unroll
        for (int a = 0; a < E; a++) dq[r][a] = u[G::pos[G::hi[a]]] - u[G::pos[G::lo[a]]];
      }
    }
    /* first and second derivatives of pphi (laplacian.hpp:102-215) */
    double th[E];
    /* in registers, or (DDL) in the sample's LDS slots (stride: samples per
     * block, G::BLK / LPS).  With two lanes per sample (LPS = 2, always DDL)
     * both lanes compute every entry and store the same value to the same
     * slot, so each lane reads back only what it stored itself. */
    static_assert(LPS == 1 || DDL, "two lanes per sample keep the second derivatives in LDS");
    double dd_regs[DDL ? 1 : E * (E + 1) / 2];
    double* const dd = DDL ? dd_slots : dd_regs;
    constexpr int DS = DDL ? G::BLK / LPS : 1;
#pragma unroll
    for (int a = 0; a < E; a++) {
      const int ka = G::pos[G::hi[a]], la = G::pos[G::lo[a]];
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < G::PR; r++) s += p.Peta[r] * (dq[r][a] * dq[r][a]);
      const double t = FTR_SYM(Gm, ka, ka) + FTR_SYM(Gm, la, la) - 2.0 * FTR_SYM(Gm, ka, la);
      const double xa = Xi[a] * Xi[a];
      const double xas = xa * s;
      th[a] = -xas;
      dd[FTR_TRI(a, a) * DS] = 2.0 * xas * (Xi[a] - xa * t);
#pragma unroll
      for (int b = 0; b < a; b++) {
        const int kb = G::pos[G::hi[b]], lb = G::pos[G::lo[b]];
        double s2 = 0.0;
#pragma unroll
        for (int r = 0; r < G::PR; r++) s2 += p.Peta[r] * (dq[r][a] * dq[r][b]);
        const double t2 = FTR_SYM(Gm, ka, kb) + FTR_SYM(Gm, la, lb) - FTR_SYM(Gm, ka, lb) - FTR_SYM(Gm, la, kb);
        const double xb = Xi[b] * Xi[b];
        dd[FTR_TRI(a, b) * DS] = -2.0 * xb * xa * s2 * t2;
      }
    }
    /* deformation Jacobian: det(I + i c H X) = det(X^-1 + i c H) prod x.
     * Left-looking (Crout) LDL^T of the complex-symmetric B = X^-1 + i c H:
     * column j needs rows >= j of the finished columns only, so the factor
     * rows die as the sweep advances; only the pivots are kept. */
    cplx jdet = cmk(1.0, 0.0);
    crout_ldlt<G, LPS, DS>(p, Xi, dd, jdet);
    if constexpr (!G::NU1) jdet = xprod * jdet;
    /* deformation: theta = def_ref (d pphi + m^2), z = x e^{i theta}; the
     * complex Laplacian La(1/z) (elimination order, v*'s diagonal raised by
     * alpha as for the real one), prod z and sum m^2 z are accumulated edge
     * by edge as each z is formed, so no array of z or 1/z is kept */
    cplx cdet, ctr;
    cplx cxp = cmk(1.0, 0.0), cdot = cmk(0.0, 0.0);
    double xne = 0.0;
    {
      cplx CM[NT], cdinv[V];
#pragma unroll
      for (int q = 0; q < NT; q++) CM[q] = cmk(0.0, 0.0);
#pragma unroll
      for (int e = 0; e < E; e++) {
        const double t = (th[e] + p.m2[e]) * p.def_ref;
        double sn, cs;
        sincos_fast(t, &sn, &cs);
        const cplx z = cmk(X[e] * cs, X[e] * sn), zi = cmk(Xi[e] * cs, -Xi[e] * sn);
        if constexpr (!G::NU1) xne += p.nu[e] * t;
        const int a = G::pos[G::lo[e]], b = G::pos[G::hi[e]];
        const int o = (a > b) ? FTR_TRI(a, b) : FTR_TRI(b, a);
        CM[FTR_TRI(a, a)] = CM[FTR_TRI(a, a)] + zi;
        CM[FTR_TRI(b, b)] = CM[FTR_TRI(b, b)] + zi;
        CM[o] = CM[o] - zi;
        cxp = cxp * z;
        cdot = cdot + p.m2[e] * z;
      }
      /* complex-symmetric LDL^T and tr(La(1/z)^-1 P) on the external block */
#pragma unroll
      for (int i = 0; i < V; i++) CM[FTR_TRI(i, i)].re += (G::ord[i] == vs) ? alpha : 0.0;
      cdet = ldlt_inplace<V>(CM, cdinv, ialpha);
      ctr = trace_inv_P<G, cplx, PT>(CM, cdinv, p);
    }
    const cplx cPsi = cxp * cdet;
    const cplx cPXi = cdot - ctr;
    cplx cR = p.IGtr * cpow_mD(p, cPsi);
    if constexpr (G::NU1) {
      /* every nu = 1: exp(i sum nu theta) = prod_e e^{i theta_e} = cxp / prod x,
       * so (prod x) jdet exp(i sum theta) = jdet cxp (jdet without the
       * prod x factor here) -- no sincos of the summed angle */
      cR = cR * (jdet * cxp);
    } else {
      double sn, cs;
      sincos_fast(xne, &sn, &cs);
      cR = cR * (jdet * cmk(cs, sn));
    }
    cR = cR * cpow_mW(p, cPXi);
    /* Branch-free tail: the finiteness flag is returned, not branched on
     * (an early return here made the gfx950 backend at -O2/-O3 produce wrong
     * values for the 10-edge ladder under high register pressure). */
    const bool ok = cfinite(cR);
    out[0] = cR;
    if (K > 1) {
      const cplx lp = clog_fast(cPsi), lx = clog_fast(cPXi);
      const cplx ef = lp - p.Lf * lx;
#pragma unroll
      for (int k = 1; k < K; k++) {
        cR = cR * cmk(ef.re / (double)k, ef.im / (double)k);
        out[k] = cR;
      }
    }
    return ok;
  }
}
#undef FTR_SYM

/* The integration kernel's parameters: read from memory (DevParams), or,
 * for kernels specialised to one kinematic point, a type whose members are
 * compile-time constants (no storage). */
template <class PT>
struct ParamsFor {
  PT c;
  __device__ const PT& get(const Launch&) const { return c; }
};
template <>
struct ParamsFor<DevParams> {
  __device__ const DevParams& get(const Launch& a) const { return *a.prm; }
};

/* -------------------------------------------------------- table views -- */
/* TMODE 0: whole table in LDS; 1: w = J/r in LDS, the rest from global
 * (L2-resident); 2: everything from global; 3: rows -- the E cut weights
 * w(g\e) of a state contiguous in HBM plus a 32-byte record per state, so a
 * cut step reads one row and one record instead of |g| scattered entries
 * (large tables); 4: folded rows -- running sums and the record's fields in
 * one row (ftr_abi.hpp frow_stride), for tables far larger than the L2. */
template <int TMODE>
struct TableView {
  static constexpr int TM = TMODE;
  const double* pw;
  const double* pjr;
  const double* pir;
  const uint8_t* pmm;
  __device__ inline double w(uint32_t g) const { return pw[g]; }
  __device__ inline double jr(uint32_t g) const { return pjr[g]; }
  __device__ inline double ir(uint32_t g) const { return pir[g]; }
  __device__ inline uint32_t mmL(uint32_t g) const { return pmm[g]; }
};
template <>
struct TableView<3> {
  static constexpr int TM = 3;
  const double* pw;
  const double* rows;
  const StateRec* rec;
  int E;
  double* xs; /* this thread's E LDS slots for X */
  __device__ inline double w(uint32_t g) const { return pw[g]; }
  __device__ inline double jr(uint32_t g) const { return rec[g].jr; }
  __device__ inline double ir(uint32_t g) const { return rec[g].ir; }
  __device__ inline uint32_t mmL(uint32_t g) const { return rec[g].mmL; }
  __device__ inline const double* row(uint32_t g) const { return rows + (size_t)g * E; }
};
template <>
struct TableView<4> {
  static constexpr int TM = 4;
  const double* rows;
  int E;
  double* xs; /* this thread's E LDS slots for X */
  __device__ inline const double* row(uint32_t g) const { return rows + (size_t)g * frow_stride(E); }
  /* the row's tail: 1/r(g) and the mm | L<<1 flags */
  __device__ inline double ir(const double* r) const { return r[E]; }
  __device__ inline uint32_t mmL(const double* r) const { return *reinterpret_cast<const uint32_t*>(r + E + 1); }
};

template <int TMODE>
__host__ __device__ constexpr size_t table_lds_bytes(int E) {
  return TMODE == 0 ? ((size_t)25 << E) + 64 : (TMODE == 1 ? ((size_t)8 << E) : 0);
}
/* With the tables in HBM (TMODE 3, 4) the LDS holds each thread's E slots of X,
 * where the sampler scatters the radius of the cut edge (one store per cut
 * instead of a select over all E edges).  Per-thread LDS slots are laid out
 * [slot][thread] (slot e of this thread at ptr[e * G::BLK], G::BLK = the
 * block size): the 64 lanes of a wave touch 64 consecutive doubles whatever
 * slot each lane picks, so the accesses are free of bank conflicts, and
 * every offset is a compile-time constant. */
template <int TMODE>
__device__ inline size_t xs_lds_bytes(int E) {
  return TMODE >= 3 ? (size_t)E * blockDim.x * 8 : 0;
}
template <int TMODE>
__device__ inline double* xs_slots(char* lds, int E) {
  return TMODE >= 3 ? reinterpret_cast<double*>(lds + table_lds_bytes<TMODE>(E)) + threadIdx.x : nullptr;
}
/* the deformation's second derivatives in LDS (G::DDL): E(E+1)/2 slots per
 * sample after the X slots, [slot][sample] layout (the block's
 * blockDim.x / G::LPS samples; a lane pair shares its sample's slots) */
template <class G>
__device__ inline size_t dd_lds_bytes() {
  return G::DDL ? (size_t)(G::E * (G::E + 1) / 2) * (blockDim.x / G::LPS) * 8 : 0;
}
template <class G, int TMODE>
__device__ inline double* dd_slots(char* lds) {
  return G::DDL ? reinterpret_cast<double*>(lds + table_lds_bytes<TMODE>(G::E) + xs_lds_bytes<TMODE>(G::E)) +
                      threadIdx.x / G::LPS
                : nullptr;
}
/* start of the block reduction's LDS area */
template <class G, int TMODE>
__device__ inline double* red_area(char* lds) {
  return reinterpret_cast<double*>(lds + table_lds_bytes<TMODE>(G::E) + xs_lds_bytes<TMODE>(G::E) +
                                   dd_lds_bytes<G>());
}

/* Stage the table into LDS and build the view. */
template <int TMODE>
__device__ inline TableView<TMODE> stage_table(const Launch& a, int E, char* lds) {
  TableView<TMODE> tv;
  const uint32_t NT = 1u << E;
  if constexpr (TMODE == 4) {
    tv.rows = a.t_rows; tv.E = E;
    tv.xs = xs_slots<4>(lds, E);
    return tv;
  } else if constexpr (TMODE == 3) {
    tv.pw = a.t_w; tv.rows = a.t_rows; tv.rec = a.t_rec; tv.E = E;
    tv.xs = xs_slots<3>(lds, E);
    return tv;
  } else if constexpr (TMODE == 2) {
    tv.pw = a.t_w; tv.pjr = a.t_jr; tv.pir = a.t_ir; tv.pmm = a.t_mmL;
    return tv;
  } else {
    double* w = reinterpret_cast<double*>(lds);
    for (uint32_t i = threadIdx.x; i < NT; i += blockDim.x) w[i] = a.t_w[i];
    tv.pw = w;
    if constexpr (TMODE == 0) {
      double* jr = w + NT;
      double* ir = jr + NT;
      uint8_t* mm = reinterpret_cast<uint8_t*>(ir + NT);
      for (uint32_t i = threadIdx.x; i < NT; i += blockDim.x) {
        jr[i] = a.t_jr[i];
        ir[i] = a.t_ir[i];
        mm[i] = a.t_mmL[i];
      }
      tv.pjr = jr; tv.pir = ir; tv.pmm = mm;
    } else {
      tv.pjr = a.t_jr; tv.pir = a.t_ir; tv.pmm = a.t_mmL;
    }
    __syncthreads();
    return tv;
  }
}

/* ------------------------------------------------------ reductions -- */
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

/* state <- q(M) state for a jump polynomial q that differs between lanes
 * (the accumulation is masked per lane instead of branched on). */
__device__ inline void apply_poly_lane(Xoshiro& g, const uint64_t* q) {
  Xoshiro acc{0, 0, 0, 0};
  for (int w = 0; w < 4; w++) {
    const uint64_t bits = q[w];
    for (int b = 0; b < 64; b++) {
      const uint64_t m = 0ull - ((bits >> b) & 1ull);
      acc.s0 ^= g.s0 & m;
      acc.s1 ^= g.s1 & m;
      acc.s2 ^= g.s2 & m;
      acc.s3 ^= g.s3 & m;
      g.next();
    }
  }
  g = acc;
}

/* Lane stream positioning: the block's first state comes from the device
 * (ftrj_block_states); sample slot t of the block (the lane, or with LPS
 * lanes per sample the lane / LPS) applies, for each base-256 digit d_k of
 * t, the jump polynomial of d_k 256^k slots (lane_poly[255 k + d_k - 1]):
 * one polynomial for blocks of up to 256 slots. */
template <int LPS = 1>
__device__ inline Xoshiro lane_stream(const Launch& a, uint32_t bid) {
  const uint64_t* bs = a.block_state + 4 * (size_t)bid;
  Xoshiro g{bs[0], bs[1], bs[2], bs[3]};
  const unsigned t = threadIdx.x / LPS;
  for (int k = 0; (1u << (8 * k)) < blockDim.x / LPS; k++) {
    const unsigned d = (t >> (8 * k)) & 255u;
    if (d) apply_poly_lane(g, a.lane_poly + 4 * (255 * k + d - 1));
  }
  return g;
}

/* Main body: each sample slot (a lane, or a pair of lanes with
 * G::LPS = 2) draws its contiguous run of samples from the single stream,
 * evaluates the integrand and accumulates (n, drops, S1, S2) per order; the
 * block writes one row of partial moments [n, drops, (s1re, s2re, s1im,
 * s2im) x K].  Both lanes of a pair draw the same samples and hold the same
 * values; the pair's second lane leaves the reduction. */
template <class G, int K, int MODE, int TMODE>
__device__ inline void integrate_body(const Launch& a, uint32_t bid) {
  extern __shared__ __align__(16) char lds[];
  constexpr int E = G::E;
  constexpr int NQ = 2 + 4 * K;
  using PT = typename G::PRM;
  const ParamsFor<PT> ps{};
  const PT& p = ps.get(a);
  TableView<TMODE> tv = stage_table<TMODE>(a, E, lds);
  double* red = red_area<G, TMODE>(lds);
  double* const dds = dd_slots<G, TMODE>(lds);

  constexpr int LPS = G::LPS;
  Xoshiro rng = lane_stream<LPS>(a, bid);
  const uint64_t lane = (uint64_t)bid * (blockDim.x / LPS) + threadIdx.x / LPS;
  const uint64_t i0 = lane * a.S;
  /* this slot's sample count (one loop register instead of two bounds) */
  const uint64_t n = (i0 < a.N) ? ((a.N - i0 < a.S) ? a.N - i0 : a.S) : 0;

  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) acc[q] = 0.0;
  /* G::PIPE: software-pipelined -- sample i+1 is drawn (sampler unrolled)
   * right after sample i's point is taken, so that its dependent table
   * loads overlap sample i's integrand; the lane's last draw is discarded.
   * The samples and their order are those of the plain loop. */
  double Xn[E], lpn = 0.0, lGn = 0.0;
  if constexpr (G::PIPE) {
    if (n > 0) sample_point<G, MODE != 2, G::PF, true>(rng, tv, Xn, lpn, lGn);
  }
  for (uint64_t j = n; j > 0; j--) {
    double X[E], lpsi, lG;
    if constexpr (G::PIPE) {
#pragma unroll
      for (int e = 0; e < E; e++) X[e] = Xn[e];
      lpsi = lpn;
      lG = lGn;
      sample_point<G, MODE != 2, G::PF, true>(rng, tv, Xn, lpn, lGn);
    } else {
      sample_point<G, MODE != 2, G::PF>(rng, tv, X, lpsi, lG);
    }
    cplx out[K];
    if (!integrand<G, K, MODE, PT, G::DDL, LPS>(p, X, log_scale(p, lpsi, lG), out, dds)) {
      acc[1] += 1.0;
      continue;
    }
    acc[0] += 1.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      acc[2 + 4 * k] += out[k].re;
      acc[3 + 4 * k] += out[k].re * out[k].re;
      if constexpr (MODE == 2) {
        acc[4 + 4 * k] += out[k].im;
        acc[5 + 4 * k] += out[k].im * out[k].im;
      }
    }
  }
  if constexpr (LPS > 1) {
    if (threadIdx.x % LPS) {
#pragma unroll
      for (int q = 0; q < NQ; q++) acc[q] = 0.0;
    }
  }
  /* wave shuffle tree, then LDS across waves, fixed order: deterministic */
  const int wid = threadIdx.x >> 6, lid = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const double v = wave_sum(acc[q]);
    if (lid == 0) red[wid * NQ + q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += red[w * NQ + threadIdx.x];
    a.partials[(size_t)bid * NQ + threadIdx.x] = s;
  }
}

/* Replay: evaluate the integrand on given points (one sample per lane;
 * with G::LPS = 2 on the deformed path, one per lane pair, the second
 * derivatives in dynamic LDS as in the integration kernel: the block must
 * be G::BLK threads).  X: n x E, psixi: n x 2 -> out: n x K x 2, dropped: n. */
template <class G, int K, int MODE>
__device__ inline void replay_body(const DevParams* prm, uint64_t n, const double* Xin,
                                   const double* psixi, double* out, uint8_t* dropped) {
  constexpr int E = G::E;
  constexpr int LPS = MODE == 2 ? G::LPS : 1;
  const uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPS;
  if (i >= n) return;
  double X[E];
#pragma unroll
  for (int e = 0; e < E; e++) X[e] = Xin[i * E + e];
  cplx r[K];
  const double ps = psixi[2 * i], xs = psixi[2 * i + 1];
  const double ls = prm->inv_dlw * (prm->mhalfD * log(ps) + (prm->W != 0.0 ? prm->W * log(ps / xs) : 0.0));
  bool ok;
  if constexpr (LPS == 1) {
    ok = integrand<G, K, MODE>(*prm, X, ls, r);
  } else {
    extern __shared__ __align__(16) char lds[];
    ok = integrand<G, K, MODE, DevParams, true, LPS>(*prm, X, ls, r,
                                                      reinterpret_cast<double*>(lds) + threadIdx.x / LPS);
    if (threadIdx.x % LPS) return;
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    out[(i * K + k) * 2] = ok ? r[k].re : 0.0;
    out[(i * K + k) * 2 + 1] = ok ? r[k].im : 0.0;
  }
  dropped[i] = ok ? 0 : 1;
}

/* Sampler only: lane l writes samples [l S, min((l+1) S, N)) as X, psixi. */
template <class G, int TMODE>
__device__ inline void sample_body(const Launch& a, double* Xout, double* pxout) {
  extern __shared__ __align__(16) char lds[];
  constexpr int E = G::E;
  TableView<TMODE> tv = stage_table<TMODE>(a, E, lds);
  Xoshiro rng = lane_stream(a, blockIdx.x);
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i0 = lane * a.S;
  const uint64_t i1 = (i0 + a.S < a.N) ? i0 + a.S : a.N;
  for (uint64_t i = i0; i < i1; i++) {
    double X[E], lps, lG;
    sample_point<G>(rng, tv, X, lps, lG);
#pragma unroll
    for (int e = 0; e < E; e++) Xout[i * E + e] = X[e];
    const double ps = exp(lps);
    pxout[2 * i] = ps;
    pxout[2 * i + 1] = exp(lG) * ps;
  }
}

/* Learned proposal: the cut and the radius distributions come from the
 * learned tables; the exact FP64 ratio p_trop/q of the drawn path multiplies
 * the integrand, so E_q[R p/q] is the same integral.  Same 2E-1 draws per
 * sample as the plain sampler.
 *
 * One sample: X, log psi_tr, log G, the weight p_trop/q, the cut order (4
 * bits per step) and, with KEEP_LY, the radial log-draws (stored to ly_out
 * with stride ly_stride).  The E cut weights
 * of the current state are loaded together (one memory latency per step,
 * not one per candidate), and the running sum over them in ascending edge
 * order selects the cut as the tropical sampler does (zero weights of
 * absent edges leave the sum unchanged).  XS (row placement): the radius
 * of the cut edge is stored to this thread's LDS slot xs[e * G::BLK]
 * instead of selected into X over all E edges. */
template <class G, bool KEEP_LY, bool XS>
__device__ inline void learn_point(const LearnLaunch& L, Xoshiro& rng, double (&X)[G::E], double& lpsi,
                                   double& lG, double& wgt, uint64_t& path, double* ly_out, uint64_t ly_stride,
                                   double* xs) {
  constexpr int E = G::E;
  uint32_t g = (1u << E) - 1u;
  LQuad st = L.lst[g];
  bool MM = st.mmL & 1;
  int oL = (int)(st.mmL >> 1);
  double lGv = 0.0, lF = 0.0, lX = 0.0, XX = 1.0, ratio = 1.0, lrad = 0.0;
  /* the cut's ratio factor is a load that depends on the cut: it is
   * multiplied in one step later (the same products in the same order) */
  double fpend = 1.0;
  path = 0;
#pragma unroll
  for (int e = 0; e < E; e++) X[e] = 0.0;
  /* the state's learned weights, all in flight together; the row of the
   * state a cut reaches is requested in the cut's own step, after its
   * record, so that it is in flight during the radial draw (measured: 21.9
   * ms per 1e8 on the triple ladder against 24.1 ms when each step requests
   * its own row, double box 12.4 against 12.9 ms) */
  double rv[E];
#pragma unroll
  for (int q = 0; q < E; q++) rv[q] = L.lw[(size_t)g * E + q];
  for (int j = 0; j < E; j++) {
    const double R = rng.uniform() * st.lZ;
    /* first q with T_q > R = the number of nondecreasing sums <= R */
    double T = 0.0;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < E; q++) {
      T += rv[q];
      cnt += (T <= R) ? 1 : 0;
    }
    const int e = (cnt < E) ? cnt : 31 - __builtin_clz(g);
    const uint32_t c = g ^ (1u << e);
    ratio *= fpend;
    const LEdge ed = L.led[(size_t)g * E + e];
    fpend = st.A * ed.tw;
    st = L.lst[c];
    if (j < E - 1) {
#pragma unroll
      for (int q = 0; q < E; q++) rv[q] = L.lw[(size_t)c * E + q];
    }
    const bool mm = st.mmL & 1;
    const int Lc = (int)(st.mmL >> 1);
    if (MM && !mm) lGv = lX;
    MM = mm;
    if constexpr (XS) {
      xs[e * G::BLK] = XX;
    } else {
#pragma unroll
      for (int k = 0; k < E; k++) X[k] = (k == e) ? XX : X[k];
    }
    if (Lc < oL) {
      lF += lX;
      oL = Lc;
    }
    path |= (uint64_t)e << (4 * j);
    if (j < E - 1) {
      const double U = rng.uniform();
      const double lY = ed.lia * log_pos(1. - U);
      lX += lY;
      XX = exp(lX);
      lrad += ed.lrad + ed.ldiff * lY;
      if (KEEP_LY) ly_out[(size_t)j * ly_stride] = lY;
    }
    g = c;
  }
  if constexpr (XS) {
#pragma unroll
    for (int k = 0; k < E; k++) X[k] = xs[k * G::BLK];
  }
  lpsi = lF;
  lG = lGv;
  wgt = (ratio * fpend) * exp(lrad);
}

/* Integration with the learned proposal.  GRAD (training) also records
 * every sample's path, c = |f| p_trop/q or its square (objective; 0 when
 * dropped) and radial
 * log-draws for learn_acc_body: gbuf[i] = c, gbuf[(1+j) N + i] = lY_j,
 * gpath[i] = cut order. */
template <class G, int K, int MODE, int TMODE, bool GRAD>
__device__ inline void learn_body(const LearnLaunch& L) {
  static_assert(G::LPS == 1, "learned kernels evaluate one sample per lane");
  extern __shared__ __align__(16) char lds[];
  constexpr int E = G::E;
  constexpr int NQ = 2 + 4 * K;
  const Launch& a = L.a;
  using PT = typename G::PRM;
  const ParamsFor<PT> ps{};
  const PT& p = ps.get(a);
  /* the learned tables are read from L2/HBM: the LDS holds the X slots
   * (TMODE 3, 4) and the reduction */
  double* red = red_area<G, TMODE>(lds);
  double* xs = xs_slots<TMODE>(lds, E);
  double* const dds = dd_slots<G, TMODE>(lds);
  Xoshiro rng = lane_stream(a, blockIdx.x);
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i0 = lane * a.S;
  const uint64_t i1 = (i0 + a.S < a.N) ? i0 + a.S : a.N;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) acc[q] = 0.0;
  for (uint64_t i = i0; i < i1; i++) {
    double X[E], lF, lGv, wgt;
    uint64_t path;
    learn_point<G, GRAD, TMODE >= 3>(L, rng, X, lF, lGv, wgt, path, GRAD ? L.gbuf + a.N + i : nullptr, a.N, xs);
    if (GRAD) L.gpath[i] = path;
    cplx out[K];
    if (!integrand<G, K, MODE, PT, G::DDL>(p, X, log_scale(p, lF, lGv), out, dds)) {
      acc[1] += 1.0;
      if (GRAD) L.gbuf[i] = 0.0;
      continue;
    }
    acc[0] += 1.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double re = out[k].re * wgt, im = out[k].im * wgt;
      acc[2 + 4 * k] += re;
      acc[3 + 4 * k] += re * re;
      acc[4 + 4 * k] += im;
      acc[5 + 4 * k] += im * im;
    }
    if (GRAD) {
      const double cf = wgt * sqrt(out[0].re * out[0].re + out[0].im * out[0].im);
      L.gbuf[i] = L.objective ? cf * cf : cf;
    }
  }
  const int wid = threadIdx.x >> 6, lid = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const double v = wave_sum(acc[q]);
    if (lid == 0) red[wid * NQ + q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double sm = 0.0;
    for (int w = 0; w < nw; w++) sm += red[w * NQ + threadIdx.x];
    a.partials[(size_t)blockIdx.x * NQ + threadIdx.x] = sm;
  }
}

/* Forward-KL statistics of the recorded samples:
 *   C(g, e) = sum of c over visits of state g that cut e,
 *   S(g, e) = sum of c (1 + a(g,e) lY) over the radii drawn after cutting
 *             e from g,
 *   gsum    = (sum of c, number of samples).
 * They are kept in fixed point: every value is rounded to an integer
 * multiple of 2^-k and added in 64-bit integers, so the sums do not depend
 * on the order in which threads add them and training is reproducible bit
 * for bit.  k is fixed per step from the largest |value| M over the step's N
 * samples (learn_max_body) so that N M 2^k < 2^62: no sum can overflow.  The
 * rounding is relative to the largest value (2^-62 N M at most); the
 * gradients are formed in FP32.  Non-finite values count as 0.
 * Layout of an accumulator (learn_hist_len entries): C(g,e) at g E + e
 * (e in g), S(g,e) at (g\e) E + e (e not in the slot's state, so no C
 * lives there), then the two sums. */
__host__ __device__ constexpr int learn_pairs(int E) { return E * (E - 1) / 2; }
/* the cut (C) and radial (S) statistics of the three top levels, each
 * indexed by (state, cut edge): E + E^2 + pairs * E per kind */
__host__ __device__ constexpr int learn_top_len(int E) { return E + E * E + learn_pairs(E) * E; }
__host__ __device__ constexpr size_t learn_acc_lds(int E) { return (size_t)(2 + 2 * learn_top_len(E)) * 8; }
__host__ __device__ constexpr size_t learn_hist_len(int E) { return (size_t)E * ((size_t)1 << E) + 2; }
__device__ inline int pair_index(int a, int b) {
  const int hi = a > b ? a : b, lo = a > b ? b : a;
  return hi * (hi - 1) / 2 + lo;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

/* c (1 + a(g,e) lY_j): the radial statistic of step j of sample i, whose
 * cut was e from g (k = g E + e) */
__device__ inline double learn_radial(const LearnLaunch& L, double c, uint64_t N, uint64_t i, int j, size_t k) {
  return c * (1.0 + L.gbuf[(size_t)(1 + j) * N + i] / L.lia[k]);
}
/* accumulator slot of S(g,e): (g\e) E + e, never a slot of C */
__device__ inline size_t learn_s_slot(uint32_t g, int e, int E) { return (size_t)(g ^ (1u << e)) * E + e; }

__device__ inline bool finite_d(double v) { return __builtin_isfinite(v); }

/* 2^k with N M 2^k < 2^62 (M = *gmax); exponents from the bit patterns
 * (frexp convention: x < 2^e), k clamped to [-1000, 1000]. */
__device__ inline double learn_fix_scale(const LearnLaunch& L, uint64_t N) {
  const long long mb = (long long)*L.gmax;
  if (mb <= 0) return 1.0;
  const int em = (int)((mb >> 52) & 0x7ff) - 1022;
  const int en = (int)((__double_as_longlong((double)N) >> 52) & 0x7ff) - 1022;
  int k = 62 - em - en;
  k = k > 1000 ? 1000 : (k < -1000 ? -1000 : k);
  return __longlong_as_double((long long)(k + 1023) << 52);
}

__device__ inline fixacc to_fix(double v, double scale) { return (fixacc)(long long)rint(v * scale); }

/* *gmax = max |value| over the step's samples, as the bits of a
 * non-negative double (whose order is that of the integers).  Grid-stride
 * over the samples; one atomic per block (atomics on one address
 * serialise). */
template <int E>
__device__ inline void learn_max_body(const LearnLaunch& L, uint64_t N) {
  __shared__ double wm[16];
  double m = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const double c = L.gbuf[i];
    if (c == 0.0 || !finite_d(c)) continue;
    m = fmax(m, fabs(c));
    const uint64_t path = L.gpath[i];
    uint32_t g = (1u << E) - 1u;
    for (int j = 0; j < E - 1; j++) {
      const int e = (int)((path >> (4 * j)) & 15u);
      const double sv = learn_radial(L, c, N, i, j, (size_t)g * E + e);
      if (finite_d(sv)) m = fmax(m, fabs(sv));
      g ^= 1u << e;
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, wm[w]);
    if (m > 0.0) atomicMax((unsigned long long*)L.gmax, (unsigned long long)__double_as_longlong(m));
  }
}

/* E > FTR_LEARN_LDS_E: one thread per sample, adding into one accumulator
 * `acc` (zeroed beforehand; slots as learn_hist_len).  The (state, cut)
 * pairs of the first three levels (full graph, one and two edges removed)
 * are shared by many samples: their C and S sums go through LDS (fixed
 * indices: removed edges, then the cut) and reach global memory once per
 * block. */
template <int E>
__device__ inline void learn_acc_body(const LearnLaunch& L, uint64_t N, fixacc* acc) {
  extern __shared__ __align__(16) char lds[];
  fixacc* sh = reinterpret_cast<fixacc*>(lds);
  constexpr uint32_t NT = 1u << E;
  constexpr int NB = learn_top_len(E);
  fixacc* s_sum = sh;          /* 2 */
  fixacc* s_c = s_sum + 2;     /* NB: level 0 (e0), 1 (e0 E + e1), 2 (pair(e0,e1) E + e2) */
  fixacc* s_s = s_c + NB;      /* NB: the same indices for S */
  constexpr int NSH = 2 + 2 * NB;
  fixacc* asum = acc + (size_t)NT * E;
  for (int q = threadIdx.x; q < NSH; q += blockDim.x) sh[q] = 0;
  __syncthreads();
  const double scale = learn_fix_scale(L, N);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    const double c = L.gbuf[i];
    atomicAdd(&s_sum[1], (fixacc)1);
    if (c != 0.0 && finite_d(c)) {
      const fixacc qc = to_fix(c, scale);
      const uint64_t path = L.gpath[i];
      atomicAdd(&s_sum[0], qc);
      uint32_t g = NT - 1u;
      int e0 = 0, e1 = 0;
      for (int j = 0; j < E; j++) {
        const int e = (int)((path >> (4 * j)) & 15u);
        fixacc qs = 0;
        if (j < E - 1) {
          const double sv = learn_radial(L, c, N, i, j, (size_t)g * E + e);
          qs = finite_d(sv) ? to_fix(sv, scale) : (fixacc)0;
        }
        int t = -1;
        if (j == 0) {
          e0 = e;
          t = e;
        } else if (j == 1) {
          e1 = e;
          t = E + e0 * E + e;
        } else if (j == 2) {
          t = E + E * E + pair_index(e0, e1) * E + e;
        }
        if (t >= 0) {
          atomicAdd(&s_c[t], qc);
          atomicAdd(&s_s[t], qs);
        } else {
          atomicAdd(&acc[(size_t)g * E + e], qc);
          if (j < E - 1) atomicAdd(&acc[learn_s_slot(g, e, E)], qs);
        }
        g ^= 1u << e;
      }
    }
  }
  __syncthreads();
  const uint32_t full = NT - 1u;
  for (int q = threadIdx.x; q < NSH; q += blockDim.x) {
    const fixacc v = sh[q];
    if (v == 0) continue;
    if (q < 2) {
      atomicAdd(&asum[q], v);
      continue;
    }
    const bool is_s = q >= 2 + NB;
    int k = q - 2 - (is_s ? NB : 0);
    uint32_t g;
    int e;
    if (k < E) {
      g = full;
      e = k;
    } else if (k < E + E * E) {
      k -= E;
      g = full ^ (1u << (k / E));
      e = k % E;
    } else {
      k -= E + E * E;
      const int pidx = k / E;
      e = k % E;
      int hi = 1;
      while ((hi + 1) * hi / 2 <= pidx) hi++;
      const int lo = pidx - hi * (hi - 1) / 2;
      g = full ^ (1u << hi) ^ (1u << lo);
    }
    atomicAdd(&acc[is_s ? learn_s_slot(g, e, E) : (size_t)g * E + e], v);
  }
}

/* E <= FTR_LEARN_LDS_E: every block holds a whole accumulator in LDS, walks
 * a strided share of the recorded samples and stores its accumulator to its
 * own slice of `part`; learn_reduce_body sums the slices. */
#define FTR_LEARN_LDS_E 10

template <int E>
__device__ inline void learn_hist_body(const LearnLaunch& L, uint64_t N, fixacc* part) {
  if constexpr (E <= FTR_LEARN_LDS_E) {
    extern __shared__ __align__(16) char lds[];
    fixacc* h = reinterpret_cast<fixacc*>(lds);
    constexpr uint32_t NT = 1u << E;
    constexpr size_t HL = learn_hist_len(E);
    fixacc* hCS = h;                  /* NT x E: C and S slots */
    fixacc* hsum = h + (size_t)NT * E; /* 2 */
    for (size_t q = threadIdx.x; q < HL; q += blockDim.x) h[q] = 0;
    __syncthreads();
    const double scale = learn_fix_scale(L, N);
    fixacc c_sum = 0, n_sum = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
      const double c = L.gbuf[i];
      n_sum += 1;
      if (c == 0.0 || !finite_d(c)) continue;
      const fixacc qc = to_fix(c, scale);
      c_sum += qc;
      const uint64_t path = L.gpath[i];
      uint32_t g = NT - 1u;
      for (int j = 0; j < E; j++) {
        const int e = (int)((path >> (4 * j)) & 15u);
        atomicAdd(&hCS[(size_t)g * E + e], qc);
        if (j < E - 1) {
          const double sv = learn_radial(L, c, N, i, j, (size_t)g * E + e);
          if (finite_d(sv)) atomicAdd(&hCS[learn_s_slot(g, e, E)], to_fix(sv, scale));
        }
        g ^= 1u << e;
      }
    }
    atomicAdd(&hsum[0], c_sum);
    atomicAdd(&hsum[1], n_sum);
    __syncthreads();
    fixacc* out = part + (size_t)blockIdx.x * HL;
    for (size_t q = threadIdx.x; q < HL; q += blockDim.x) out[q] = h[q];
  }
}

/* gC, gS, gsum from the nblk accumulators (one thread per entry): the
 * integer sum, times 2^-k (the sample count unscaled). */
template <int E>
__device__ inline void learn_reduce_body(const LearnLaunch& L, const fixacc* part, int nblk, uint64_t N) {
  constexpr uint32_t NT = 1u << E;
  constexpr size_t HL = learn_hist_len(E);
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= HL) return;
  fixacc sb[8] = {0, 0, 0, 0, 0, 0, 0, 0}; /* 8 loads in flight */
  int b = 0;
  for (; b + 8 <= nblk; b += 8)
#pragma unroll
    for (int t = 0; t < 8; t++) sb[t] += part[(size_t)(b + t) * HL + q];
  for (; b < nblk; b++) sb[0] += part[(size_t)b * HL + q];
  const fixacc s = ((sb[0] + sb[1]) + (sb[2] + sb[3])) + ((sb[4] + sb[5]) + (sb[6] + sb[7]));
  if (q == HL - 1) {
    L.gsum[1] = (double)s;
    return;
  }
  const double v = (double)(long long)s / learn_fix_scale(L, N);
  if (q < (size_t)NT * E) {
    const uint32_t g = (uint32_t)(q / E);
    const int e = (int)(q % E);
    if ((g >> e) & 1u) {
      L.gC[q] = v; /* gS[q] is written by the thread of slot (g\e) E + e */
    } else {
      L.gC[q] = 0.0;
      L.gS[q] = 0.0;
      L.gS[(size_t)(g | (1u << e)) * E + e] = v;
    }
  } else {
    L.gsum[0] = v;
  }
}

/* Learned sampler only (parity): X, the weight p/q and the cut order of
 * samples [first, first+n) of the stream. */
template <class G, int TMODE>
__device__ inline void learn_sample_body(const LearnLaunch& L, double* Xout, double* wout, uint64_t* pout) {
  extern __shared__ __align__(16) char lds[];
  constexpr int E = G::E;
  const Launch& a = L.a;
  Xoshiro rng = lane_stream(a, blockIdx.x);
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i0 = lane * a.S;
  const uint64_t i1 = (i0 + a.S < a.N) ? i0 + a.S : a.N;
  for (uint64_t i = i0; i < i1; i++) {
    double X[E], lF, lGv, wgt;
    uint64_t path;
    learn_point<G, false, TMODE >= 3>(L, rng, X, lF, lGv, wgt, path, nullptr, 0, xs_slots<TMODE>(lds, E));
#pragma unroll
    for (int e = 0; e < E; e++) Xout[i * E + e] = X[e];
    wout[i] = wgt;
    pout[i] = path;
  }
}

/* Persistent batched launch (no reference counterpart): the next unit of
 * work for this block, or nwork when the list is exhausted.  Thread 0 takes
 * the unit from the counter; the two barriers publish it to the block and
 * keep the slot from being overwritten before every thread has read it.
 * Every block leaves the loop once the counter passes nwork, so the grid
 * drains. */
__device__ inline uint32_t persistent_next(uint32_t* counter, uint32_t nwork) {
  __shared__ uint32_t slot;
  if (threadIdx.x == 0) {
    const uint32_t w = atomicAdd(counter, 1u);
    slot = w < nwork ? w : nwork;
  }
  __syncthreads();
  const uint32_t w = slot;
  __syncthreads();
  return w;
}

}  // namespace ftr

/* A bucket kernel: graph specs FG0..FGn-1 sharing (K, MODE, TMODE, BLOCK),
 * one persistent grid over BatchWork units; the generated source lists one
 * FTR_BUCKET_CASE per graph between FTR_BUCKET_BEGIN and FTR_BUCKET_END. */
#define FTR_BUCKET_BEGIN(NAME, BLOCK)                                                                 \
  extern "C" __global__ void __launch_bounds__(BLOCK)                                                 \
      NAME(const ftr::BatchItem* items, const ftr::BatchWork* work, uint32_t nwork, uint32_t* counter) { \
    for (;;) {                                                                                        \
      const uint32_t w_ = ftr::persistent_next(counter, nwork);                                       \
      if (w_ >= nwork) return;                                                                        \
      const ftr::BatchWork wk_ = work[w_];                                                            \
      switch (wk_.topo) {
#define FTR_BUCKET_CASE(I, G, K, MODE, TMODE)                                                         \
  case I:                                                                                             \
    ftr::integrate_body<G, K, MODE, TMODE>(items[wk_.item].a, wk_.blk);                               \
    break;
#define FTR_BUCKET_END                                                                                \
  default:                                                                                            \
    break;                                                                                            \
    }                                                                                                 \
    }                                                                                                 \
    }

/* Learned-proposal kernels (compiled only for problems that use them). */
#define FTR_INSTANTIATE_LEARNED(NAME, G, K, MODE, TMODE, BLOCK)                                   \
  extern "C" __global__ void __launch_bounds__(BLOCK) NAME##_lev(ftr::LearnLaunch L) {              \
    ftr::learn_body<G, K, MODE, TMODE, false>(L);                                                 \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(BLOCK) NAME##_lgr(ftr::LearnLaunch L) {              \
    ftr::learn_body<G, K, MODE, TMODE, true>(L);                                                  \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(BLOCK) NAME##_lsm(ftr::LearnLaunch L, double* X,     \
                                                                   double* w, uint64_t* path) {   \
    ftr::learn_sample_body<G, TMODE>(L, X, w, path);                                              \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(1024) NAME##_hmx(ftr::LearnLaunch L, uint64_t N) {   \
    ftr::learn_max_body<G::E>(L, N);                                                              \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(256) NAME##_acc(ftr::LearnLaunch L, uint64_t N,      \
                                                              ftr::fixacc* acc) {                 \
    ftr::learn_acc_body<G::E>(L, N, acc);                                                         \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(1024) NAME##_hst(ftr::LearnLaunch L, uint64_t N,     \
                                                                ftr::fixacc* part) {              \
    ftr::learn_hist_body<G::E>(L, N, part);                                                       \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(256) NAME##_hrd(ftr::LearnLaunch L,                  \
                                                               const ftr::fixacc* part, int nblk, \
                                                               uint64_t N) {                      \
    ftr::learn_reduce_body<G::E>(L, part, nblk, N);                                               \
  }

/* Instantiation macro used by the generated kernel units:
 *   FTR_INSTANTIATE(name, Graph, K, MODE, TMODE, BLOCK)
 * defines extern "C" kernels  name_int, name_rep, name_smp. */
#ifndef FTR_INT_ATTR
#define FTR_INT_ATTR /* extra attributes of the integration kernel (compile option) */
#endif
#define FTR_INSTANTIATE(NAME, G, K, MODE, TMODE, BLOCK)                                           \
  extern "C" __global__ void __launch_bounds__(BLOCK) FTR_INT_ATTR NAME##_int(ftr::Launch a) {    \
    ftr::integrate_body<G, K, MODE, TMODE>(a, blockIdx.x);                                        \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(G::LPS == 2 ? (BLOCK) : 256)                       \
      NAME##_rep(const ftr::DevParams* prm, uint64_t n, const double* X, const double* px,        \
                 double* out, uint8_t* dropped) {                                                 \
    ftr::replay_body<G, K, MODE>(prm, n, X, px, out, dropped);                                    \
  }                                                                                               \
  extern "C" __global__ void __launch_bounds__(BLOCK) NAME##_smp(ftr::Launch a, double* X,        \
                                                                   double* px) {                  \
    ftr::sample_body<G, TMODE>(a, X, px);                                                         \
  }
