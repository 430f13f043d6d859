/*
 * ftr_oracle.cpp -- CPU FP64 restatement of feyntrop's sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Built three
 * times from this file: strict IEEE double (-O2 -ffp-contract=off, no
 * fast-math) as libftr_oracle.so, FMA-contracted double as
 * libftr_oracle_fma.so, and x87 long double (ORACLE_REAL) as
 * libftr_oracle_ld.so.  The long-double build is the ground truth that
 * defines the well-conditioned mask: samples on which the double evaluation
 * of the reference algorithm is itself accurate to 1e-13.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   xoshiro256+/splitmix/uniform  -- published xoshiro256+ test stream (first
 *       outputs for seed 0 from the compiled reference, tests/golden) and the
 *       definition; or_xoshiro_advance == d x next() by brute force; the
 *       polynomial x^(2^128) mod charpoly equals the published 2^128 jump
 *       constant and x^(2^192) the long-jump constant (random.hpp:148,178).
 *   J-table / IGtr                -- closed-form IGtr of the fixtures (84 for
 *       the K4 wheel, 1.5/3/3/14/117/7.3333/50.25), brute force over edge
 *       orderings for E <= 7, and the compiled reference table (oracle/_ref).
 *   sampler                      -- bitwise vs the compiled reference on the
 *       same stream; ordering probabilities vs exact ones (chi^2).
 *   integrand                    -- Kirchhoff / 2-forest brute force for psi
 *       and pphi, finite differences for the derivatives, 6 zeta(3) for the
 *       wheel eps^0, the compiled reference per sample, the 7 golden fixtures.
 *   fp_summer / stats            -- exact rational sums (fractions) in tests.
 */
#include "ftr_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
/* The arithmetic type: double for the strict / FMA builds, long double for
 * the extended-precision build that serves as ground truth. */
typedef ORACLE_REAL real;
typedef std::complex<real> cd;

/* ------------------------------------------------------------------ RNG -- */
/* random.hpp:16-19, :61-66, :86-92, :125-140 */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t splitmix_next(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Xo {
  uint64_t s[4];
  uint64_t next() {
    const uint64_t res = s[0] + s[3];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return res;
  }
  real uniform() { return (real)(next() >> 11) * (real(1) / (real)(UINT64_C(1) << 53)); }
};

static void xo_seed(uint64_t seed, uint64_t* st) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) st[i] = splitmix_next(x);
}

/* GF(2) polynomials of degree < 256 as 4 words; the characteristic polynomial
 * has the implicit x^256 term plus 4 low words. */
struct Poly4 { uint64_t w[4]; };

static Poly4 g_charpoly;
static bool g_charpoly_ready = false;

static void compute_charpoly() {
  if (g_charpoly_ready) return;
  /* Berlekamp-Massey over GF(2) on bit 0 of s[0] (a linear functional of the
   * 256-bit state).  The minimal polynomial of xoshiro256's transition has
   * degree 256, so 512 terms determine it. */
  const int NS = 1024;
  std::vector<uint8_t> s(NS);
  Xo g;
  xo_seed(0x123456789abcdefULL, g.s);
  for (int t = 0; t < NS; t++) { s[t] = (uint8_t)(g.s[0] & 1); g.next(); }
  std::vector<uint8_t> C(NS + 1, 0), B(NS + 1, 0), T;
  C[0] = B[0] = 1;
  int L = 0, m = 1;
  for (int n = 0; n < NS; n++) {
    uint8_t d = s[n];
    for (int i = 1; i <= L; i++) d ^= (uint8_t)(C[i] & s[n - i]);
    if (d == 0) { m++; continue; }
    if (2 * L <= n) {
      T = C;
      for (int i = 0; i + m <= NS; i++) C[i + m] ^= B[i];
      L = n + 1 - L; B = T; m = 1;
    } else {
      for (int i = 0; i + m <= NS; i++) C[i + m] ^= B[i];
      m++;
    }
  }
  if (L != 256) throw std::runtime_error("xoshiro charpoly: unexpected degree");
  /* p(x) = x^L C(1/x): coefficient of x^(L-i) is C_i. */
  Poly4 p; memset(&p, 0, sizeof p);
  for (int i = 1; i <= L; i++)
    if (C[i]) { int j = L - i; p.w[j >> 6] |= UINT64_C(1) << (j & 63); }
  g_charpoly = p;
  g_charpoly_ready = true;
}

/* a*b mod p */
static Poly4 poly_mulmod(const Poly4& a, const Poly4& b) {
  uint64_t r[8] = {0};
  for (int i = 0; i < 256; i++) {
    if (!((a.w[i >> 6] >> (i & 63)) & 1)) continue;
    /* r ^= b << i */
    int ws = i >> 6, bs = i & 63;
    for (int k = 0; k < 4; k++) {
      r[k + ws] ^= b.w[k] << bs;
      if (bs) r[k + ws + 1] ^= b.w[k] >> (64 - bs);
    }
  }
  /* reduce: x^256 == charpoly low part */
  for (int k = 511; k >= 256; k--) {
    if (!((r[k >> 6] >> (k & 63)) & 1)) continue;
    r[k >> 6] ^= UINT64_C(1) << (k & 63);
    int sh = k - 256, ws = sh >> 6, bs = sh & 63;
    for (int q = 0; q < 4; q++) {
      r[q + ws] ^= g_charpoly.w[q] << bs;
      if (bs && q + ws + 1 < 8) r[q + ws + 1] ^= g_charpoly.w[q] >> (64 - bs);
    }
  }
  Poly4 out;
  for (int k = 0; k < 4; k++) out.w[k] = r[k];
  return out;
}

/* x^d mod p for a 128-bit d */
static Poly4 poly_xpow(uint64_t lo, uint64_t hi) {
  compute_charpoly();
  Poly4 res; memset(&res, 0, sizeof res); res.w[0] = 1;
  Poly4 base; memset(&base, 0, sizeof base); base.w[0] = 2; /* x */
  for (int bit = 0; bit < 128; bit++) {
    uint64_t wv = bit < 64 ? lo : hi;
    if ((wv >> (bit & 63)) & 1) res = poly_mulmod(res, base);
    base = poly_mulmod(base, base);
  }
  return res;
}

static void poly_apply(const Poly4& q, uint64_t* st) {
  Xo g; memcpy(g.s, st, 32);
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int i = 0; i < 256; i++) {
    if ((q.w[i >> 6] >> (i & 63)) & 1)
      for (int k = 0; k < 4; k++) acc[k] ^= g.s[k];
    g.next();
  }
  memcpy(st, acc, 32);
}

/* 128-bit multiply helper for jump distances */
static void mul_u64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
  unsigned __int128 r = (unsigned __int128)a * b;
  *lo = (uint64_t)r; *hi = (uint64_t)(r >> 64);
}

/* --------------------------------------------------------------- graph -- */
struct Graph {
  int E = 0, V = 0;
  std::vector<int> hi, lo;   /* edge j: (hi[j], lo[j]) with hi > lo (graph.hpp:181-185) */
  std::vector<real> nu;
  std::vector<std::vector<std::pair<int, int>>> adj; /* (edge, other vertex) */
};

/* graph.hpp:169-214 */
static void build_graph(Graph& g, int E, const int* edges, const real* nu) {
  g.E = E;
  g.hi.resize(E); g.lo.resize(E); g.nu.assign(nu, nu + E);
  for (int j = 0; j < E; j++) {
    int k = edges[2 * j], l = edges[2 * j + 1];
    real c = nu[j];
    g.hi[j] = std::max(k, l); g.lo[j] = std::min(k, l);
    if (c <= 1e-12) throw std::invalid_argument("graph: edge weights must be larger than 0");
    if (k < 0 || l < 0) throw std::invalid_argument("graph: vertex numbers must start from 0");
    if (k == l) throw std::invalid_argument("graph: has a tadpole (selfloop) edge. This program cannot take care of this. ");
    size_t Vp = (size_t)std::max(k, l) + 1;
    if (Vp > g.adj.size()) g.adj.resize(Vp);
    g.adj[k].push_back({j, l});
    g.adj[l].push_back({j, k});
  }
  g.V = (int)g.adj.size();
  for (int i = 0; i < g.V; i++)
    if (g.adj[i].empty()) {
      std::stringstream s; s << "graph: the graph has a singleton vertex " << i;
      throw std::invalid_argument(s.str());
    }
}

/* components.hpp:92-120: component labels in order of the lowest vertex */
static int components(const Graph& g, uint64_t sub, std::vector<int>& cmap) {
  if (g.V > 32) throw std::runtime_error("Error: stacksize for BFS stack too small. Compile with larger MAX_BFS_STACKSIZE.");
  cmap.assign(g.V, -1);
  int C = 0;
  std::vector<int> stack;
  for (int i = 0; i < g.V; i++) {
    if (cmap[i] >= 0) continue;
    stack.clear(); stack.push_back(i); cmap[i] = C;
    while (!stack.empty()) {
      int v = stack.back(); stack.pop_back();
      for (auto& a : g.adj[v]) {
        if (!((sub >> a.first) & 1)) continue;
        if (cmap[a.second] < 0) { cmap[a.second] = C; stack.push_back(a.second); }
      }
    }
    C++;
  }
  return C;
}

/* ------------------------------------------------------- small matrices -- */
template <class T> struct Mat {
  int r = 0, c = 0;
  std::vector<T> a;
  void resize(int rr, int cc) { r = rr; c = cc; a.assign((size_t)rr * cc, T(0)); }
  T& operator()(int i, int j) { return a[(size_t)i * c + j]; }
  const T& operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

/* Eigen LDLT<MatrixXd> (Cholesky/LDLT.h:299-389): lower storage, diagonal
 * pivoting on the largest |diagonal| (first index on ties). */
struct LDLT {
  Mat<real> m;
  std::vector<int> tr;
  int n = 0;
  void compute(const Mat<real>& A) {
    n = A.r; m = A; tr.assign(n, 0);
    std::vector<real> temp(n);
    if (n <= 1) { for (int k = 0; k < n; k++) tr[k] = k; return; }
    for (int k = 0; k < n; k++) {
      int big = k; real bv = std::fabs(m(k, k));
      for (int i = k + 1; i < n; i++) if (std::fabs(m(i, i)) > bv) { bv = std::fabs(m(i, i)); big = i; }
      tr[k] = big;
      if (k != big) {
        for (int j = 0; j < k; j++) std::swap(m(k, j), m(big, j));
        for (int i = big + 1; i < n; i++) std::swap(m(i, k), m(i, big));
        std::swap(m(k, k), m(big, big));
        for (int i = k + 1; i < big; i++) { real t = m(i, k); m(i, k) = m(big, i); m(big, i) = t; }
      }
      int rs = n - k - 1;
      if (k > 0) {
        for (int j = 0; j < k; j++) temp[j] = m(j, j) * m(k, j);
        real s = 0; for (int j = 0; j < k; j++) s += m(k, j) * temp[j];
        m(k, k) -= s;
        for (int i = k + 1; i < n; i++) {
          real t = 0; for (int j = 0; j < k; j++) t += m(i, j) * temp[j];
          m(i, k) -= t;
        }
      }
      real akk = m(k, k);
      bool valid = std::fabs(akk) > 0.0;
      if (k == 0 && !valid) { for (int j = 0; j < n; j++) tr[j] = j; return; }
      if (rs > 0 && valid) for (int i = k + 1; i < n; i++) m(i, k) /= akk;
    }
  }
  real prodD() const { real p = 1; for (int i = 0; i < n; i++) p *= m(i, i); return p; }
  /* LDLT.h:574-610 */
  void solve(Mat<real>& B) const {
    int nc = B.c;
    for (int k = 0; k < n; k++) if (tr[k] != k) for (int j = 0; j < nc; j++) std::swap(B(k, j), B(tr[k], j));
    for (int j = 0; j < nc; j++)
      for (int i = 0; i < n; i++) { real s = B(i, j); for (int q = 0; q < i; q++) s -= m(i, q) * B(q, j); B(i, j) = s; }
    const real tol = std::numeric_limits<real>::min();
    for (int i = 0; i < n; i++) {
      real d = m(i, i);
      for (int j = 0; j < nc; j++) B(i, j) = (std::fabs(d) > tol) ? B(i, j) / d : 0.0;
    }
    for (int j = 0; j < nc; j++)
      for (int i = n - 1; i >= 0; i--) { real s = B(i, j); for (int q = i + 1; q < n; q++) s -= m(q, i) * B(q, j); B(i, j) = s; }
    for (int k = n - 1; k >= 0; k--) if (tr[k] != k) for (int j = 0; j < nc; j++) std::swap(B(k, j), B(tr[k], j));
  }
};

/* Eigen FullPivLU<MatrixXcd> (LU/FullPivLU.h:488-587, solve :745-780):
 * complete pivoting on the largest |z| (column-major scan, first max). */
struct FullPivLUc {
  Mat<cd> lu;
  std::vector<int> rt, ct;
  int n = 0, nonzero = 0, ntr = 0;
  real maxpivot = 0;
  void compute(const Mat<cd>& A) {
    n = A.r; lu = A; rt.assign(n, 0); ct.assign(n, 0); ntr = 0; nonzero = n; maxpivot = 0;
    for (int k = 0; k < n; k++) {
      int br = k, bc = k; real bv = -1;
      for (int j = k; j < n; j++)
        for (int i = k; i < n; i++) {
          real v = std::abs(lu(i, j));
          if (bv < 0 || v > bv) { bv = v; br = i; bc = j; }
        }
      if (bv == 0.0) { nonzero = k; for (int i = k; i < n; i++) { rt[i] = i; ct[i] = i; } break; }
      if (bv > maxpivot) maxpivot = bv;
      rt[k] = br; ct[k] = bc;
      if (k != br) { for (int j = 0; j < n; j++) std::swap(lu(k, j), lu(br, j)); ntr++; }
      if (k != bc) { for (int i = 0; i < n; i++) std::swap(lu(i, k), lu(i, bc)); ntr++; }
      if (k < n - 1) for (int i = k + 1; i < n; i++) lu(i, k) /= lu(k, k);
      if (k < n - 1)
        for (int i = k + 1; i < n; i++)
          for (int j = k + 1; j < n; j++) lu(i, j) -= lu(i, k) * lu(k, j);
    }
  }
  cd det() const {
    cd p(1.0, 0.0);
    for (int i = 0; i < n; i++) p *= lu(i, i);
    return (ntr % 2 ? real(-1) : real(1)) * p;
  }
  int rank() const {
    real thr = std::fabs(maxpivot) * (std::numeric_limits<real>::epsilon() * n);
    int r = 0;
    for (int i = 0; i < nonzero; i++) r += (std::abs(lu(i, i)) > thr);
    return r;
  }
  /* returns X with A X = B */
  Mat<cd> solve(const Mat<cd>& B) const {
    int nc = B.c, rk = rank();
    Mat<cd> X; X.resize(n, nc);
    if (rk == 0) return X;
    /* P from the row transpositions: applied in reverse order (FullPivLU.h:566-568) */
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    /* m_p.applyTranspositionOnTheRight(k, rt[k]) for k = n-1..0; permutationP()*rhs
       gives c.row(i) = rhs.row(pinv(i)).  Build by applying the row swaps to rhs
       directly in forward order, which is the same permutation. */
    Mat<cd> c = B;
    for (int k = 0; k < n; k++) if (rt[k] != k) for (int j = 0; j < nc; j++) std::swap(c(k, j), c(rt[k], j));
    for (int j = 0; j < nc; j++)
      for (int i = 0; i < n; i++) { cd s = c(i, j); for (int q = 0; q < i; q++) s -= lu(i, q) * c(q, j); c(i, j) = s; }
    for (int j = 0; j < nc; j++)
      for (int i = rk - 1; i >= 0; i--) {
        cd s = c(i, j);
        for (int q = i + 1; q < rk; q++) s -= lu(i, q) * c(q, j);
        c(i, j) = s / lu(i, i);
      }
    /* Q: column transpositions applied in forward order build q; dst.row(q(i)) = c.row(i) */
    std::vector<int> q(n);
    for (int i = 0; i < n; i++) q[i] = i;
    for (int k = 0; k < n; k++) std::swap(q[k], q[ct[k]]);
    for (int i = 0; i < rk; i++) for (int j = 0; j < nc; j++) X(q[i], j) = c(i, j);
    for (int i = rk; i < n; i++) for (int j = 0; j < nc; j++) X(q[i], j) = cd(0, 0);
    (void)perm;
    return X;
  }
};

/* Eigen PartialPivLU<MatrixXcd> unblocked (LU/PartialPivLU.h:358-423),
 * determinant :552-556.  (Eigen switches to a blocked variant above 16
 * columns; the pivots are the same row choices, rounding may differ.) */
static cd partial_piv_det(Mat<cd> lu) {
  int n = lu.r, ntr = 0;
  for (int k = 0; k < n; k++) {
    int br = k; real bv = std::abs(lu(k, k));
    for (int i = k + 1; i < n; i++) { real v = std::abs(lu(i, k)); if (v > bv) { bv = v; br = i; } }
    if (bv != 0.0) {
      if (k != br) { for (int j = 0; j < n; j++) std::swap(lu(k, j), lu(br, j)); ntr++; }
      for (int i = k + 1; i < n; i++) lu(i, k) /= lu(k, k);
    }
    for (int i = k + 1; i < n; i++)
      for (int j = k + 1; j < n; j++) lu(i, j) -= lu(i, k) * lu(k, j);
  }
  cd p(1.0, 0.0);
  for (int i = 0; i < n; i++) p *= lu(i, i);
  return (ntr % 2 ? real(-1) : real(1)) * p;
}

/* libstdc++ pow(complex<real>, real) (complex:1028-1039) */
static cd cpow_real(const cd& z, real y) {
  if (z.imag() == 0.0 && z.real() > 0.0) return cd(std::pow(z.real(), y), 0.0);
  cd t = std::log(z);
  return std::polar(std::exp(y * t.real()), y * t.imag());
}

/* laplacian.hpp:14-67: lower triangle of the reduced, contracted Laplacian */
template <class F>
static void reduced_laplacian(Mat<F>& La, const Graph& g, uint64_t contracted, int nC,
                              const std::vector<int>& cmap, const F* Xinv) {
  La.resize(nC - 1, nC - 1);
  for (int j = 0; j < g.E; j++) {
    if ((contracted >> j) & 1) continue;
    int k = cmap[g.hi[j]], l = cmap[g.lo[j]];
    if (k == l) continue;
    if (k < l) std::swap(k, l);
    F x = Xinv[j];
    La(l, l) += x;
    if (k == nC - 1) continue;
    La(k, k) += x;
    La(k, l) += -x;
  }
}

/* laplacian.hpp:70-99 */
static void pgp_matrix(Mat<real>& PGP, const Graph& g, const std::vector<real>& P, int nC,
                       const std::vector<int>& cmap) {
  PGP.resize(nC - 1, nC - 1);
  for (int i = 0; i < g.V; i++)
    for (int j = 0; j < g.V; j++) {
      int v = cmap[i], w = cmap[j];
      if (v == nC - 1 || w == nC - 1) continue;
      PGP(v, w) += P[(size_t)i * g.V + j];
    }
}

static real frob(const Mat<real>& A) { real s = 0; for (real v : A.a) s += v * v; return std::sqrt(s); }

/* ------------------------------------------------------------- problem -- */
struct or_problem {
  Graph g;
  std::vector<real> P, m2;
  real D = 0, lambda = 0;
  int K = 1;
  int L = 0;
  real W = 0, W_ref = 0, IGtr = 0;
  bool bEuclidean = true, bPseudoEuclidean = true, bGeneric = true, bGP = true;
  real lambda_eff = 0, def_ref = 0;
  bool W_zero = false, complex_path = false;
  std::vector<real> Jr, r;
  std::vector<uint8_t> mm, Lt;
};

/* tropical_sampling.hpp:26-44 */
static real omega(const Graph& g, real D, int L, uint64_t sub) {
  real m = 0.0;
  for (int j = 0; j < g.E; j++) if ((sub >> j) & 1) m += g.nu[j];
  return m - D / 2.0 * L;
}

/* Cyclic Jacobi eigen-decomposition of a symmetric matrix (row-major n x n).
 * Returns eigenvalues in w and eigenvectors as columns of Q. */
static void jacobi_eig(int n, std::vector<real> A, std::vector<real>& w, std::vector<real>& Q) {
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) Q[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    real off = 0;
    for (int i = 0; i < n; i++) for (int j = i + 1; j < n; j++) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off == 0.0) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        real apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        real app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        real theta = (aqq - app) / (2 * apq);
        real t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        real c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          real akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          real apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          real qkp = Q[(size_t)k * n + p], qkq = Q[(size_t)k * n + q];
          Q[(size_t)k * n + p] = c * qkp - s * qkq;
          Q[(size_t)k * n + q] = s * qkp + c * qkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; i++) w[i] = A[(size_t)i * n + i];
}

static bool is_approx(const std::vector<real>& a, const std::vector<real>& b, real prec = 1e-12) {
  real d = 0, na = 0, nb = 0;
  for (size_t i = 0; i < a.size(); i++) { d += (a[i] - b[i]) * (a[i] - b[i]); na += a[i] * a[i]; nb += b[i] * b[i]; }
  return d <= prec * prec * std::min(na, nb);
}

static void preprocess(or_problem* pr) {
  Graph& g = pr->g;
  const real D = pr->D;
  const int V = g.V, E = g.E;
  const std::vector<real>& P = pr->P;
  /* feyntrop.cpp:98-112 */
  {
    std::vector<real> Pt((size_t)V * V);
    for (int i = 0; i < V; i++) for (int j = 0; j < V; j++) Pt[(size_t)j * V + i] = P[(size_t)i * V + j];
    if (!is_approx(P, Pt)) throw std::domain_error("integrate_graph: The matrix of invariants was given as input and it is not symmetric or not a VxV matrix.");
    for (int j = 0; j < V; j++) {
      real s = 0; for (int i = 0; i < V; i++) s += P[(size_t)i * V + j];
      if (!(std::fabs(s) <= 1e-12)) throw std::domain_error("Scalar product (P) matrix has to fulfill momentum conservation.");
    }
  }
  /* feyntrop.cpp:128-133 */
  {
    std::vector<real> w, Q;
    jacobi_eig(V, P, w, Q);
    std::vector<real> Pabs((size_t)V * V, 0.0);
    for (int i = 0; i < V; i++)
      for (int j = 0; j < V; j++) {
        real s = 0; for (int k = 0; k < V; k++) s += Q[(size_t)i * V + k] * std::fabs(w[k]) * Q[(size_t)j * V + k];
        Pabs[(size_t)i * V + j] = -s;
      }
    bool zero = true;
    for (real v : P) if (!(std::fabs(v) <= 1e-12)) zero = false;
    pr->bEuclidean = true;
    if (!zero && !is_approx(P, Pabs)) pr->bEuclidean = false;
  }
  /* tropical_sampling.hpp:284-458 */
  if (E > 63) { std::stringstream s; s << "E = " << E << " this graph is too large for this implementation!"; throw std::domain_error(s.str()); }
  const bool eps_expansion = pr->K > 1, bNonEuclidean = !pr->bEuclidean;
  std::vector<int> cmap;
  const uint64_t full = (E == 64) ? ~0ULL : ((UINT64_C(1) << E) - 1);
  int C = components(g, full, cmap);
  int L = E - V + C;
  if (C != 1) throw std::domain_error("Error: the graph is not connected.");
  real W = omega(g, D, L, full), W_ref = omega(g, -D, L, full);
  pr->L = L; pr->W = W; pr->W_ref = W_ref;
  Mat<real> PGP;
  int cV = components(g, 0, cmap);
  pgp_matrix(PGP, g, P, cV, cmap);
  real m2n = 0; for (real v : pr->m2) m2n += v * v;
  real pmsqr_ref = frob(PGP) + std::sqrt(m2n);
  bool bPhiTr = eps_expansion || std::fabs(W / W_ref) > 1e-12;
  if (bPhiTr && std::fabs(pmsqr_ref) < 1e-12) throw std::domain_error("kinematics are necessary: the incoming momenta/masses are too small.");
  const int n_tests = 100;
  std::vector<std::vector<real>> Xs(n_tests), Xinvs(n_tests);
  if (bPhiTr) {
    Xo rg; xo_seed(0x5eedULL, rg.s);
    for (int i = 0; i < n_tests; i++) {
      Xs[i].resize(E); Xinvs[i].resize(E);
      for (int j = 0; j < E; j++) { Xs[i][j] = 1.0 + 2.0 * rg.uniform(); Xinvs[i][j] = 1.0 / Xs[i][j]; }
    }
  }
  size_t NT = (size_t)1 << E;
  pr->Jr.assign(NT, 0); pr->r.assign(NT, 0); pr->mm.assign(NT, 0); pr->Lt.assign(NT, 0);
  bool bGeneric = true, bPseudo = true, bGP = true;
  Mat<real> La;
  LDLT ldlt;
  for (int n = 0; n <= E; n++) {
    uint64_t sub = (n == 0) ? 0 : ((n == 64) ? ~0ULL : ((UINT64_C(1) << n) - 1));
    while (true) {
      int cVs = components(g, sub, cmap);
      bool mm = false;
      if (bPhiTr) {
        pgp_matrix(PGP, g, P, cVs, cmap);
        if (cVs == 2 && PGP(0, 0) / pmsqr_ref >= 1e-12) bPseudo = false;
        if (cVs == 3) {
          /* tropical_sampling.hpp:84-148 */
          bool zero = true;
          for (real v : PGP.a) if (!(std::fabs(v) <= 1e-12)) zero = false;
          bool gen = true;
          if (!zero) {
            real a = PGP(0, 0), b = PGP(1, 1), gm = PGP(0, 1), c = a + b + 2 * gm;
            real A = 0, B = 0, Cc = 0;
            for (int j = 0; j < E; j++) {
              if ((sub >> j) & 1) continue;
              int kc = cmap[g.hi[j]], lc = cmap[g.lo[j]];
              if (kc == lc) continue;
              if (kc == 0 || lc == 0) A += pr->m2[j];
              if (kc == 1 || lc == 1) B += pr->m2[j];
              if (kc == 2 || lc == 2) Cc += pr->m2[j];
            }
            const real th = 1e-5;
            if (std::fabs(a - A) / pmsqr_ref > 1e-10 && std::fabs(b - B) / pmsqr_ref < th && std::fabs(c - Cc) / pmsqr_ref < th) gen = false;
            if (std::fabs(b - B) / pmsqr_ref > 1e-10 && std::fabs(a - A) / pmsqr_ref < th && std::fabs(c - Cc) / pmsqr_ref < th) gen = false;
            if (std::fabs(c - Cc) / pmsqr_ref > 1e-10 && std::fabs(a - A) / pmsqr_ref < th && std::fabs(b - B) / pmsqr_ref < th) gen = false;
          }
          bGeneric = bGeneric && gen;
        }
        /* tropical_sampling.hpp:46-82 */
        real sum = 0;
        for (int i = 0; i < n_tests; i++) {
          reduced_laplacian<real>(La, g, sub, cVs, cmap, Xinvs[i].data());
          ldlt.compute(La);
          Mat<real> S = PGP;
          ldlt.solve(S);
          real tr = 0; for (int q = 0; q < S.r; q++) tr += S(q, q);
          real pxi = -tr;
          for (int j = 0; j < E; j++) if (!((sub >> j) & 1)) pxi += pr->m2[j] * Xs[i][j];
          sum += std::fabs(pxi);
        }
        mm = sum / pmsqr_ref < 1e-12;
      } else if (n == E) {
        mm = true;
      }
      int Ls = n - V + cVs;
      real Jr, r;
      if (n == 0) { Jr = 1.0; r = 1.0; }
      else {
        Jr = 0.0;
        for (int j = 0; j < E; j++) if ((sub >> j) & 1) { uint64_t c2 = sub & ~(UINT64_C(1) << j); Jr += pr->Jr[c2] / pr->r[c2]; }
        r = omega(g, D, Ls, sub);
        if (mm) r -= W;
      }
      if (r / W_ref < 1e-12 && !(n == E)) {
        std::stringstream s;
        s << "Error: the graph has a UV or IR subdivergence. E.g. the subgraph with edge numbers { ";
        for (int j = 0; j < E; j++) if ((sub >> j) & 1) s << j << " ";
        s << "} has an (effective) superficial degree of divergence of " << r << ".";
        throw std::domain_error(s.str());
      }
      if (bPhiTr && bNonEuclidean && n >= 2) {
        /* tropical_sampling.hpp:152-228 */
        int z = mm ? Ls + 1 : Ls;
        bool ok = true;
        for (int a = 0; a < E && ok; a++) {
          if (!((sub >> a) & 1)) continue;
          uint64_t A = sub & ~(UINT64_C(1) << a);
          int zA = pr->mm[A] ? pr->Lt[A] + 1 : pr->Lt[A];
          for (int b = a + 1; b < E; b++) {
            if (!((sub >> b) & 1)) continue;
            uint64_t AB = A & ~(UINT64_C(1) << b), B = sub & ~(UINT64_C(1) << b);
            int zB = pr->mm[B] ? pr->Lt[B] + 1 : pr->Lt[B];
            int zAB = pr->mm[AB] ? pr->Lt[AB] + 1 : pr->Lt[AB];
            if (zA + zB > z + zAB) { ok = false; break; }
          }
        }
        bGP = bGP && ok;
      }
      pr->Jr[sub] = Jr; pr->r[sub] = r; pr->mm[sub] = mm; pr->Lt[sub] = (uint8_t)Ls;
      if (n == 0 || n == E) break;
      /* Gosper's next subset of the same popcount (graph.hpp:56-81) */
      uint64_t t = sub | (sub - 1);
      uint64_t nx = (t + 1) | (((~t & -~t) - 1) >> (__builtin_ctzll(sub) + 1));
      if ((nx >> E) != 0) break;
      sub = nx;
    }
  }
  pr->IGtr = pr->Jr[full];
  pr->bGeneric = bGeneric; pr->bPseudoEuclidean = bPseudo; pr->bGP = bGP;
  /* feyntrop.cpp:182-231: fatal combinations */
  if (bGeneric && !bGP) throw std::domain_error("The momenta seem generic, but the generalized permutahedron property is not fulfilled.");
  if (pr->bEuclidean && !bGP) throw std::domain_error("The momenta seem Euclidean, but the generalized permutahedron property is not fulfilled.");
  if (!(!bGeneric && !pr->bEuclidean && W > 0) && !bGP && W > 0)
    throw std::domain_error("The momenta seem Euclidean or generic, but the generalized permutahedron property seems not fulfilled.");
  /* feyntrop.cpp:233-239, feynman_integral.hpp:125-140 */
  pr->lambda_eff = pr->bPseudoEuclidean ? 0.0 : pr->lambda;
  pr->W_zero = std::fabs(W / W_ref) < 1e-12;
  pr->def_ref = 0.0;
  if (pr->lambda_eff != 0 && !(pr->W_zero && pr->K == 1)) pr->def_ref = -pr->lambda_eff;
  pr->complex_path = !(pr->def_ref == 0.0 || (pr->W_zero && pr->K == 1));
}

/* ------------------------------------------------------------- sampler -- */
/* tropical_sampling.hpp:468-564; returns (psi_tr, xi_tr) */
static void sample_point(const or_problem* pr, Xo& gen, real* X, real* psixi) {
  const int E = pr->g.E;
  uint64_t sub = (UINT64_C(1) << E) - 1;
  real JJr = pr->Jr[sub];
  bool MM = pr->mm[sub];
  int oL = pr->Lt[sub];
  real G = 0.0, F = 1.0, XX = 1.0;
  for (int j = 0;; j++) {
    real R = gen.uniform() * JJr;
    real T = 0.0;
    int e = -1;
    for (int q = 0; q < E; q++) {
      if (!((sub >> q) & 1)) continue;
      uint64_t c2 = sub & ~(UINT64_C(1) << q);
      T += pr->Jr[c2] / pr->r[c2];
      e = q;
      if (T > R) break;
    }
    uint64_t c2 = sub & ~(UINT64_C(1) << e);
    JJr = pr->Jr[c2];
    real r = pr->r[c2];
    bool mm = pr->mm[c2];
    int L = pr->Lt[c2];
    sub = c2;
    if (MM && !mm) G = XX;
    MM = mm;
    X[e] = XX;
    if (L < oL) { F *= XX; oL = L; }
    if (j >= E - 1) break;
    real U = gen.uniform();
    real Y = std::pow(1. - U, 1. / r);
    XX *= Y;
  }
  psixi[0] = F;
  psixi[1] = G * F;
}

/* ----------------------------------------------------------- integrand -- */
/* feynman_integral.hpp:166-362 for one sample; out: K*2 values.
 * Returns false when the sample is dropped (non-finite R0). */
static bool integrand(const or_problem* pr, const real* Xin, const real* psixi, real* out, int* vstar_out) {
  const Graph& g = pr->g;
  const int E = g.E, V = g.V, K = pr->K;
  const real D = pr->D, W = pr->W;
  const int L = pr->L;
  std::vector<real> X(Xin, Xin + E), XV(V, 0.0);
  for (int j = 0; j < E; j++) { XV[g.hi[j]] += X[j]; XV[g.lo[j]] += X[j]; }
  int minV = 0;
  for (int i = 1; i < V; i++) if (XV[minV] > XV[i]) minV = i;
  if (vstar_out) *vstar_out = minV;
  std::vector<int> cmap(V);
  for (int i = 0; i < V; i++) cmap[i] = i < minV ? i : (i > minV ? i - 1 : V - 1);
  real psi_tr = psixi[0], xi_tr = psixi[1];
  real tgt = std::pow(psi_tr, -D / 2.) * std::pow(psi_tr / xi_tr, W);
  real scale = std::pow(tgt, 1. / (D / 2. * L + W));
  for (int j = 0; j < E; j++) X[j] *= scale;
  std::vector<real> Xinv(E);
  for (int j = 0; j < E; j++) Xinv[j] = 1.0 / X[j];
  Mat<real> La;
  reduced_laplacian<real>(La, g, 0, V, cmap, Xinv.data());
  LDLT ldlt; ldlt.compute(La);
  for (int q = 0; q < 2 * K; q++) out[q] = 0.0;
  if (!pr->complex_path) {
    real detD = ldlt.prodD();
    real xp = 1; for (int j = 0; j < E; j++) xp *= X[j];
    real psi = detD * xp;
    real R = pr->IGtr * std::pow(psi, -D / 2.);
    if (!std::isfinite(R)) return false;
    if (pr->W_zero && K == 1) { out[0] = R; return true; }
    Mat<real> PGP; pgp_matrix(PGP, g, pr->P, V, cmap);
    Mat<real> S = PGP; ldlt.solve(S);
    real tr = 0; for (int q = 0; q < S.r; q++) tr += S(q, q);
    real pphi = -tr;
    real dot = 0; for (int j = 0; j < E; j++) dot += pr->m2[j] * X[j];
    real pxi = pphi + dot;
    R *= std::pow(pxi, -W);
    if (!std::isfinite(R)) return false;
    out[0] = R;
    if (K > 1) {
      real ef = std::log(psi) - L * std::log(pxi);
      for (int k = 1; k < K; k++) { R *= ef / (real)k; out[2 * k] = R; }
    }
    return true;
  }
  /* analytic continuation (feynman_integral.hpp:278-362) */
  const int n = V - 1;
  Mat<real> PGP; pgp_matrix(PGP, g, pr->P, V, cmap);
  Mat<real> LaInv; LaInv.resize(n, n);
  for (int i = 0; i < n; i++) LaInv(i, i) = 1.0;
  ldlt.solve(LaInv);
  Mat<real> T1; T1.resize(n, n);
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) { real s = 0; for (int k = 0; k < n; k++) s += LaInv(i, k) * PGP(k, j); T1(i, j) = s; }
  Mat<real> LP; LP.resize(n, n);
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) { real s = 0; for (int k = 0; k < n; k++) s += T1(i, k) * LaInv(j, k); LP(i, j) = s; }
  std::vector<real> Xs2(E), dp(E);
  for (int j = 0; j < E; j++) Xs2[j] = Xinv[j] * Xinv[j];
  Mat<real> dd; dd.resize(E, E);
  /* laplacian.hpp:102-215 */
  for (int a = 0; a < E; a++) {
    int ka = cmap[g.hi[a]], la = cmap[g.lo[a]];
    if (la > ka) std::swap(la, ka);
    real s, t;
    if (ka == V - 1) { s = LP(la, la); t = LaInv(la, la); }
    else { s = LP(ka, ka) + LP(la, la) - 2. * LP(ka, la); t = LaInv(ka, ka) + LaInv(la, la) - 2. * LaInv(ka, la); }
    real xa = Xs2[a], xas = xa * s;
    dp[a] = -xas;
    dd(a, a) = 2. * xas * (Xinv[a] - xa * t);
    for (int b = a + 1; b < E; b++) {
      int kb = cmap[g.hi[b]], lb = cmap[g.lo[b]];
      if (lb > kb) std::swap(lb, kb);
      real s2, t2;
      if (kb == V - 1) {
        if (ka == V - 1) { s2 = LP(la, lb); t2 = LaInv(la, lb); }
        else { s2 = LP(la, lb) - LP(ka, lb); t2 = LaInv(la, lb) - LaInv(ka, lb); }
      } else {
        if (ka == V - 1) { s2 = LP(la, lb) - LP(la, kb); t2 = LaInv(la, lb) - LaInv(la, kb); }
        else {
          s2 = LP(ka, kb) + LP(la, lb) - LP(ka, lb) - LP(la, kb);
          t2 = LaInv(ka, kb) + LaInv(la, lb) - LaInv(ka, lb) - LaInv(la, kb);
        }
      }
      real xb = Xs2[b];
      dd(a, b) = dd(b, a) = -2. * xa * xb * s2 * t2;
    }
  }
  for (int j = 0; j < E; j++) { dp[j] += pr->m2[j]; dp[j] *= pr->def_ref; }
  std::vector<cd> cX(E), cXinv(E);
  for (int j = 0; j < E; j++) cX[j] = X[j] * cd(std::cos(dp[j]), std::sin(dp[j]));
  for (int j = 0; j < E; j++) cXinv[j] = real(1) / cX[j];
  Mat<cd> cLa;
  reduced_laplacian<cd>(cLa, g, 0, V, cmap, cXinv.data());
  for (int k = 0; k < n; k++) for (int l = k + 1; l < n; l++) cLa(k, l) = cLa(l, k);
  FullPivLUc lu; lu.compute(cLa);
  cd xp(1.0, 0.0); for (int j = 0; j < E; j++) xp *= cX[j];
  cd cPsi = xp * lu.det();
  cd cR = pr->IGtr * cpow_real(cPsi, -D / 2.);
  Mat<cd> J; J.resize(E, E);
  const cd ci(0.0, pr->def_ref);
  for (int a = 0; a < E; a++) for (int b = 0; b < E; b++) J(a, b) = (ci * dd(a, b)) * X[b] + (a == b ? real(1) : real(0));
  cd detJ = partial_piv_det(J);
  real xne = 0; for (int j = 0; j < E; j++) xne += g.nu[j] * dp[j];
  cR *= detJ * cd(std::cos(xne), std::sin(xne));
  Mat<cd> cPGP; cPGP.resize(n, n);
  for (size_t q = 0; q < PGP.a.size(); q++) cPGP.a[q] = PGP.a[q];
  Mat<cd> Sx = lu.solve(cPGP);
  cd tr(0.0, 0.0); for (int q = 0; q < n; q++) tr += Sx(q, q);
  cd cPPhi = -tr;
  cd dot(0.0, 0.0); for (int j = 0; j < E; j++) dot += pr->m2[j] * cX[j];
  cd cPXi = cPPhi + dot;
  cR *= cpow_real(cPXi, -W);
  if (!std::isfinite(cR.real()) || !std::isfinite(cR.imag())) return false;
  out[0] = cR.real(); out[1] = cR.imag();
  if (K > 1) {
    cd ef = std::log(cPsi) - (real)L * std::log(cPXi);
    for (int k = 1; k < K; k++) { cR *= ef / (real)k; out[2 * k] = cR.real(); out[2 * k + 1] = cR.imag(); }
  }
  return true;
}

/* --------------------------------------------------------------- stats -- */
/* stats.hpp:13-57 (64-limb binary-carry summation) and :59-134 */
struct Summer {
  real s[64];
  Summer() { for (real& v : s) v = 0.0; }
  void add(real v) {
    if (v == 0.0) return;
    for (real& x : s) {
      if (x == 0.0) { x = v; return; }
      v += x; x = 0.0;
    }
    throw std::runtime_error("ERROR: More limbs needed for fp_summer");
  }
  real value() const { real v = 0.0; for (real x : s) v += x; return v; }
};
static Summer merge_summers(const Summer& a, const Summer& b) {
  Summer m;
  for (int i = 0; i < 63; i++) m.s[i + 1] = a.s[i] + b.s[i];
  m.s[0] = 0.0;
  return m;
}
struct Stats {
  uint64_t N = 0; Summer S1, S2;
  void update(real v) { N++; S1.add(v); S2.add(v * v); }
};
static Stats merge_stats(const Stats& a, const Stats& b) {
  Stats m; m.N = a.N + b.N; m.S1 = merge_summers(a.S1, b.S1); m.S2 = merge_summers(a.S2, b.S2); return m;
}
static Stats merge_vec(std::vector<Stats> v) {
  for (size_t e = 0; (1ULL << e) < v.size(); e++)
    for (size_t i = 0; i + (1ULL << e) < v.size(); i += 1ULL << (e + 1)) v[i] = merge_stats(v[i], v[i + (1ULL << e)]);
  return v[0];
}

/* ------------------------------------------------------------- C API -- */
extern "C" {

or_problem* or_create(int E, const int* edges, const double* nu, int V_in, const double* P,
                      const double* m2, double D, int K, double lambda, char* err, int errlen) {
  or_problem* pr = new or_problem();
  try {
    std::vector<real> nur(nu, nu + E);
    build_graph(pr->g, E, edges, nur.data());
    if (V_in != pr->g.V) throw std::invalid_argument("input scalarproducts must be a (V x V)-matrix");
    pr->P.assign(P, P + (size_t)V_in * V_in);
    pr->m2.assign(m2, m2 + E);
    pr->D = D; pr->K = K; pr->lambda = lambda;
    preprocess(pr);
  } catch (std::exception& ex) {
    if (err && errlen > 0) { strncpy(err, ex.what(), errlen - 1); err[errlen - 1] = 0; }
    delete pr;
    return nullptr;
  }
  return pr;
}

void or_destroy(or_problem* p) { delete p; }

void or_info(const or_problem* p, double* o) {
  o[0] = p->g.E; o[1] = p->g.V; o[2] = p->L; o[3] = (double)p->W; o[4] = (double)p->IGtr;
  o[5] = p->bEuclidean; o[6] = p->bPseudoEuclidean; o[7] = p->bGeneric; o[8] = p->bGP;
  o[9] = (double)p->lambda_eff; o[10] = (double)p->def_ref; o[11] = p->complex_path; o[12] = p->K;
}

void or_jtable(const or_problem* p, double* Jr, double* r, uint8_t* mm, uint8_t* L) {
  size_t NT = p->Jr.size();
  for (size_t i = 0; i < NT; i++) { Jr[i] = (double)p->Jr[i]; r[i] = (double)p->r[i]; }
  memcpy(mm, p->mm.data(), NT); memcpy(L, p->Lt.data(), NT);
}

void or_sample(const or_problem* p, uint64_t* st, uint64_t n, double* X, double* psixi) {
  Xo g; memcpy(g.s, st, 32);
  const int E = p->g.E;
  std::vector<real> x(E);
  real px[2];
  for (uint64_t i = 0; i < n; i++) {
    sample_point(p, g, x.data(), px);
    for (int e = 0; e < E; e++) X[i * E + e] = (double)x[e];
    psixi[2 * i] = (double)px[0]; psixi[2 * i + 1] = (double)px[1];
  }
  memcpy(st, g.s, 32);
}

void or_replay(const or_problem* p, uint64_t n, const double* X, const double* psixi, double* out,
               uint8_t* dropped, int32_t* vstar) {
  const int E = p->g.E, K = p->K;
  std::vector<real> x(E), o(2 * K);
  for (uint64_t i = 0; i < n; i++) {
    int vs = 0;
    for (int e = 0; e < E; e++) x[e] = X[i * E + e];
    real px[2] = {psixi[2 * i], psixi[2 * i + 1]};
    bool ok = integrand(p, x.data(), px, o.data(), &vs);
    for (int q = 0; q < 2 * K; q++) out[i * 2 * K + q] = ok ? (double)o[q] : 0.0;
    if (dropped) dropped[i] = ok ? 0 : 1;
    if (vstar) vstar[i] = vs;
  }
}

uint64_t or_estimate(const or_problem* p, uint64_t N, uint64_t seed, int chunks, int threads, double* moments) {
  const int E = p->g.E, K = p->K;
  if (chunks < 1) chunks = 1;
  std::vector<std::vector<Stats>> re(K, std::vector<Stats>(chunks)), im(K, std::vector<Stats>(chunks));
  if (!p->complex_path || p->def_ref == 0.0)
    for (int c = 0; c < chunks; c++) for (int k = 0; k < K; k++) { im[k][c].update(0); im[k][c].update(0); }
  uint64_t st0[4]; xo_seed(seed, st0);
  std::vector<uint64_t> drops(chunks, 0);
  const uint64_t per = 2 * (uint64_t)E - 1;
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int c = 0; c < chunks; c++) {
    uint64_t a = (uint64_t)((unsigned __int128)N * c / chunks), b = (uint64_t)((unsigned __int128)N * (c + 1) / chunks);
    Xo g; memcpy(g.s, st0, 32);
    uint64_t lo, hi; mul_u64(a, per, &lo, &hi);
    if (lo | hi) poly_apply(poly_xpow(lo, hi), g.s);
    std::vector<real> X(E), out(2 * K);
    real px[2];
    for (uint64_t i = a; i < b; i++) {
      sample_point(p, g, X.data(), px);
      if (!integrand(p, X.data(), px, out.data(), nullptr)) { drops[c]++; continue; }
      if (!p->complex_path) {
        re[0][c].update(out[0]);
        for (int k = 1; k < K; k++) re[k][c].update(out[2 * k]);
      } else {
        for (int k = 0; k < K; k++) { re[k][c].update(out[2 * k]); im[k][c].update(out[2 * k + 1]); }
      }
    }
  }
  uint64_t dropped = 0;
  for (uint64_t d : drops) dropped += d;
  for (int k = 0; k < K; k++) {
    Stats a = merge_vec(re[k]), b = merge_vec(im[k]);
    double* m = moments + 6 * k;
    m[0] = (double)a.N; m[1] = (double)a.S1.value(); m[2] = (double)a.S2.value();
    m[3] = (double)b.N; m[4] = (double)b.S1.value(); m[5] = (double)b.S2.value();
  }
  return dropped;
}

void or_finalize(const or_problem* p, const double* moments, double* res) {
  for (int k = 0; k < p->K; k++)
    for (int q = 0; q < 2; q++) {
      const double* m = moments + 6 * k + 3 * q;
      double N = m[0], avg = m[1] / N;
      double var = (N / (N - 1)) * (m[2] / N - std::pow(avg, 2));
      res[4 * k + 2 * q] = avg;
      res[4 * k + 2 * q + 1] = std::sqrt(var / N);
    }
}

void or_xoshiro_seed(uint64_t seed, uint64_t* st) { xo_seed(seed, st); }

void or_xoshiro_next(uint64_t* st, uint64_t n, uint64_t* out) {
  Xo g; memcpy(g.s, st, 32);
  for (uint64_t i = 0; i < n; i++) { uint64_t v = g.next(); if (out) out[i] = v; }
  memcpy(st, g.s, 32);
}

void or_xoshiro_advance(uint64_t* st, uint64_t lo, uint64_t hi) {
  if (!(lo | hi)) return;
  poly_apply(poly_xpow(lo, hi), st);
}

void or_xoshiro_charpoly(uint64_t* p5) {
  compute_charpoly();
  for (int i = 0; i < 4; i++) p5[i] = g_charpoly.w[i];
  p5[4] = 1;
}

void or_xoshiro_jumppoly_pow2(int e, uint64_t* p4) {
  compute_charpoly();
  Poly4 b; memset(&b, 0, sizeof b); b.w[0] = 2;
  for (int i = 0; i < e; i++) b = poly_mulmod(b, b);
  for (int i = 0; i < 4; i++) p4[i] = b.w[i];
}

void or_xoshiro_jump_pow2(int e, uint64_t* st) {
  compute_charpoly();
  Poly4 b; memset(&b, 0, sizeof b); b.w[0] = 2;
  for (int i = 0; i < e; i++) b = poly_mulmod(b, b);
  poly_apply(b, st);
}

double or_fp_sum(const double* v, uint64_t n) {
  Summer s;
  for (uint64_t i = 0; i < n; i++) s.add(v[i]);
  return (double)s.value();
}

} /* extern "C" */
