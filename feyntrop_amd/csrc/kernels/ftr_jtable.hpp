/*
 * ftr_jtable.hpp -- the subgraph J-table on the GPU (tropical_sampling.hpp:
 * 284-458), one launch per subset size n, one thread per subset.
 *
 * Subsets of size n are visited in increasing numeric order, which is the
 * reference's Gosper order, so thread t of level n works on the subset of
 * colex rank t.  Per subset:
 *   - components of (V, edges in the subset), labelled in order of their
 *     lowest vertex;
 *   - loop number;
 *   - the mass-momentum-spanning test (pxi of the contracted graph at the
 *     fixed random points, pivoted LDL^T);
 *   - J = sum over removed edges e of J(s\e)/r(s\e), and r = omega (- W if mm);
 *   - the genericity, pseudo-Euclidean and generalized-permutahedron tests,
 *     and the subdivergence check.
 * Arithmetic follows the host implementation operation by operation with FP
 * contraction off, so the tables are bitwise those of the host.
 */
#pragma once
#pragma clang fp contract(off)

#ifdef __HIPCC_RTC__
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace ftrj {

struct JtParams {
  int E, V, n_tests, phi, noneucl, pad;
  double D, W, W_ref, pmsqr;
  int hi[30], lo[30];
  double nu[30], m2[30];
  double P[24 * 24];
};

/* subset of size n with colex rank t (binom[c * 31 + i] = C(c, i)) */
__device__ inline uint64_t unrank(uint64_t t, int n, int E, const uint64_t* binom) {
  uint64_t s = 0;
  int c = E - 1;
  for (int i = n; i >= 1; i--) {
    while (binom[c * 31 + i] > t) c--;
    s |= (uint64_t)1 << c;
    t -= binom[c * 31 + i];
    c--;
  }
  return s;
}

/* tr(A^-1 B) by LDL^T with largest-diagonal pivoting; A (n x n, row-major,
 * destroyed), B symmetric.  Columns of B that are zero contribute exactly
 * zero and are skipped. */
template <int NB>
__device__ inline double ldlt_trace_solve(double* A, const double* B, int n) {
  int perm[NB];
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int k = 0; k < n; k++) {
    int big = k;
    double bv = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i * n + i]) > bv) {
        bv = fabs(A[i * n + i]);
        big = i;
      }
    if (big != k) {
      const int tp = perm[k];
      perm[k] = perm[big];
      perm[big] = tp;
      for (int j = 0; j < n; j++) {
        const double t = A[k * n + j];
        A[k * n + j] = A[big * n + j];
        A[big * n + j] = t;
      }
      for (int i = 0; i < n; i++) {
        const double t = A[i * n + k];
        A[i * n + k] = A[i * n + big];
        A[i * n + big] = t;
      }
    }
    const double d = A[k * n + k];
    if (!(fabs(d) > 0.0)) continue;
    for (int i = k + 1; i < n; i++) {
      const double l = A[i * n + k] / d;
      for (int j = k + 1; j <= i; j++) A[i * n + j] -= l * A[j * n + k];
      for (int j = k + 1; j <= i; j++) A[j * n + i] = A[i * n + j];
    }
    for (int i = k + 1; i < n; i++) A[i * n + k] /= d;
  }
  double tr = 0.0;
  double y[NB];
  const double tol = 2.2250738585072014e-308; /* smallest normal double */
  for (int c = 0; c < n; c++) {
    bool nz = false;
    for (int i = 0; i < n; i++) nz = nz || (B[i * n + c] != 0.0);
    if (!nz) continue;
    for (int i = 0; i < n; i++) y[i] = B[perm[i] * n + c];
    for (int i = 0; i < n; i++)
      for (int q = 0; q < i; q++) y[i] -= A[i * n + q] * y[q];
    for (int i = 0; i < n; i++) y[i] = (fabs(A[i * n + i]) > tol) ? y[i] / A[i * n + i] : 0.0;
    for (int i = n - 1; i >= 0; i--)
      for (int q = i + 1; q < n; q++) y[i] -= A[q * n + i] * y[q];
    for (int i = 0; i < n; i++)
      if (perm[i] == c) tr += y[i];
  }
  return tr;
}

/* flags[0..2]: generic, pseudo-Euclidean, GP property (cleared to 0 by any
 * subset that violates them); divs[n]: smallest subset of size n with a
 * vanishing effective degree of divergence (all ones if none). */
template <int VB>
__device__ inline void jt_level_body(const JtParams* __restrict__ prm, int n, uint64_t count,
                                     const uint64_t* __restrict__ binom, const double* __restrict__ Xs,
                                     const double* __restrict__ Xi, double* Jr, double* rr, uint8_t* mmv, uint8_t* Lt,
                                     uint32_t* flags, unsigned long long* divs) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const JtParams& p = *prm;
  const int E = p.E, V = p.V;
  const uint64_t s = unrank(t, n, E, binom);
  /* components: union-find, then labels in order of the lowest vertex */
  int par[VB], cmap[VB];
  for (int i = 0; i < V; i++) par[i] = i;
  for (int j = 0; j < E; j++) {
    if (!((s >> j) & 1)) continue;
    int a = p.hi[j], b = p.lo[j];
    while (par[a] != a) a = par[a];
    while (par[b] != b) b = par[b];
    if (a != b) par[a > b ? a : b] = a > b ? b : a;
  }
  int cV = 0;
  for (int i = 0; i < V; i++) {
    int a = i;
    while (par[a] != a) a = par[a];
    if (a == i) cmap[i] = cV++;
    else cmap[i] = cmap[a];
  }
  bool mm = false, lvGeneric = true, lvPseudo = true, lvGP = true;
  if (p.phi) {
    constexpr int NB = VB - 1;
    double PGP[NB * NB], La[NB * NB];
    const int m = cV - 1;
    for (int q = 0; q < m * m; q++) PGP[q] = 0.0;
    for (int i = 0; i < V; i++)
      for (int j = 0; j < V; j++) {
        const int v = cmap[i], w = cmap[j];
        if (v == cV - 1 || w == cV - 1) continue;
        PGP[v * m + w] += p.P[i * V + j];
      }
    if (cV == 2 && PGP[0] / p.pmsqr >= 1e-12) lvPseudo = false;
    if (cV == 3) {
      bool zero = true;
      for (int q = 0; q < 4; q++)
        if (!(fabs(PGP[q]) <= 1e-12)) zero = false;
      if (!zero) {
        const double a = PGP[0], b = PGP[1 * m + 1], gm = PGP[0 * m + 1], c = a + b + 2 * gm;
        double A = 0, B = 0, Cc = 0;
        for (int j = 0; j < E; j++) {
          if ((s >> j) & 1) continue;
          const int kc = cmap[p.hi[j]], lc = cmap[p.lo[j]];
          if (kc == lc) continue;
          if (kc == 0 || lc == 0) A += p.m2[j];
          if (kc == 1 || lc == 1) B += p.m2[j];
          if (kc == 2 || lc == 2) Cc += p.m2[j];
        }
        const double th = 1e-5;
        const double da = fabs(a - A) / p.pmsqr, db = fabs(b - B) / p.pmsqr, dc = fabs(c - Cc) / p.pmsqr;
        if ((da > 1e-10 && db < th && dc < th) || (db > 1e-10 && da < th && dc < th) ||
            (dc > 1e-10 && da < th && db < th))
          lvGeneric = false;
      }
    }
    double sum = 0;
    for (int tt = 0; tt < p.n_tests; tt++) {
      for (int q = 0; q < m * m; q++) La[q] = 0.0;
      for (int j = 0; j < E; j++) {
        if ((s >> j) & 1) continue;
        int k = cmap[p.hi[j]], l = cmap[p.lo[j]];
        if (k == l) continue;
        if (k < l) {
          const int x = k;
          k = l;
          l = x;
        }
        const double x = Xi[(size_t)tt * E + j];
        La[l * m + l] += x;
        if (k == cV - 1) continue;
        La[k * m + k] += x;
        La[k * m + l] -= x;
        La[l * m + k] -= x;
      }
      double pxi = -ldlt_trace_solve<NB>(La, PGP, m);
      for (int j = 0; j < E; j++)
        if (!((s >> j) & 1)) pxi += p.m2[j] * Xs[(size_t)tt * E + j];
      sum += fabs(pxi);
      /* the sum only grows: once over the threshold the flag is settled
       * (the reference sums all the test points; same flag) */
      if (!(sum / p.pmsqr < 1e-12)) break;
    }
    mm = sum / p.pmsqr < 1e-12;
  } else if (n == E) {
    mm = true;
  }
  const int Ls = n - V + cV;
  double J, r;
  if (n == 0) {
    J = 1.0;
    r = 1.0;
  } else {
    J = 0.0;
    for (int j = 0; j < E; j++)
      if ((s >> j) & 1) {
        const uint64_t c2 = s & ~((uint64_t)1 << j);
        J += Jr[c2] / rr[c2];
      }
    double nus = 0.0;
    for (int j = 0; j < E; j++)
      if ((s >> j) & 1) nus += p.nu[j];
    r = nus - p.D / 2.0 * Ls;
    if (mm) r -= p.W;
  }
  if (r / p.W_ref < 1e-12 && n != E) atomicMin(&divs[n], (unsigned long long)s);
  if (p.phi && p.noneucl && n >= 2) {
    const int z = mm ? Ls + 1 : Ls;
    bool ok = true;
    for (int a = 0; a < E && ok; a++) {
      if (!((s >> a) & 1)) continue;
      const uint64_t A = s & ~((uint64_t)1 << a);
      const int zA = mmv[A] ? Lt[A] + 1 : Lt[A];
      for (int b = a + 1; b < E; b++) {
        if (!((s >> b) & 1)) continue;
        const uint64_t AB = A & ~((uint64_t)1 << b), B = s & ~((uint64_t)1 << b);
        const int zB = mmv[B] ? Lt[B] + 1 : Lt[B];
        const int zAB = mmv[AB] ? Lt[AB] + 1 : Lt[AB];
        if (zA + zB > z + zAB) {
          ok = false;
          break;
        }
      }
    }
    lvGP = ok;
  }
  Jr[s] = J;
  rr[s] = r;
  mmv[s] = mm ? 1 : 0;
  Lt[s] = (uint8_t)Ls;
  if (!lvGeneric) atomicAnd(&flags[0], 0u);
  if (!lvPseudo) atomicAnd(&flags[1], 0u);
  if (!lvGP) atomicAnd(&flags[2], 0u);
}

}  // namespace ftrj

/* Stream start states of the blocks of a launch: state of block b = the
 * base state jumped by b blocks.  b is written in base 256 (one digit for up
 * to 256 blocks); digit k with value d > 0 applies the jump polynomial of
 * d 256^k blocks, polys[(255 k + d - 1) 4 ...] (a jump by polynomial q maps
 * the state s to q(M) s, M the
 * xoshiro256+ transition over GF(2); jumps commute, so any decomposition
 * gives the same state).  One thread per block. */
__device__ inline void xo_apply_dev(uint64_t (&s)[4], const uint64_t* q) {
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int w = 0; w < 4; w++) {
    const uint64_t bits = q[w];
    for (int b = 0; b < 64; b++) {
      if ((bits >> b) & 1) {
        acc[0] ^= s[0];
        acc[1] ^= s[1];
        acc[2] ^= s[2];
        acc[3] ^= s[3];
      }
      const uint64_t t = s[1] << 17;
      s[2] ^= s[0];
      s[3] ^= s[1];
      s[1] ^= s[2];
      s[0] ^= s[3];
      s[2] ^= t;
      s[3] = (s[3] << 45) | (s[3] >> 19);
    }
  }
  s[0] = acc[0];
  s[1] = acc[1];
  s[2] = acc[2];
  s[3] = acc[3];
}

extern "C" __global__ void __launch_bounds__(256)
    ftrj_block_states(uint64_t b0, uint64_t b1, uint64_t b2, uint64_t b3, const uint64_t* polys, int ndig,
                      uint64_t nblocks, uint64_t* out) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  uint64_t s[4] = {b0, b1, b2, b3};
  for (int k = 0; k < ndig; k++) {
    const int d = (int)((b >> (8 * k)) & 255);
    if (d) xo_apply_dev(s, polys + 4 * (255 * k + d - 1));
  }
  out[4 * b + 0] = s[0];
  out[4 * b + 1] = s[1];
  out[4 * b + 2] = s[2];
  out[4 * b + 3] = s[3];
}


/* zero n counters (the work counters of persistent batched launches; a
 * kernel rather than a memset so that the launch graph holds no memset node) */
extern "C" __global__ void __launch_bounds__(64) ftrj_zero_u32(uint32_t* c, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) c[i] = 0u;
}

#define FTR_JT_KERNEL(VB)                                                                                       \
  extern "C" __global__ void __launch_bounds__(64)                                                              \
      ftrj_level_##VB(const ftrj::JtParams* prm, int n, uint64_t count, const uint64_t* binom, const double* Xs, \
                      const double* Xi, double* Jr, double* rr, uint8_t* mmv, uint8_t* Lt, uint32_t* flags,      \
                      unsigned long long* divs) {                                                               \
    ftrj::jt_level_body<VB>(prm, n, count, binom, Xs, Xi, Jr, rr, mmv, Lt, flags, divs);                        \
  }

FTR_JT_KERNEL(8)
FTR_JT_KERNEL(16)
FTR_JT_KERNEL(24)

/* sampler tables from the J-table: w = J/r, 1/r, mm | L << 1 */
extern "C" __global__ void ftrj_sampler_tables(uint64_t NT, const double* Jr, const double* rr, const uint8_t* mmv,
                                               const uint8_t* Lt, double* w, double* ir, uint8_t* mml) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT) return;
  w[i] = Jr[i] / rr[i];
  ir[i] = 1. / rr[i];
  mml[i] = (uint8_t)((mmv[i] ? 1 : 0) | (Lt[i] << 1));
}

/* Folded rows (sampler table mode 4; the layout of ftr_abi.hpp's
 * frow_stride, (E + 3) & ~1 doubles per row): the running sums of w(g\e)
 * over e = 0..E-1 (absent edges add 0; J(g) itself is summed in this
 * order, so T_{E-1} = J(g) for g != 0), 1/r(g), mm | L<<1 as the low 32
 * bits of a double, zero padding. */
extern "C" __global__ void ftrj_frows(uint64_t NT, int E, const double* w, const double* ir, const uint8_t* mml,
                                      double* rows) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= NT) return;
  const int RS = (E + 3) & ~1;
  double* r = rows + g * RS;
  double T = 0.0;
  for (int e = 0; e < E; e++) {
    T += ((g >> e) & 1) ? w[g ^ ((uint64_t)1 << e)] : 0.0;
    r[e] = T;
  }
  r[E] = ir[g];
  uint32_t* m = reinterpret_cast<uint32_t*>(r + E + 1);
  m[0] = mml[g];
  m[1] = 0;
  if (RS > E + 2) r[E + 2] = 0.0;
}

/* Row-table placement (sampler table mode 3): rows[g E + e] = w(g\e) for e in
 * g, else 0; rec[g] = {J(g), 1/r(g), mm | L<<1 (as the low 32 bits of the
 * third double), 0}. */
extern "C" __global__ void ftrj_rows(uint64_t NT, int E, const double* w, const double* Jr, const double* ir,
                                     const uint8_t* mml, double* rows, double* rec) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= NT) return;
  for (int e = 0; e < E; e++) rows[g * E + e] = ((g >> e) & 1) ? w[g ^ ((uint64_t)1 << e)] : 0.0;
  double* r = rec + g * 4;
  r[0] = Jr[g];
  r[1] = ir[g];
  uint32_t* m = reinterpret_cast<uint32_t*>(r + 2);
  m[0] = mml[g];
  m[1] = 0;
  r[3] = 0.0;
}
