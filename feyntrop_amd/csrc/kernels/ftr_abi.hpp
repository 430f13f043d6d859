/*
 * ftr_abi.hpp -- data layout shared by the host runtime and the HIP kernels
 * (plain C++: no HIP types).  Limits: E <= FTR_MAX_E edges (the 2^E J-table
 * is indexed by 32-bit masks), V <= FTR_MAX_V vertices.
 */
#pragma once

#ifdef __HIPCC_RTC__
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#else
#include <stdint.h>
#endif

#define FTR_MAX_E 30
#define FTR_MAX_V 24
/* Deformed-path kernels with the row-table placement, blocks of at most 256
 * threads and E <= FTR_DD_LDS_MAX_E keep the second derivatives of the
 * deformation (E(E+1)/2 doubles per thread) in LDS, laid out [entry][thread],
 * instead of registers.  Measured per 1e8 samples against registers: 3-loop
 * ladder (E = 10) specialised 27.8 -> 22.8 ms, generic 23.6 -> 23.5 ms,
 * learned 26.9 -> 25.6 ms; double box generic 14.3 -> 13.7 ms (the
 * 10-edge kernels no longer spill to scratch).  E = 11 would need 132 KiB
 * per 256-thread block on top of the X slots. */
#define FTR_DD_LDS_MAX_E 10

namespace ftr {

/* ----------------------------------------------------------- parameters -- */
/* Uniform per-problem data, read by every lane (scalar loads). */
struct DevParams {
  double D;          /* dimension */
  double mhalfD;     /* -D/2 */
  double W;          /* superficial degree of divergence */
  double mW;         /* -W */
  double Lf;         /* loop number as double */
  double IGtr;       /* J(G) */
  double def_ref;    /* -lambda_eff (0 on the real path) */
  double inv_dlw;    /* 1 / (D/2 L + W) */
  double iD, iW;     /* -D/2 and -W when they are integers (|n| <= 64), else 0 */
  double use_iD, use_iW; /* 1 when the integer powers apply */
  double nu[FTR_MAX_E];
  double m2[FTR_MAX_E];
  double P[FTR_MAX_V * FTR_MAX_V];
  /* P's block on the vertices with external momentum (in elimination
   * order) as sum_r Peta[r] q_r q_r^T, q_r = Pq[r * FTR_MAX_V + a] over the
   * external positions a, Peta = +-1; r < the graph spec's PR */
  double Pq[FTR_MAX_V * FTR_MAX_V];
  double Peta[FTR_MAX_V];
};

/* Folded rows (table mode 4, for tables far larger than the L2): state g's
 * row of frow_stride(E) doubles holds the running sums
 * T_q = sum_{e <= q, e in g} w(g\e) of its cut weights in edge order
 * (absent edges add 0, so T_{E-1} = J(g) for g != 0), then 1/r(g), then
 * mm | L<<1 in the low 32 bits of the next double: a cut step reads one row
 * and no record.  The stride is even, so every row starts on 16 bytes. */
constexpr int frow_stride(int E) { return (E + 3) & ~1; }
/* smallest E whose default row placement is table mode 4.  Generic
 * kernels, ms per 1e8 samples, mode 3 against 4: 14-edge cubic graph 37.4
 * against 50.5; 4-point ladders (and diagonals) of 16, 17, 18 edges 24.4
 * against 20.9, 33.7 against 25.4, 40.0 against 28.7; 22 edges 83.7
 * against 56.4 (mode-3 tables of 10 MB at E = 16, 0.87 GB at 22).  The
 * kernel specialised to a kinematic point loses with mode 4 at 16 edges
 * (22.6 against 29.9 ms) and ties from 19; the autotuner weighs both. */
#define FTR_FROWS_MIN_E 16

/* Per-state record of the row-table placement (table mode 3). */
struct alignas(16) StateRec {
  double jr;      /* J(g) */
  double ir;      /* 1/r(g) */
  uint32_t mmL;   /* mm | L<<1 */
  uint32_t pad[3];
};

/* Kernel arguments for one problem. */
struct Launch {
  const DevParams* prm;
  const double* t_w;       /* J(g)/r(g), 2^E */
  const double* t_jr;      /* J(g), 2^E */
  const double* t_ir;      /* 1/r(g), 2^E */
  const uint8_t* t_mmL;    /* mm | L<<1, 2^E */
  const double* t_rows;    /* table mode 3: 2^E x E, row g holds w(g\e) for e in g, 0 else;
                            * table mode 4: 2^E folded rows of frow_stride(E) */
  const StateRec* t_rec;   /* table mode 3: 2^E records */
  const uint64_t* block_state; /* nblocks x 4: stream state at the block's first sample */
  const uint64_t* lane_poly;   /* 255 per base-256 digit of the lane index, x 4 words: entry
                                * 255 k + d - 1 jumps d 256^k lanes (S (2E-1) draws each) */
  uint64_t N;              /* total samples */
  uint64_t S;              /* samples per lane */
  double* partials;        /* nblocks x (2 + 4K) */
};

/* One problem of a batched launch: its launch data (block0 and nblocks
 * describe its range of virtual blocks). */
struct BatchItem {
  Launch a;
  uint32_t block0;
  uint32_t nblocks;
};

/* One unit of a persistent batched launch: virtual block `blk` of batch item
 * `item`, whose graph is case `topo` of the bucket kernel.  A bucket kernel
 * runs a grid sized to the device; its blocks take units from an atomic
 * counter until the list is exhausted. */
struct BatchWork {
  uint32_t item;
  uint32_t blk;
  uint32_t topo;
  uint32_t pad;
};

/* Learned proposal tables (2^E states g; E <= 16; [g*E+e] entries are zero
 * for e not in g):
 *   lw[g*E+e]     cut weight of e from g: w(g\e) exp(Delta_e(g))
 *   lZ[g]         sum of lw[g*E+e] over e in g, ascending e
 *   lratio[g*E+e] p_trop(e|g) / q(e|g) = lZ[g] w(g\e) / (J(g) lw[g*E+e])
 *   lia[g*E+e]    1 / a(g,e), a(g,e) = r(g\e) exp(sigma(g,e)): exponent of
 *                 the radius drawn after cutting e from g (plain sampling:
 *                 a = r(g\e)); zero when g\e is empty (no radius follows)
 *   lrad[g*E+e]   log(r/a) = -sigma(g,e);   ldiff[g*E+e]  r(g\e) - a(g,e)
 * Training statistics (doubles, sums over a step's samples, accumulated in
 * fixed point): gC[g*E+e] = sum of c over visits of g that cut e; gS[g*E+e]
 * = sum of c (1 + a(g,e) log Y) over the radii drawn after cutting e from g;
 * gsum = (sum of c, number of samples), with c = weight * |R_0|.  The
 * fixed-point accumulator packs both into 2^E x E slots: C(g,e) at g*E+e
 * (e in g) and S(g,e) at (g\e)*E+e (a slot whose edge is not in its state,
 * unused by C), then the two sums. */
/* Per-state record of the learned tables, one 32-byte load per visited
 * state: lZ(g), A(g) = lZ(g)/J(g) and the state's mm | L<<1 flags. */
struct alignas(32) LQuad {
  double lZ, A;
  uint32_t mmL, pad0;
  double pad1;
};
/* Per-cut record, one 32-byte load per cut: the ratio factor
 * w(g\e)/lw(g,e) and the radial parameters lia, lrad, ldiff of (g,e). */
struct alignas(32) LEdge {
  double tw, lia, lrad, ldiff;
};

/* Fixed-point accumulator of the training statistics (two's complement). */
typedef unsigned long long fixacc;

struct LearnLaunch {
  Launch a;
  const LQuad* lst;
  const LEdge* led;    /* 2^E x E per-cut records */
  const double* lw;
  const double* lZ;
  const double* lratio;
  const double* lia;
  const double* lrad;
  const double* ldiff;
  double* gC;
  double* gS;
  double* gsum;
  unsigned long long* gmax; /* largest |statistic| of the step, double bits */
  double* gbuf;      /* training record: c (N), then lY_j (N each, j < E-1) */
  int objective;     /* 0: forward KL, c = w|f|;  1: variance, c = (w|f|)^2 */
  uint64_t* gpath;   /* training record: cut order per sample */
};

}  // namespace ftr
