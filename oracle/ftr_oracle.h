/*
 * ftr_oracle.h -- CPU FP64 restatement of feyntrop's tropical sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU product is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (feyntrop_amd) never links,
 * loads or falls back to it.
 *
 * What it restates (reference file:line):
 *   - xoshiro256+ / splitmix64 / uniform      srcs/random.hpp:16-140
 *   - arbitrary-distance jump (GF(2) polynomial; generalises random.hpp:147-199)
 *   - graph normalisation + validation        srcs/graph.hpp:169-214
 *   - connected components (BFS)              srcs/components.hpp:92-120
 *   - J-table generation + regime flags       srcs/tropical_sampling.hpp:26-458
 *   - regime detection on P                    srcs/feyntrop.cpp:98-133
 *   - sampler                                  srcs/tropical_sampling.hpp:468-564
 *   - per-sample integrand, real + deformed    srcs/feynman_integral.hpp:166-362
 *       with Eigen 3.4 LDLT / FullPivLU / PartialPivLU pivot rules
 *       (Cholesky/LDLT.h:299-389,567-610; LU/FullPivLU.h:488-587,745-780;
 *        LU/PartialPivLU.h:358-423,552-556)
 *   - fp_summer / stats / tree merge           srcs/stats.hpp:13-134
 *
 * Parity pinning (which check pins which function) is listed in the header
 * comment of ftr_oracle.cpp.
 */
#ifndef FTR_ORACLE_H
#define FTR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_problem or_problem;

/* Create a problem: validates like srcs/feyntrop.cpp:45-119 and graph.hpp,
 * runs the regime detection and the J-table preprocessing.
 * edges: 2*E ints (as given in the JSON), nu: E, P: V*V row-major (V is
 * derived from the edges), m2: E.  Returns NULL on error; the message is
 * copied to err (errlen bytes). */
or_problem* or_create(int E, const int* edges, const double* nu, int V_in,
                      const double* P, const double* m2, double D, int K,
                      double lambda, char* err, int errlen);
void or_destroy(or_problem* p);

/* Scalars: out[0]=E out[1]=V out[2]=L out[3]=W out[4]=IGtr
 * out[5]=bEuclidean out[6]=bPseudoEuclidean out[7]=bGeneric out[8]=bGPprop
 * out[9]=lambda_eff out[10]=def_ref out[11]=complex_path (1/0) out[12]=K */
void or_info(const or_problem* p, double* out);

/* J-table: arrays of length 2^E. */
void or_jtable(const or_problem* p, double* Jr, double* r, uint8_t* mm,
               uint8_t* L);

/* Draw n samples from the single stream whose state is st (advanced in
 * place).  X: n*E, psixi: n*2 (psi_tr, xi_tr). */
void or_sample(const or_problem* p, uint64_t* st, uint64_t n, double* X,
               double* psixi);

/* Per-sample integrand for given points.  out: n*K*2 (re, im per order),
 * dropped: n (1 if the sample is dropped by the non-finite rule).
 * vstar: n (chosen deleted vertex, may be NULL). */
void or_replay(const or_problem* p, uint64_t n, const double* X,
               const double* psixi, double* out, uint8_t* dropped,
               int32_t* vstar);

/* Single-stream estimate: samples i in [0, N) use positions
 * [i(2E-1), (i+1)(2E-1)) of the xoshiro256+ stream seeded with seed.
 * The range is split into `chunks` contiguous pieces (jump-ahead per piece,
 * OpenMP-parallel when threads > 1); the multiset of samples does not depend
 * on chunks.  moments: K*2*3 doubles (n, s1, s2) for (order, re/im);
 * the imaginary slots of the real path are filled like
 * feynman_integral.hpp:341-356 with 2 zero updates per chunk.
 * Returns dropped-sample count. */
uint64_t or_estimate(const or_problem* p, uint64_t N, uint64_t seed,
                     int chunks, int threads, double* moments);

/* Final avg/acc exactly like stats.hpp:86-104; res: K*4 (re, dre, im, dim). */
void or_finalize(const or_problem* p, const double* moments, double* res);

/* RNG helpers. */
void or_xoshiro_seed(uint64_t seed, uint64_t* st);
void or_xoshiro_next(uint64_t* st, uint64_t n, uint64_t* out);
/* Advance st by d steps (polynomial jump). */
void or_xoshiro_advance(uint64_t* st, uint64_t d_lo, uint64_t d_hi);
/* Characteristic polynomial of the xoshiro256 transition (257 bits: 5 words,
 * bit i = coefficient of x^i), computed by Berlekamp-Massey. */
void or_xoshiro_charpoly(uint64_t* poly5);
/* x^(2^e) mod charpoly, 4 words (the jump polynomial for distance 2^e). */
void or_xoshiro_jumppoly_pow2(int e, uint64_t* poly4);

/* Advance st by 2^e steps (e may exceed 127). */
void or_xoshiro_jump_pow2(int e, uint64_t* st);

/* fp_summer restatement (stats.hpp:13-57) on a stream, for tests. */
double or_fp_sum(const double* v, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
