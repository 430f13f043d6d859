/*
 * feyntrop_amd.h -- C ABI of the MI355X tropical-sampling engine.
 *
 * Plain C types only (no torch, no STL): the CLI binary `feyntrop`, Python
 * (ctypes) and any FFI bind these symbols.  Every call returns FTR_OK (0) or
 * a negative status; ftr_last_error() gives the message of the last failure
 * on that context.  All calls are synchronous with respect to the host unless
 * stated otherwise; a context is re-entrant only from one host thread at a
 * time.
 *
 * Reference interfaces replaced (michibo/feyntrop):
 *   ftr_prepare     -- feyntrop.cpp:45-239: graph construction and validation
 *                      (graph.hpp:169-214), scalar-product checks, regime
 *                      detection (feyntrop.cpp:128-133) and
 *                      generate_subgraph_table (tropical_sampling.hpp:284-458)
 *   ftr_integrate   -- feynman_integral_estimate<Generator>(N, g, D, P, m2, K,
 *                      lambda_eff, J, gen)  (feynman_integral.hpp:30-41,
 *                      called at feyntrop.cpp:251) returning the raw moments
 *                      that stats.hpp:73-134 holds
 *   ftr_sample      -- get_random_psi_xi_tropical_sample
 *                      (tropical_sampling.hpp:503-564) over a stream range
 *   ftr_replay      -- the per-sample loop body of feynman_integral_estimate
 *                      (feynman_integral.hpp:166-362) on given points
 *   ftr_run_json    -- the feyntrop binary's stdin -> stdout JSON contract
 *                      (feyntrop.cpp:20-281)
 *   ftr_finalize    -- stats::avg/acc (stats.hpp:86-104)
 */
#ifndef FEYNTROP_AMD_H
#define FEYNTROP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTR_OK 0
#define FTR_ERR_INVALID -1   /* invalid input (reference: std::invalid_argument) */
#define FTR_ERR_DOMAIN -2    /* mathematically unsupported input (std::domain_error) */
#define FTR_ERR_DEVICE -3    /* HIP runtime failure */
#define FTR_ERR_COMPILE -4   /* kernel specialisation failed */
#define FTR_ERR_LIMIT -5     /* outside this engine's size limits */

typedef struct ftr_ctx ftr_ctx;
typedef struct ftr_problem ftr_problem;

/* One Feynman-integral problem, as in the low-level JSON input. */
typedef struct {
  int E;                  /* number of edges */
  int V;                  /* number of rows/cols of P (must equal the vertex count) */
  const int32_t* edges;   /* 2*E vertex ids, as given */
  const double* nu;       /* E edge weights */
  const double* P;        /* V*V scalar products, row-major */
  const double* m2;       /* E squared masses */
  double D;               /* dimension */
  int K;                  /* num_eps_terms */
  double lambda;          /* deformation parameter */
} ftr_input;

/* Preprocessing results (J-table scalars and regime flags). */
typedef struct {
  int E, V, L, K;
  double W, IGtr;
  int euclidean, pseudo_euclidean, generic, gp_property;
  double lambda_eff;      /* lambda if the deformation is active, else 0 */
  int path;               /* 0 real (psi, phi), 1 real (psi only), 2 deformed complex */
  double seconds_preprocessing;
} ftr_info;

/* Raw moments of one integration: per order k and part q (0 re, 1 im):
 * n[k][q], s1[k][q], s2[k][q]; plus dropped samples and device timing. */
#define FTR_MAX_K 16
typedef struct {
  int K;
  double n[FTR_MAX_K][2];
  double s1[FTR_MAX_K][2];
  double s2[FTR_MAX_K][2];
  uint64_t dropped;
  double kernel_ms;       /* HIP-event time of the sampling kernel(s) */
  double seconds;         /* wall time of the call */
  uint64_t lanes;         /* sample slots launched (lanes, or lane pairs) */
  uint64_t samples_per_lane;
} ftr_moments;

/* Launch options.  The kernel variant (J-table placement, block size) is
 * fixed when the problem is prepared (ftr_set_variant, ftr_autotune change
 * it); these fields can only confirm it, and a call whose block or
 * table_mode differs from it fails with FTR_ERR_INVALID. */
typedef struct {
  int block;              /* 0 = the compiled one, else must equal it (256, 512 or 1024) */
  int table_mode;         /* -1 = the compiled one, else must equal it: 0 whole table in LDS,
                           * 1 w in LDS + rest global, 2 all global, 3 rows in HBM,
                           * 4 folded rows in HBM (running sums and the state's record
                           * in one row: graphs of 16 or more edges) */
  uint64_t lanes;         /* total sample slots (lanes, or lane pairs at two lanes per sample);
                           * 0 = fill the device */
} ftr_opts;

int ftr_create(int device, ftr_ctx** ctx);
void ftr_destroy(ftr_ctx* ctx);
const char* ftr_last_error(const ftr_ctx* ctx);

/* Validate + preprocess + specialise kernels.  *out owned by the caller.
 * ctx may be NULL: host-only preparation (validation, regime, J-table) on
 * any host; such a problem can be inspected but not integrated.  A problem
 * must be freed before its context is destroyed. */
int ftr_prepare(ftr_ctx* ctx, const ftr_input* in, ftr_problem** out);
void ftr_problem_free(ftr_problem* p);
/* Prepare n problems at once (no reference counterpart; for scans): the host
 * phases (validation, regime, J-table) run in parallel over the problems,
 * then each problem gets its kernel and one device allocation.  out[i] is
 * NULL and status[i] nonzero for a rejected problem, whose message
 * ftr_batch_error(ctx, i) returns (valid until the next batch call). */
int ftr_prepare_batch(ftr_ctx* ctx, int n, const ftr_input* in, ftr_problem** out, int* status);
const char* ftr_batch_error(const ftr_ctx* ctx, int i);
int ftr_problem_info(const ftr_problem* p, ftr_info* info);
/* J-table: arrays of 2^E entries (Jr, r, mm, L). */
int ftr_jtable(const ftr_problem* p, double* Jr, double* r, uint8_t* mm, uint8_t* L);

/* Integrate N samples of the single xoshiro256+ stream seeded with `seed`:
 * sample i uses stream positions [i(2E-1), (i+1)(2E-1)), so the multiset of
 * samples equals the reference run with one OpenMP thread. */
int ftr_integrate(ftr_ctx* ctx, ftr_problem* p, uint64_t N, uint64_t seed, const ftr_opts* opts,
                  ftr_moments* out);

/* ftr_integrate without any device allocation (no reference counterpart;
 * for callers that cannot call hipMalloc: code under HIP graph capture, tight
 * serving loops, harnesses that forbid it).  Same kernels, samples and
 * moments as ftr_integrate (bit for bit at the same launch geometry).  The
 * block start states, jump polynomials and partial moments live in a fixed
 * device array of FTR_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; a launch that needs more (32 + 8 (2 + 4K) bytes per block: about
 * 7000 blocks at K = 16, 37000 at K = 2; only a large opts->lanes asks for
 * that many) is rejected with FTR_ERR_LIMIT.  Allocates no device or
 * pinned memory (the problem must have been prepared, which allocates; the
 * host keeps its cache of jump polynomials).  All calls on one GPU share the
 * array: calls that overlap on one GPU (other threads, other contexts) race. */
#define FTR_NOALLOC_SCRATCH_BYTES (4u << 20)
int ftr_integrate_noalloc(ftr_ctx* ctx, ftr_problem* p, uint64_t N, uint64_t seed, const ftr_opts* opts,
                          ftr_moments* out);

/* Final estimates from moments like stats.hpp:86-104 (real path: the
 * imaginary parts are exactly 0, as with the reference's zero updates).
 * res: K x 4 = (re, dre, im, dim). */
int ftr_finalize(const ftr_problem* p, const ftr_moments* m, double* res);

/* Device-pointer entry points for parity work (caller-owned HBM buffers):
 * samples [first, first+n) of the stream seeded with seed -> X (n*E), psixi (n*2). */
int ftr_sample(ftr_ctx* ctx, ftr_problem* p, uint64_t first, uint64_t n, uint64_t seed,
               double* X_dev, double* psixi_dev);
/* Integrand on given points: X (n*E), psixi (n*2) -> out (n*K*2), dropped (n). */
int ftr_replay(ftr_ctx* ctx, ftr_problem* p, uint64_t n, const double* X_dev,
               const double* psixi_dev, double* out_dev, uint8_t* dropped_dev);

/* Full CLI contract: JSON text in -> JSON text out (caller frees with
 * ftr_free_string); log lines as the reference writes to std::clog. */
int ftr_run_json(ftr_ctx* ctx, const char* json_in, char** json_out, char** log_out);
void ftr_free_string(char* s);

/* Batched scan (no reference counterpart): n problems of uneven size in one
 * hipGraph.  The problems are bucketed by (kernel path, eps orders, J-table
 * placement, block size, edge and vertex counts); each bucket is one persistent
 * launch of a kernel that holds every topology of the bucket as a
 * compile-time graph spec and whose blocks take (problem, block of samples)
 * units from a work queue, costliest first.  Every problem gets the samples
 * and moments of its own ftr_integrate call (up to summation order).  A call
 * with the same launch shapes as the previous one (same problems and sample
 * counts; seeds may differ) replays that graph. */
int ftr_integrate_batch(ftr_ctx* ctx, int n, ftr_problem* const* probs, const uint64_t* N,
                        const uint64_t* seeds, ftr_moments* out);
/* Compile (no GPU needed) the bucket kernels that ftr_integrate_batch would
 * use for these n problems into the kernel cache (problems the
 * preprocessing rejects are skipped). */
int ftr_precompile_batch(int n, const ftr_input* in, char* err, int errlen);
/* Number of batch calls served by replaying the previous graph. */
int ftr_batch_graph_replays(const ftr_ctx* ctx);

/* Autotuner (no reference counterpart; replaces the reference's fixed OpenMP
 * chunking): times every feasible (J-table placement, block size, lanes per
 * sample) kernel variant on n_probe samples, keeps the fastest for the
 * problem and records it in the cache directory, where later ftr_prepare
 * calls pick it up; then times the kernel specialised to the problem's
 * kinematic point (ftr_set_specialized) and keeps it for this problem if it
 * is faster.  table_mode: 0 whole J-table in LDS, 1 J/r in LDS + rest from
 * L2/HBM, 2 all from L2/HBM, 3 rows in HBM, 4 folded rows in HBM (tried
 * from 16 edges on).  The chosen lanes per sample: ftr_get_lanes_per_sample. */
int ftr_autotune(ftr_ctx* ctx, ftr_problem* p, uint64_t n_probe, int* table_mode, int* block,
                 double* best_kernel_ms);
/* Select the (table_mode, block) variant at the problem's current lanes per
 * sample; FTR_ERR_LIMIT when it is infeasible for the problem. */
int ftr_set_variant(ftr_ctx* ctx, ftr_problem* p, int table_mode, int block);
/* Lanes per sample (no reference counterpart): 1, one sample per lane, or,
 * on the deformed (complex) path only, 2: a lane pair evaluates one sample,
 * splitting the rows of the deformation Jacobian's LDL^T between its lanes
 * (half the live factor rows per lane) with the second derivatives in LDS
 * per pair; table mode 3 or 4, blocks of 128, 256 or 512 threads.  Same samples
 * and, per sample, the same values as one lane per sample; ftr_opts.lanes
 * and ftr_moments.lanes then count sample slots (lane pairs).  lps = 1
 * returns to the default variant.  FTR_ERR_LIMIT on the real paths or
 * when no variant fits. */
int ftr_set_lanes_per_sample(ftr_ctx* ctx, ftr_problem* p, int lps);
int ftr_get_lanes_per_sample(const ftr_problem* p, int* lps);
/* Specialise the problem's integration kernel to its kinematic point (no
 * reference counterpart): D, P, m^2, nu, lambda and the exponents become
 * compile-time constants of the generated kernel (one hipRTC compile per
 * kinematic point, cached), on = 0 returns to the generic kernel.  Same
 * samples, same estimator; for long runs of one point. */
int ftr_set_specialized(ftr_ctx* ctx, ftr_problem* p, int on);
int ftr_get_specialized(const ftr_problem* p, int* on);
/* The drop-in binary's kernel choice (no reference counterpart; the feyntrop
 * CLI calls it after ftr_prepare): switches p to its specialised kernel
 * when that kernel's code object is already in the cache, so no compile is
 * paid; environment FEYNTROP_AMD_SPECIALIZE=1 specialises always (compiling
 * if needed), =0 never.  FTR_OK, or the error of ftr_set_specialized. */
int ftr_use_cached_specialization(ftr_ctx* ctx, ftr_problem* p);
/* Compile the specialised kernel of a problem into the cache without a GPU. */
int ftr_precompile_specialized(const ftr_input* in, char* err, int errlen);
/* Compile one kernel variant of a problem (as ftr_set_variant and
 * ftr_set_lanes_per_sample select it; specialized: the kernel specialised
 * to the kinematic point) into the cache without a GPU, for the LDS size of
 * gfx950 (160 KiB).  FTR_ERR_LIMIT when the variant is infeasible. */
int ftr_precompile_variant(const ftr_input* in, int table_mode, int block, int lps, int specialized, char* err,
                           int errlen);
int ftr_get_variant(const ftr_problem* p, int* table_mode, int* block);

/* Where ftr_prepare builds the 2^E-entry J-table (tropical_sampling.hpp:
 * 284-458): 0 auto (the device from E = 13 on, the host below), 1 host
 * (OpenMP), 2 device (one launch per subset size, one thread per subset).
 * The tables and flags are bitwise the same either way.  The environment
 * variable FEYNTROP_AMD_JTABLE=host|device sets the default of new contexts. */
int ftr_set_jtable_mode(ftr_ctx* ctx, int mode);

/* Kernel code-object cache directory (default: the package's kcache/). */
int ftr_set_cache_dir(const char* dir);

/* Compile (hipRTC) the kernels for a problem into the cache without a GPU. */
int ftr_precompile(const ftr_input* in, char* err, int errlen);

/* ---- Learned proposal (no reference counterpart: an addition on top of
 * the tropical sampler of feynman_integral.hpp:31-77) ----
 * A small MLP (input 32, 3 x 128 SiLU, output 32; BF16 weights on the MFMA
 * units, FP32 masters, Adam) maps every subgraph state g of the J-table to
 * a tilt Delta_e(g) of each cut weight and a radial exponent shift
 * sigma(g).  Sampling from the tilted proposal q and weighting every sample
 * by the exact FP64 ratio p_trop(path)/q(path) keeps the estimator unbiased
 * for any network; training minimises the forward KL divergence between
 * |integrand| x p_trop and q.  Needs E <= 16.  An untrained learner
 * (zero output layer) reproduces the plain sampler sample for sample.
 * Network input per state: edge bits, popcount/E, 1, r(g)/E and the scalar
 * products among the vertices with external momentum (the kinematics). */
typedef struct ftr_learner ftr_learner;
#define FTR_LEARN_HIDDEN 128
#define FTR_LEARN_NPARAMS (128 * 32 + 128 + 2 * (128 * 128 + 128) + 32 * 128 + 32)

int ftr_learner_create(ftr_ctx* ctx, ftr_problem* p, uint64_t init_seed, ftr_learner** out);
void ftr_learner_free(ftr_learner* l);
/* FP32 master parameters (FTR_LEARN_NPARAMS floats: W1[128x32] b1 W2[128x128]
 * b2 W3[128x128] b3 W4[32x128] b4); get and/or set (either may be NULL). */
int ftr_learner_params(ftr_learner* l, float* get, const float* set);
/* Network output for every state: Y (max(32, 2^E) x 32 floats, host). */
int ftr_learner_forward(ftr_ctx* ctx, ftr_learner* l, float* Y);
/* Parameter gradient of sum(dY * Y) for a given dY (same shape as Y). */
int ftr_learner_backward(ftr_ctx* ctx, ftr_learner* l, const float* dY, float* grad);
/* `steps` Adam steps, each on `samples` proposal samples (stream seed
 * seed + step).  stats (optional, steps x 2): per step the relative
 * second moment N sum(w f)^2 / (sum w f)^2 of the epsilon^0 real part and
 * the kernel milliseconds.  Reproducible: the same initial parameters,
 * samples, seed and lr give the same network bit for bit (statistics in
 * fixed point, gradient sums in a fixed order). */
int ftr_learner_train(ftr_ctx* ctx, ftr_learner* l, int steps, uint64_t samples, uint64_t seed, double lr,
                      double* stats);
/* One training step's statistics and table gradient, without the update
 * (for checking them): draws `samples` proposal samples (stream seed `seed`)
 * and returns, each output optional (NULL skips it):
 *   cw[N]         c = |f| p_trop/q per sample (its square with objective 1;
 *                 0 for a dropped sample)
 *   ly[(E-1) N]   the radial log-draws, ly[j N + i] for step j of sample i
 *   path[N]       the cut order, 4 bits per step
 *   gC[2^E E], gS[2^E E], gsum[2]   the forward-KL statistics per (state,
 *                 cut edge) (ftr_abi.hpp LearnLaunch), summed in fixed point
 *   dY[max(32, 2^E) x 32]         dL/dY of the network outputs
 * No reference counterpart (learned proposal). */
int ftr_learner_statistics(ftr_ctx* ctx, ftr_learner* l, uint64_t samples, uint64_t seed, double* cw, double* ly,
                           uint64_t* path, double* gC, double* gS, double* gsum, float* dY);
/* Training objective: 0 forward KL between |f| p_trop and q (default; the
 * statistics weight each path by c = |f| p_trop/q), 1 the estimator's
 * variance E_q[(f p_trop/q)^2] (c = (|f| p_trop/q)^2). */
int ftr_learner_set_objective(ftr_learner* l, int objective);
/* Current tables, host arrays: lw, lratio, lia, lrad, ldiff (2^E x E, per
 * state and cut edge); lZ (2^E). */
int ftr_learner_tables(ftr_ctx* ctx, ftr_learner* l, double* lw, double* lZ, double* lratio, double* lia,
                       double* lrad, double* ldiff);
/* Integrate with the learned proposal: moments of R * p_trop/q. */
int ftr_learner_integrate(ftr_ctx* ctx, ftr_learner* l, uint64_t N, uint64_t seed, const ftr_opts* opts,
                          ftr_moments* out);
/* Plain sampling as the judge of the network: runs both proposals on
 * n_probe samples and compares their efficiencies (relative variance of the
 * eps^0 coefficient x kernel time per sample, lower is better); when the
 * learned one is not better, ftr_learner_integrate falls back to the plain
 * sampler (ftr_integrate).  Outputs are optional. */
int ftr_learner_select(ftr_ctx* ctx, ftr_learner* l, uint64_t n_probe, uint64_t seed, int* use_learned,
                       double* eff_plain, double* eff_learned);
/* Proposal samples [first, first+n): X (n*E), weight p/q (n), cut order
 * (n, 4 bits per step, first cut lowest) into device buffers. */
int ftr_learner_sample(ftr_ctx* ctx, ftr_learner* l, uint64_t first, uint64_t n, uint64_t seed, double* X_dev,
                       double* w_dev, uint64_t* path_dev);
/* Compile the learned-proposal kernels of a problem (and the network's)
 * into the cache without a GPU. */
int ftr_precompile_learned(const ftr_input* in, char* err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
