/*
 * learn.cpp -- learned proposal for the tropical sampler (no reference
 * counterpart; it sits on top of the sampler of feynman_integral.hpp:31-77).
 *
 * Per problem: an MLP over the 2^E subgraph states (one row per state) whose
 * output tilts the cut weights and shifts the radial exponents of the
 * tropical proposal.  The network runs on the MFMA units in BF16 (module
 * ftr_mlp.hpp, compiled once); the tables it produces are FP64, and so is the
 * importance weight p_trop/q of every sample, so the estimate is unbiased
 * whatever the network outputs.  Training: forward KL between |f| p_trop and
 * q, estimated from proposal samples (self-normalised), Adam.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ftr_runtime.hpp"

namespace {

constexpr int H = FTR_LEARN_HIDDEN;
constexpr int NIN = 32, NOUT = 32;
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + (size_t)H * NIN;
constexpr size_t OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + (size_t)H * H;
constexpr size_t OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + (size_t)H * H;
constexpr size_t OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + (size_t)NOUT * H;
constexpr size_t NPAR = OFF_B4 + NOUT;
static_assert(NPAR == FTR_LEARN_NPARAMS, "parameter layout");
constexpr int MAX_E = 16;          /* cut order packed 4 bits per step */
constexpr double DMAX = 8.0;       /* |Delta| bound (cut-weight tilt) */
constexpr int HIST_E = 10;         /* histogram in LDS up to this E (ftr_device.hpp) */
constexpr size_t hist_len(int E) { return (size_t)E * ((size_t)1 << E) + 2; } /* = ftr::learn_hist_len */
constexpr int LEARN_BLOCK = 256;  /* one wave per SIMD per block: the full register budget per lane */
constexpr double SMAX = 0.6;       /* |sigma| bound: a = r e^sigma < 2r keeps p/q of finite variance */
constexpr int MAX_SPLITS = 64;     /* split-K blocks of a weight-gradient GEMM */
constexpr int MAX_CHUNKS = 256;    /* row chunks of a bias-gradient column sum */
constexpr size_t WS_FLOATS = (size_t)MAX_SPLITS * H * H; /* their workspace */

struct MlpModule {
  hipModule_t mod = nullptr;
  hipFunction_t gbb, gfb, colsum, split_sum, adam, fill, fill_f, to_bf16, tables, tgrad;
};

MlpModule load_mlp(ftr_ctx* c) {
  MlpModule m;
  m.mod = load_module_raw(c, ftr_embedded_mlp_source);
  HIPCHK(hipModuleGetFunction(&m.gbb, m.mod, "mlp_gemm_bb"));
  HIPCHK(hipModuleGetFunction(&m.gfb, m.mod, "mlp_gemm_fb"));
  HIPCHK(hipModuleGetFunction(&m.colsum, m.mod, "mlp_colsum"));
  HIPCHK(hipModuleGetFunction(&m.split_sum, m.mod, "mlp_split_sum"));
  HIPCHK(hipModuleGetFunction(&m.adam, m.mod, "mlp_adam"));
  HIPCHK(hipModuleGetFunction(&m.fill, m.mod, "mlp_fill"));
  HIPCHK(hipModuleGetFunction(&m.fill_f, m.mod, "mlp_fill_f"));
  HIPCHK(hipModuleGetFunction(&m.to_bf16, m.mod, "mlp_to_bf16"));
  HIPCHK(hipModuleGetFunction(&m.tables, m.mod, "mlp_build_tables"));
  HIPCHK(hipModuleGetFunction(&m.tgrad, m.mod, "mlp_table_grad"));
  return m;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

}  // namespace

struct ftr_learner {
  ftr_problem* p = nullptr;
  int device = 0;
  Variant var; /* the problem's kernel variant when the learner was made */
  size_t lds = 0;
  int E = 0;
  uint32_t NT = 0;
  int M = 0; /* network rows: max(32, 2^E) */
  MlpModule mlp;
  hipFunction_t f_lev = nullptr, f_lgr = nullptr, f_lsm = nullptr, f_hmx = nullptr, f_acc = nullptr,
                f_hst = nullptr, f_hrd = nullptr;
  ftr::fixacc* d_part = nullptr;     /* fixed-point statistics: per block (E <= HIST_E) or one */
  unsigned long long* d_gmax = nullptr; /* the step's largest |statistic| (fixed-point scale) */
  float* d_ws = nullptr;             /* split-K and column-sum partials */
  int hist_blocks = 0;
  double* d_gbuf = nullptr; /* training record (see ftr_abi.hpp LearnLaunch) */
  uint64_t* d_gpath = nullptr;
  uint64_t rec_cap = 0;     /* samples the record holds */
  float *d_par = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
  uint16_t *d_parb = nullptr, *d_X0b = nullptr, *d_H1 = nullptr, *d_H2 = nullptr, *d_H3 = nullptr;
  float *d_Z1 = nullptr, *d_Z2 = nullptr, *d_Z3 = nullptr, *d_Y = nullptr, *d_dY = nullptr, *d_dH = nullptr,
        *d_dZ = nullptr;
  double *d_lw = nullptr, *d_lZ = nullptr, *d_lratio = nullptr, *d_lia = nullptr, *d_lrad = nullptr,
         *d_ldiff = nullptr, *d_r = nullptr, *d_gC = nullptr, *d_gS = nullptr, *d_gsum = nullptr;
  ftr::LQuad* d_lst = nullptr; /* per-state records read by the sampler */
  ftr::LEdge* d_led = nullptr;  /* per-cut records read by the sampler */
  int adam_t = 0;
  bool tables_valid = false;
  bool use_plain = false; /* set by ftr_learner_select when the plain sampler wins */
  int objective = 0;      /* 0 forward KL, 1 variance */
  std::vector<void*> allocs;
};

namespace {

template <class T>
T* dalloc(ftr_learner* l, size_t n) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256)));
  l->allocs.push_back(p);
  return (T*)p;
}

void launch(hipFunction_t f, unsigned gx, unsigned gy, unsigned block, hipStream_t s, void** args) {
  HIPCHK(hipModuleLaunchKernel(f, gx, gy, 1, block, 1, 1, 0, s, args, nullptr));
}

/* out[i] = sum over z < splits of ws[z n + i] (fixed order) */
void split_sum(ftr_ctx* c, ftr_learner* l, float* out, uint64_t n, int splits) {
  void* args[] = {&l->d_ws, &out, &n, &splits};
  launch(l->mlp.split_sum, (unsigned)((n + 255) / 256), 1, 256, c->stream, args);
}

/* C = act(A B + bias), operands addressed by strides (see ftr_mlp.hpp);
 * act 2: C = (A B) * silu'(zs) */
void gemm(ftr_ctx* c, ftr_learner* l, bool a_float, const void* A, uint64_t sai, uint64_t sak, const uint16_t* B,
          uint64_t sbk, uint64_t sbj, const float* bias, float* C, uint64_t ldc, uint16_t* Cb, uint64_t ldcb, int M,
          int N, int K, int act, bool split_k = false, const float* zs = nullptr) {
  if (M % 32 || N % 32 || K % 32) throw Error(FTR_ERR_INVALID, "network GEMM shape not tiled");
  /* the kernel stages operands in 8-element vectors along their unit stride */
  auto vec_ok = [](const void* p, uint64_t s0, uint64_t s1) {
    return ((uintptr_t)p % 16 == 0) && ((s0 == 1 && s1 % 8 == 0) || (s1 == 1 && s0 % 8 == 0));
  };
  if (!vec_ok(A, sai, sak) || !vec_ok(B, sbk, sbj)) throw Error(FTR_ERR_INVALID, "network GEMM operand layout");
  if ((act == 2) != (zs != nullptr)) throw Error(FTR_ERR_INVALID, "network GEMM: SiLU backward needs Z");
  /* long reductions (weight gradients over all states) split K over at
   * most MAX_SPLITS blocks; their partial products go to the workspace and
   * are added in a fixed order (reproducible, no atomics) */
  int splits = 1, kc = K;
  if (split_k && K >= 256 && !Cb && !bias && !act) {
    if (ldc != (uint64_t)N || (size_t)MAX_SPLITS * M * N > WS_FLOATS)
      throw Error(FTR_ERR_INVALID, "network GEMM split does not fit the workspace");
    kc = std::max(64, (((K + MAX_SPLITS - 1) / MAX_SPLITS) + 63) / 64 * 64);
    splits = (K + kc - 1) / kc;
  }
  float* dst = splits > 1 ? l->d_ws : C;
  void* args[] = {(void*)&A, &sai, &sak, (void*)&B, &sbk, &sbj, (void*)&bias, &dst, &ldc, &Cb, &ldcb, &K, &act, &kc,
                  (void*)&zs};
  HIPCHK(hipModuleLaunchKernel(a_float ? l->mlp.gfb : l->mlp.gbb, (unsigned)(N / 32), (unsigned)(M / 32),
                               (unsigned)splits, 64, 1, 1, 0, c->stream, args, nullptr));
  if (splits > 1) split_sum(c, l, C, (uint64_t)M * N, splits);
}

void fill_d(ftr_ctx* c, ftr_learner* l, double* p, uint64_t n, double v) {
  void* args[] = {&p, &n, &v};
  launch(l->mlp.fill, (unsigned)((n + 255) / 256), 1, 256, c->stream, args);
}

void fill_f(ftr_ctx* c, ftr_learner* l, float* p, uint64_t n, float v) {
  void* args[] = {&p, &n, &v};
  launch(l->mlp.fill_f, (unsigned)((n + 255) / 256), 1, 256, c->stream, args);
}

void refresh_bf16(ftr_ctx* c, ftr_learner* l) {
  uint64_t n = NPAR;
  void* args[] = {&l->d_par, &l->d_parb, &n};
  launch(l->mlp.to_bf16, (unsigned)((n + 255) / 256), 1, 256, c->stream, args);
}

void forward(ftr_ctx* c, ftr_learner* l) {
  const uint16_t* W = l->d_parb;
  const float* P = l->d_par;
  const int M = l->M;
  /* Z = X W^T + b: B[k][j] = W[j][k] */
  gemm(c, l, false, l->d_X0b, NIN, 1, W + OFF_W1, 1, NIN, P + OFF_B1, l->d_Z1, H, l->d_H1, H, M, H, NIN, 1);
  gemm(c, l, false, l->d_H1, H, 1, W + OFF_W2, 1, H, P + OFF_B2, l->d_Z2, H, l->d_H2, H, M, H, H, 1);
  gemm(c, l, false, l->d_H2, H, 1, W + OFF_W3, 1, H, P + OFF_B3, l->d_Z3, H, l->d_H3, H, M, H, H, 1);
  gemm(c, l, false, l->d_H3, H, 1, W + OFF_W4, 1, H, P + OFF_B4, l->d_Y, NOUT, nullptr, 0, M, NOUT, H, 0);
}

/* db = column sums of dZ: sums over row chunks, then the chunks added in a
 * fixed order */
void colsum(ftr_ctx* c, ftr_learner* l, const float* dZ, float* db, int rows, int cols) {
  int chunk = std::max(32, (rows + MAX_CHUNKS - 1) / MAX_CHUNKS);
  const int nchunks = (rows + chunk - 1) / chunk;
  if ((size_t)nchunks * cols > WS_FLOATS) throw Error(FTR_ERR_INVALID, "network column sum does not fit the workspace");
  void* args[] = {(void*)&dZ, &l->d_ws, &rows, &cols, &chunk};
  launch(l->mlp.colsum, (unsigned)((cols + 63) / 64), (unsigned)nchunks, 64, c->stream, args);
  split_sum(c, l, db, (uint64_t)cols, nchunks);
}

/* parameter gradient of sum(dY * Y) into d_grad (forward must have run) */
void backward(ftr_ctx* c, ftr_learner* l) {
  const uint16_t* W = l->d_parb;
  float* G = l->d_grad;
  const int M = l->M;
  /* layer 4: dW4 = dY^T H3, db4 = colsum dY, dZ3 = (dY W4) silu'(Z3) */
  gemm(c, l, true, l->d_dY, 1, NOUT, l->d_H3, H, 1, nullptr, G + OFF_W4, H, nullptr, 0, NOUT, H, M, 0, true);
  colsum(c, l, l->d_dY, G + OFF_B4, M, NOUT);
  gemm(c, l, true, l->d_dY, NOUT, 1, W + OFF_W4, H, 1, nullptr, l->d_dZ, H, nullptr, 0, M, H, NOUT, 2, false,
       l->d_Z3);
  /* layer 3 (dZ3 in d_dZ; dZ2 to d_dH) */
  gemm(c, l, true, l->d_dZ, 1, H, l->d_H2, H, 1, nullptr, G + OFF_W3, H, nullptr, 0, H, H, M, 0, true);
  colsum(c, l, l->d_dZ, G + OFF_B3, M, H);
  gemm(c, l, true, l->d_dZ, H, 1, W + OFF_W3, H, 1, nullptr, l->d_dH, H, nullptr, 0, M, H, H, 2, false, l->d_Z2);
  /* layer 2 (dZ2 in d_dH; dZ1 to d_dZ) */
  gemm(c, l, true, l->d_dH, 1, H, l->d_H1, H, 1, nullptr, G + OFF_W2, H, nullptr, 0, H, H, M, 0, true);
  colsum(c, l, l->d_dH, G + OFF_B2, M, H);
  gemm(c, l, true, l->d_dH, H, 1, W + OFF_W2, H, 1, nullptr, l->d_dZ, H, nullptr, 0, M, H, H, 2, false, l->d_Z1);
  /* layer 1 (dZ1 in d_dZ) */
  gemm(c, l, true, l->d_dZ, 1, H, l->d_X0b, NIN, 1, nullptr, G + OFF_W1, NIN, nullptr, 0, H, NIN, M, 0, true);
  colsum(c, l, l->d_dZ, G + OFF_B1, M, H);
}

void build_tables(ftr_ctx* c, ftr_learner* l) {
  int E = l->E;
  double dmax = DMAX, smax = SMAX;
  void* args[] = {&l->d_Y,    &E,          &l->p->d_w, &l->p->d_jr, &l->d_r,    &l->d_lw,          &l->d_lZ,
                  &l->d_lratio, &l->d_lia, &l->d_lrad, &l->d_ldiff, &dmax, &smax, &l->d_lst, &l->p->d_mml,
                  &l->d_led};
  launch(l->mlp.tables, (l->NT + 255) / 256, 1, 256, c->stream, args);
}

void ensure_tables(ftr_ctx* c, ftr_learner* l) {
  if (l->tables_valid) return;
  forward(c, l);
  build_tables(c, l);
  l->tables_valid = true;
}

ftr::LearnLaunch learn_launch(ftr_learner* l, const Geometry& G, uint64_t N) {
  ftr_problem* p = l->p;
  ftr::LearnLaunch L;
  memset(&L, 0, sizeof L);
  fill_tables(L.a, p);
  L.a.block_state = G.d_states;
  L.a.lane_poly = G.d_polys;
  L.a.N = N;
  L.a.S = G.S;
  L.a.partials = G.d_partials;
  L.lst = l->d_lst;
  L.led = l->d_led;
  L.lw = l->d_lw;
  L.lZ = l->d_lZ;
  L.lratio = l->d_lratio;
  L.lia = l->d_lia;
  L.lrad = l->d_lrad;
  L.ldiff = l->d_ldiff;
  L.gC = l->d_gC;
  L.gS = l->d_gS;
  L.gsum = l->d_gsum;
  L.gmax = l->d_gmax;
  L.gbuf = l->d_gbuf;
  L.gpath = l->d_gpath;
  L.objective = l->objective;
  return L;
}

/* moments of the block partials (host, block order) */
void reduce_partials(ftr_ctx* c, const Geometry& G, int K, int path, ftr_moments* out) {
  const int NQ = 2 + 4 * K;
  c->h_partials.resize((size_t)G.nblocks * NQ);
  HIPCHK(hipMemcpyDtoHAsync(c->h_partials.data(), G.d_partials, c->h_partials.size() * 8, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<double> tot(NQ, 0.0);
  for (uint64_t b = 0; b < G.nblocks; b++)
    for (int q = 0; q < NQ; q++) tot[q] += c->h_partials[b * NQ + q];
  out->K = K;
  out->dropped = (uint64_t)tot[1];
  out->lanes = G.lanes;
  out->samples_per_lane = G.S;
  for (int k = 0; k < K; k++) {
    out->n[k][0] = out->n[k][1] = tot[0];
    out->s1[k][0] = tot[2 + 4 * k];
    out->s2[k][0] = tot[3 + 4 * k];
    out->s1[k][1] = tot[4 + 4 * k];
    out->s2[k][1] = tot[5 + 4 * k];
    if (path != 2) {
      out->n[k][1] = 2;
      out->s1[k][1] = 0;
      out->s2[k][1] = 0;
    }
  }
}

int fail_exc(ftr_ctx* c, std::exception& e) {
  const Error* fe = dynamic_cast<const Error*>(&e);
  return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
}

}  // namespace

namespace {

/* the training record (c, radial log-draws, cut order) for `samples` */
void ensure_records(ftr_learner* l, uint64_t samples) {
  if (samples <= l->rec_cap) return;
  if (l->d_gbuf) HIPCHK(hipFree(l->d_gbuf));
  if (l->d_gpath) HIPCHK(hipFree(l->d_gpath));
  l->d_gbuf = nullptr;
  l->d_gpath = nullptr;
  l->rec_cap = 0;
  HIPCHK(hipMalloc((void**)&l->d_gbuf, (size_t)samples * l->E * 8));
  HIPCHK(hipMalloc((void**)&l->d_gpath, (size_t)samples * 8));
  l->rec_cap = samples;
}

/* One training step up to dL/dY (d_dY): network forward and tables, the
 * proposal samples with their record, the fixed-point forward-KL statistics
 * (gC, gS, gsum) and the table gradient.  ev0/ev1 bracket the sampling and
 * the statistics. */
Geometry step_statistics(ftr_ctx* c, ftr_learner* l, uint64_t samples, uint64_t seed) {
  ftr_problem* p = l->p;
  const int E = l->E, NQ = 2 + 4 * p->prep.K;
  /* = ftr::learn_acc_lds: the sums and the C and S statistics of the three
   * top levels, E + E^2 + E (E-1)/2 E entries each */
  const size_t acc_lds = (size_t)(2 + 2 * (E + E * E + E * (E - 1) / 2 * E)) * 8;
  forward(c, l);
  build_tables(c, l);
  /* fixed-point statistics: the scale's maximum and (E > HIST_E) the one
   * accumulator start at zero (zero bits) */
  fill_d(c, l, reinterpret_cast<double*>(l->d_gmax), 1, 0.0);
  if (E > HIST_E) fill_d(c, l, reinterpret_cast<double*>(l->d_part), hist_len(E), 0.0);
  Geometry G = setup_stream(p, 0, samples, seed, 0, l->var.block, NQ, l->f_lgr);
  ftr::LearnLaunch L = learn_launch(l, G, samples);
  void* args[] = {&L};
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  HIPCHK(hipModuleLaunchKernel(l->f_lgr, (unsigned)G.nblocks, 1, 1, l->var.block, 1, 1, (unsigned)l->lds,
                               c->stream, args, nullptr));
  uint64_t n = samples;
  void* margs[] = {&L, &n};
  const unsigned mblocks = (unsigned)std::min<uint64_t>((samples + 1023) / 1024, c->prop.multiProcessorCount);
  HIPCHK(hipModuleLaunchKernel(l->f_hmx, mblocks, 1, 1, 1024, 1, 1, 0, c->stream, margs, nullptr));
  void* hargs[] = {&L, &n, &l->d_part};
  if (E <= HIST_E)
    HIPCHK(hipModuleLaunchKernel(l->f_hst, (unsigned)l->hist_blocks, 1, 1, 1024, 1, 1,
                                 (unsigned)(hist_len(E) * 8), c->stream, hargs, nullptr));
  else
    HIPCHK(hipModuleLaunchKernel(l->f_acc, (unsigned)((samples + 255) / 256), 1, 1, 256, 1, 1, (unsigned)acc_lds,
                                 c->stream, hargs, nullptr));
  int nb = l->hist_blocks;
  void* rargs[] = {&L, &l->d_part, &nb, &n};
  HIPCHK(hipModuleLaunchKernel(l->f_hrd, (unsigned)((hist_len(E) + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream,
                               rargs, nullptr));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  fill_f(c, l, l->d_dY, (uint64_t)l->M * NOUT, 0.0f);
  int e = E;
  double dmax = DMAX, smax = SMAX;
  void* targs[] = {&l->d_Y, &e, &l->d_lw, &l->d_lZ, &l->d_gC, &l->d_gS, &l->d_gsum, &l->d_dY, &dmax, &smax};
  launch(l->mlp.tgrad, (l->NT + 255) / 256, 1, 256, c->stream, targs);
  return G;
}

}  // namespace

extern "C" {

int ftr_learner_create(ftr_ctx* c, ftr_problem* p, uint64_t init_seed, ftr_learner** out) {
  if (!out) return FTR_ERR_INVALID;
  *out = nullptr;
  if (!c || !p || !p->m.f_int) return fail(c, FTR_ERR_INVALID, "learner needs a problem prepared on a device");
  ftr_learner* l = new ftr_learner();
  try {
    HIPCHK(hipSetDevice(c->device));
    const int E = p->prep.g.E;
    if (E > MAX_E) throw Error(FTR_ERR_LIMIT, "the learned proposal supports at most 16 edges");
    l->p = p;
    l->device = c->device;
    l->E = E;
    l->NT = 1u << E;
    l->M = std::max<int>(32, (int)l->NT);
    l->mlp = load_mlp(c);
    l->var = p->var;
    l->var.block = LEARN_BLOCK;
    l->var.lps = 1; /* the learned kernels evaluate one sample per lane */
    /* the learned kernels are specialised to the problem's kinematic point
     * (its parameters compile-time constants, as ftr_set_specialized does
     * for the plain kernel): a learner belongs to one point anyway */
    l->var.cparams = 1;
    l->lds = variant_lds(l->var, E);
    const hipModule_t gm = load_module_raw(c, full_source(p->prep, l->var, true));
    HIPCHK(hipModuleGetFunction(&l->f_lev, gm, "ftrl_lev"));
    HIPCHK(hipModuleGetFunction(&l->f_lgr, gm, "ftrl_lgr"));
    HIPCHK(hipModuleGetFunction(&l->f_lsm, gm, "ftrl_lsm"));
    HIPCHK(hipModuleGetFunction(&l->f_hmx, gm, "ftrl_hmx"));
    HIPCHK(hipModuleGetFunction(&l->f_acc, gm, "ftrl_acc"));
    HIPCHK(hipModuleGetFunction(&l->f_hst, gm, "ftrl_hst"));
    HIPCHK(hipModuleGetFunction(&l->f_hrd, gm, "ftrl_hrd"));
    if (E <= HIST_E) {
      /* one 1024-thread block per CU, each with the whole histogram in LDS */
      const size_t hb = hist_len(E) * 8;
      if (hb > 64 * 1024)
        HIPCHK(hipFuncSetAttribute((const void*)l->f_hst, hipFuncAttributeMaxDynamicSharedMemorySize, (int)hb));
      l->hist_blocks = c->prop.multiProcessorCount;
      l->d_part = dalloc<ftr::fixacc>(l, (size_t)l->hist_blocks * hist_len(E));
    } else {
      l->hist_blocks = 1;
      l->d_part = dalloc<ftr::fixacc>(l, hist_len(E));
    }
    l->d_gmax = dalloc<unsigned long long>(l, 1);
    l->d_ws = dalloc<float>(l, WS_FLOATS);
    if (l->lds > 64 * 1024)
      for (hipFunction_t f : {l->f_lev, l->f_lgr, l->f_lsm})
        (void)hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l->lds);
    const size_t M = l->M, NT = l->NT;
    l->d_par = dalloc<float>(l, NPAR);
    l->d_grad = dalloc<float>(l, NPAR);
    l->d_m = dalloc<float>(l, NPAR);
    l->d_v = dalloc<float>(l, NPAR);
    l->d_parb = dalloc<uint16_t>(l, NPAR);
    l->d_X0b = dalloc<uint16_t>(l, M * NIN);
    l->d_H1 = dalloc<uint16_t>(l, M * H);
    l->d_H2 = dalloc<uint16_t>(l, M * H);
    l->d_H3 = dalloc<uint16_t>(l, M * H);
    l->d_Z1 = dalloc<float>(l, M * H);
    l->d_Z2 = dalloc<float>(l, M * H);
    l->d_Z3 = dalloc<float>(l, M * H);
    l->d_dH = dalloc<float>(l, M * H);
    l->d_dZ = dalloc<float>(l, M * H);
    l->d_Y = dalloc<float>(l, M * NOUT);
    l->d_dY = dalloc<float>(l, M * NOUT);
    l->d_lw = dalloc<double>(l, NT * E);
    l->d_lratio = dalloc<double>(l, NT * E);
    l->d_gC = dalloc<double>(l, NT * E);
    l->d_lZ = dalloc<double>(l, NT);
    l->d_lia = dalloc<double>(l, NT * E);
    l->d_lrad = dalloc<double>(l, NT * E);
    l->d_ldiff = dalloc<double>(l, NT * E);
    l->d_r = dalloc<double>(l, NT);
    l->d_gS = dalloc<double>(l, NT * E);
    l->d_gsum = dalloc<double>(l, 2);
    l->d_lst = dalloc<ftr::LQuad>(l, NT);
    l->d_led = dalloc<ftr::LEdge>(l, NT * E);
    /* initial parameters: hidden layers uniform(+-sqrt(3/fan_in)), zero
     * biases, zero output layer (q = p_trop) */
    std::vector<float> par(NPAR, 0.0f);
    uint64_t s = init_seed;
    auto init = [&](size_t off, int rows, int cols) {
      const double a = std::sqrt(3.0 / cols);
      for (size_t i = 0; i < (size_t)rows * cols; i++)
        par[off + i] = (float)(a * (2.0 * ((splitmix(s) >> 11) * 0x1.0p-53) - 1.0));
    };
    init(OFF_W1, H, NIN);
    init(OFF_W2, H, H);
    init(OFF_W3, H, H);
    /* inputs per state g: its edge bits, popcount/E, 1, r(g)/E, then the
     * kinematics: the scalar products among the vertices with external
     * momentum (upper triangle, row-major, at most 13), scaled by
     * max(1, their largest magnitude) */
    std::vector<double> kin;
    const std::vector<int>& ext = p->prep.ext;
    const int V = p->prep.g.V;
    for (size_t a = 0; a < ext.size(); a++)
      for (size_t b = a; b < ext.size(); b++)
        if (kin.size() < (size_t)(NIN - MAX_E - 3)) kin.push_back(p->prep.P[(size_t)ext[a] * V + ext[b]]);
    double ks = 1.0;
    for (double v : kin) ks = std::max(ks, std::fabs(v));
    std::vector<float> X0((size_t)M * NIN, 0.0f);
    for (uint32_t g = 0; g < l->NT; g++) {
      float* x = &X0[(size_t)g * NIN];
      for (int e = 0; e < E; e++) x[e] = (float)((g >> e) & 1u);
      x[MAX_E] = (float)__builtin_popcount(g) / (float)E;
      x[MAX_E + 1] = 1.0f;
      x[MAX_E + 2] = (float)(p->prep.r[g] / E);
      for (size_t k = 0; k < kin.size(); k++) x[MAX_E + 3 + k] = (float)(kin[k] / ks);
    }
    std::vector<float> zeros(NPAR, 0.0f);
    float* d_x0f = nullptr;
    HIPCHK(hipMalloc((void**)&d_x0f, X0.size() * 4));
    HIPCHK(hipMemcpyHtoDAsync(d_x0f, X0.data(), X0.size() * 4, c->stream));
    {
      uint64_t n = X0.size();
      void* args[] = {&d_x0f, &l->d_X0b, &n};
      launch(l->mlp.to_bf16, (unsigned)((n + 255) / 256), 1, 256, c->stream, args);
    }
    HIPCHK(hipMemcpyHtoDAsync(l->d_par, par.data(), NPAR * 4, c->stream));
    HIPCHK(hipMemcpyHtoDAsync(l->d_m, zeros.data(), NPAR * 4, c->stream));
    HIPCHK(hipMemcpyHtoDAsync(l->d_v, zeros.data(), NPAR * 4, c->stream));
    HIPCHK(hipMemcpyHtoDAsync(l->d_r, p->prep.r.data(), NT * 8, c->stream));
    refresh_bf16(c, l);
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(d_x0f);
  } catch (std::exception& e) {
    ftr_learner_free(l);
    return fail_exc(c, e);
  }
  *out = l;
  return FTR_OK;
}

void ftr_learner_free(ftr_learner* l) {
  if (!l) return;
  (void)hipSetDevice(l->device);
  for (void* p : l->allocs) (void)hipFree(p);
  if (l->d_gbuf) (void)hipFree(l->d_gbuf);
  if (l->d_gpath) (void)hipFree(l->d_gpath);
  delete l;
}

int ftr_learner_params(ftr_learner* l, float* get, const float* set) {
  if (!l) return FTR_ERR_INVALID;
  ftr_ctx* c = l->p->ctx;
  try {
    HIPCHK(hipSetDevice(l->device));
    if (get) {
      HIPCHK(hipMemcpyDtoHAsync(get, l->d_par, NPAR * 4, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (set) {
      HIPCHK(hipMemcpyHtoDAsync(l->d_par, set, NPAR * 4, c->stream));
      refresh_bf16(c, l);
      HIPCHK(hipStreamSynchronize(c->stream));
      l->tables_valid = false;
    }
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_forward(ftr_ctx* c, ftr_learner* l, float* Y) {
  if (!c || !l || !Y) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    forward(c, l);
    build_tables(c, l);
    l->tables_valid = true;
    HIPCHK(hipMemcpyDtoHAsync(Y, l->d_Y, (size_t)l->M * NOUT * 4, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_backward(ftr_ctx* c, ftr_learner* l, const float* dY, float* grad) {
  if (!c || !l || !dY || !grad) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    forward(c, l);
    HIPCHK(hipMemcpyHtoDAsync(l->d_dY, dY, (size_t)l->M * NOUT * 4, c->stream));
    backward(c, l);
    HIPCHK(hipMemcpyDtoHAsync(grad, l->d_grad, NPAR * 4, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_train(ftr_ctx* c, ftr_learner* l, int steps, uint64_t samples, uint64_t seed, double lr,
                      double* stats) {
  if (!c || !l || steps < 0 || samples == 0 || !(lr > 0)) return FTR_ERR_INVALID;
  ftr_problem* p = l->p;
  try {
    HIPCHK(hipSetDevice(c->device));
    const int K = p->prep.K;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ensure_records(l, samples);
    for (int st = 0; st < steps; st++) {
      Geometry G = step_statistics(c, l, samples, seed + (uint64_t)st);
      backward(c, l);
      l->adam_t++;
      const float c1 = 1.0f - std::pow(b1, (float)l->adam_t), c2 = 1.0f - std::pow(b2, (float)l->adam_t);
      {
        uint64_t n = NPAR;
        float flr = (float)lr;
        void* aargs[] = {&l->d_par, &l->d_parb, &l->d_grad, &l->d_m, &l->d_v, &n, &flr,
                         (void*)&b1, (void*)&b2, (void*)&eps, (void*)&c1, (void*)&c2};
        launch(l->mlp.adam, (unsigned)((n + 255) / 256), 1, 256, c->stream, aargs);
      }
      if (stats) {
        ftr_moments m;
        memset(&m, 0, sizeof m);
        reduce_partials(c, G, K, p->prep.path, &m);
        const double n = m.n[0][0], s1 = m.s1[0][0], s2 = m.s2[0][0];
        stats[2 * st] = (s1 != 0.0) ? n * s2 / (s1 * s1) : 0.0;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats[2 * st + 1] = ms;
      }
    }
    l->tables_valid = false;
    ensure_tables(c, l);
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_statistics(ftr_ctx* c, ftr_learner* l, uint64_t samples, uint64_t seed, double* cw, double* ly,
                           uint64_t* path, double* gC, double* gS, double* gsum, float* dY) {
  if (!c || !l || samples == 0) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    ensure_records(l, samples);
    step_statistics(c, l, samples, seed);
    const size_t NT = l->NT, E = l->E, N = samples;
    if (cw) HIPCHK(hipMemcpyDtoHAsync(cw, l->d_gbuf, N * 8, c->stream));
    if (ly && E > 1) HIPCHK(hipMemcpyDtoHAsync(ly, l->d_gbuf + N, (E - 1) * N * 8, c->stream));
    if (path) HIPCHK(hipMemcpyDtoHAsync(path, l->d_gpath, N * 8, c->stream));
    if (gC) HIPCHK(hipMemcpyDtoHAsync(gC, l->d_gC, NT * E * 8, c->stream));
    if (gS) HIPCHK(hipMemcpyDtoHAsync(gS, l->d_gS, NT * E * 8, c->stream));
    if (gsum) HIPCHK(hipMemcpyDtoHAsync(gsum, l->d_gsum, 2 * 8, c->stream));
    if (dY) HIPCHK(hipMemcpyDtoHAsync(dY, l->d_dY, (size_t)l->M * NOUT * 4, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_tables(ftr_ctx* c, ftr_learner* l, double* lw, double* lZ, double* lratio, double* lia,
                       double* lrad, double* ldiff) {
  if (!c || !l) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    ensure_tables(c, l);
    const size_t NT = l->NT, E = l->E;
    if (lw) HIPCHK(hipMemcpyDtoHAsync(lw, l->d_lw, NT * E * 8, c->stream));
    if (lratio) HIPCHK(hipMemcpyDtoHAsync(lratio, l->d_lratio, NT * E * 8, c->stream));
    if (lZ) HIPCHK(hipMemcpyDtoHAsync(lZ, l->d_lZ, NT * 8, c->stream));
    if (lia) HIPCHK(hipMemcpyDtoHAsync(lia, l->d_lia, NT * E * 8, c->stream));
    if (lrad) HIPCHK(hipMemcpyDtoHAsync(lrad, l->d_lrad, NT * E * 8, c->stream));
    if (ldiff) HIPCHK(hipMemcpyDtoHAsync(ldiff, l->d_ldiff, NT * E * 8, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_integrate(ftr_ctx* c, ftr_learner* l, uint64_t N, uint64_t seed, const ftr_opts* opts,
                          ftr_moments* out) {
  if (!c || !l || !out) return FTR_ERR_INVALID;
  ftr_problem* p = l->p;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(c->device));
    const int K = p->prep.K, NQ = 2 + 4 * K;
    memset(out, 0, sizeof *out);
    out->K = K;
    if (l->use_plain) return ftr_integrate(c, p, N, seed, opts, out);
    if (opts && opts->block && opts->block != l->var.block)
      throw Error(FTR_ERR_INVALID, "block size differs from the compiled variant");
    if (opts && opts->table_mode != -1 && opts->table_mode != l->var.tmode)
      throw Error(FTR_ERR_INVALID, "table_mode " + std::to_string(opts->table_mode) +
                                       " differs from the compiled placement " + std::to_string(l->var.tmode));
    ensure_tables(c, l);
    if (N > 0) {
      Geometry G = setup_stream(p, 0, N, seed, opts ? opts->lanes : 0, l->var.block, NQ, l->f_lev);
      ftr::LearnLaunch L = learn_launch(l, G, N);
      void* args[] = {&L};
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      HIPCHK(hipModuleLaunchKernel(l->f_lev, (unsigned)G.nblocks, 1, 1, l->var.block, 1, 1, (unsigned)l->lds,
                                   c->stream, args, nullptr));
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      reduce_partials(c, G, K, p->prep.path, out);
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      out->kernel_ms = ms;
    } else if (p->prep.path != 2) {
      for (int k = 0; k < K; k++) out->n[k][1] = 2;
    }
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_learner_set_objective(ftr_learner* l, int objective) {
  if (!l || objective < 0 || objective > 1) return FTR_ERR_INVALID;
  l->objective = objective;
  return FTR_OK;
}

int ftr_learner_select(ftr_ctx* c, ftr_learner* l, uint64_t n_probe, uint64_t seed, int* use_learned,
                       double* eff_plain, double* eff_learned) {
  if (!c || !l || n_probe < 2) return FTR_ERR_INVALID;
  ftr_problem* p = l->p;
  /* efficiency = relative second moment of the eps^0 coefficient x kernel
   * time per sample (lower is better) */
  auto eff = [&](const ftr_moments& m) {
    double s1 = m.s1[0][0], s2 = m.s2[0][0];
    double n = m.n[0][0];
    double v = s2 / n - (s1 / n) * (s1 / n);
    if (p->prep.path == 2) {
      const double t1 = m.s1[0][1], t2 = m.s2[0][1];
      v += t2 / n - (t1 / n) * (t1 / n);
      s1 = std::hypot(s1, t1);
    }
    const double mean2 = (s1 / n) * (s1 / n);
    return mean2 > 0 ? v / mean2 * (m.kernel_ms / (double)n_probe) : HUGE_VAL;
  };
  ftr_moments mp, ml;
  l->use_plain = false;
  int rc = ftr_integrate(c, p, n_probe, seed, nullptr, &mp);
  if (rc) return rc;
  rc = ftr_learner_integrate(c, l, n_probe, seed + 1, nullptr, &ml);
  if (rc) return rc;
  const double ep = eff(mp), el = eff(ml);
  l->use_plain = !(el < ep);
  if (use_learned) *use_learned = l->use_plain ? 0 : 1;
  if (eff_plain) *eff_plain = ep;
  if (eff_learned) *eff_learned = el;
  return FTR_OK;
}

int ftr_learner_sample(ftr_ctx* c, ftr_learner* l, uint64_t first, uint64_t n, uint64_t seed, double* X_dev,
                       double* w_dev, uint64_t* path_dev) {
  if (!c || !l || (n && (!X_dev || !w_dev || !path_dev))) return FTR_ERR_INVALID;
  ftr_problem* p = l->p;
  try {
    HIPCHK(hipSetDevice(c->device));
    ensure_tables(c, l);
    if (n == 0) return FTR_OK;
    Geometry G = setup_stream(p, first, n, seed, 0, l->var.block, 2 + 4 * p->prep.K, l->f_lsm);
    ftr::LearnLaunch L = learn_launch(l, G, n);
    void* args[] = {&L, &X_dev, &w_dev, &path_dev};
    HIPCHK(hipModuleLaunchKernel(l->f_lsm, (unsigned)G.nblocks, 1, 1, l->var.block, 1, 1, (unsigned)l->lds,
                                 c->stream, args, nullptr));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    return fail_exc(c, e);
  }
  return FTR_OK;
}

int ftr_precompile_learned(const ftr_input* in, char* err, int errlen) {
  try {
    ftrh::Prep pr;
    prep_from_input(pr, in);
    if (pr.g.E > MAX_E) throw Error(FTR_ERR_LIMIT, "the learned proposal supports at most 16 edges");
    code_object(ftr_embedded_mlp_source, nullptr);
    for (size_t lim : {(size_t)64 * 1024, (size_t)160 * 1024}) {
      Variant v = choose_variant(pr, lim);
      v.block = LEARN_BLOCK;
      v.cparams = 1;
      code_object(full_source(pr, v, true), nullptr);
    }
  } catch (std::exception& e) {
    if (err && errlen > 0) {
      strncpy(err, e.what(), errlen - 1);
      err[errlen - 1] = 0;
    }
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fe ? fe->code : FTR_ERR_INVALID;
  }
  return FTR_OK;
}

}  // extern "C"
