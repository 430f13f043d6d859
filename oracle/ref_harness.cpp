/*
 * ref_harness.cpp -- exposes the compiled reference (michibo/feyntrop srcs/,
 * included from where it lies under the reference checkout, never copied)
 * through a small C API so the oracle can be checked against it.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ with
 * the reference's own compiler flags (Makefile:11-17 of the reference); it is
 * only loaded by tests that run in the build container.
 *
 * Entry points mirror ftr_oracle.h:
 *   ref_create   -- feyntrop.cpp:64-239 (graph, validation, regime via
 *                   Eigen::SelfAdjointEigenSolver, generate_subgraph_table,
 *                   GP checks, lambda_eff)
 *   ref_sample   -- get_random_psi_xi_tropical_sample (tropical_sampling.hpp:503)
 *   ref_values   -- per-sample integrand: feynman_integral_estimate with N=1
 *                   on one thread, the generator positioned at the sample's
 *                   stream offset (an N=1 estimate's average is the sample)
 *   ref_estimate -- feynman_integral_estimate (feynman_integral.hpp:30)
 *   ref_xoshiro  -- raw xoshiro256+ outputs and the 2^128 / 2^192 jumps
 */
#include <omp.h>
#include <cstring>
#include <string>
#include <Eigen/Dense>
#include <Eigen/Eigenvalues>

#include "graph.hpp"
#include "tropical_sampling.hpp"
#include "feynman_integral.hpp"

struct ref_problem {
  graph g;
  int V, E, K;
  double D, lambda, W, IGtr, lambda_eff;
  bool bEuclidean, bPseudoEuclidean, bGeneric, bGP;
  Eigen::MatrixXd P;
  Eigen::VectorXd m2;
  J_vector table;
  ref_problem(const graph& gg) : g(gg) {}
};

static void set_state(true_random::xoshiro256& gen, const uint64_t* st) {
  static_assert(sizeof(true_random::xoshiro256) == 32, "layout");
  memcpy((void*)&gen, st, 32);
}
static void get_state(const true_random::xoshiro256& gen, uint64_t* st) { memcpy(st, (const void*)&gen, 32); }

extern "C" {

ref_problem* ref_create(int E, const int* edges, const double* nu, int V_in, const double* Pin,
                        const double* m2, double D, int K, double lambda, char* err, int errlen) {
  try {
    graph::edge_vector ev;
    for (int j = 0; j < E; j++) ev.push_back(std::make_pair(std::make_pair(edges[2 * j], edges[2 * j + 1]), nu[j]));
    graph g(ev);
    ref_problem* pr = new ref_problem(g);
    pr->V = g._V; pr->E = g._E; pr->K = K; pr->D = D; pr->lambda = lambda;
    if (V_in != g._V) { delete pr; throw std::invalid_argument("input scalarproducts must be a (V x V)-matrix"); }
    pr->P.resize(g._V, g._V);
    for (int i = 0; i < g._V; i++) for (int j = 0; j < g._V; j++) pr->P(i, j) = Pin[i * g._V + j];
    pr->m2.resize(E);
    for (int j = 0; j < E; j++) pr->m2[j] = m2[j];
    Eigen::MatrixXd& scalarproducts = pr->P;
    if (!scalarproducts.isApprox(scalarproducts.transpose())) { delete pr; throw std::domain_error("not symmetric"); }
    if (!scalarproducts.colwise().sum().isZero()) { delete pr; throw std::domain_error("momentum conservation"); }
    Eigen::SelfAdjointEigenSolver<Eigen::MatrixXd> es(scalarproducts);
    Eigen::MatrixXd pabs = -es.eigenvectors() * es.eigenvalues().cwiseAbs().asDiagonal() * es.eigenvectors().inverse();
    pr->bEuclidean = true;
    if (!scalarproducts.isZero() && !scalarproducts.isApprox(pabs)) pr->bEuclidean = false;
    try {
      std::tie(pr->table, pr->W, pr->IGtr, pr->bGP, pr->bPseudoEuclidean, pr->bGeneric) =
          generate_subgraph_table(pr->g, D, pr->P, pr->m2, K > 1, !pr->bEuclidean);
    } catch (...) { delete pr; throw; }
    bool bG = pr->bGeneric, bGP = pr->bGP;
    if ((bG && !bGP) || (pr->bEuclidean && !bGP) || (!(!bG && !pr->bEuclidean && pr->W > 0) && !bGP && pr->W > 0)) {
      delete pr; throw std::domain_error("generalized permutahedron property");
    }
    pr->lambda_eff = pr->bPseudoEuclidean ? 0.0 : lambda;
    return pr;
  } catch (std::exception& ex) {
    if (err && errlen > 0) { strncpy(err, ex.what(), errlen - 1); err[errlen - 1] = 0; }
    return nullptr;
  }
}

void ref_destroy(ref_problem* p) { delete p; }

void ref_info(const ref_problem* p, double* o) {
  int L = p->E - p->V + 1;
  o[0] = p->E; o[1] = p->V; o[2] = L; o[3] = p->W; o[4] = p->IGtr;
  o[5] = p->bEuclidean; o[6] = p->bPseudoEuclidean; o[7] = p->bGeneric; o[8] = p->bGP;
  o[9] = p->lambda_eff; o[10] = 0; o[11] = 0; o[12] = p->K;
}

void ref_jtable(const ref_problem* p, double* Jr, double* r, uint8_t* mm, uint8_t* L) {
  for (size_t i = 0; i < p->table.size(); i++) {
    Jr[i] = std::get<0>(p->table[i]); r[i] = std::get<1>(p->table[i]);
    mm[i] = (uint8_t)std::get<2>(p->table[i]); L[i] = (uint8_t)std::get<3>(p->table[i]);
  }
}

void ref_sample(const ref_problem* p, uint64_t* st, uint64_t n, double* X, double* psixi) {
  true_random::xoshiro256 gen(0);
  set_state(gen, st);
  Eigen::VectorXd x(p->E);
  for (uint64_t i = 0; i < n; i++) {
    auto pr = get_random_psi_xi_tropical_sample(x, p->g, p->table, gen);
    for (int j = 0; j < p->E; j++) X[i * p->E + j] = x[j];
    psixi[2 * i] = pr.first; psixi[2 * i + 1] = pr.second;
  }
  get_state(gen, st);
}

void ref_values(const ref_problem* p, uint64_t* st, uint64_t n, double* out, uint8_t* dropped) {
  omp_set_num_threads(1);
  const int K = p->K, E = p->E;
  std::stringstream devnull;
  std::streambuf* old = std::clog.rdbuf(devnull.rdbuf());
  for (uint64_t i = 0; i < n; i++) {
    true_random::xoshiro256 gen(0);
    set_state(gen, st);
    auto res = feynman_integral_estimate(1, p->g, p->D, p->P, p->m2, K, p->lambda_eff, p->table, gen);
    bool drop = std::get<0>(res[0]).N == 0;
    for (int k = 0; k < K; k++) {
      out[(i * K + k) * 2] = drop ? 0.0 : std::get<0>(res[k]).S1.value();
      out[(i * K + k) * 2 + 1] = drop ? 0.0 : std::get<1>(res[k]).S1.value();
    }
    if (dropped) dropped[i] = drop;
    true_random::xoshiro256 adv(0);
    set_state(adv, st);
    for (int q = 0; q < 2 * E - 1; q++) adv.next();
    get_state(adv, st);
  }
  std::clog.rdbuf(old);
}

void ref_estimate(const ref_problem* p, uint64_t N, uint64_t seed, int threads, double* moments) {
  omp_set_num_threads(threads);
  true_random::xoshiro256 gen(seed);
  std::stringstream devnull;
  std::streambuf* old = std::clog.rdbuf(devnull.rdbuf());
  auto res = feynman_integral_estimate(N, p->g, p->D, p->P, p->m2, p->K, p->lambda_eff, p->table, gen);
  std::clog.rdbuf(old);
  for (int k = 0; k < p->K; k++) {
    double* m = moments + 6 * k;
    m[0] = (double)std::get<0>(res[k]).N; m[1] = std::get<0>(res[k]).S1.value(); m[2] = std::get<0>(res[k]).S2.value();
    m[3] = (double)std::get<1>(res[k]).N; m[4] = std::get<1>(res[k]).S1.value(); m[5] = std::get<1>(res[k]).S2.value();
  }
}

/* raw generator outputs and the fixed jumps (random.hpp:86-199) */
void ref_xoshiro(uint64_t seed, uint64_t n, uint64_t* out, uint64_t* state_after_jump,
                 uint64_t* state_after_long_jump) {
  true_random::xoshiro256 g(seed);
  for (uint64_t i = 0; i < n; i++) out[i] = g.next();
  true_random::xoshiro256 j(seed);
  j.jump();
  get_state(j, state_after_jump);
  true_random::xoshiro256 l(seed);
  l.long_jump();
  get_state(l, state_after_long_jump);
}

} /* extern "C" */
