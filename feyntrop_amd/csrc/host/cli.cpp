/*
 * cli.cpp -- the feyntrop low-level JSON contract on top of the C ABI.
 *
 * Input and checks follow srcs/feyntrop.cpp:22-119; logging follows
 * :165-257 (stderr); output is one line
 *   {"IGtr":..,"integral":[[[re,dre],[im,dim]],..],"seconds preprocessing":..,
 *    "seconds sampling":..}
 * (:264-278), keys sorted, numbers formatted like the reference's writer.
 * Additions: {"batch": [problem, ...]} integrates many problems in one
 * batched launch and answers {"batch": [result, ...]}; "proposal": "learned"
 * in a problem integrates with the learned proposal (E <= 16).
 */
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/feyntrop_amd.h"
#include "ftr_host.hpp"

using ftrh::Error;
using ftrh::Json;

const std::vector<std::string>& ftr_problem_log(const ftr_problem* p);

static char* dupstr(const std::string& s) {
  char* p = (char*)malloc(s.size() + 1);
  memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

static std::string g_str(double v) {
  std::ostringstream s;
  s << v;
  return s.str();
}

/* One problem of the low-level contract (feyntrop.cpp:22-112). */
struct CliProblem {
  uint64_t N = 0, seed = 0;
  double lambda = 0, D = 0;
  int K = 0, E = 0, V = 0;
  std::vector<int32_t> edges;
  std::vector<double> nu, P, m2;
  ftr_input in() const {
    ftr_input r;
    r.E = E;
    r.V = V;
    r.edges = edges.data();
    r.nu = nu.data();
    r.P = P.data();
    r.m2 = m2.data();
    r.D = D;
    r.K = K;
    r.lambda = lambda;
    return r;
  }
};

static CliProblem parse_problem(const Json& data) {
  CliProblem c;
  static const char* req[] = {"N", "graph", "dimension", "scalarproducts", "masses_sqr", "num_eps_terms", "lambda", "seed"};
  for (const char* k : req)
    if (!data.get(k))
      throw Error(FTR_ERR_INVALID, "input json must contain the data fields: N, graph, dimension, scalarproducts, "
                                   "masses_sqr, num_eps_terms, lambda, seed");
  if (!data.get("N")->is_number()) throw Error(FTR_ERR_INVALID, "input N must be a number");
  if (!data.get("lambda")->is_number()) throw Error(FTR_ERR_INVALID, "input lambda must be a number");
  if (!data.get("dimension")->is_number()) throw Error(FTR_ERR_INVALID, "input dimension must be a number");
  if (!data.get("num_eps_terms")->is_number()) throw Error(FTR_ERR_INVALID, "input num_eps_terms must be a number");
  if (!data.get("seed")->is_number()) throw Error(FTR_ERR_INVALID, "input seed must be a number");
  c.N = data.get("N")->as_u64();
  c.seed = data.get("seed")->as_u64();
  c.lambda = data.get("lambda")->as_double();
  c.D = data.get("dimension")->as_double();
  c.K = (int)data.get("num_eps_terms")->as_i64();
  const Json& gr = *data.get("graph");
  const std::string gmsg = "input graph must be a list of the form [[[v1_1,v2_1],nu_1],[[v1_2,v2_2],nu_2],...]";
  if (!gr.is_array()) throw Error(FTR_ERR_INVALID, gmsg);
  for (const Json& e : gr.a) {
    if (e.size() != 2 || !e.is_array() || e.a[0].size() != 2 || !e.a[0].is_array()) throw Error(FTR_ERR_INVALID, gmsg);
    if (!e.a[0].a[0].is_number() || !e.a[0].a[1].is_number() || !e.a[1].is_number()) throw Error(FTR_ERR_INVALID, gmsg);
    c.edges.push_back((int32_t)e.a[0].a[0].as_i64());
    c.edges.push_back((int32_t)e.a[0].a[1].as_i64());
    c.nu.push_back(e.a[1].as_double());
  }
  c.E = (int)c.nu.size();
  ftrh::Graph g;
  ftrh::build_graph(g, c.E, c.edges.data(), c.nu.data());
  c.V = g.V;
  const Json& sp = *data.get("scalarproducts");
  const std::string pmsg = "input scalarproducts must be a (V x V)-matrix";
  if (!sp.is_array() || (int)sp.size() != c.V) throw Error(FTR_ERR_INVALID, pmsg);
  c.P.assign((size_t)c.V * c.V, 0.0);
  for (int k = 0; k < c.V; k++) {
    if (!sp.a[k].is_array() || (int)sp.a[k].size() != c.V) throw Error(FTR_ERR_INVALID, pmsg);
    for (int l = 0; l < c.V; l++) {
      if (!sp.a[k].a[l].is_number()) throw Error(FTR_ERR_INVALID, pmsg);
      c.P[(size_t)k * c.V + l] = sp.a[k].a[l].as_double();
    }
  }
  const Json& ms = *data.get("masses_sqr");
  const std::string mmsg = "input masses_sqr must be an array of size E";
  if (!ms.is_array() || (int)ms.size() != c.E) throw Error(FTR_ERR_INVALID, mmsg);
  c.m2.resize(c.E);
  for (int e = 0; e < c.E; e++) {
    if (!ms.a[e].is_number()) throw Error(FTR_ERR_INVALID, mmsg);
    c.m2[e] = ms.a[e].as_double();
  }
  return c;
}

/* The reference's output line (feyntrop.cpp:264-278) and closing log lines
 * (:253-257). */
static std::string format_result(ftr_problem* prob, const ftr_moments& m, double tsam, std::string& log) {
  ftr_info info;
  ftr_problem_info(prob, &info);
  const int K = info.K;
  std::vector<double> res(4 * K);
  ftr_finalize(prob, &m, res.data());
  const double tpre = info.seconds_preprocessing;
  log += "Finished in " + g_str(tpre + tsam) + " seconds = " + g_str((tpre + tsam) / 3600) + " hours.\n";
  if (m.dropped) {
    log += "Warning: " + std::to_string(m.dropped) +
           " sampled values were not finite - floating point accuracy or numerical stabilty seem to be "
           "insufficient - dropping these points. If this happens often, the result will not be reliable.\n";
  }
  std::string o = "{\"IGtr\":" + ftrh::json_number(info.IGtr) + ",\"integral\":[";
  for (int k = 0; k < K; k++) {
    if (k) o += ",";
    o += "[[" + ftrh::json_number(res[4 * k]) + "," + ftrh::json_number(res[4 * k + 1]) + "],[" +
         ftrh::json_number(res[4 * k + 2]) + "," + ftrh::json_number(res[4 * k + 3]) + "]]";
  }
  o += "],\"seconds preprocessing\":" + ftrh::json_number(tpre) + ",\"seconds sampling\":" + ftrh::json_number(tsam) +
       "}";
  return o;
}

/* Batch mode (an addition; a single problem keeps the reference contract):
 * {"batch": [problem, problem, ...]} -> {"batch": [result, result, ...]},
 * every result exactly the single-problem output for that problem.  All
 * problems are integrated by one ftr_integrate_batch (one hipGraph launch);
 * an invalid item fails the whole call with "batch item i: message". */
static std::string run_batch(ftr_ctx* ctx, const Json& items, std::string& log,
                             std::vector<ftr_problem*>& probs) {
  if (!items.is_array() || items.a.empty()) throw Error(FTR_ERR_INVALID, "input batch must be a non-empty list of problems");
  const size_t n = items.a.size();
  std::vector<uint64_t> Ns(n), seeds(n);
  for (size_t i = 0; i < n; i++) {
    const std::string tag = "batch item " + std::to_string(i) + ": ";
    try {
      const CliProblem c = parse_problem(items.a[i]);
      const ftr_input in = c.in();
      ftr_problem* p = nullptr;
      const int rc = ftr_prepare(ctx, &in, &p);
      if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
      probs.push_back(p);
      Ns[i] = c.N;
      seeds[i] = c.seed;
    } catch (Error& e) {
      throw Error(e.code ? e.code : FTR_ERR_INVALID, tag + e.what());
    }
    for (const std::string& l : ftr_problem_log(probs.back())) log += "[" + std::to_string(i) + "] " + l + "\n";
  }
  if (!ctx) throw Error(FTR_ERR_DEVICE, "feyntrop: no usable HIP device");
  double tot = 0;
  for (uint64_t v : Ns) tot += (double)v;
  log += "Started integrating " + std::to_string(n) + " problems on 1 GPU and N = " + g_str(tot) + " points in total.\n";
  std::vector<ftr_moments> ms(n);
  const int rc = ftr_integrate_batch(ctx, (int)n, probs.data(), Ns.data(), seeds.data(), ms.data());
  if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
  std::string o = "{\"batch\":[";
  for (size_t i = 0; i < n; i++) {
    std::string l;
    if (i) o += ",";
    o += format_result(probs[i], ms[i], ms[i].seconds, l);
    log += "[" + std::to_string(i) + "] " + l;
  }
  return o + "]}";
}

extern "C" int ftr_run_json(ftr_ctx* ctx, const char* json_in, char** json_out, char** log_out) {
  std::string log;
  if (json_out) *json_out = nullptr;
  if (log_out) *log_out = nullptr;
  ftr_problem* prob = nullptr;
  std::vector<ftr_problem*> probs;
  int rc = FTR_OK;
  try {
    const Json data = ftrh::json_parse(json_in ? json_in : "");
    if (const Json* b = data.get("batch")) {
      const std::string o = run_batch(ctx, *b, log, probs);
      if (json_out) *json_out = dupstr(o);
    } else {
      const CliProblem c = parse_problem(data);
      const ftr_input in = c.in();
      rc = ftr_prepare(ctx, &in, &prob);
      if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
      /* the kernel specialised to this kinematic point when it is cached */
      rc = ftr_use_cached_specialization(ctx, prob);
      if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
      /* regime / GP log lines were collected during preprocessing */
      for (const std::string& l : ftr_problem_log(prob)) log += l + "\n";
      /* optional addition: "proposal": "learned" trains the learned proposal
       * first (the plain sampler judges it); absent or "plain" keeps the
       * reference's sampler */
      std::string proposal = "plain";
      if (const Json* pj = data.get("proposal")) {
        if (pj->t != Json::Str || (pj->s != "plain" && pj->s != "learned"))
          throw Error(FTR_ERR_INVALID, "input proposal must be \"plain\" or \"learned\"");
        proposal = pj->s;
      }
      log += "Started integrating on 1 GPU and N = " + g_str((double)c.N) + " points.\n";
      ftr_moments m;
      if (!ctx) throw Error(FTR_ERR_DEVICE, "feyntrop: no usable HIP device");
      double t_learn = 0.0;
      if (proposal == "learned") {
        const auto t0 = std::chrono::steady_clock::now();
        ftr_learner* l = nullptr;
        rc = ftr_learner_create(ctx, prob, 1, &l);
        if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
        struct Guard {
          ftr_learner* l;
          ~Guard() { ftr_learner_free(l); }
        } guard{l};
        rc = ftr_learner_train(ctx, l, 400, 1u << 18, c.seed + 100, 3e-3, nullptr);
        if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
        int use = 0;
        double ep = 0, el = 0;
        const uint64_t probe = std::min<uint64_t>(4000000, std::max<uint64_t>(c.N / 10, 100000));
        rc = ftr_learner_select(ctx, l, probe, c.seed + 7, &use, &ep, &el);
        if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
        t_learn = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        log += std::string("Learned proposal: ") + (use ? "used" : "not better than the plain sampler, not used") +
               " (relative variance x time per sample: plain " + g_str(ep) + ", learned " + g_str(el) +
               "; training " + g_str(t_learn) + " s).\n";
        rc = ftr_learner_integrate(ctx, l, c.N, c.seed, nullptr, &m);
        if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
      } else {
        rc = ftr_integrate(ctx, prob, c.N, c.seed, nullptr, &m);
        if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
      }
      const std::string o = format_result(prob, m, m.seconds + t_learn, log);
      if (json_out) *json_out = dupstr(o);
    }
    rc = FTR_OK;
  } catch (Error& e) {
    log += e.what();
    log += "\n";
    rc = e.code ? e.code : FTR_ERR_INVALID;
  } catch (std::exception& e) {
    log += e.what();
    log += "\n";
    rc = FTR_ERR_INVALID;
  }
  if (prob) ftr_problem_free(prob);
  for (ftr_problem* p : probs) ftr_problem_free(p);
  if (log_out) *log_out = dupstr(log);
  return rc;
}
