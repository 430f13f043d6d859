/*
 * cli.cpp -- the feyntrop low-level JSON contract on top of the C ABI.
 *
 * Input and checks follow srcs/feyntrop.cpp:22-119; logging follows
 * :165-257 (stderr); output is one line
 *   {"IGtr":..,"integral":[[[re,dre],[im,dim]],..],"seconds preprocessing":..,
 *    "seconds sampling":..}
 * (:264-278), keys sorted, numbers formatted like the reference's writer.
 */
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

extern "C" int ftr_run_json(ftr_ctx* ctx, const char* json_in, char** json_out, char** log_out) {
  std::string log;
  if (json_out) *json_out = nullptr;
  if (log_out) *log_out = nullptr;
  ftr_problem* prob = nullptr;
  int rc = FTR_OK;
  try {
    const Json data = ftrh::json_parse(json_in ? json_in : "");
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
    const uint64_t N = data.get("N")->as_u64();
    const uint64_t seed = data.get("seed")->as_u64();
    const double lambda = data.get("lambda")->as_double();
    const double D = data.get("dimension")->as_double();
    const int K = (int)data.get("num_eps_terms")->as_i64();
    const Json& gr = *data.get("graph");
    const std::string gmsg = "input graph must be a list of the form [[[v1_1,v2_1],nu_1],[[v1_2,v2_2],nu_2],...]";
    if (!gr.is_array()) throw Error(FTR_ERR_INVALID, gmsg);
    std::vector<int32_t> edges;
    std::vector<double> nu;
    for (const Json& e : gr.a) {
      if (e.size() != 2 || !e.is_array() || e.a[0].size() != 2 || !e.a[0].is_array()) throw Error(FTR_ERR_INVALID, gmsg);
      if (!e.a[0].a[0].is_number() || !e.a[0].a[1].is_number() || !e.a[1].is_number()) throw Error(FTR_ERR_INVALID, gmsg);
      edges.push_back((int32_t)e.a[0].a[0].as_i64());
      edges.push_back((int32_t)e.a[0].a[1].as_i64());
      nu.push_back(e.a[1].as_double());
    }
    const int E = (int)nu.size();
    ftrh::Graph g;
    ftrh::build_graph(g, E, edges.data(), nu.data());
    const int V = g.V;
    const Json& sp = *data.get("scalarproducts");
    const std::string pmsg = "input scalarproducts must be a (V x V)-matrix";
    if (!sp.is_array() || (int)sp.size() != V) throw Error(FTR_ERR_INVALID, pmsg);
    std::vector<double> P((size_t)V * V);
    for (int k = 0; k < V; k++) {
      if (!sp.a[k].is_array() || (int)sp.a[k].size() != V) throw Error(FTR_ERR_INVALID, pmsg);
      for (int l = 0; l < V; l++) {
        if (!sp.a[k].a[l].is_number()) throw Error(FTR_ERR_INVALID, pmsg);
        P[(size_t)k * V + l] = sp.a[k].a[l].as_double();
      }
    }
    const Json& ms = *data.get("masses_sqr");
    const std::string mmsg = "input masses_sqr must be an array of size E";
    if (!ms.is_array() || (int)ms.size() != E) throw Error(FTR_ERR_INVALID, mmsg);
    std::vector<double> m2(E);
    for (int e = 0; e < E; e++) {
      if (!ms.a[e].is_number()) throw Error(FTR_ERR_INVALID, mmsg);
      m2[e] = ms.a[e].as_double();
    }
    ftr_input in;
    in.E = E;
    in.V = V;
    in.edges = edges.data();
    in.nu = nu.data();
    in.P = P.data();
    in.m2 = m2.data();
    in.D = D;
    in.K = K;
    in.lambda = lambda;
    rc = ftr_prepare(ctx, &in, &prob);
    if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
    ftr_info info;
    ftr_problem_info(prob, &info);
    /* regime / GP log lines were collected during preprocessing */
    for (const std::string& l : ftr_problem_log(prob)) log += l + "\n";
    log += "Started integrating on 1 GPU and N = " + g_str((double)N) + " points.\n";
    ftr_moments m;
    if (!ctx) throw Error(FTR_ERR_DEVICE, "feyntrop: no usable HIP device");
    rc = ftr_integrate(ctx, prob, N, seed, nullptr, &m);
    if (rc != FTR_OK) throw Error(rc, ftr_last_error(ctx));
    std::vector<double> res(4 * K);
    ftr_finalize(prob, &m, res.data());
    const double tpre = info.seconds_preprocessing, tsam = m.seconds;
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
    o += "],\"seconds preprocessing\":" + ftrh::json_number(tpre) +
         ",\"seconds sampling\":" + ftrh::json_number(tsam) + "}";
    if (json_out) *json_out = dupstr(o);
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
  if (log_out) *log_out = dupstr(log);
  return rc;
}
