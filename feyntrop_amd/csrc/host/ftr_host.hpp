/* Internal host-side declarations of the feyntrop_amd runtime. */
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace ftrh {

/* Errors carry the C-ABI status code. */
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Graph {
  int E = 0, V = 0;
  std::vector<int> hi, lo;      /* endpoints, hi > lo (graph.hpp:181-185) */
  std::vector<int> raw;         /* edges as given, 2E */
  std::vector<double> nu;
  std::vector<std::vector<std::pair<int, int>>> adj;
};

/* Inputs of the J-table construction shared by the host and device
 * implementations (tropical_sampling.hpp:284-458). */
struct JtSetup {
  int n_tests = 0;
  const double* Xs = nullptr; /* n_tests x E random points in [1, 3) */
  const double* Xi = nullptr; /* their inverses */
  double pmsqr_ref = 0, W = 0, W_ref = 0;
  bool phi = false, noneucl = false;
};
/* Flags of the J-table pass; div_level >= 0 names the first subset size
 * with a vanishing effective degree of divergence and div_subset its first
 * subset in increasing numeric (Gosper) order. */
struct JtResult {
  bool generic = true, pseudo = true, gp = true;
  int div_level = -1;
  uint64_t div_subset = 0;
};
struct Prep;
void host_jtable(Prep& pr, const JtSetup& js, JtResult& res);

/* Host-side preprocessing results for one problem. */
struct Prep {
  Graph g;
  std::vector<double> P, m2;
  double D = 0, lambda = 0;
  int K = 1, L = 0;
  double W = 0, W_ref = 0, IGtr = 0;
  bool euclidean = true, pseudo = true, generic = true, gp = true;
  double lambda_eff = 0, def_ref = 0;
  bool W_zero = false;
  int path = 0;                 /* 0 real, 1 real psi-only, 2 complex */
  std::vector<int> ext;         /* vertices with a nonzero row of P */
  std::vector<double> Jr, r;    /* 2^E */
  std::vector<uint8_t> mm, Lt;  /* 2^E */
  std::vector<std::string> log; /* regime / GP log lines, in reference order */
  double seconds = 0;
  bool jtable_on_device = false; /* set when jtable_fn built the table */
  /* optional replacement of host_jtable (the device J-table) */
  std::function<void(Prep&, const JtSetup&, JtResult&)> jtable_fn;
};

void build_graph(Graph& g, int E, const int32_t* edges, const double* nu);
/* Full preprocessing: validation, regime, J-table (throws Error). */
void preprocess(Prep& pr);

/* xoshiro256+ jump-ahead over GF(2). */
struct Poly { uint64_t w[4]; };
Poly xo_jump_poly(uint64_t lo, uint64_t hi);   /* x^d mod charpoly */
Poly xo_poly_square(const Poly& a);
Poly xo_poly_mul(const Poly& a, const Poly& b); /* product mod charpoly: the composed jump */
void xo_apply(const Poly& q, uint64_t* st);
void xo_seed(uint64_t seed, uint64_t* st);
void xo_charpoly(uint64_t* w5);

/* Minimal JSON value. */
struct Json {
  enum Type { Null, Bool, Num, Str, Arr, Obj } t = Null;
  bool b = false;
  double num = 0;
  bool is_int = false, is_neg = false;
  uint64_t u = 0;               /* magnitude for integers */
  std::string s;
  std::vector<Json> a;
  std::vector<std::pair<std::string, Json>> o;
  const Json* get(const std::string& k) const;
  bool is_number() const { return t == Num; }
  bool is_array() const { return t == Arr; }
  size_t size() const { return t == Arr ? a.size() : (t == Obj ? o.size() : (t == Null ? 0 : 1)); }
  double as_double() const { return num; }
  uint64_t as_u64() const;
  int64_t as_i64() const;
};
Json json_parse(const std::string& text);
std::string json_number(double v);   /* nlohmann-compatible formatting */

}  // namespace ftrh
