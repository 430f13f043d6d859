/*
 * runtime.cpp -- C ABI of the MI355X engine: device context, per-graph kernel
 * specialisation (hipRTC, cached code objects), launches and moments.
 *
 * Hot path: one kernel per problem (ftrk_int).  Each lane owns a contiguous
 * run of S samples of the single xoshiro256+ stream; the host supplies the
 * stream state at every block's first sample (chained jump polynomials) and
 * the in-block jump polynomials, so the samples are exactly those of the
 * reference run with one OpenMP thread (feynman_integral.hpp:85-93, :148-164
 * with max_threads = 1).  Blocks write partial moments; the host sums them in
 * block order (deterministic).
 */
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "ftr_runtime.hpp"

thread_local std::string g_host_err; /* errors of ctx-less calls */
static std::string g_cache_dir;
static std::mutex g_cache_mu;

static std::string default_cache_dir() {
  if (const char* e = getenv("FEYNTROP_AMD_KCACHE")) return e;
  Dl_info info;
  if (dladdr((void*)&default_cache_dir, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t s = p.rfind('/');
    std::string dir = (s == std::string::npos) ? "." : p.substr(0, s);
    return dir + "/../kcache";
  }
  return "kcache";
}

static std::string cache_dir() {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  if (g_cache_dir.empty()) g_cache_dir = default_cache_dir();
  return g_cache_dir;
}

static uint64_t fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

/* deformed-path kernels specialised to a kinematic point run the
 * pipelined sample loop with the in-step row request for graphs of
 * FTR_SPEC_PIPE_MIN_E to FTR_DD_LDS_MAX_E edges (measured there; larger
 * one-lane kernels spill already) */
constexpr int FTR_SPEC_PIPE_MIN_E = 8;
static bool spec_pipe(const Variant& v, int E) {
  return v.mode == 2 && v.cparams && v.lps == 1 && E >= FTR_SPEC_PIPE_MIN_E && E <= FTR_DD_LDS_MAX_E;
}

/* Kernel variant for a problem at one lane per sample: path -> MODE,
 * table size -> placement. */
static Variant choose_one_lane(const ftrh::Prep& pr, size_t lds_limit, size_t mem_limit) {
  Variant v;
  v.K = pr.K;
  v.mode = pr.path;
  const int E = pr.g.E;
  const size_t full = ((size_t)25 << E) + 64, wonly = (size_t)8 << E;
  const size_t cap0 = std::min<size_t>(lds_limit, 64 * 1024);
  v.block = 256;
  if (rows_bytes(E) <= mem_limit) {
    /* the row layout (all E weights of a state in flight at once, the cut
     * found by counting) beats the LDS scans: deformed path +5..9 % on 5- to
     * 10-edge graphs, real path +2 % (K4 wheel) and +7 % (14-edge cubic
     * graph, against w in LDS with 1024-thread blocks); folded rows for
     * tables far beyond the L2 */
    v.tmode = E >= FTR_FROWS_MIN_E ? 4 : 3;
  } else if (full <= cap0) {
    v.tmode = 0;
  } else if (wonly <= lds_limit) {
    /* one table copy fills most of the LDS, so one block per CU: make it a
     * big one on the real path (the deformed path needs > 128 VGPRs) */
    v.tmode = 1;
    if (pr.path != 2 && wonly + 16 * (2 + 4 * pr.K) * 8 <= lds_limit) v.block = 1024;
  } else if (frows_bytes(E) <= mem_limit) {
    v.tmode = 4; /* folded rows in HBM: one contiguous read per cut step */
  } else {
    v.tmode = 2;
  }
  return v;
}

/* Kernel variant for a problem: the one-lane choice, or for deformed graphs
 * whose one-lane kernel spills (E >= 11: the second derivatives no longer
 * fit in LDS per lane) a lane pair per sample, measured 13.3 -> 11.2 ms
 * (generic) and 10.9 -> 8.4 ms (specialised) per 1e7 samples on the
 * 13-edge ladder. */
Variant choose_variant(const ftrh::Prep& pr, size_t lds_limit, size_t mem_limit) {
  const Variant v = choose_one_lane(pr, lds_limit, mem_limit);
  const int E = pr.g.E;
  if (pr.path == 2 && E > FTR_DD_LDS_MAX_E && v.tmode >= 3) {
    Variant w = v;
    w.lps = 2;
    for (int blk : {256, 128}) {
      w.block = blk;
      if (variant_lds(w, E) + 16 <= lds_limit) return w;
    }
  }
  return v;
}

/* A problem's variant inside a batch: bucket kernels read the parameters
 * from memory and evaluate one sample per lane (a lane-pair variant runs
 * with the one-lane block size of its row placement, 256). */
static Variant batch_member_variant(Variant v) {
  v.cparams = 0;
  if (v.lps != 1) {
    v.lps = 1;
    v.block = 256;
  }
  return v;
}

/* Feasible (table placement, block size) pairs.  The autotuner's list
 * (tuning) leaves out the folded rows below FTR_FROWS_MIN_E edges, where
 * they lose; ftr_set_variant accepts them. */
static std::vector<Variant> candidate_variants(const ftrh::Prep& pr, size_t lds_limit,
                                               size_t mem_limit = (size_t)128 << 30, bool tuning = false) {
  std::vector<Variant> out;
  const int E = pr.g.E;
  const size_t full = ((size_t)25 << E) + 64, wonly = (size_t)8 << E;
  for (int tm = 0; tm < 5; tm++)
    for (int blk : {256, 512, 1024}) {
      const size_t t = tm == 0 ? full : (tm == 1 ? wonly : (tm >= 3 ? (size_t)E * blk * 8 : 0));
      if (t + (size_t)(blk / 64) * (2 + 4 * pr.K) * 8 + 16 > lds_limit) continue;
      if (tm == 0 && full > 64 * 1024 && blk == 256) continue;
      if (tm >= 3 && table_rows_bytes(tm, E) > mem_limit) continue;
      if (tm == 4 && tuning && E < FTR_FROWS_MIN_E) continue;
      Variant v;
      v.K = pr.K;
      v.mode = pr.path;
      v.tmode = tm;
      v.block = blk;
      out.push_back(v);
    }
  /* two lanes per sample: deformed path, rows in HBM, second derivatives
   * in LDS per lane pair */
  for (int tm : {3, 4}) {
    if (pr.path != 2 || table_rows_bytes(tm, E) > mem_limit || (tm == 4 && tuning && E < FTR_FROWS_MIN_E))
      continue;
    for (int blk : {128, 256, 512}) {
      Variant v;
      v.K = pr.K;
      v.mode = pr.path;
      v.tmode = tm;
      v.block = blk;
      v.lps = 2;
      if (variant_lds(v, E) + 16 <= lds_limit) out.push_back(v);
    }
  }
  return out;
}

size_t variant_lds(const Variant& v, int E) {
  size_t t = v.tmode == 0 ? ((size_t)25 << E) + 64 : (v.tmode == 1 ? ((size_t)8 << E) : 0);
  t = (t + 15) & ~(size_t)15;
  /* table modes 3, 4: each thread's E slots for the sampler's scatter of X */
  if (v.tmode >= 3) t += (size_t)E * v.block * 8;
  if (dd_in_lds(v, E)) t += (size_t)(E * (E + 1) / 2) * (v.block / v.lps) * 8;
  return t + (size_t)(v.block / 64) * (2 + 4 * v.K) * 8;
}

int dd_lds_override() {
  static const int o = [] {
    const char* e = getenv("FEYNTROP_AMD_DDL");
    return e ? atoi(e) : -1;
  }();
  return o;
}

static int p_lowrank(const ftrh::Prep& pr, double* q, double* eta);

/* The graph spec of a problem's kernel (struct `gname`, and `gname`P for the
 * parameters of a kernel specialised to the kinematic point), followed by the
 * kernel instantiations unless `instantiate` is false (bucket kernels list
 * several specs and instantiate them themselves). */
static std::string spec_source(const ftrh::Prep& pr, const Variant& v, bool learned = false,
                               const std::string& gname = "FG", bool instantiate = true) {
  const ftrh::Graph& g = pr.g;
  /* elimination order: internal vertices, then those with nonzero P rows */
  std::vector<int> ord, pos(g.V);
  std::vector<char> isext(g.V, 0);
  for (int x : pr.ext) isext[x] = 1;
  for (int u = 0; u < g.V; u++)
    if (!isext[u]) ord.push_back(u);
  for (int u = 0; u < g.V; u++)
    if (isext[u]) ord.push_back(u);
  for (int q = 0; q < g.V; q++) pos[ord[q]] = q;
  auto arr = [](std::ostringstream& s, const char* name, const std::vector<int>& a) {
    s << "  static constexpr int " << name << "[" << a.size() << "] = {";
    for (size_t i = 0; i < a.size(); i++) s << (i ? "," : "") << a[i];
    s << "};\n";
  };
  std::ostringstream s;
  if (v.cparams) {
    /* the kinematic point as compile-time constants (exact hex literals) */
    const ftr::DevParams d = make_params(pr);
    auto lit = [](double x) -> std::string {
      if (std::isnan(x)) return "__builtin_nan(\"\")";
      if (std::isinf(x)) return x > 0 ? "__builtin_inf()" : "-__builtin_inf()";
      char b[64];
      snprintf(b, sizeof b, "%a", x);
      return b;
    };
    auto darr = [&](const char* name, const double* a, int n) {
      s << "  static constexpr double " << name << "[" << n << "] = {";
      for (int i = 0; i < n; i++) s << (i ? "," : "") << lit(a[i]);
      s << "};\n";
    };
    s << "\nstruct " << gname << "P {\n";
    s << "  static constexpr double D = " << lit(d.D) << ", mhalfD = " << lit(d.mhalfD) << ", W = " << lit(d.W)
      << ", mW = " << lit(d.mW) << ", Lf = " << lit(d.Lf) << ", IGtr = " << lit(d.IGtr) << ", def_ref = "
      << lit(d.def_ref) << ", inv_dlw = " << lit(d.inv_dlw) << ", iD = " << lit(d.iD) << ", iW = " << lit(d.iW)
      << ", use_iD = " << lit(d.use_iD) << ", use_iW = " << lit(d.use_iW) << ";\n";
    darr("nu", d.nu, FTR_MAX_E);
    darr("m2", d.m2, FTR_MAX_E);
    darr("P", d.P, FTR_MAX_V * FTR_MAX_V);
    darr("Pq", d.Pq, FTR_MAX_V * FTR_MAX_V);
    darr("Peta", d.Peta, FTR_MAX_V);
    s << "};\n";
  }
  s << "\nstruct " << gname << " {\n  static constexpr int E = " << g.E << ", V = " << g.V << ", NEXT = " << pr.ext.size() << ";\n";
  s << "  static constexpr bool DDL = " << (dd_in_lds(v, g.E) ? "true" : "false") << ";\n";
  s << "  static constexpr int BLK = " << v.block << ", LPS = " << v.lps << ";\n";
  /* rank of P's external block (trace and derivatives in low-rank form) */
  s << "  static constexpr int PR = " << p_lowrank(pr, nullptr, nullptr) << ";\n";
  {
    static const bool allow = [] {
      const char* e = getenv("FEYNTROP_AMD_NU1");   /* =0: always the summed-angle sincos */
      return !(e && !strcmp(e, "0"));
    }();
    bool nu1 = allow;
    for (int e = 0; e < g.E; e++) nu1 = nu1 && g.nu[e] == 1.0;
    /* every edge weight 1: the deformation's phase is the product of the
     * edges' phases (ftr_device.hpp integrand) */
    s << "  static constexpr bool NU1 = " << (nu1 ? "true" : "false") << ";\n";
  }
  {
    /* in-step row request of the sampler (ftr_device.hpp sample_point):
     * generic deformed-path kernels, lane-pair kernels (13-edge deformed
     * ladder, specialised: 8.31 -> 7.95 ms per 1e7) and folded rows
     * (22-edge ladder: 57.9 -> 56.5 ms per 1e8; 19 edges 41.4 -> 40.4) */
    static const int pf = [] {
      const char* e = getenv("FEYNTROP_AMD_PF");  /* =0 / =1 override */
      return e ? atoi(e) : -1;
    }();
    const bool on =
        pf >= 0 ? pf == 1 : (v.tmode == 4 || (v.mode == 2 && (!v.cparams || v.lps == 2)) || spec_pipe(v, g.E));
    s << "  static constexpr bool PF = " << (on ? "true" : "false") << ";\n";
  }
  {
    /* software-pipelined sample loop (ftr_device.hpp integrate_body) */
    static const int pipe = [] {
      const char* e = getenv("FEYNTROP_AMD_PIPE");  /* =0 / =1 override */
      return e ? atoi(e) : -1;
    }();
    /* measured per 1e8 samples: real path cubic5 -2.2 % (specialised) and
     * -2.6 % (generic), K4 wheel -3.4 % (specialised); generic deformed
     * kernels 0 to -2.7 % (triple ladder); deformed kernels specialised to
     * a kinematic point, with the in-step row request: 8 and 10 edges -5 %
     * (box-pentagon 15.5 -> 14.7 ms, triple ladder 20.0 -> 19.0), 7 edges
     * +4 % (double box 10.65 -> 11.1), 5 edges +-1 % -- on for 8 to 10
     * edges (spec_pipe);
     * the batched bucket kernels (many graphs in one kernel, not
     * instantiated here) +3 % on the 4096-point scan -- off; two lanes
     * per sample (kernels for graphs whose one-lane kernel spills) -- off:
     * the next sample's state adds to the register demand (13-edge
     * deformed ladder: 144 B of scratch per lane without, 400 B with) */
    const bool on = pipe >= 0 ? pipe == 1
                              : (instantiate && v.lps == 1 && (v.mode != 2 || !v.cparams || spec_pipe(v, g.E)));
    s << "  static constexpr bool PIPE = " << (on ? "true" : "false") << ";\n";
  }
  s << (v.cparams ? "  using PRM = " + gname + "P;\n" : std::string("  using PRM = ftr::DevParams;\n"));
  arr(s, "hi", g.hi);
  arr(s, "lo", g.lo);
  arr(s, "ord", ord);
  arr(s, "pos", pos);
  s << "};\n";
  if (!instantiate) return s.str();
  /* folded rows: at most one wave per SIMD (the 16-edge ladder's
   * specialised kernel, which fits two, ran 30.0 ms per 1e8 against 21.7
   * at one; the 22-edge one holds one anyway; mode-3 kernels that fit two
   * lose at one: cubic5 36.6 -> 42.4 ms per 3e8, K4 wheel +10 %) */
  if (v.tmode == 4)
    s << "#undef FTR_INT_ATTR\n#define FTR_INT_ATTR __attribute__((amdgpu_waves_per_eu(1, 1)))\n";
  s << "FTR_INSTANTIATE(ftrk, FG, " << v.K << ", " << v.mode << ", " << v.tmode << ", " << v.block << ")\n";
  if (learned)
    s << "FTR_INSTANTIATE_LEARNED(ftrl, FG, " << v.K << ", " << v.mode << ", " << v.tmode << ", " << v.block << ")\n";
  return s.str();
}

static const std::vector<std::string>& rtc_options() {
  static const std::vector<std::string> o = [] {
    const char* opt = getenv("FEYNTROP_AMD_RTC_OPT");   /* e.g. -O1, for experiments */
    std::vector<std::string> v{"--offload-arch=gfx950", opt ? opt : "-O3", "-std=c++17",
                               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    if (const char* x = getenv("FEYNTROP_AMD_RTC_EXTRA")) {  /* extra options, space separated */
      std::istringstream is(x);
      std::string t;
      while (is >> t) v.push_back(t);
    }
    return v;
  }();
  return o;
}

/* The cache path of a source's code object: the source, the compile
 * options and the hipRTC version hashed. */
static std::string code_object_path(const std::string& src) {
  std::string key = src;
  for (auto& o : rtc_options()) key += "\n" + o;
  int rtcver = 0, rtcmin = 0;
  hiprtcVersion(&rtcver, &rtcmin);
  key += "\nhiprtc " + std::to_string(rtcver) + "." + std::to_string(rtcmin);
  char name[64];
  snprintf(name, sizeof name, "ftr_%016llx.co", (unsigned long long)fnv1a(key));
  return cache_dir() + "/" + name;
}

/* Compile (or fetch from the cache) the code object for a source. */
std::string code_object(const std::string& src, std::string* path_out) {
  const std::string dir = cache_dir();
  const std::string path = code_object_path(src);
  if (path_out) *path_out = path;
  {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      std::stringstream ss;
      ss << f.rdbuf();
      std::string co = ss.str();
      if (!co.empty()) return co;
    }
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "ftr_kernel.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw Error(FTR_ERR_COMPILE, "hiprtcCreateProgram failed");
  std::vector<const char*> opts;
  for (auto& o : rtc_options()) opts.push_back(o.c_str());
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, 0);
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(FTR_ERR_COMPILE, "kernel specialisation failed: " + log.substr(0, 4000));
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::string co(cs, 0);
  hiprtcGetCode(prog, &co[0]);
  hiprtcDestroyProgram(&prog);
  /* store atomically when the cache directory is usable */
  mkdir(dir.c_str(), 0755);
  const std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (f) f.write(co.data(), (std::streamsize)co.size());
  }
  if (rename(tmp.c_str(), path.c_str()) != 0) unlink(tmp.c_str());
  return co;
}

/* Tuning cache: the autotuner's choice per (graph, K, path), next to the
 * code objects. */
static std::string tune_path(const ftrh::Prep& pr) {
  Variant v;
  v.K = pr.K;
  v.mode = pr.path;
  v.tmode = 9;
  v.block = 0;
  char name[64];
  snprintf(name, sizeof name, "/tune_%016llx.txt", (unsigned long long)fnv1a(spec_source(pr, v)));
  return cache_dir() + name;
}

/* The autotuner's recorded variant of a graph, through the context's memo of
 * tuning files (one file read per graph and process; the autotuner drops the
 * entry it rewrites). */
static bool load_tuned_memo(ftr_ctx* c, const ftrh::Prep& pr, size_t lds_limit, Variant* out) {
  const std::string path = tune_path(pr);
  auto it = c->tune_memo.find(path);
  if (it == c->tune_memo.end()) {
    std::ifstream f(path);
    int tm = -1, blk = 0, lps = 1;
    if (!(f >> tm >> blk)) tm = -1;
    else if (!(f >> lps)) lps = 1; /* files of two fields: one lane per sample */
    it = c->tune_memo.emplace(path, std::array<int, 3>{tm, blk, lps}).first;
  }
  const std::array<int, 3>& t = it->second;
  if (t[0] < 0) return false;
  for (const Variant& v : candidate_variants(pr, lds_limit))
    if (v.tmode == t[0] && v.block == t[1] && v.lps == t[2]) {
      *out = v;
      return true;
    }
  return false;
}

std::string full_source(const ftrh::Prep& pr, const Variant& v, bool learned) {
  std::string src = std::string(ftr_embedded_device_source) + spec_source(pr, v, learned);
  if (const char* d = getenv("FEYNTROP_AMD_DUMP_SRC")) {
    char name[64];
    snprintf(name, sizeof name, "/ftr_%016llx.hip", (unsigned long long)fnv1a(src));
    std::ofstream(std::string(d) + name) << src;
  }
  return src;
}

/* The vertices with external momentum in elimination order (the order of
 * the trailing block of the kernels' matrices: ascending vertex number). */
static std::vector<int> ext_order(const ftrh::Prep& pr) {
  std::vector<char> isext(pr.g.V, 0);
  for (int x : pr.ext) isext[x] = 1;
  std::vector<int> o;
  for (int u = 0; u < pr.g.V; u++)
    if (isext[u]) o.push_back(u);
  return o;
}

/* P's external block B (symmetric, rows summing to zero by momentum
 * conservation) as sum_r eta_r q_r q_r^T: cyclic Jacobi eigen-decomposition
 * in long double, q_r = sqrt|lambda_r| v_r, eta_r = sign lambda_r, with the
 * eigenvalues below 1e-14 of the largest (the null direction 1 of
 * conservation) dropped.  q (rows of FTR_MAX_V) and eta may be null. */
static int p_lowrank(const ftrh::Prep& pr, double* q, double* eta) {
  const std::vector<int> ex = ext_order(pr);
  const int n = (int)ex.size();
  std::vector<long double> B((size_t)n * n), Vm((size_t)n * n, 0.0L);
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) B[(size_t)a * n + b] = pr.P[(size_t)ex[a] * pr.g.V + ex[b]];
  for (int a = 0; a < n; a++) Vm[(size_t)a * n + a] = 1.0L;
  for (int sweep = 0; sweep < 60; sweep++) {
    long double off = 0;
    for (int a = 0; a < n; a++)
      for (int b = a + 1; b < n; b++) off += B[(size_t)a * n + b] * B[(size_t)a * n + b];
    if (off < 1e-60L) break;
    for (int a = 0; a < n; a++)
      for (int b = a + 1; b < n; b++) {
        const long double apq = B[(size_t)a * n + b];
        if (apq == 0) continue;
        const long double th = (B[(size_t)b * n + b] - B[(size_t)a * n + a]) / (2 * apq);
        const long double t = (th >= 0 ? 1 : -1) / (fabsl(th) + sqrtl(th * th + 1));
        const long double c = 1 / sqrtl(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) { /* B <- J^T B J */
          const long double bka = B[(size_t)k * n + a], bkb = B[(size_t)k * n + b];
          B[(size_t)k * n + a] = c * bka - s * bkb;
          B[(size_t)k * n + b] = s * bka + c * bkb;
        }
        for (int k = 0; k < n; k++) {
          const long double bak = B[(size_t)a * n + k], bbk = B[(size_t)b * n + k];
          B[(size_t)a * n + k] = c * bak - s * bbk;
          B[(size_t)b * n + k] = s * bak + c * bbk;
        }
        for (int k = 0; k < n; k++) {
          const long double vka = Vm[(size_t)k * n + a], vkb = Vm[(size_t)k * n + b];
          Vm[(size_t)k * n + a] = c * vka - s * vkb;
          Vm[(size_t)k * n + b] = s * vka + c * vkb;
        }
      }
  }
  long double lmax = 0;
  for (int a = 0; a < n; a++) lmax = std::max(lmax, fabsl(B[(size_t)a * n + a]));
  int R = 0;
  for (int a = 0; a < n; a++) {
    const long double lam = B[(size_t)a * n + a];
    if (lmax == 0 || fabsl(lam) <= 1e-14L * lmax) continue;
    const long double sc = sqrtl(fabsl(lam));
    if (q)
      for (int b = 0; b < n; b++) q[(size_t)R * FTR_MAX_V + b] = (double)(sc * Vm[(size_t)b * n + a]);
    if (eta) eta[R] = lam > 0 ? 1.0 : -1.0;
    R++;
  }
  return R;
}

ftr::DevParams make_params(const ftrh::Prep& pr) {
  ftr::DevParams d;
  memset(&d, 0, sizeof d);
  d.D = pr.D;
  d.mhalfD = -pr.D / 2.;
  d.W = pr.W;
  d.mW = -pr.W;
  d.Lf = (double)pr.L;
  d.IGtr = pr.IGtr;
  d.def_ref = pr.def_ref;
  d.inv_dlw = 1. / (pr.D / 2. * pr.L + pr.W);
  const double mD = -pr.D / 2., mW = -pr.W;
  if (mD == std::nearbyint(mD) && std::fabs(mD) <= 64) { d.iD = mD; d.use_iD = 1; }
  if (mW == std::nearbyint(mW) && std::fabs(mW) <= 64) { d.iW = mW; d.use_iW = 1; }
  for (int e = 0; e < pr.g.E; e++) {
    d.nu[e] = pr.g.nu[e];
    d.m2[e] = pr.m2[e];
  }
  for (int i = 0; i < pr.g.V; i++)
    for (int j = 0; j < pr.g.V; j++) d.P[i * FTR_MAX_V + j] = pr.P[(size_t)i * pr.g.V + j];
  p_lowrank(pr, d.Pq, d.Peta);
  return d;
}

static void fill_params(ftr_problem* p) { p->hprm = make_params(p->prep); }

void prep_from_input(ftrh::Prep& pr, const ftr_input* in) {
  if (!in || in->E < 1) throw Error(FTR_ERR_INVALID, "input graph must have at least one edge");
  ftrh::build_graph(pr.g, in->E, in->edges, in->nu);
  if (in->V != pr.g.V) throw Error(FTR_ERR_INVALID, "input scalarproducts must be a (V x V)-matrix");
  if (in->K < 1 || in->K > FTR_MAX_K) throw Error(FTR_ERR_LIMIT, "num_eps_terms must be in [1, 16]");
  if (!(in->D > 0)) throw Error(FTR_ERR_DOMAIN, "dimension must be positive");
  pr.P.assign(in->P, in->P + (size_t)in->V * in->V);
  pr.m2.assign(in->m2, in->m2 + in->E);
  pr.D = in->D;
  pr.K = in->K;
  pr.lambda = in->lambda;
  if (pr.g.V > FTR_MAX_V) throw Error(FTR_ERR_LIMIT, "more than 24 vertices are not supported on the GPU path");
  ftrh::preprocess(pr);
  if (pr.g.E > FTR_MAX_E) throw Error(FTR_ERR_LIMIT, "more than 30 edges are not supported");
}

void ensure_scratch(ftr_ctx* c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return;
  if (c->d_scratch) HIPCHK(hipFree(c->d_scratch));
  c->d_scratch = nullptr;
  const size_t nb = std::max(bytes, (size_t)1 << 20);
  HIPCHK(hipMalloc(&c->d_scratch, nb));
  c->scratch_bytes = nb;
}

hipModule_t load_module_raw(ftr_ctx* c, const std::string& src) {
  auto it = c->raw_modules.find(src);
  if (it != c->raw_modules.end()) return it->second;
  const std::string co = code_object(src, nullptr);
  hipModule_t mod = nullptr;
  HIPCHK(hipModuleLoadData(&mod, co.data()));
  c->raw_modules[src] = mod;
  return mod;
}

static Module load_module(ftr_ctx* c, const std::string& src) {
  auto it = c->modules.find(src);
  if (it != c->modules.end()) return it->second;
  Module m;
  m.mod = load_module_raw(c, src);
  HIPCHK(hipModuleGetFunction(&m.f_int, m.mod, "ftrk_int"));
  HIPCHK(hipModuleGetFunction(&m.f_rep, m.mod, "ftrk_rep"));
  HIPCHK(hipModuleGetFunction(&m.f_smp, m.mod, "ftrk_smp"));
  c->modules[src] = m;
  return m;
}

/* The integration module of a problem's variant, looked up by the graph spec
 * (a few hundred bytes) instead of the full kernel source. */
static Module module_for(ftr_ctx* c, const ftrh::Prep& pr, const Variant& v) {
  std::string key = spec_source(pr, v);
  auto it = c->spec_modules.find(key);
  if (it != c->spec_modules.end()) return it->second;
  const Module m = load_module(c, full_source(pr, v));
  c->spec_modules.emplace(std::move(key), m);
  return m;
}

/* Lane jump polynomials in base 256 from the jump of one lane (q1): entry
 * 255 k + d - 1 of `out` jumps d 256^k lanes (d = 1..255, k < (lb+7)/8;
 * the top digit of a block's lanes uses the first entries only).
 * Returns the jump of 2^lb lanes (one block). */
static size_t lane_poly_count(int lb) { return 255 * (size_t)((lb + 7) / 8); }
static ftrh::Poly lane_digit_polys(const ftrh::Poly& q1, int lb, uint64_t* out) {
  const int ldig = (lb + 7) / 8;
  ftrh::Poly q = q1; /* 256^k lanes */
  for (int k = 0; k < ldig; k++) {
    const int top = (k == ldig - 1) ? (1 << (lb - 8 * k)) - 1 : 255; /* largest digit used */
    ftrh::Poly r = q;
    for (int d = 1; d <= top; d++) {
      memcpy(out + 4 * (255 * (size_t)k + d - 1), r.w, 32);
      if (d < top) r = ftrh::xo_poly_mul(r, q);
    }
    for (int sq = 0; sq < 8; sq++) q = ftrh::xo_poly_square(q);
  }
  ftrh::Poly blk = q1;
  for (int j = 0; j < lb; j++) blk = ftrh::xo_poly_square(blk);
  return blk;
}

/* Jump polynomials of a launch geometry: the lane digit jumps
 * (lane_digit_polys), then, for base-256 digits k < ndig of the block index,
 * the jumps of d 256^k blocks (d = 1..255); cached per context. */
static const std::vector<uint64_t>& jump_polys(ftr_ctx* c, uint64_t S, uint64_t per, int lb, int ndig) {
  const auto key = std::make_tuple(S, per, lb, ndig);
  auto it = c->jump_cache.find(key);
  if (it != c->jump_cache.end()) return it->second;
  const size_t nlane = lane_poly_count(lb);
  std::vector<uint64_t> polys(4 * (nlane + 255 * (size_t)ndig));
  const unsigned __int128 dl = (unsigned __int128)S * per;
  /* q: the jump polynomial for one block (S * per * block) */
  ftrh::Poly q = lane_digit_polys(ftrh::xo_jump_poly((uint64_t)dl, (uint64_t)(dl >> 64)), lb, polys.data());
  for (int k = 0; k < ndig; k++) {
    ftrh::Poly r = q;
    for (int d = 1; d <= 255; d++) {
      memcpy(&polys[4 * (nlane + 255 * (size_t)k + d - 1)], r.w, 32);
      if (d < 255) r = ftrh::xo_poly_mul(r, q);
    }
    for (int sq = 0; sq < 8; sq++) q = ftrh::xo_poly_square(q); /* 256^(k+1) blocks */
  }
  return c->jump_cache.emplace(key, std::move(polys)).first->second;
}

hipFunction_t block_states_fn(ftr_ctx* c) {
  if (!c->f_block_states) {
    hipModule_t mod = load_module_raw(c, ftr_embedded_jtable_source);
    HIPCHK(hipModuleGetFunction(&c->f_block_states, mod, "ftrj_block_states"));
  }
  return c->f_block_states;
}

Geometry setup_stream(ftr_problem* p, uint64_t first, uint64_t n, uint64_t seed, uint64_t lanes_req, int block,
                      int nq, hipFunction_t occ_fn, char* ext, size_t ext_bytes, int lps) {
  ftr_ctx* c = p->ctx;
  const int E = p->prep.g.E;
  Geometry G;
  const int slots = block / lps; /* sample slots per block */
  uint64_t lanes_target = lanes_req;
  if (!lanes_target) {
    int nbpc = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nbpc, occ_fn ? occ_fn : p->m.f_int, block, p->lds) != hipSuccess || nbpc < 1)
      nbpc = 1;
    lanes_target = (uint64_t)c->prop.multiProcessorCount * nbpc * slots;
  }
  const uint64_t Smin = 4;
  uint64_t S = (n + lanes_target - 1) / lanes_target;
  if (S < Smin) S = Smin;
  const uint64_t lanes = (n + S - 1) / S;
  const uint64_t nblocks = (lanes + slots - 1) / slots;
  G.nblocks = nblocks;
  G.S = S;
  G.lanes = lanes;
  const uint64_t per = 2 * (uint64_t)E - 1;
  int lb = 0;
  while ((1 << lb) < slots) lb++;
  int ndig = 0;
  while (ndig < 8 && ((uint64_t)1 << (8 * ndig)) < nblocks) ndig++;
  const size_t bytes_states = (size_t)nblocks * 32;
  const size_t bytes_polys = 32 * (lane_poly_count(lb) + 255 * (size_t)ndig);
  const size_t bytes_part = (size_t)nblocks * nq * 8;
  const size_t need = ((bytes_states + 255) & ~(size_t)255) + ((bytes_polys + 255) & ~(size_t)255) + bytes_part +
                      (size_t)nq * 8 + 512;
  char* d;
  if (ext) {
    if (need > ext_bytes)
      throw Error(FTR_ERR_LIMIT, "launch of " + std::to_string(nblocks) + " blocks needs " + std::to_string(need) +
                                     " bytes of scratch, more than the fixed " + std::to_string(ext_bytes) +
                                     " bytes of the no-allocation entry point (request fewer lanes)");
    d = ext;
  } else {
    ensure_scratch(c, need);
    d = (char*)c->d_scratch;
  }
  const std::vector<uint64_t>& polys = jump_polys(c, S, per, lb, ndig);
  /* the stream state at sample `first`; block b starts b blocks later
   * (ftrj_block_states on the device, which takes it by value) */
  uint64_t base[4];
  ftrh::xo_seed(seed, base);
  {
    const unsigned __int128 d0 = (unsigned __int128)first * per;
    if (d0) ftrh::xo_apply(ftrh::xo_jump_poly((uint64_t)d0, (uint64_t)(d0 >> 64)), base);
  }
  G.d_states = (uint64_t*)d;
  G.d_polys = (uint64_t*)(d + ((bytes_states + 255) & ~(size_t)255));
  G.d_partials = (double*)((char*)G.d_polys + ((bytes_polys + 255) & ~(size_t)255));
  G.d_totals = G.d_partials + (size_t)nblocks * nq;
  /* the polynomials live in the context's cache, so the asynchronous copy
   * reads memory that outlives it */
  if (!polys.empty()) HIPCHK(hipMemcpyHtoDAsync(G.d_polys, (void*)polys.data(), polys.size() * 8, c->stream));
  {
    hipFunction_t f = block_states_fn(c);
    const uint64_t* bp = G.d_polys + 4 * lane_poly_count(lb);
    int nb = ndig;
    uint64_t nbl = nblocks;
    uint64_t b0 = base[0], b1 = base[1], b2 = base[2], b3 = base[3];
    void* args[] = {&b0, &b1, &b2, &b3, &bp, &nb, &nbl, &G.d_states};
    HIPCHK(hipModuleLaunchKernel(f, (unsigned)((nblocks + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, args,
                                 nullptr));
  }
  return G;
}

int fail(ftr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

void fill_tables(ftr::Launch& a, const ftr_problem* p) {
  a.prm = p->d_prm;
  a.t_w = p->d_w;
  a.t_jr = p->d_jr;
  a.t_ir = p->d_ir;
  a.t_mmL = p->d_mml;
  a.t_rows = p->d_rows;
  a.t_rec = p->d_rec;
}

/* device memory available for table mode 3 (half of the device) */
static size_t rows_mem_limit(const ftr_ctx* c) { return c->prop.totalGlobalMem / 2; }

/* --------------------------------------------------------------- C ABI -- */
extern "C" {

int ftr_create(int device, ftr_ctx** out) {
  if (!out) return FTR_ERR_INVALID;
  *out = nullptr;
  ftr_ctx* c = new ftr_ctx();
  try {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(FTR_ERR_DEVICE, "no such HIP device");
    c->device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipGetDeviceProperties(&c->prop, device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    if (const char* m = getenv("FEYNTROP_AMD_JTABLE")) {
      if (!strcmp(m, "host")) c->jtable_mode = 1;
      else if (!strcmp(m, "device")) c->jtable_mode = 2;
    }
    /* the block start-state kernel, loaded here so that no integration
     * call has to load a module (ftr_integrate_noalloc allocates nothing) */
    block_states_fn(c);
    /* the fixed scratch of ftr_integrate_noalloc, resolved once per device
     * with the device current; its failure only disables that entry point */
    {
      const hipError_t e = ftrn_scratch_init(device, &c->na_scratch);
      if (e != hipSuccess) {
        c->na_scratch = nullptr;
        c->na_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
      }
    }
  } catch (Error& e) {
    fprintf(stderr, "feyntrop_amd: %s\n", e.what());
    delete c;
    return e.code;
  }
  *out = c;
  return FTR_OK;
}

void ftr_destroy(ftr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  for (auto& kv : c->raw_modules) (void)hipModuleUnload(kv.second);
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
  if (c->d_batch) (void)hipFree(c->d_batch);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  for (int k = 0; k < FTR_SIDE_STREAMS; k++) {
    if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* ftr_last_error(const ftr_ctx* c) { return c ? c->err.c_str() : g_host_err.c_str(); }

int ftr_set_cache_dir(const char* dir) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_cache_dir = dir ? dir : "";
  return FTR_OK;
}

int ftr_precompile(const ftr_input* in, char* err, int errlen) {
  try {
    ftrh::Prep pr;
    prep_from_input(pr, in);
    code_object(ftr_embedded_jtable_source, nullptr); /* the device J-table module */
    /* both LDS limits a gfx950 part may report */
    for (size_t lim : {(size_t)64 * 1024, (size_t)160 * 1024}) {
      const Variant v = choose_variant(pr, lim);
      code_object(full_source(pr, v), nullptr);
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

int ftr_precompile_variant(const ftr_input* in, int table_mode, int block, int lps, int specialized, char* err,
                           int errlen) {
  try {
    ftrh::Prep pr;
    prep_from_input(pr, in);
    bool found = false;
    for (Variant v : candidate_variants(pr, (size_t)160 * 1024))
      if (v.tmode == table_mode && v.block == block && v.lps == lps) {
        v.cparams = specialized ? 1 : 0;
        code_object(full_source(pr, v), nullptr);
        found = true;
        break;
      }
    if (!found) throw Error(FTR_ERR_LIMIT, "infeasible (table_mode, block, lanes per sample) for this problem");
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

int ftr_precompile_specialized(const ftr_input* in, char* err, int errlen) {
  try {
    ftrh::Prep pr;
    prep_from_input(pr, in);
    for (size_t lim : {(size_t)64 * 1024, (size_t)160 * 1024}) {
      Variant v = choose_variant(pr, lim);
      v.cparams = 1;
      code_object(full_source(pr, v), nullptr);
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

/* Host phase of preparation: validation, regime, J-table (host or device). */
static void prepare_host(ftr_ctx* c, ftr_problem* p, const ftr_input* in) {
  if (c && c->jtable_mode != 1) {
    /* the J-table on the device: always (mode 2) or from E = 13 on (auto) */
    const int mode = c->jtable_mode;
    p->prep.jtable_fn = [c, mode](ftrh::Prep& pr, const ftrh::JtSetup& js, ftrh::JtResult& r) {
      if (mode == 2 || pr.g.E >= FTR_DEVICE_JTABLE_MIN_E) device_jtable(c, pr, js, r);
      else ftrh::host_jtable(pr, js, r);
    };
  }
  prep_from_input(p->prep, in);
}

/* Device phase: kernel module, and the parameters and sampler tables in one
 * allocation (DevParams | w | J | 1/r | mm|L<<1), uploaded with one copy
 * from a host staging buffer. */
/* Variant, kernel module and parameters of a prepared problem. */
static void choose_and_load(ftr_ctx* c, ftr_problem* p) {
  if (!load_tuned_memo(c, p->prep, c->prop.sharedMemPerBlock, &p->var))
    p->var = choose_variant(p->prep, c->prop.sharedMemPerBlock, rows_mem_limit(c));
  p->lds = variant_lds(p->var, p->prep.g.E);
  p->m = module_for(c, p->prep, p->var);
  if (p->lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)p->m.f_int, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds);
  fill_params(p);
}

/* Layout of a problem's slab: DevParams | w | J | 1/r | mm|L<<1. */
struct SlabLayout {
  size_t NT, o_w, o_jr, o_ir, o_mml, total;
  explicit SlabLayout(int E) {
    NT = (size_t)1 << E;
    o_w = (sizeof(ftr::DevParams) + 255) & ~(size_t)255;
    o_jr = o_w + NT * 8;
    o_ir = o_jr + NT * 8;
    o_mml = o_ir + NT * 8;
    total = o_mml + NT;
  }
};

/* host image of the slab */
static void stage_slab(const ftr_problem* p, const SlabLayout& L, char* h) {
  memcpy(h, &p->hprm, sizeof(ftr::DevParams));
  double* w = (double*)(h + L.o_w);
  double* ir = (double*)(h + L.o_ir);
  uint8_t* mml = (uint8_t*)(h + L.o_mml);
  memcpy(h + L.o_jr, p->prep.Jr.data(), L.NT * 8);
  for (size_t i = 0; i < L.NT; i++) {
    w[i] = p->prep.Jr[i] / p->prep.r[i];
    ir[i] = 1. / p->prep.r[i];
    mml[i] = (uint8_t)((p->prep.mm[i] ? 1 : 0) | (p->prep.Lt[i] << 1));
  }
}

/* device pointers of a slab at d */
static void bind_slab(ftr_problem* p, const SlabLayout& L, char* d) {
  p->d_prm = (ftr::DevParams*)d;
  p->d_w = (double*)(d + L.o_w);
  p->d_jr = (double*)(d + L.o_jr);
  p->d_ir = (double*)(d + L.o_ir);
  p->d_mml = (uint8_t*)(d + L.o_mml);
}

static void prepare_device(ftr_ctx* c, ftr_problem* p) {
  HIPCHK(hipSetDevice(c->device));
  choose_and_load(c, p);
  const SlabLayout L(p->prep.g.E);
  std::vector<char> h(L.total);
  stage_slab(p, L, h.data());
  char* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, L.total));
  p->d_slab = d;
  bind_slab(p, L, d);
  HIPCHK(hipMemcpyHtoDAsync(d, h.data(), L.total, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (p->var.tmode >= 3) device_rows(c, p, p->var.tmode);
}

/* Device phase of many problems at once: one allocation shared by all
 * slabs and row tables, one copy of the slabs, the row kernels back to
 * back and one wait. */
static void prepare_device_batch(ftr_ctx* c, const std::vector<ftr_problem*>& ps) {
  if (ps.empty()) return;
  HIPCHK(hipSetDevice(c->device));
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  std::vector<size_t> o_slab(ps.size()), o_rows(ps.size(), 0);
  size_t slabs = 0;
  for (size_t i = 0; i < ps.size(); i++) {
    choose_and_load(c, ps[i]);
    o_slab[i] = slabs;
    slabs += al(SlabLayout(ps[i]->prep.g.E).total);
  }
  size_t total = slabs;
  for (size_t i = 0; i < ps.size(); i++)
    if (ps[i]->var.tmode >= 3) {
      o_rows[i] = total;
      total += al(rows_region_bytes(ps[i]->var.tmode, ps[i]->prep.g.E));
    }
  std::vector<char> h(slabs);
  for (size_t i = 0; i < ps.size(); i++) stage_slab(ps[i], SlabLayout(ps[i]->prep.g.E), h.data() + o_slab[i]);
  char* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, total));
  const int dev = c->device;
  std::shared_ptr<void> own(d, [dev](void* q) {
    (void)hipSetDevice(dev);
    (void)hipFree(q);
  });
  HIPCHK(hipMemcpyHtoDAsync(d, h.data(), slabs, c->stream));
  for (size_t i = 0; i < ps.size(); i++) {
    ps[i]->d_shared = own;
    bind_slab(ps[i], SlabLayout(ps[i]->prep.g.E), d + o_slab[i]);
    if (ps[i]->var.tmode >= 3) launch_rows(c, ps[i], ps[i]->var.tmode, d + o_rows[i]);
  }
  HIPCHK(hipStreamSynchronize(c->stream));
}

int ftr_prepare(ftr_ctx* c, const ftr_input* in, ftr_problem** out) {
  if (!out) return FTR_ERR_INVALID;
  *out = nullptr;
  ftr_problem* p = new ftr_problem();
  p->ctx = c;
  p->device = c ? c->device : 0;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    prepare_host(c, p, in);
    if (c) prepare_device(c, p);
    p->prep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    int code = fe ? fe->code : FTR_ERR_INVALID;
    ftr_problem_free(p);
    if (!c) g_host_err = e.what();
    return fail(c, code, e.what());
  }
  *out = p;
  return FTR_OK;
}

int ftr_prepare_batch(ftr_ctx* c, int n, const ftr_input* in, ftr_problem** out, int* status) {
  if (!c || n < 0 || (n > 0 && (!in || !out || !status))) return FTR_ERR_INVALID;
  c->batch_errors.assign(n, std::string());
  std::vector<ftr_problem*> ps(n, nullptr);
  for (int i = 0; i < n; i++) {
    out[i] = nullptr;
    status[i] = FTR_OK;
    ps[i] = new ftr_problem();
    ps[i]->ctx = c;
    ps[i]->device = c->device;
  }
  /* host phases in parallel over problems (each problem's own J-table pass
   * then runs on one thread); the device J-table, when chosen, is issued
   * from one thread at a time */
  std::mutex dev_mu;
  const int mode = c->jtable_mode;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n; i++) {
    const auto t0 = std::chrono::steady_clock::now();
    try {
      ftr_problem* p = ps[i];
      if (mode != 1)
        p->prep.jtable_fn = [c, mode, &dev_mu](ftrh::Prep& pr, const ftrh::JtSetup& js, ftrh::JtResult& r) {
          if (mode == 2 || pr.g.E >= FTR_DEVICE_JTABLE_MIN_E) {
            std::lock_guard<std::mutex> lk(dev_mu);
            device_jtable(c, pr, js, r);
          } else {
            ftrh::host_jtable(pr, js, r);
          }
        };
      prep_from_input(p->prep, &in[i]);
      p->prep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } catch (std::exception& e) {
      const Error* fe = dynamic_cast<const Error*>(&e);
      status[i] = fe ? fe->code : FTR_ERR_INVALID;
      c->batch_errors[i] = e.what();
    }
  }
  std::vector<ftr_problem*> ok;
  for (int i = 0; i < n; i++) {
    if (status[i] != FTR_OK) {
      ftr_problem_free(ps[i]);
      ps[i] = nullptr;
    } else {
      ok.push_back(ps[i]);
    }
  }
  try {
    const auto t0 = std::chrono::steady_clock::now();
    prepare_device_batch(c, ok);
    const double share = ok.empty() ? 0.0
                                    : std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() /
                                          (double)ok.size();
    for (ftr_problem* p : ok) p->prep.seconds += share;
  } catch (std::exception& e) {
    /* the device phase failed as a whole: every prepared problem fails */
    const Error* fe = dynamic_cast<const Error*>(&e);
    for (int i = 0; i < n; i++)
      if (ps[i]) {
        status[i] = fe ? fe->code : FTR_ERR_DEVICE;
        c->batch_errors[i] = e.what();
        ftr_problem_free(ps[i]);
        ps[i] = nullptr;
      }
  }
  for (int i = 0; i < n; i++) out[i] = ps[i];
  return FTR_OK;
}

const char* ftr_batch_error(const ftr_ctx* c, int i) {
  if (!c || i < 0 || i >= (int)c->batch_errors.size()) return "";
  return c->batch_errors[i].c_str();
}

void ftr_problem_free(ftr_problem* p) {
  if (!p) return;
  /* the context may already be gone: only the device id is used */
  if (p->d_slab || p->d_rows_slab) (void)hipSetDevice(p->device);
  if (p->d_slab) (void)hipFree(p->d_slab);
  if (p->d_rows_slab) (void)hipFree(p->d_rows_slab);
  delete p;
}

int ftr_problem_info(const ftr_problem* p, ftr_info* o) {
  if (!p || !o) return FTR_ERR_INVALID;
  const ftrh::Prep& pr = p->prep;
  o->E = pr.g.E;
  o->V = pr.g.V;
  o->L = pr.L;
  o->K = pr.K;
  o->W = pr.W;
  o->IGtr = pr.IGtr;
  o->euclidean = pr.euclidean;
  o->pseudo_euclidean = pr.pseudo;
  o->generic = pr.generic;
  o->gp_property = pr.gp;
  o->lambda_eff = pr.lambda_eff;
  o->path = pr.path;
  o->seconds_preprocessing = pr.seconds;
  return FTR_OK;
}

int ftr_jtable(const ftr_problem* p, double* Jr, double* r, uint8_t* mm, uint8_t* L) {
  if (!p) return FTR_ERR_INVALID;
  const size_t NT = p->prep.Jr.size();
  if (Jr) memcpy(Jr, p->prep.Jr.data(), NT * 8);
  if (r) memcpy(r, p->prep.r.data(), NT * 8);
  if (mm) memcpy(mm, p->prep.mm.data(), NT);
  if (L) memcpy(L, p->prep.Lt.data(), NT);
  return FTR_OK;
}

/* ftr_integrate and ftr_integrate_noalloc: one launch of the problem's
 * integration kernel over N samples.  noalloc: all scratch in the device's
 * fixed array, the block partials summed on the device in block order (the
 * host loop's additions, in its order) and only the totals copied back. */
static int integrate_impl(ftr_ctx* c, ftr_problem* p, uint64_t N, uint64_t seed, const ftr_opts* opts,
                          ftr_moments* out, bool noalloc) {
  if (!c) {
    g_host_err = "no HIP device context: the problem was prepared host-only";
    return FTR_ERR_DEVICE;
  }
  if (!p || !out) return FTR_ERR_INVALID;
  if (!p->m.f_int) return fail(c, FTR_ERR_INVALID, "problem was prepared without a device context");
  if (p->ctx != c) return fail(c, FTR_ERR_INVALID, "problem belongs to another context");
  try {
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(c->device));
    const int K = p->prep.K, NQ = 2 + 4 * K;
    memset(out, 0, sizeof *out);
    out->K = K;
    if (opts && opts->block && opts->block != p->var.block)
      throw Error(FTR_ERR_INVALID, "block size differs from the compiled variant");
    if (opts && opts->table_mode != -1 && opts->table_mode != p->var.tmode)
      throw Error(FTR_ERR_INVALID, "table_mode " + std::to_string(opts->table_mode) +
                                       " differs from the compiled placement " + std::to_string(p->var.tmode) +
                                       " (choose it with ftr_set_variant)");
    if (noalloc && !c->na_scratch)
      throw Error(FTR_ERR_DEVICE, "the no-allocation scratch of this device is unavailable: " + c->na_err);
    if (N > 0) {
      Geometry G = setup_stream(p, 0, N, seed, opts ? opts->lanes : 0, p->var.block, NQ, nullptr,
                                noalloc ? (char*)c->na_scratch : nullptr, noalloc ? FTR_NOALLOC_SCRATCH_BYTES : 0,
                                p->var.lps);
      ftr::Launch a;
      fill_tables(a, p);
      a.block_state = G.d_states;
      a.lane_poly = G.d_polys;
      a.N = N;
      a.S = G.S;
      a.partials = G.d_partials;
      void* args[] = {&a};
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      HIPCHK(hipModuleLaunchKernel(p->m.f_int, (unsigned)G.nblocks, 1, 1, p->var.block, 1, 1, (unsigned)p->lds,
                                   c->stream, args, nullptr));
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      double tot[2 + 4 * FTR_MAX_K];
      if (noalloc) {
        HIPCHK(ftrn_sum_partials_launch(G.d_partials, G.nblocks, NQ, G.d_totals, c->stream));
        HIPCHK(hipMemcpyDtoHAsync(tot, G.d_totals, (size_t)NQ * 8, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      } else {
        c->h_partials.resize((size_t)G.nblocks * NQ);
        HIPCHK(hipMemcpyDtoHAsync(c->h_partials.data(), G.d_partials, c->h_partials.size() * 8, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < NQ; q++) tot[q] = 0.0;
        for (uint64_t b = 0; b < G.nblocks; b++)
          for (int q = 0; q < NQ; q++) tot[q] += c->h_partials[b * NQ + q];
      }
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      out->kernel_ms = ms;
      out->lanes = G.lanes;
      out->samples_per_lane = G.S;
      out->dropped = (uint64_t)tot[1];
      for (int k = 0; k < K; k++) {
        out->n[k][0] = tot[0];
        out->s1[k][0] = tot[2 + 4 * k];
        out->s2[k][0] = tot[3 + 4 * k];
        out->n[k][1] = tot[0];
        out->s1[k][1] = tot[4 + 4 * k];
        out->s2[k][1] = tot[5 + 4 * k];
      }
    }
    if (p->prep.path != 2)
      for (int k = 0; k < K; k++) {
        /* the reference's two zero updates of the imaginary stats */
        out->n[k][1] = 2;
        out->s1[k][1] = 0;
        out->s2[k][1] = 0;
      }
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return FTR_OK;
}

int ftr_integrate(ftr_ctx* c, ftr_problem* p, uint64_t N, uint64_t seed, const ftr_opts* opts, ftr_moments* out) {
  return integrate_impl(c, p, N, seed, opts, out, false);
}

int ftr_integrate_noalloc(ftr_ctx* c, ftr_problem* p, uint64_t N, uint64_t seed, const ftr_opts* opts,
                          ftr_moments* out) {
  return integrate_impl(c, p, N, seed, opts, out, true);
}

int ftr_finalize(const ftr_problem* p, const ftr_moments* m, double* res) {
  if (!p || !m || !res) return FTR_ERR_INVALID;
  for (int k = 0; k < m->K; k++)
    for (int q = 0; q < 2; q++) {
      const double N = m->n[k][q];
      const double avg = m->s1[k][q] / N;
      const double var = (N / (N - 1)) * (m->s2[k][q] / N - std::pow(avg, 2));
      res[4 * k + 2 * q] = avg;
      res[4 * k + 2 * q + 1] = std::sqrt(var / N);
    }
  return FTR_OK;
}

int ftr_sample(ftr_ctx* c, ftr_problem* p, uint64_t first, uint64_t n, uint64_t seed, double* X_dev,
               double* psixi_dev) {
  if (!c || !p) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    if (n == 0) return FTR_OK;
    Geometry G = setup_stream(p, first, n, seed, 0, p->var.block, 2 + 4 * p->prep.K, nullptr);
    ftr::Launch a;
    memset(&a, 0, sizeof a);
    fill_tables(a, p);
    a.block_state = G.d_states;
    a.lane_poly = G.d_polys;
    a.N = n;
    a.S = G.S;
    a.partials = G.d_partials;
    void* args[] = {&a, &X_dev, &psixi_dev};
    HIPCHK(hipModuleLaunchKernel(p->m.f_smp, (unsigned)G.nblocks, 1, 1, p->var.block, 1, 1, (unsigned)p->lds,
                                 c->stream, args, nullptr));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return FTR_OK;
}

int ftr_replay(ftr_ctx* c, ftr_problem* p, uint64_t n, const double* X_dev, const double* psixi_dev,
               double* out_dev, uint8_t* dropped_dev) {
  if (!c || !p) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    if (n == 0) return FTR_OK;
    const ftr::DevParams* prm = p->d_prm;
    void* args[] = {(void*)&prm, &n, (void*)&X_dev, (void*)&psixi_dev, &out_dev, &dropped_dev};
    /* two lanes per sample: blocks of the variant's size, the second
     * derivatives of its block's samples in dynamic LDS */
    const bool pair = p->var.mode == 2 && p->var.lps == 2;
    const int E = p->prep.g.E;
    const unsigned threads = pair ? (unsigned)p->var.block : 256u;
    const uint64_t lanes = pair ? 2 * n : n;
    const size_t lds = pair ? (size_t)(E * (E + 1) / 2) * (p->var.block / 2) * 8 : 0;
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)p->m.f_rep, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const uint64_t nb = (lanes + threads - 1) / threads;
    HIPCHK(hipModuleLaunchKernel(p->m.f_rep, (unsigned)nb, 1, 1, threads, 1, 1, (unsigned)lds, c->stream, args,
                                 nullptr));
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return FTR_OK;
}

/* Batched scan: one persistent launch per bucket of problems.
 *
 * A bucket gathers the problems whose kernels share (path, eps orders, table
 * placement, block size) and the graph size (E, V).  Its kernel is generated from the graph specs of all of
 * its topologies (each one compile-time specialised, dispatched by a switch)
 * and runs a grid sized to the device; the blocks take units of work --
 * (problem, virtual block) -- from an atomic counter, the costliest first,
 * until the list is exhausted.  Every unit is the block of one-call
 * integration it stands for: the same stream positions and the same
 * partial moments, so every problem's result is that of its own call up to
 * the summation order.  The counter reset, all bucket kernels and the
 * copy-back of the partials are captured in one hipGraph; a call with the
 * same launch shapes replays it. */
namespace {

/* size key of a bucket: the edge and vertex counts, so that the topologies
 * that share a bucket kernel have the same matrix sizes and register and LDS
 * demand (a kernel's occupancy is that of its largest case) */
int size_class(const ftrh::Graph& g) { return g.E * 64 + g.V; }

/* relative cost of one sample (ordering of the work list only): the
 * algorithmic FP64 flops plus the draws and transcendentals of the sampler */
double sample_cost(const ftrh::Prep& pr) {
  const double E = pr.g.E, n = pr.g.V - 1, K = pr.K;
  double f = E * (E + 1) + 4 * E + 40 * (2 * E - 1);
  if (pr.path == 2)
    f += n * n * n / 3 + 2 * n * n * n + 4 * n * n * n + 5.5 * E * E + 8.0 / 3 * n * n * n + 8 * n * n * n +
         8.0 / 3 * E * E * E + 2 * E * E + 40 * E + 8 * K;
  else
    f += n * n * n / 3 + 2 * n * n * n + n * n + (n + 1) * (n + 1) + 12 * E + 3 * K;
  return f;
}

struct Bucket {
  int mode = 0, K = 1, tmode = 0, block = 256;
  std::vector<int> idx;                  /* problems with N > 0 */
  std::vector<int> topo;                 /* case of each problem in the bucket kernel */
  std::vector<std::string> specs;        /* one graph spec per case */
  std::vector<std::pair<int, size_t>> polys; /* (E, offset in the polynomial area) */
  hipFunction_t f = nullptr;
  size_t lds = 0;
  uint64_t S = 0;
  uint32_t grid = 0, nwork = 0;
  size_t item_off = 0, work_off = 0;
};

}  // namespace

/* Buckets of a batch: members[i] = (preparation, generic variant) of every
 * problem with N > 0, in order; the cases of a bucket are numbered in order
 * of first appearance.  lds_of[i]: the member's LDS bytes. */
static std::vector<Bucket> plan_buckets(const std::vector<std::pair<const ftrh::Prep*, Variant>>& members,
                                        const std::vector<int>& index, std::vector<size_t>& lds_of) {
  std::map<std::tuple<int, int, int, int, int>, Bucket> bmap;
  lds_of.assign(members.size(), 0);
  for (size_t m = 0; m < members.size(); m++) {
    const ftrh::Prep& pr = *members[m].first;
    const Variant& v = members[m].second;
    Bucket& b = bmap[std::make_tuple(v.mode, v.K, v.tmode, v.block, size_class(pr.g))];
    b.mode = v.mode;
    b.K = v.K;
    b.tmode = v.tmode;
    b.block = v.block;
    const std::string spec = spec_source(pr, v, false, "@", false);
    int t = -1;
    for (size_t q = 0; q < b.specs.size(); q++)
      if (b.specs[q] == spec) t = (int)q;
    if (t < 0) {
      t = (int)b.specs.size();
      b.specs.push_back(spec);
    }
    b.idx.push_back(index[m]);
    b.topo.push_back(t);
    lds_of[m] = variant_lds(v, pr.g.E);
    b.lds = std::max(b.lds, lds_of[m]);
  }
  std::vector<Bucket> bs;
  for (auto& kv : bmap) {
    Bucket b = std::move(kv.second);
    /* canonical case order (sorted specs): a set of topologies always gives
     * the same kernel source, whichever problems brought them */
    {
      std::vector<int> perm(b.specs.size());
      for (size_t t = 0; t < perm.size(); t++) perm[t] = (int)t;
      std::sort(perm.begin(), perm.end(), [&](int x, int y) { return b.specs[x] < b.specs[y]; });
      std::vector<int> rank(perm.size());
      std::vector<std::string> sorted(perm.size());
      for (size_t r = 0; r < perm.size(); r++) {
        rank[perm[r]] = (int)r;
        sorted[r] = b.specs[perm[r]];
      }
      b.specs = std::move(sorted);
      for (int& t : b.topo) t = rank[t];
    }
    for (size_t t = 0; t < b.specs.size(); t++) {
      std::string& sp = b.specs[t];
      const std::string nm = "FG" + std::to_string(t);
      for (size_t pos; (pos = sp.find('@')) != std::string::npos;) sp.replace(pos, 1, nm);
    }
    bs.push_back(std::move(b));
  }
  return bs;
}

/* kernel source of a bucket: every case's graph spec and the persistent loop */
static std::string bucket_source(const Bucket& b) {
  std::ostringstream src;
  src << ftr_embedded_device_source;
  for (size_t t = 0; t < b.specs.size(); t++) src << b.specs[t];
  src << "\nFTR_BUCKET_BEGIN(ftrb_run, " << b.block << ")\n";
  for (size_t t = 0; t < b.specs.size(); t++)
    src << "FTR_BUCKET_CASE(" << t << ", FG" << t << ", " << b.K << ", " << b.mode << ", " << b.tmode << ")\n";
  src << "FTR_BUCKET_END\n";
  return src.str();
}

/* side streams and events of the batch graph's fork and join (created once
 * per context, default priority) */
static void ensure_side_streams(ftr_ctx* c) {
  if (c->side[0]) return;
  for (int k = 0; k < FTR_SIDE_STREAMS; k++) {
    HIPCHK(hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
}

static hipFunction_t bucket_kernel(ftr_ctx* c, const Bucket& b) {
  hipModule_t mod = load_module_raw(c, bucket_source(b));
  hipFunction_t f = nullptr;
  HIPCHK(hipModuleGetFunction(&f, mod, "ftrb_run"));
  if (b.lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.lds);
  return f;
}

int ftr_precompile_batch(int n, const ftr_input* in, char* err, int errlen) {
  try {
    if (n < 0 || (n > 0 && !in)) throw Error(FTR_ERR_INVALID, "invalid batch");
    std::vector<ftrh::Prep> prs(n);
    std::vector<char> ok(n, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; i++) {
      try {
        prep_from_input(prs[i], &in[i]);
        ok[i] = 1;
      } catch (std::exception&) {
        ok[i] = 0; /* rejected problems are not part of a batch */
      }
    }
    /* both LDS limits a gfx950 part may report */
    for (size_t lim : {(size_t)64 * 1024, (size_t)160 * 1024}) {
      std::vector<std::pair<const ftrh::Prep*, Variant>> members;
      std::vector<int> index;
      for (int i = 0; i < n; i++) {
        if (!ok[i]) continue;
        Variant v = batch_member_variant(choose_variant(prs[i], lim));
        members.emplace_back(&prs[i], v);
        index.push_back(i);
      }
      std::vector<size_t> lds_of;
      std::vector<Bucket> bs = plan_buckets(members, index, lds_of);
      std::vector<std::string> srcs;
      for (auto& b : bs) srcs.push_back(bucket_source(b));
#pragma omp parallel for schedule(dynamic, 1)
      for (long q = 0; q < (long)srcs.size(); q++) code_object(srcs[q], nullptr);
    }
    code_object(ftr_embedded_jtable_source, nullptr);
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

int ftr_integrate_batch(ftr_ctx* c, int n, ftr_problem* const* probs, const uint64_t* Ns, const uint64_t* seeds,
                        ftr_moments* out) {
  if (!c || n < 0 || (n > 0 && (!probs || !Ns || !seeds || !out))) return FTR_ERR_INVALID;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(c->device));
    for (int i = 0; i < n; i++) {
      if (!probs[i] || probs[i]->ctx != c || !probs[i]->m.f_int)
        throw Error(FTR_ERR_INVALID, "batch: every problem must be prepared on this context");
      memset(&out[i], 0, sizeof out[i]);
      out[i].K = probs[i]->prep.K;
    }
    /* buckets; within a bucket one case per distinct graph spec */
    std::vector<std::pair<const ftrh::Prep*, Variant>> members;
    std::vector<int> index;
    for (int i = 0; i < n; i++) {
      if (Ns[i] == 0) continue;
      const Variant v = batch_member_variant(probs[i]->var);
      members.emplace_back(&probs[i]->prep, v);
      index.push_back(i);
    }
    std::vector<size_t> lds_of;
    std::vector<Bucket> bs = plan_buckets(members, index, lds_of);
    for (auto& b : bs) b.f = bucket_kernel(c, b); /* loaded once per context and bucket composition */
    size_t n_items = 0, n_states = 0, n_polys = 0, n_part = 0, n_work = 0;
    std::vector<size_t> state_off(n, 0), part_off(n, 0);
    std::vector<uint32_t> nb(n, 0);
    for (auto& b : bs) {
      int nbpc = 0;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nbpc, b.f, b.block, b.lds) != hipSuccess || nbpc < 1)
        nbpc = 1;
      const uint64_t lanes_target = (uint64_t)c->prop.multiProcessorCount * nbpc * b.block;
      uint64_t tot = 0;
      for (int i : b.idx) tot += Ns[i];
      /* enough samples per lane to amortise the stream positioning, enough
       * lanes for every problem */
      b.S = std::max<uint64_t>(16, (tot + lanes_target - 1) / lanes_target);
      b.item_off = n_items;
      n_items += b.idx.size();
      int lb = 0;
      while ((1 << lb) < b.block) lb++;
      for (int i : b.idx) {
        const int E = probs[i]->prep.g.E;
        bool have = false;
        for (auto& pe : b.polys) have = have || pe.first == E;
        if (!have) {
          b.polys.emplace_back(E, n_polys);
          n_polys += std::max<size_t>(lane_poly_count(lb), 1);
        }
        const uint64_t lanes = (Ns[i] + b.S - 1) / b.S;
        nb[i] = (uint32_t)((lanes + b.block - 1) / b.block);
        state_off[i] = n_states;
        n_states += nb[i];
        part_off[i] = n_part;
        n_part += (size_t)nb[i] * (2 + 4 * probs[i]->prep.K);
        b.nwork += nb[i];
      }
      b.work_off = n_work;
      n_work += b.nwork;
      b.grid = (uint32_t)std::min<uint64_t>((uint64_t)c->prop.multiProcessorCount * nbpc, b.nwork);
    }
    /* device layout: items | work | counters | states | polys | partials */
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_items = al(n_items * sizeof(ftr::BatchItem)), b_work = al(n_work * sizeof(ftr::BatchWork) + 16),
                 b_cnt = al(bs.size() * 4 + 4), b_states = al(n_states * 32), b_polys = al(n_polys * 32 + 32),
                 b_part = al(n_part * 8 + 8);
    const size_t need = b_items + b_work + b_cnt + b_states + b_polys + b_part;
    if (need > c->batch_bytes) {
      if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
      c->batch_exec = nullptr;
      if (c->d_batch) HIPCHK(hipFree(c->d_batch));
      c->d_batch = nullptr;
      HIPCHK(hipMalloc(&c->d_batch, need));
      c->batch_bytes = need;
    }
    char* base = (char*)c->d_batch;
    ftr::BatchItem* d_items = (ftr::BatchItem*)base;
    ftr::BatchWork* d_work = (ftr::BatchWork*)(base + b_items);
    uint32_t* d_cnt = (uint32_t*)(base + b_items + b_work);
    uint64_t* d_states = (uint64_t*)(base + b_items + b_work + b_cnt);
    uint64_t* d_polys = (uint64_t*)(base + b_items + b_work + b_cnt + b_states);
    double* d_part = (double*)(base + b_items + b_work + b_cnt + b_states + b_polys);
    std::vector<ftr::BatchItem> items(n_items);
    std::vector<ftr::BatchWork> work(n_work);
    std::vector<uint64_t> states(4 * n_states), polys(4 * n_polys, 0);
    for (auto& b : bs) {
      int lb = 0;
      while ((1 << lb) < b.block) lb++;
      /* lane polynomials per edge count, and the jump of one block */
      std::map<int, ftrh::Poly> qblock;
      std::map<int, size_t> poff;
      for (auto& pe : b.polys) {
        const uint64_t per = 2 * (uint64_t)pe.first - 1;
        const unsigned __int128 dl = (unsigned __int128)b.S * per;
        qblock[pe.first] =
            lane_digit_polys(ftrh::xo_jump_poly((uint64_t)dl, (uint64_t)(dl >> 64)), lb, &polys[4 * pe.second]);
        poff[pe.first] = pe.second;
      }
      /* the work list, costliest units first */
      std::vector<std::pair<double, ftr::BatchWork>> units;
      units.reserve(b.nwork);
      for (size_t t = 0; t < b.idx.size(); t++) {
        const int i = b.idx[t];
        ftr_problem* p = probs[i];
        ftr::BatchItem& it = items[b.item_off + t];
        memset(&it, 0, sizeof it);
        fill_tables(it.a, p);
        it.a.block_state = d_states + 4 * state_off[i];
        it.a.lane_poly = d_polys + 4 * poff[p->prep.g.E];
        it.a.N = Ns[i];
        it.a.S = b.S;
        it.a.partials = d_part + part_off[i];
        it.block0 = 0;
        it.nblocks = nb[i];
        const double cs = sample_cost(p->prep);
        const uint64_t per_block = b.S * (uint64_t)b.block;
        for (uint32_t k = 0; k < nb[i]; k++) {
          const uint64_t first = (uint64_t)k * per_block;
          const uint64_t ns = std::min<uint64_t>(per_block, Ns[i] - first);
          ftr::BatchWork w;
          w.item = (uint32_t)(b.item_off + t);
          w.blk = k;
          w.topo = (uint32_t)b.topo[t];
          w.pad = 0;
          units.emplace_back(cs * (double)ns, w);
        }
      }
      std::stable_sort(units.begin(), units.end(),
                       [](const std::pair<double, ftr::BatchWork>& x, const std::pair<double, ftr::BatchWork>& y) {
                         return x.first > y.first;
                       });
      for (size_t u = 0; u < units.size(); u++) work[b.work_off + u] = units[u].second;
      /* stream start state of every block of every problem */
#pragma omp parallel for schedule(dynamic, 1)
      for (long t = 0; t < (long)b.idx.size(); t++) {
        const int i = b.idx[t];
        const ftrh::Poly& qb = qblock[probs[i]->prep.g.E];
        uint64_t st[4];
        ftrh::xo_seed(seeds[i], st);
        for (uint32_t k = 0; k < nb[i]; k++) {
          memcpy(&states[4 * (state_off[i] + k)], st, 32);
          if (k + 1 < nb[i]) ftrh::xo_apply(qb, st);
        }
      }
    }
    if (n_items) {
      HIPCHK(hipMemcpyHtoDAsync(d_items, items.data(), items.size() * sizeof(ftr::BatchItem), c->stream));
      HIPCHK(hipMemcpyHtoDAsync(d_work, work.data(), work.size() * sizeof(ftr::BatchWork), c->stream));
      HIPCHK(hipMemcpyHtoDAsync(d_states, states.data(), states.size() * 8, c->stream));
      HIPCHK(hipMemcpyHtoDAsync(d_polys, polys.data(), polys.size() * 8, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if (n_part * 8 > c->pinned_bytes) {
        if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
        c->batch_exec = nullptr;
        if (c->h_pinned) HIPCHK(hipHostFree(c->h_pinned));
        c->h_pinned = nullptr;
        HIPCHK(hipHostMalloc((void**)&c->h_pinned, n_part * 8, hipHostMallocDefault));
        c->pinned_bytes = n_part * 8;
      }
      double* h_part = c->h_pinned;
      /* one graph: the counter reset, every bucket kernel, then the
       * copy-back.  Its nodes depend only on the launch shapes and buffer
       * addresses (items, work lists, states and polynomials are uploaded
       * above, outside the graph), so a call with the same shapes replays
       * the previous graph */
      std::ostringstream key;
      key << (const void*)c->d_batch << ' ' << (const void*)h_part << ' ' << n_part << ' ' << bs.size();
      for (auto& b : bs)
        key << ' ' << (const void*)b.f << ':' << b.grid << ':' << b.nwork << ':' << b.block << ':' << b.lds << ':'
            << b.work_off;
      if (!c->batch_exec || c->batch_key != key.str()) {
        if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
        c->batch_exec = nullptr;
        c->batch_key.clear();
        hipFunction_t fz = nullptr;
        {
          hipModule_t mz = load_module_raw(c, ftr_embedded_jtable_source);
          HIPCHK(hipModuleGetFunction(&fz, mz, "ftrj_zero_u32"));
        }
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        {
          int nc = (int)bs.size();
          void* args[] = {&d_cnt, &nc};
          HIPCHK(hipModuleLaunchKernel(fz, (unsigned)((nc + 63) / 64), 1, 1, 64, 1, 1, 0, c->stream, args, nullptr));
        }
        /* the buckets run concurrently: forked over the side streams
         * (default priority) and joined before the copy-back, so that one
         * bucket's tail overlaps the next bucket's work */
        ensure_side_streams(c);
        HIPCHK(hipEventRecord(c->ev_fork, c->stream));
        for (int k = 0; k < FTR_SIDE_STREAMS; k++) HIPCHK(hipStreamWaitEvent(c->side[k], c->ev_fork, 0));
        for (size_t q = 0; q < bs.size(); q++) {
          Bucket& b = bs[q];
          const ftr::BatchItem* it = d_items;
          const ftr::BatchWork* wk = d_work + b.work_off;
          uint32_t nw = b.nwork;
          uint32_t* cnt = d_cnt + q;
          void* args[] = {(void*)&it, (void*)&wk, &nw, &cnt};
          HIPCHK(hipModuleLaunchKernel(b.f, b.grid, 1, 1, b.block, 1, 1, (unsigned)b.lds,
                                       c->side[q % FTR_SIDE_STREAMS], args, nullptr));
        }
        for (int k = 0; k < FTR_SIDE_STREAMS; k++) {
          HIPCHK(hipEventRecord(c->ev_join[k], c->side[k]));
          HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join[k], 0));
        }
        HIPCHK(hipMemcpyDtoHAsync(h_part, d_part, n_part * 8, c->stream));
        HIPCHK(hipStreamEndCapture(c->stream, &graph));
        const hipError_t ie = hipGraphInstantiate(&c->batch_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(ie);
        c->batch_key = key.str();
      } else {
        c->batch_replays++;
      }
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      HIPCHK(hipGraphLaunch(c->batch_exec, c->stream));
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      for (auto& b : bs)
        for (int i : b.idx) {
          const int K = probs[i]->prep.K, NQ = 2 + 4 * K;
          std::vector<double> tot(NQ, 0.0);
          for (uint32_t k = 0; k < nb[i]; k++)
            for (int q = 0; q < NQ; q++) tot[q] += h_part[part_off[i] + (size_t)k * NQ + q];
          ftr_moments& m = out[i];
          m.dropped = (uint64_t)tot[1];
          for (int k = 0; k < K; k++) {
            m.n[k][0] = m.n[k][1] = tot[0];
            m.s1[k][0] = tot[2 + 4 * k];
            m.s2[k][0] = tot[3 + 4 * k];
            m.s1[k][1] = tot[4 + 4 * k];
            m.s2[k][1] = tot[5 + 4 * k];
          }
          m.kernel_ms = ms; /* the whole graph */
          m.lanes = (uint64_t)nb[i] * b.block;
          m.samples_per_lane = b.S;
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; i++) {
      if (probs[i]->prep.path != 2)
        for (int k = 0; k < probs[i]->prep.K; k++) {
          out[i].n[k][1] = 2;
          out[i].s1[k][1] = 0;
          out[i].s2[k][1] = 0;
        }
      out[i].seconds = secs;
    }
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return FTR_OK;
}

/* Switch p to the feasible variant (table_mode, block, lps), keeping its
 * specialisation; false when no candidate matches. */
static bool apply_variant(ftr_ctx* c, ftr_problem* p, int table_mode, int block, int lps) {
  for (Variant v : candidate_variants(p->prep, c->prop.sharedMemPerBlock, rows_mem_limit(c)))
    if (v.tmode == table_mode && v.block == block && v.lps == lps) {
      v.cparams = p->var.cparams;
      Module m = module_for(c, p->prep, v);
      if (v.tmode >= 3) device_rows(c, p, v.tmode);
      p->m = m;
      p->var = v;
      p->lds = variant_lds(v, p->prep.g.E);
      if (p->lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)p->m.f_int, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds);
      return true;
    }
  return false;
}

int ftr_set_variant(ftr_ctx* c, ftr_problem* p, int table_mode, int block) {
  if (!c || !p || !p->m.f_int) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    if (apply_variant(c, p, table_mode, block, p->var.lps)) return FTR_OK;
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return fail(c, FTR_ERR_LIMIT, "infeasible (table_mode, block) for this problem at " + std::to_string(p->var.lps) +
                                    " lane(s) per sample");
}

int ftr_set_lanes_per_sample(ftr_ctx* c, ftr_problem* p, int lps) {
  if (!c || !p || !p->m.f_int) return FTR_ERR_INVALID;
  if (lps != 1 && lps != 2) return fail(c, FTR_ERR_INVALID, "lanes per sample must be 1 or 2");
  if (lps == p->var.lps) return FTR_OK;
  if (lps == 2 && p->prep.path != 2)
    return fail(c, FTR_ERR_LIMIT, "two lanes per sample apply to the deformed (complex) path only");
  try {
    HIPCHK(hipSetDevice(c->device));
    if (lps == 1) {
      const Variant d = choose_one_lane(p->prep, c->prop.sharedMemPerBlock, rows_mem_limit(c));
      if (apply_variant(c, p, d.tmode, d.block, 1)) return FTR_OK;
    } else {
      /* rows in HBM (the current row placement, else table mode 3); the
       * current block size when its LDS fits, else 256, 128, 512 */
      const int tm = p->var.tmode >= 3 ? p->var.tmode : 3;
      for (int blk : {p->var.block, 256, 128, 512})
        if (apply_variant(c, p, tm, blk, 2)) return FTR_OK;
    }
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return fail(c, FTR_ERR_LIMIT, "no feasible variant at " + std::to_string(lps) + " lane(s) per sample");
}

int ftr_get_lanes_per_sample(const ftr_problem* p, int* lps) {
  if (!p || !lps) return FTR_ERR_INVALID;
  *lps = p->var.lps;
  return FTR_OK;
}

int ftr_autotune(ftr_ctx* c, ftr_problem* p, uint64_t n_probe, int* table_mode, int* block, double* best_ms) {
  if (!c || !p || !p->m.f_int) return FTR_ERR_INVALID;
  const Variant keep = p->var;
  auto probe = [&]() {
    ftr_moments m;
    double t = 0;
    for (int rep = 0; rep < 2; rep++) {
      if (ftr_integrate(c, p, n_probe, 12345, nullptr, &m) != FTR_OK) return 1e300;
      t = m.kernel_ms;
    }
    return t;
  };
  double best = 1e300;
  Variant bestv = keep;
  /* (table placement, block) with the generic kernel ... */
  if (ftr_set_specialized(c, p, 0) != FTR_OK) return FTR_ERR_DEVICE;
  for (const Variant& v : candidate_variants(p->prep, c->prop.sharedMemPerBlock, rows_mem_limit(c), true)) {
    try {
      HIPCHK(hipSetDevice(c->device));
      if (!apply_variant(c, p, v.tmode, v.block, v.lps)) continue;
    } catch (std::exception&) {
      continue;
    }
    const double t = probe();
    if (t < best) {
      best = t;
      bestv = p->var;
    }
  }
  try {
    HIPCHK(hipSetDevice(c->device));
    if (best >= 1e300) {
      apply_variant(c, p, keep.tmode, keep.block, keep.lps);
      ftr_set_specialized(c, p, keep.cparams);
      return fail(c, FTR_ERR_DEVICE, "autotune: no variant ran");
    }
    if (!apply_variant(c, p, bestv.tmode, bestv.block, bestv.lps))
      return fail(c, FTR_ERR_DEVICE, "autotune: the chosen variant is no longer feasible");
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  {
    /* the placement, block and lanes per sample are recorded per graph;
     * specialisation to a kinematic point is not, since it costs a compile
     * per point */
    mkdir(cache_dir().c_str(), 0755);
    std::ofstream f(tune_path(p->prep));
    if (f) f << bestv.tmode << " " << bestv.block << " " << bestv.lps << "\n";
    c->tune_memo.erase(tune_path(p->prep));
  }
  /* ... then the kernel specialised to this kinematic point */
  if (ftr_set_specialized(c, p, 1) == FTR_OK) {
    const double t = probe();
    if (t < best) best = t;
    else ftr_set_specialized(c, p, 0);
  }
  if (table_mode) *table_mode = bestv.tmode;
  if (block) *block = bestv.block;
  if (best_ms) *best_ms = best;
  return FTR_OK;
}

int ftr_batch_graph_replays(const ftr_ctx* c) { return c ? c->batch_replays : 0; }

int ftr_get_specialized(const ftr_problem* p, int* on) {
  if (!p || !on) return FTR_ERR_INVALID;
  *on = p->var.cparams;
  return FTR_OK;
}

int ftr_get_variant(const ftr_problem* p, int* table_mode, int* block) {
  if (!p) return FTR_ERR_INVALID;
  if (table_mode) *table_mode = p->var.tmode;
  if (block) *block = p->var.block;
  return FTR_OK;
}

void ftr_free_string(char* s) { free(s); }

int ftr_set_specialized(ftr_ctx* c, ftr_problem* p, int on) {
  if (!c || !p || !p->m.f_int) return FTR_ERR_INVALID;
  try {
    HIPCHK(hipSetDevice(c->device));
    Variant v = p->var;
    v.cparams = on ? 1 : 0;
    p->m = module_for(c, p->prep, v);
    p->var = v;
    p->lds = variant_lds(v, p->prep.g.E);
    if (p->lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)p->m.f_int, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds);
  } catch (std::exception& e) {
    const Error* fe = dynamic_cast<const Error*>(&e);
    return fail(c, fe ? fe->code : FTR_ERR_DEVICE, e.what());
  }
  return FTR_OK;
}

int ftr_use_cached_specialization(ftr_ctx* c, ftr_problem* p) {
  if (!c || !p || !p->m.f_int || p->var.cparams) return FTR_OK;
  const char* e = getenv("FEYNTROP_AMD_SPECIALIZE");
  if (e && !strcmp(e, "0")) return FTR_OK;
  const bool always = e && !strcmp(e, "1");
  if (!always) {
    Variant v = p->var;
    v.cparams = 1;
    std::ifstream f(code_object_path(full_source(p->prep, v)), std::ios::binary);
    if (!f) return FTR_OK;
  }
  return ftr_set_specialized(c, p, 1);
}

int ftr_set_jtable_mode(ftr_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 2) return FTR_ERR_INVALID;
  c->jtable_mode = mode;
  return FTR_OK;
}

}  // extern "C"

const std::vector<std::string>& ftr_problem_log(const ftr_problem* p) { return p->prep.log; }
