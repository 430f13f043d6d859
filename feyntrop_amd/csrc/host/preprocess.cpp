/*
 * preprocess.cpp -- host-side problem preparation of the MI355X engine:
 * graph validation, kinematic-regime detection and the J-table.
 *
 * Follows (michibo/feyntrop):
 *   graph construction + checks          srcs/graph.hpp:169-214
 *   scalar-product checks, regime        srcs/feyntrop.cpp:98-133
 *   J-table, mm/genericity/GP flags      srcs/tropical_sampling.hpp:26-458
 *   fatal flag combinations, lambda_eff  srcs/feyntrop.cpp:165-239
 *   real/complex path selection          srcs/feynman_integral.hpp:125-140
 * The table is built level-synchronously by edge count (every subset of
 * size n depends only on subsets of size n-1), OpenMP-parallel per level.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>

#include "ftr_host.hpp"

namespace ftrh {

static std::string graph_str(const Graph& g) {
  std::stringstream s;
  s << "G(";
  for (int j = 0; j < g.E; j++) {
    double c = g.nu[j];
    s << "[{" << g.hi[j] << "," << g.lo[j] << "},";
    if (std::fabs(std::nearbyint(c) / c - 1.0) < 1e-12) s << std::lrint(c);
    else s << c;
    s << "]";
    if (j != g.E - 1) s << ",";
  }
  s << ")";
  return s.str();
}

void build_graph(Graph& g, int E, const int32_t* edges, const double* nu) {
  g.E = E;
  g.raw.assign(edges, edges + 2 * E);
  g.hi.resize(E);
  g.lo.resize(E);
  g.nu.assign(nu, nu + E);
  g.adj.clear();
  for (int j = 0; j < E; j++) {
    const int k = edges[2 * j], l = edges[2 * j + 1];
    g.hi[j] = std::max(k, l);
    g.lo[j] = std::min(k, l);
    if (nu[j] <= 1e-12) throw Error(-1, "graph: edge weights must be larger than 0");
    if (k < 0 || l < 0) throw Error(-1, "graph: vertex numbers must start from 0");
    if (k == l) throw Error(-1, "graph: has a tadpole (selfloop) edge. This program cannot take care of this. ");
    const size_t Vp = (size_t)std::max(k, l) + 1;
    if (Vp > g.adj.size()) g.adj.resize(Vp);
    g.adj[k].push_back({j, l});
    g.adj[l].push_back({j, k});
  }
  g.V = (int)g.adj.size();
  for (int i = 0; i < g.V; i++)
    if (g.adj[i].empty()) {
      std::stringstream s;
      s << "graph: the graph has a singleton vertex " << i;
      throw Error(-1, s.str());
    }
}

/* Components of (V, edges in sub); labels in order of the lowest vertex
 * (components.hpp:92-120). */
static int components(const Graph& g, uint64_t sub, int* cmap) {
  int stack[64];
  for (int i = 0; i < g.V; i++) cmap[i] = -1;
  int C = 0;
  for (int i = 0; i < g.V; i++) {
    if (cmap[i] >= 0) continue;
    int sp = 0;
    stack[sp++] = i;
    cmap[i] = C;
    while (sp) {
      const int v = stack[--sp];
      for (auto& a : g.adj[v])
        if (((sub >> a.first) & 1) && cmap[a.second] < 0) {
          cmap[a.second] = C;
          stack[sp++] = a.second;
        }
    }
    C++;
  }
  return C;
}

/* dense symmetric helpers on small matrices (row-major, n x n) */
struct Small {
  int n = 0;
  double a[32 * 32];
  double& operator()(int i, int j) { return a[i * n + j]; }
  double operator()(int i, int j) const { return a[i * n + j]; }
};

/* LDL^T with largest-diagonal pivoting (the same pivot rule as the
 * reference's Eigen LDLT), used for the mass-momentum-spanning test.
 * Returns tr(A^-1 B) for symmetric B. */
static double ldlt_trace_solve(Small& A, const Small& B) {
  const int n = A.n;
  int perm[32];
  for (int i = 0; i < n; i++) perm[i] = i;
  /* symmetric pivoted factorisation on the full matrix */
  for (int k = 0; k < n; k++) {
    int big = k;
    double bv = std::fabs(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A(i, i)) > bv) { bv = std::fabs(A(i, i)); big = i; }
    if (big != k) {
      std::swap(perm[k], perm[big]);
      for (int j = 0; j < n; j++) std::swap(A(k, j), A(big, j));
      for (int i = 0; i < n; i++) std::swap(A(i, k), A(i, big));
    }
    const double d = A(k, k);
    if (!(std::fabs(d) > 0.0)) continue;
    for (int i = k + 1; i < n; i++) {
      const double l = A(i, k) / d;
      for (int j = k + 1; j <= i; j++) A(i, j) -= l * A(j, k);
      for (int j = k + 1; j <= i; j++) A(j, i) = A(i, j);
    }
    for (int i = k + 1; i < n; i++) A(i, k) /= d;
  }
  /* tr(A^-1 B) = sum_j (A^-1 B)_jj, columns of B permuted */
  double tr = 0.0;
  double y[32];
  const double tol = std::numeric_limits<double>::min();
  for (int c = 0; c < n; c++) {
    /* a zero column of B adds exactly zero */
    bool nz = false;
    for (int i = 0; i < n; i++) nz = nz || (B(i, c) != 0.0);
    if (!nz) continue;
    for (int i = 0; i < n; i++) y[i] = B(perm[i], c);
    for (int i = 0; i < n; i++)
      for (int q = 0; q < i; q++) y[i] -= A(i, q) * y[q];
    for (int i = 0; i < n; i++) y[i] = (std::fabs(A(i, i)) > tol) ? y[i] / A(i, i) : 0.0;
    for (int i = n - 1; i >= 0; i--)
      for (int q = i + 1; q < n; q++) y[i] -= A(q, i) * y[q];
    for (int i = 0; i < n; i++)
      if (perm[i] == c) tr += y[i];
  }
  return tr;
}

static bool is_approx(const std::vector<double>& a, const std::vector<double>& b) {
  const double prec = 1e-12;
  double d = 0, na = 0, nb = 0;
  for (size_t i = 0; i < a.size(); i++) {
    d += (a[i] - b[i]) * (a[i] - b[i]);
    na += a[i] * a[i];
    nb += b[i] * b[i];
  }
  return d <= prec * prec * std::min(na, nb);
}

/* cyclic Jacobi eigen-decomposition of a symmetric V x V matrix */
static void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& Q) {
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) Q[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off == 0.0) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          const double x = A[(size_t)k * n + p], y = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * x - s * y;
          A[(size_t)k * n + q] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = A[(size_t)p * n + k], y = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * x - s * y;
          A[(size_t)q * n + k] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = Q[(size_t)k * n + p], y = Q[(size_t)k * n + q];
          Q[(size_t)k * n + p] = c * x - s * y;
          Q[(size_t)k * n + q] = s * x + c * y;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; i++) w[i] = A[(size_t)i * n + i];
}

static double omega(const Graph& g, double D, int L, uint64_t sub) {
  double m = 0.0;
  for (int j = 0; j < g.E; j++)
    if ((sub >> j) & 1) m += g.nu[j];
  return m - D / 2.0 * L;
}

static std::string matrix_str(const std::vector<double>& P, int V) {
  std::stringstream s;
  for (int i = 0; i < V; i++) {
    for (int j = 0; j < V; j++) s << (j ? " " : "") << P[(size_t)i * V + j];
    s << "\n";
  }
  return s.str();
}

/* The J-table on the host (tropical_sampling.hpp:284-458): subsets level by
 * level (size n), each level in Gosper order, OpenMP over a level. */
void host_jtable(Prep& pr, const JtSetup& js, JtResult& res) {
  const Graph& g = pr.g;
  const int V = g.V, E = g.E;
  const std::vector<double>& P = pr.P;
  const double D = pr.D, W = js.W, W_ref = js.W_ref, pmsqr_ref = js.pmsqr_ref;
  const bool bPhiTr = js.phi, bNonEuclidean = js.noneucl;
  const int n_tests = js.n_tests;
  const double* Xs = js.Xs;
  const double* Xi = js.Xi;
  const size_t NT = (size_t)1 << E;
  pr.Jr.assign(NT, 0.0);
  pr.r.assign(NT, 0.0);
  pr.mm.assign(NT, 0);
  pr.Lt.assign(NT, 0);
  bool bGeneric = true, bPseudo = true, bGP = true;
  for (int n = 0; n <= E && res.div_level < 0; n++) {
    /* subsets of size n in Gosper order */
    std::vector<uint64_t> subs;
    uint64_t sub = (n == 0) ? 0 : ((UINT64_C(1) << n) - 1);
    while (true) {
      subs.push_back(sub);
      if (n == 0 || n == E) break;
      const uint64_t t = sub | (sub - 1);
      const uint64_t nx = (t + 1) | (((~t & -~t) - 1) >> (__builtin_ctzll(sub) + 1));
      if (nx >> E) break;
      sub = nx;
    }
    const long ns = (long)subs.size();
    long first_div = ns;
    bool lvGeneric = true, lvPseudo = true, lvGP = true;
#pragma omp parallel for schedule(dynamic, 16) reduction(&& : lvGeneric, lvPseudo, lvGP) reduction(min : first_div)
    for (long si = 0; si < ns; si++) {
      const uint64_t s = subs[si];
      int cmap[64];
      const int cV = components(g, s, cmap);
      bool mm = false;
      if (bPhiTr) {
        Small PGP;
        PGP.n = cV - 1;
        for (int q = 0; q < PGP.n * PGP.n; q++) PGP.a[q] = 0.0;
        for (int i = 0; i < V; i++)
          for (int j = 0; j < V; j++) {
            const int v = cmap[i], w = cmap[j];
            if (v == cV - 1 || w == cV - 1) continue;
            PGP(v, w) += P[(size_t)i * V + j];
          }
        if (cV == 2 && PGP(0, 0) / pmsqr_ref >= 1e-12) lvPseudo = false;
        if (cV == 3) {
          bool zero = true;
          for (int q = 0; q < 4; q++)
            if (!(std::fabs(PGP.a[q]) <= 1e-12)) zero = false;
          if (!zero) {
            const double a = PGP(0, 0), b = PGP(1, 1), gm = PGP(0, 1), c = a + b + 2 * gm;
            double A = 0, B = 0, Cc = 0;
            for (int j = 0; j < E; j++) {
              if ((s >> j) & 1) continue;
              const int kc = cmap[g.hi[j]], lc = cmap[g.lo[j]];
              if (kc == lc) continue;
              if (kc == 0 || lc == 0) A += pr.m2[j];
              if (kc == 1 || lc == 1) B += pr.m2[j];
              if (kc == 2 || lc == 2) Cc += pr.m2[j];
            }
            const double th = 1e-5;
            const double da = std::fabs(a - A) / pmsqr_ref, db = std::fabs(b - B) / pmsqr_ref,
                         dc = std::fabs(c - Cc) / pmsqr_ref;
            if ((da > 1e-10 && db < th && dc < th) || (db > 1e-10 && da < th && dc < th) ||
                (dc > 1e-10 && da < th && db < th))
              lvGeneric = false;
          }
        }
        /* mass-momentum spanning: pxi(X) == 0 for the random points
         * (tropical_sampling.hpp:46-82) */
        double sum = 0;
        for (int t = 0; t < n_tests; t++) {
          Small La;
          La.n = cV - 1;
          for (int q = 0; q < La.n * La.n; q++) La.a[q] = 0.0;
          for (int j = 0; j < E; j++) {
            if ((s >> j) & 1) continue;
            int k = cmap[g.hi[j]], l = cmap[g.lo[j]];
            if (k == l) continue;
            if (k < l) std::swap(k, l);
            const double x = Xi[(size_t)t * E + j];
            La(l, l) += x;
            if (k == cV - 1) continue;
            La(k, k) += x;
            La(k, l) -= x;
            La(l, k) -= x;
          }
          double pxi = -ldlt_trace_solve(La, PGP);
          for (int j = 0; j < E; j++)
            if (!((s >> j) & 1)) pxi += pr.m2[j] * Xs[(size_t)t * E + j];
          sum += std::fabs(pxi);
          /* the sum only grows: once over the threshold the flag is settled
           * (the reference sums all the test points; same flag) */
          if (!(sum / pmsqr_ref < 1e-12)) break;
        }
        mm = sum / pmsqr_ref < 1e-12;
      } else if (n == E) {
        mm = true;
      }
      const int Ls = n - V + cV;
      double Jr, r;
      if (n == 0) {
        Jr = 1.0;
        r = 1.0;
      } else {
        Jr = 0.0;
        for (int j = 0; j < E; j++)
          if ((s >> j) & 1) {
            const uint64_t c2 = s & ~(UINT64_C(1) << j);
            Jr += pr.Jr[c2] / pr.r[c2];
          }
        r = omega(g, D, Ls, s);
        if (mm) r -= W;
      }
      if (r / W_ref < 1e-12 && n != E) first_div = std::min(first_div, si);
      if (bPhiTr && bNonEuclidean && n >= 2) {
        /* tropical_sampling.hpp:152-228 */
        const int z = mm ? Ls + 1 : Ls;
        bool ok = true;
        for (int a = 0; a < E && ok; a++) {
          if (!((s >> a) & 1)) continue;
          const uint64_t A = s & ~(UINT64_C(1) << a);
          const int zA = pr.mm[A] ? pr.Lt[A] + 1 : pr.Lt[A];
          for (int b = a + 1; b < E; b++) {
            if (!((s >> b) & 1)) continue;
            const uint64_t AB = A & ~(UINT64_C(1) << b), B = s & ~(UINT64_C(1) << b);
            const int zB = pr.mm[B] ? pr.Lt[B] + 1 : pr.Lt[B];
            const int zAB = pr.mm[AB] ? pr.Lt[AB] + 1 : pr.Lt[AB];
            if (zA + zB > z + zAB) {
              ok = false;
              break;
            }
          }
        }
        lvGP = lvGP && ok;
      }
      pr.Jr[s] = Jr;
      pr.r[s] = r;
      pr.mm[s] = mm;
      pr.Lt[s] = (uint8_t)Ls;
    }
    bGeneric = bGeneric && lvGeneric;
    bPseudo = bPseudo && lvPseudo;
    bGP = bGP && lvGP;
    if (first_div < ns) {
      res.div_level = n;
      res.div_subset = subs[first_div];
    }
  }
  res.generic = bGeneric;
  res.pseudo = bPseudo;
  res.gp = bGP;
}

void preprocess(Prep& pr) {
  const auto t0 = std::chrono::steady_clock::now();
  const Graph& g = pr.g;
  const int V = g.V, E = g.E;
  const std::vector<double>& P = pr.P;
  const double D = pr.D;
  if ((int)pr.m2.size() != E) throw Error(-1, "input masses_sqr must be an array of size E");
  /* feyntrop.cpp:98-112 */
  {
    std::vector<double> Pt((size_t)V * V);
    for (int i = 0; i < V; i++)
      for (int j = 0; j < V; j++) Pt[(size_t)j * V + i] = P[(size_t)i * V + j];
    if (!is_approx(P, Pt))
      throw Error(-2, "integrate_graph: The matrix \n" + matrix_str(P, V) +
                          "\nof invariants was given as input and it is not symmetric or not a VxV matrix. One row and "
                          "column must be specified for each vertex - representing the momentum flowing into this vertex.");
    for (int j = 0; j < V; j++) {
      double s = 0;
      for (int i = 0; i < V; i++) s += P[(size_t)i * V + j];
      if (!(std::fabs(s) <= 1e-12))
        throw Error(-2, "Scalar product (P) matrix has to fulfill momentum conservation. That means all row and column "
                        "sums need to be 0. The matrix \n" + matrix_str(P, V) + "\ndoes not fulfill this requirement.");
    }
  }
  /* feyntrop.cpp:128-133: Euclidean iff P == -|P| */
  {
    std::vector<double> w, Q;
    sym_eig(V, P, w, Q);
    std::vector<double> Pabs((size_t)V * V, 0.0);
    for (int i = 0; i < V; i++)
      for (int j = 0; j < V; j++) {
        double s = 0;
        for (int k = 0; k < V; k++) s += Q[(size_t)i * V + k] * std::fabs(w[k]) * Q[(size_t)j * V + k];
        Pabs[(size_t)i * V + j] = -s;
      }
    bool zero = true;
    for (double v : P)
      if (!(std::fabs(v) <= 1e-12)) zero = false;
    pr.euclidean = zero || is_approx(P, Pabs);
  }
  pr.ext.clear();
  for (int i = 0; i < V; i++) {
    bool nz = false;
    for (int j = 0; j < V; j++)
      if (P[(size_t)i * V + j] != 0.0) nz = true;
    if (nz) pr.ext.push_back(i);
  }
  /* tropical_sampling.hpp:284-458 */
  if (E > 63) {
    std::stringstream s;
    s << "E = " << E << " this graph is too large for this implementation!";
    throw Error(-2, s.str());
  }
  if (V > 32) throw Error(-5, "Error: stacksize for BFS stack too small. Compile with larger MAX_BFS_STACKSIZE.");
  if (E > 30) throw Error(-5, "E > 30: the 2^E subgraph table does not fit this engine");
  std::vector<int> cm(V);
  const uint64_t full = (UINT64_C(1) << E) - 1;
  const int C = components(g, full, cm.data());
  const int L = E - V + C;
  if (C != 1) throw Error(-2, "Error: the graph " + graph_str(g) + " is not connected.");
  const double W = omega(g, D, L, full), W_ref = omega(g, -D, L, full);
  pr.L = L;
  pr.W = W;
  pr.W_ref = W_ref;
  double pgp0 = 0;
  for (int i = 0; i < V - 1; i++)
    for (int j = 0; j < V - 1; j++) pgp0 += P[(size_t)i * V + j] * P[(size_t)i * V + j];
  double m2n = 0;
  for (double v : pr.m2) m2n += v * v;
  const double pmsqr_ref = std::sqrt(pgp0) + std::sqrt(m2n);
  const bool bPhiTr = (pr.K > 1) || std::fabs(W / W_ref) > 1e-12;
  if (bPhiTr && std::fabs(pmsqr_ref) < 1e-12) {
    std::stringstream s;
    s << "The graph " << graph_str(g) << " has superficial degree of divergence " << W
      << " and/or we need higher orders in epsilon. So, kinematics are necessary. The incoming momenta/masses appear "
         "to be too small to get an (numerically effective) non-zero F/Phi polynomial.";
    throw Error(-2, s.str());
  }
  const bool bNonEuclidean = !pr.euclidean;
  /* random test points X in [1, 3) for the mass-momentum-spanning test */
  const int n_tests = 100;
  std::vector<double> Xs((size_t)n_tests * E), Xi((size_t)n_tests * E);
  {
    uint64_t st[4];
    xo_seed(0x6d6d7465737473ULL, st);
    for (size_t q = 0; q < Xs.size(); q++) {
      const uint64_t res = st[0] + st[3], t = st[1] << 17;
      st[2] ^= st[0]; st[3] ^= st[1]; st[1] ^= st[2]; st[0] ^= st[3]; st[2] ^= t;
      st[3] = (st[3] << 45) | (st[3] >> 19);
      Xs[q] = 1.0 + 2.0 * ((double)(res >> 11) * (1.0 / 9007199254740992.0));
      Xi[q] = 1.0 / Xs[q];
    }
  }
  JtSetup js;
  js.n_tests = n_tests;
  js.Xs = Xs.data();
  js.Xi = Xi.data();
  js.pmsqr_ref = pmsqr_ref;
  js.W = W;
  js.W_ref = W_ref;
  js.phi = bPhiTr;
  js.noneucl = bNonEuclidean;
  JtResult jres;
  if (pr.jtable_fn) pr.jtable_fn(pr, js, jres);
  else host_jtable(pr, js, jres);
  const bool bGeneric = jres.generic, bPseudo = jres.pseudo, bGP = jres.gp;
  std::string div_err;
  if (jres.div_level >= 0) {
    const uint64_t s = jres.div_subset;
    std::stringstream m;
    m << "Error: the graph " << graph_str(g) << " has a UV or IR subdivergence. \nE.g. the subgraph with edge numbers { ";
    for (int j = 0; j < E; j++)
      if ((s >> j) & 1) m << j << " ";
    m << "} or, explicitly, with the edges { ";
    for (int j = 0; j < E; j++)
      if ((s >> j) & 1) m << "[" << g.hi[j] << "," << g.lo[j] << "] ";
    m << "} has an (effective) superficial degree of divergence of " << pr.r[s] << ".";
    div_err = m.str();
  }
  if (!div_err.empty()) throw Error(-2, div_err);
  pr.IGtr = pr.Jr[full];
  pr.generic = bGeneric;
  pr.pseudo = bPseudo;
  pr.gp = bGP;

  /* feyntrop.cpp:165-239 */
  std::string regime = "(Effective) kinematic regime: ";
  regime += pr.euclidean ? "Euclidean" : (bPseudo ? "Pseudo-Euclidean" : "Minkowski");
  regime += bGeneric ? " (generic)." : " (exceptional).";
  pr.log.push_back(regime);
  std::stringstream dump;
  dump << graph_str(g) << "\n";
  for (double v : pr.m2) dump << v << "\n";
  dump << matrix_str(P, V);
  const std::string bug =
      " the generalized permutahedron property is not fulfilled. This is likely a bug and the result of the "
      "integration shouldn't be trusted. Please contact the developers and send them the following data such that "
      "they can fix this bug. Thank you!\n";
  if (bGeneric && !bGP) throw Error(-2, "The momenta seem generic, but" + bug + dump.str());
  if (pr.euclidean && !bGP) throw Error(-2, "The momenta seem Euclidean, but" + bug + dump.str());
  if (!bGeneric && !pr.euclidean && W > 0) {
    pr.log.push_back(
        "Warning: Kinematics are non-Euclidean and (very) exceptional. Detailed info on the N[F] polytope is needed. "
        "The integration might fail or be unstable. Check the result by 1) varying the number of sample points 2) "
        "evaluating at multiple kinematic points close to the current one (i.e. by making the kinematics more generic) "
        "3) increasing the spacetime dimension D0.");
    pr.log.push_back(bGP ? "The generalized permutahedron property seems fulfilled. This is a good sign that the "
                           "integration might succeed anyway."
                         : "The generalized permutahedron property does NOT seem fulfilled. Likely, there are singular "
                           "points in the integration domain that impeed convergence.");
  } else {
    if (bGP) {
      pr.log.push_back("Generalized permutahedron property seems fulfilled.");
    } else {
      if (W > 0)
        throw Error(-2, "The momenta seem Euclidean or generic, but the generalized permutahedron property seems not "
                        "fulfilled. This is likely a bug and the result of the integration shouldn't be trusted. Please "
                        "contact the developers and send them the following data such that they can fix this bug. "
                        "Thank you!\n" + dump.str());
      pr.log.push_back("Generalized permutahedron property seems NOT fulfilled. However, the integration should work "
                       "fine as the N[F] polytope is not relevant for the given dimension and edge weights.");
    }
  }
  pr.lambda_eff = 0.0;
  if (!bPseudo) {
    std::stringstream s;
    s << "Analytic continuation: activated. Lambda = " << pr.lambda << ".";
    pr.log.push_back(s.str());
    pr.lambda_eff = pr.lambda;
  }
  pr.W_zero = std::fabs(W / W_ref) < 1e-12;
  pr.def_ref = 0.0;
  if (pr.lambda_eff != 0 && !(pr.W_zero && pr.K == 1)) pr.def_ref = -pr.lambda_eff;
  if (pr.def_ref != 0.0) pr.path = 2;
  else if (pr.W_zero && pr.K == 1) pr.path = 1;
  else pr.path = 0;
  pr.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace ftrh
