/*
 * jtable_gpu.cpp -- host side of the device J-table (kernels/ftr_jtable.hpp),
 * a replacement for host_jtable (tropical_sampling.hpp:284-458) that builds
 * the 2^E-entry table level by level on the GPU: E + 1 launches, one thread
 * per subset.  The tables and flags are bitwise those of the host pass.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "ftr_runtime.hpp"

namespace {

struct JtParams { /* mirror of ftrj::JtParams (kernels/ftr_jtable.hpp) */
  int E, V, n_tests, phi, noneucl, pad;
  double D, W, W_ref, pmsqr;
  int hi[30], lo[30];
  double nu[30], m2[30];
  double P[24 * 24];
};

struct JtModule {
  hipModule_t mod = nullptr;
  hipFunction_t lv8 = nullptr, lv16 = nullptr, lv24 = nullptr;
};

JtModule load_jt(ftr_ctx* c) {
  JtModule m;
  m.mod = load_module_raw(c, ftr_embedded_jtable_source);
  HIPCHK(hipModuleGetFunction(&m.lv8, m.mod, "ftrj_level_8"));
  HIPCHK(hipModuleGetFunction(&m.lv16, m.mod, "ftrj_level_16"));
  HIPCHK(hipModuleGetFunction(&m.lv24, m.mod, "ftrj_level_24"));
  return m;
}

/* device buffers of one J-table build, freed on every exit */
struct DevBufs {
  std::vector<void*> p;
  template <class T>
  T* alloc(size_t n) {
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(n * sizeof(T), 64)));
    p.push_back(q);
    return (T*)q;
  }
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
};

}  // namespace

void device_jtable(ftr_ctx* c, ftrh::Prep& pr, const ftrh::JtSetup& js, ftrh::JtResult& res) {
  const ftrh::Graph& g = pr.g;
  const int E = g.E, V = g.V;
  if (V > 24 || E > 30) throw Error(FTR_ERR_LIMIT, "device J-table: graph too large");
  HIPCHK(hipSetDevice(c->device));
  const JtModule m = load_jt(c);
  hipFunction_t f = V <= 8 ? m.lv8 : (V <= 16 ? m.lv16 : m.lv24);
  JtParams hp;
  memset(&hp, 0, sizeof hp);
  hp.E = E;
  hp.V = V;
  hp.n_tests = js.n_tests;
  hp.phi = js.phi ? 1 : 0;
  hp.noneucl = js.noneucl ? 1 : 0;
  hp.D = pr.D;
  hp.W = js.W;
  hp.W_ref = js.W_ref;
  hp.pmsqr = js.pmsqr_ref;
  for (int e = 0; e < E; e++) {
    hp.hi[e] = g.hi[e];
    hp.lo[e] = g.lo[e];
    hp.nu[e] = g.nu[e];
    hp.m2[e] = pr.m2[e];
  }
  for (int i = 0; i < V * V; i++) hp.P[i] = pr.P[i];
  std::vector<uint64_t> binom(31 * 31, 0);
  for (int n = 0; n <= 30; n++) {
    binom[n * 31 + 0] = 1;
    for (int k = 1; k <= n; k++) binom[n * 31 + k] = binom[(n - 1) * 31 + k - 1] + (k <= n - 1 ? binom[(n - 1) * 31 + k] : 0);
  }
  const size_t NT = (size_t)1 << E;
  const size_t nx = (size_t)js.n_tests * E;
  DevBufs b;
  JtParams* d_prm = b.alloc<JtParams>(1);
  uint64_t* d_binom = b.alloc<uint64_t>(binom.size());
  double* d_Xs = b.alloc<double>(nx);
  double* d_Xi = b.alloc<double>(nx);
  double* d_J = b.alloc<double>(NT);
  double* d_r = b.alloc<double>(NT);
  uint8_t* d_mm = b.alloc<uint8_t>(NT);
  uint8_t* d_L = b.alloc<uint8_t>(NT);
  uint32_t* d_flags = b.alloc<uint32_t>(3);
  unsigned long long* d_divs = b.alloc<unsigned long long>(E + 1);
  const uint32_t flags0[3] = {1, 1, 1};
  std::vector<unsigned long long> divs0(E + 1, ~0ULL);
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyHtoDAsync(d_prm, &hp, sizeof hp, st));
  HIPCHK(hipMemcpyHtoDAsync(d_binom, binom.data(), binom.size() * 8, st));
  if (nx) {
    HIPCHK(hipMemcpyHtoDAsync(d_Xs, (void*)js.Xs, nx * 8, st));
    HIPCHK(hipMemcpyHtoDAsync(d_Xi, (void*)js.Xi, nx * 8, st));
  }
  HIPCHK(hipMemcpyHtoDAsync(d_flags, (void*)flags0, sizeof flags0, st));
  HIPCHK(hipMemcpyHtoDAsync(d_divs, divs0.data(), divs0.size() * 8, st));
  for (int n = 0; n <= E; n++) {
    uint64_t count = binom[E * 31 + n];
    int nn = n;
    void* args[] = {&d_prm, &nn, &count, &d_binom, &d_Xs, &d_Xi, &d_J, &d_r, &d_mm, &d_L, &d_flags, &d_divs};
    const unsigned blocks = (unsigned)((count + 63) / 64);
    HIPCHK(hipModuleLaunchKernel(f, blocks, 1, 1, 64, 1, 1, 0, st, args, nullptr));
  }
  pr.Jr.resize(NT);
  pr.r.resize(NT);
  pr.mm.resize(NT);
  pr.Lt.resize(NT);
  uint32_t flags[3];
  std::vector<unsigned long long> divs(E + 1);
  HIPCHK(hipMemcpyDtoHAsync(pr.Jr.data(), d_J, NT * 8, st));
  HIPCHK(hipMemcpyDtoHAsync(pr.r.data(), d_r, NT * 8, st));
  HIPCHK(hipMemcpyDtoHAsync(pr.mm.data(), d_mm, NT, st));
  HIPCHK(hipMemcpyDtoHAsync(pr.Lt.data(), d_L, NT, st));
  HIPCHK(hipMemcpyDtoHAsync(flags, d_flags, sizeof flags, st));
  HIPCHK(hipMemcpyDtoHAsync(divs.data(), d_divs, divs.size() * 8, st));
  HIPCHK(hipStreamSynchronize(st));
  res.generic = flags[0] != 0;
  res.pseudo = flags[1] != 0;
  res.gp = flags[2] != 0;
  res.div_level = -1;
  for (int n = 0; n <= E; n++)
    if (divs[n] != ~0ULL) {
      res.div_level = n;
      res.div_subset = divs[n];
      break;
    }
  pr.jtable_on_device = true;
}

size_t rows_region_bytes(int tm, int E) {
  const size_t NT = (size_t)1 << E;
  if (tm == 4) return frows_bytes(E);
  return ((NT * E * 8 + 255) & ~(size_t)255) + NT * sizeof(ftr::StateRec);
}

void launch_rows(ftr_ctx* c, ftr_problem* p, int tm, char* d) {
  const int E = p->prep.g.E;
  const uint64_t NT = (uint64_t)1 << E;
  hipModule_t mod = load_module_raw(c, ftr_embedded_jtable_source);
  hipFunction_t f = nullptr;
  uint64_t nt = NT;
  int e = E;
  p->d_rows = (double*)d;
  p->rows_tm = tm;
  if (tm == 4) {
    HIPCHK(hipModuleGetFunction(&f, mod, "ftrj_frows"));
    p->d_rec = nullptr;
    void* args[] = {&nt, &e, &p->d_w, &p->d_ir, &p->d_mml, &p->d_rows};
    HIPCHK(hipModuleLaunchKernel(f, (unsigned)((NT + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, args, nullptr));
    return;
  }
  HIPCHK(hipModuleGetFunction(&f, mod, "ftrj_rows"));
  p->d_rec = (ftr::StateRec*)(d + ((NT * E * 8 + 255) & ~(size_t)255));
  void* args[] = {&nt, &e, &p->d_w, &p->d_jr, &p->d_ir, &p->d_mml, &p->d_rows, &p->d_rec};
  HIPCHK(hipModuleLaunchKernel(f, (unsigned)((NT + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, args, nullptr));
}

void device_rows(ftr_ctx* c, ftr_problem* p, int tm) {
  if (p->d_rows && p->rows_tm == tm) return;
  HIPCHK(hipSetDevice(c->device));
  char* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, rows_region_bytes(tm, p->prep.g.E)));
  /* rows of the other placement are replaced (rows inside a batch's shared
   * allocation stay with it) */
  if (p->d_rows_slab) (void)hipFree(p->d_rows_slab);
  p->d_rows_slab = d;
  launch_rows(c, p, tm, d);
  HIPCHK(hipStreamSynchronize(c->stream));
}
