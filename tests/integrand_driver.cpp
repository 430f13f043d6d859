// Host driver for tests/test_host_integrand.py: the kernels' per-sample
// integrand (feyntrop_amd/csrc/kernels/ftr_device.hpp, ftr::integrand) built
// with g++ for one graph spec (FTR_SPEC: the structs FG and FGP the runtime
// generates, FGP holding the kinematic point), evaluated on sample points
// read from stdin -- "<n>" then per point E + 2 hex doubles (X, psi_tr,
// xi_tr) -- as the replay kernel does, printing per point the drop flag and
// K complex values as hex doubles.  Built with the undefined-behaviour and
// address sanitizers by the test.
#include <cstdio>

#include "kernel_host_shim.hpp"
#include "ftr_device.hpp"
#include FTR_SPEC

template <class P>
static ftr::DevParams params_from() {
  ftr::DevParams d;
  std::memset(&d, 0, sizeof d);
  d.D = P::D; d.mhalfD = P::mhalfD; d.W = P::W; d.mW = P::mW; d.Lf = P::Lf; d.IGtr = P::IGtr;
  d.def_ref = P::def_ref; d.inv_dlw = P::inv_dlw; d.iD = P::iD; d.iW = P::iW;
  d.use_iD = P::use_iD; d.use_iW = P::use_iW;
  for (int i = 0; i < FTR_MAX_E; i++) { d.nu[i] = P::nu[i]; d.m2[i] = P::m2[i]; }
  for (int i = 0; i < FTR_MAX_V * FTR_MAX_V; i++) {
    d.P[i] = P::P[i];
    d.Pq[i] = P::Pq[i];
  }
  for (int i = 0; i < FTR_MAX_V; i++) d.Peta[i] = P::Peta[i];
  return d;
}

int main() {
  constexpr int E = FG::E;
  const ftr::DevParams prm = params_from<FGP>();
  unsigned long n = 0;
  if (std::scanf("%lu", &n) != 1) return 2;
  for (unsigned long i = 0; i < n; i++) {
    double X[E], ps, xs;
    for (int e = 0; e < E; e++)
      if (std::scanf("%la", &X[e]) != 1) return 2;
    if (std::scanf("%la %la", &ps, &xs) != 2) return 2;
    /* the replay kernel's projective rescale from psi_tr and xi_tr */
    const double ls = prm.inv_dlw * (prm.mhalfD * std::log(ps) + (prm.W != 0.0 ? prm.W * std::log(ps / xs) : 0.0));
    ftr::cplx out[FTR_K];
    const bool ok = ftr::integrand<FG, FTR_K, FTR_MODE>(prm, X, ls, out);
    std::printf("%d", ok ? 0 : 1);
    for (int k = 0; k < FTR_K; k++) std::printf(" %a %a", ok ? out[k].re : 0.0, ok ? out[k].im : 0.0);
    std::printf("\n");
  }
  return 0;
}
