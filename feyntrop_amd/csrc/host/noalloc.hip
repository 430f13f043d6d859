/*
 * noalloc.hip -- the fixed device scratch of ftr_integrate_noalloc (no
 * reference counterpart; for callers that cannot allocate device memory).
 *
 * `scratch` is a file-scope device array: every GPU the library runs on holds
 * its own copy (FTRN_SCRATCH_BYTES) until the process ends.  In host code
 * the name `scratch` is only a host-side stand-in, never a device address:
 * ftrn_scratch_init() resolves the real address once per device with
 * hipGetSymbolAddress (with that device current) and every later use goes
 * through that pointer.  All calls on one GPU share the array; calls that
 * overlap on one GPU race.  The integration writes every byte it reads back
 * (block states, jump polynomials, all partial moments, the totals) before
 * reading it, so nothing has to start at zero.
 *
 * Also here: the device-side sum of the per-block partial moments, in block
 * order, so that the totals are the ones ftr_integrate sums on the host.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../../include/feyntrop_amd.h"

static __device__ __align__(256) unsigned char scratch[FTR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

/* One thread per moment q: tot[q] = sum over blocks b (in order) of
 * partials[b nq + q] -- the additions of the host loop, in its order. */
__global__ void __launch_bounds__(128) ftrn_sum_partials(const double* partials, uint64_t nblocks, int nq,
                                                        double* tot) {
  const int q = threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  for (uint64_t b = 0; b < nblocks; b++) s += partials[b * (uint64_t)nq + q];
  tot[q] = s;
}

/* The device address of this device's scratch; call with `device` current.
 * Resolved once per device and kept.  Returns hipSuccess or the error of
 * hipGetSymbolAddress. */
extern "C" hipError_t ftrn_scratch_init(int device, void** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
  }
  *ptr = g_ptr[device];
  return hipSuccess;
}

/* tot[0..nq) = block-ordered sums of partials (nblocks x nq), on stream s. */
extern "C" hipError_t ftrn_sum_partials_launch(const double* partials, uint64_t nblocks, int nq, double* tot,
                                               hipStream_t s) {
  if (nq < 1 || nq > 128) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ftrn_sum_partials, dim3(1), dim3(128), 0, s, partials, nblocks, nq, tot);
  return hipGetLastError();
}
