// Host build of the kernel header ftr_device.hpp (test infrastructure): the
// HIP qualifiers become plain C++ and the few device intrinsics the header
// names get host stand-ins, so that the integrand can be compiled with g++
// and its undefined-behaviour and address sanitizers.  Only the per-sample
// integrand is run on the host; the block-level code is compiled, never
// called.
#pragma once
#define FTR_HOST_BUILD 1
#include <cmath>
#include <cstdint>
#include <cstring>
#define __host__
#define __device__
#define __global__
#define __shared__
#define __launch_bounds__(...)
#define __align__(n)
struct ftr_host_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
static ftr_host_dim3 threadIdx, blockIdx, blockDim, gridDim;
inline double __shfl_xor(double v, int, int = 64) { return v; }
inline void __syncthreads() {}
inline long long __double_as_longlong(double d) {
  long long r;
  std::memcpy(&r, &d, sizeof r);
  return r;
}
inline double __longlong_as_double(long long v) {
  double r;
  std::memcpy(&r, &v, sizeof r);
  return r;
}
inline unsigned atomicAdd(unsigned* p, unsigned v) {
  const unsigned o = *p;
  *p += v;
  return o;
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  const unsigned long long o = *p;
  *p += v;
  return o;
}
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  const unsigned long long o = *p;
  if (v > o) *p = v;
  return o;
}
