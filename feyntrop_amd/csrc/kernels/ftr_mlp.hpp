/*
 * ftr_mlp.hpp -- the learned proposal's network on gfx950 MFMA (no reference
 * counterpart).  A 4-layer MLP (input 32, three hidden layers of 128 SiLU
 * units, output 32) in BF16 with FP32 accumulation:
 *   Z1 = X W1^T + b1, H1 = silu(Z1); Z2, H2; Z3, H3; Y = H3 W4^T + b4.
 * Every dense product is gfx950's v_mfma_f32_32x32x16_bf16 (one wave per
 * 32x32 output tile, operands staged through LDS); forward and backward use
 * the same kernel with strided operands.
 *
 * MFMA lane layout used (checked on the device against a numpy product):
 * for lane l,
 *   A (32x16): A[l%32][8(l/32)+t], t = 0..7
 *   B (16x32): B[8(l/32)+t][l%32]
 *   C (32x32): acc v -> C[8(v/4) + 4(l/32) + v%4][l%32]
 */
#pragma once

#ifdef __HIPCC_RTC__
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace ftrm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

/* round-to-nearest-even float -> bf16 bits */
__device__ inline unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);  /* NaN */
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

/* Operands: element (i, k) of A lives at A[i*si + k*sk], with one of the
 * two strides 1 and the other a multiple of 8; T is the storage type
 * (unsigned short = bf16 bits, or float rounded to bf16).  Raw<T> holds 8
 * consecutive elements as loaded (16-byte aligned); pack8 turns them into 8
 * bf16 in a uint4.  Loading and packing are separate so that a lane's loads
 * can all be in flight before the first conversion. */
template <class T>
struct Raw;
template <>
struct Raw<unsigned short> {
  uint4 v;
  __device__ inline void load(const unsigned short* p) { v = *reinterpret_cast<const uint4*>(p); }
  __device__ inline uint4 pack8() const { return v; }
};
template <>
struct Raw<float> {
  float4 a, b;
  __device__ inline void load(const float* p) {
    a = reinterpret_cast<const float4*>(p)[0];
    b = reinterpret_cast<const float4*>(p)[1];
  }
  __device__ inline uint4 pack8() const {
    uint4 r;
    r.x = (uint32_t)f2bf(a.x) | ((uint32_t)f2bf(a.y) << 16);
    r.y = (uint32_t)f2bf(a.z) | ((uint32_t)f2bf(a.w) << 16);
    r.z = (uint32_t)f2bf(b.x) | ((uint32_t)f2bf(b.y) << 16);
    r.w = (uint32_t)f2bf(b.z) | ((uint32_t)f2bf(b.w) << 16);
    return r;
  }
};

constexpr int KC = 64;       /* K chunk staged through LDS */
constexpr int KP = KC + 8;   /* padded LDS row (144 B: 16-byte reads spread over the banks) */

/* Staging of a K chunk: dst[r][k] = src(r, k0 + k) for r < 32,
 * k < 16 NV (NV = 1, 2 or 4 vectors per lane), element (r, k) at
 * src[r*sr + k*sk] with sk == 1 or sr == 1.  load_n reads, store_n writes
 * LDS: a chunk issues the loads of both operands before its first LDS
 * write, so it costs one memory round trip. */
template <int NV, class T>
__device__ inline void load_n(Raw<T> (&v)[NV], const T* src, size_t sr, size_t sk, int k0) {
  const int l = threadIdx.x;
  if (sk == 1) {
    /* vectors along k: vector q covers row q / (2 NV), k = 8 (q % (2 NV)) */
    constexpr int SH = NV == 4 ? 3 : (NV == 2 ? 2 : 1);
#pragma unroll
    for (int t = 0; t < NV; t++) {
      const int q = l + 64 * t;
      v[t].load(src + (size_t)(q >> SH) * sr + (size_t)(k0 + 8 * (q & ((1 << SH) - 1))));
    }
  } else {
    /* vectors along the rows: vector q covers rows 8 (q % 4) .. +8 at k = q / 4 */
#pragma unroll
    for (int t = 0; t < NV; t++) {
      const int q = l + 64 * t;
      v[t].load(src + (size_t)(8 * (q & 3)) + (size_t)(k0 + (q >> 2)) * sk);
    }
  }
}

template <int NV, class T>
__device__ inline void store_n(unsigned short (*dst)[KP], const Raw<T> (&v)[NV], size_t sk) {
  const int l = threadIdx.x;
  if (sk == 1) {
    constexpr int SH = NV == 4 ? 3 : (NV == 2 ? 2 : 1);
#pragma unroll
    for (int t = 0; t < NV; t++) {
      const int q = l + 64 * t;
      *reinterpret_cast<uint4*>(&dst[q >> SH][8 * (q & ((1 << SH) - 1))]) = v[t].pack8();
    }
  } else {
#pragma unroll
    for (int t = 0; t < NV; t++) {
      const int q = l + 64 * t;
      const int r0 = 8 * (q & 3), k = q >> 2;
      const uint4 pv = v[t].pack8();
      const uint32_t w[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
      for (int u = 0; u < 4; u++) {
        dst[r0 + 2 * u][k] = (unsigned short)(w[u] & 0xffffu);
        dst[r0 + 2 * u + 1][k] = (unsigned short)(w[u] >> 16);
      }
    }
  }
}

/* one K chunk of 16 NV: stage both operands, then NV MFMA steps */
template <int NV, class AT, class BT>
__device__ inline void gemm_chunk(f32x16& acc, unsigned short (*sA)[KP], unsigned short (*sB)[KP], const AT* Ab,
                                  size_t sai, size_t sak, const BT* Bb, size_t sbj, size_t sbk, int k0) {
  Raw<AT> ra[NV];
  Raw<BT> rb[NV];
  load_n<NV>(ra, Ab, sai, sak, k0);
  load_n<NV>(rb, Bb, sbj, sbk, k0);
  store_n<NV>(sA, ra, sak);
  store_n<NV>(sB, rb, sbk);
  __syncthreads();
  const int l = threadIdx.x;
  const int r = l & 31, kq = 8 * (l >> 5);
#pragma unroll
  for (int kk = 0; kk < 16 * NV; kk += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[r][kk + kq]);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[r][kk + kq]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  __syncthreads();
}

/* C[M x N] = act(A[M x K] B[K x N] + bias)  (M, N, K multiples of 32).
 * Writes the FP32 result to C (if non-null) and its bf16 rounding after the
 * activation to Cb (if non-null).  act: 0 identity, 1 SiLU, 2 the SiLU
 * backward: C = (A B) * silu'(Z) with Z = zs (same layout as C).  One wave
 * per output tile; grid = (N/32, M/32, splits) blocks of 64 threads.  With
 * splits > 1 block z covers K rows [z kc, (z+1) kc) and stores its partial
 * product to slice z of C (slices of M x ldc floats; bias, act and Cb
 * unused); mlp_split_sum then adds the slices in a fixed order. */
template <class AT, class BT>
__device__ inline void gemm_tile(const AT* A, size_t sai, size_t sak, const BT* B, size_t sbk, size_t sbj,
                                 const float* bias, float* C, size_t ldc, unsigned short* Cb, size_t ldcb,
                                 int K, int act, int kc, const float* zs) {
  __shared__ __align__(16) unsigned short sA[32][KP]; /* [row i][k] */
  __shared__ __align__(16) unsigned short sB[32][KP]; /* [column j][k] */
  const int l = threadIdx.x;
  const int ti = blockIdx.y, tj = blockIdx.x;
  const bool split = gridDim.z > 1;
  const int kbeg = split ? (int)blockIdx.z * kc : 0;
  const int kend = split ? (kbeg + kc < K ? kbeg + kc : K) : K;
  const AT* Ab = A + (size_t)ti * 32 * sai;
  const BT* Bb = B + (size_t)tj * 32 * sbj;
  const int r = l & 31;
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += KC) {
    const int kn = (kend - k0 < KC) ? kend - k0 : KC;   /* 16, 32 or 64 */
    if (kn == 64) gemm_chunk<4>(acc, sA, sB, Ab, sai, sak, Bb, sbj, sbk, k0);
    else if (kn == 32) gemm_chunk<2>(acc, sA, sB, Ab, sai, sak, Bb, sbj, sbk, k0);
    else gemm_chunk<1>(acc, sA, sB, Ab, sai, sak, Bb, sbj, sbk, k0);
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int i = ti * 32 + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
    const int j = tj * 32 + r;
    if (split) {
      C[(size_t)blockIdx.z * gridDim.y * 32 * ldc + (size_t)i * ldc + j] = acc[v];
      continue;
    }
    if (act == 2) {
      C[(size_t)i * ldc + j] = acc[v] * dsilu(zs[(size_t)i * ldc + j]);
      continue;
    }
    const float z = acc[v] + (bias ? bias[j] : 0.0f);
    if (C) C[(size_t)i * ldc + j] = z;
    if (Cb) Cb[(size_t)i * ldcb + j] = f2bf(act == 1 ? silu(z) : z);
  }
}

}  // namespace ftrm

extern "C" __global__ void __launch_bounds__(64)
    mlp_gemm_bb(const unsigned short* A, uint64_t sai, uint64_t sak, const unsigned short* B, uint64_t sbk,
                uint64_t sbj, const float* bias, float* C, uint64_t ldc, unsigned short* Cb, uint64_t ldcb, int K,
                int act, int kc, const float* zs) {
  ftrm::gemm_tile<unsigned short, unsigned short>(A, sai, sak, B, sbk, sbj, bias, C, ldc, Cb, ldcb, K, act, kc, zs);
}

extern "C" __global__ void __launch_bounds__(64)
    mlp_gemm_fb(const float* A, uint64_t sai, uint64_t sak, const unsigned short* B, uint64_t sbk, uint64_t sbj,
                const float* bias, float* C, uint64_t ldc, unsigned short* Cb, uint64_t ldcb, int K, int act,
                int kc, const float* zs) {
  ftrm::gemm_tile<float, unsigned short>(A, sai, sak, B, sbk, sbj, bias, C, ldc, Cb, ldcb, K, act, kc, zs);
}

/* part[y][j] = sum of dZ[i][j] over the rows i of chunk y (rows x cols,
 * row-major): one thread per column and chunk of `chunk` rows (grid.y =
 * row chunks); mlp_split_sum adds the chunks. */
extern "C" __global__ void mlp_colsum(const float* dZ, float* part, int rows, int cols, int chunk) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const int i0 = blockIdx.y * chunk;
  const int i1 = i0 + chunk < rows ? i0 + chunk : rows;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f}; /* 4 loads in flight */
  int i = i0;
  for (; i + 4 <= i1; i += 4)
#pragma unroll
    for (int t = 0; t < 4; t++) s[t] += dZ[(size_t)(i + t) * cols + j];
  for (; i < i1; i++) s[0] += dZ[(size_t)i * cols + j];
  part[(size_t)blockIdx.y * cols + j] = (s[0] + s[1]) + (s[2] + s[3]);
}

/* out[i] = sum over z < splits of ws[z n + i], in a fixed order. */
extern "C" __global__ void mlp_split_sum(const float* ws, float* out, uint64_t n, int splits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f}; /* 4 loads in flight */
  int z = 0;
  for (; z + 4 <= splits; z += 4)
#pragma unroll
    for (int t = 0; t < 4; t++) s[t] += ws[(size_t)(z + t) * n + i];
  for (; z < splits; z++) s[0] += ws[(size_t)z * n + i];
  out[i] = (s[0] + s[1]) + (s[2] + s[3]);
}

/* Adam on FP32 master parameters, refreshing their bf16 copies. */
extern "C" __global__ void mlp_adam(float* w, unsigned short* wb, const float* g, float* m, float* v, uint64_t n,
                                    float lr, float b1, float b2, float eps, float c1, float c2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = ftrm::f2bf(wi);
}

/* fill (instead of hipMemset, so the sequence can live in a graph) */
extern "C" __global__ void mlp_fill(double* p, uint64_t n, double val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = val;
}

extern "C" __global__ void mlp_fill_f(float* p, uint64_t n, float val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = val;
}

/* float -> bf16 */
extern "C" __global__ void mlp_to_bf16(const float* src, unsigned short* dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = ftrm::f2bf(src[i]);
}

/* Learned tables from the network outputs Y (rows = states g, 32 columns:
 * Delta_e(g) in columns 0..E-1, sigma(g,e) of the radius drawn after
 * cutting e from g in column E+e), squashed to Delta = dmax tanh(y/dmax),
 * sigma = smax tanh(y/smax). */
extern "C" __global__ void mlp_build_tables(const float* Y, int E, const double* w, const double* jr,
                                            const double* r, double* lw, double* lZ, double* lratio, double* lia,
                                            double* lrad, double* ldiff, double dmax, double smax, double* lst,
                                            const uint8_t* mml, double* led) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (1u << E)) return;
  const float* y = Y + (size_t)g * 32;
  double Z = 0.0;
  for (int e = 0; e < E; e++) {
    double v = 0.0;
    if ((g >> e) & 1u) v = w[g ^ (1u << e)] * exp(dmax * tanh((double)y[e] / dmax));
    lw[(size_t)g * E + e] = v;
    Z += v;
  }
  lZ[g] = Z;
  for (int e = 0; e < E; e++) {
    const size_t k = (size_t)g * E + e;
    const double v = lw[k];
    const uint32_t c = g ^ (1u << e);
    const bool in = (g >> e) & 1u;
    lratio[k] = in ? (Z * w[c]) / (jr[g] * v) : 0.0;
    /* the radius drawn after cutting e from g (none after the last cut) */
    double ia = 0.0, sr = 0.0, df = 0.0;
    if (in && c != 0u) {
      const double sg = smax * tanh((double)y[E + e] / smax);
      const double aa = r[c] * exp(sg);
      ia = 1.0 / aa;
      sr = -sg;
      df = r[c] - aa;
    }
    lia[k] = ia;
    lrad[k] = sr;
    ldiff[k] = df;
    /* the sampler's per-cut record (ftr_abi.hpp LEdge) */
    double* d = led + k * 4;
    d[0] = in ? w[c] / v : 0.0;
    d[1] = ia;
    d[2] = sr;
    d[3] = df;
  }
  /* the sampler's per-state record (ftr_abi.hpp LQuad) */
  double* q = lst + (size_t)g * 4;
  q[0] = Z;
  q[1] = Z / jr[g];
  uint32_t* qi = reinterpret_cast<uint32_t*>(q + 2);
  qi[0] = mml[g];
  qi[1] = 0;
  q[3] = 0.0;
}

/* dL/dY of the forward KL  L = -sum_paths c log q(path) / sum c  from the
 * accumulated statistics (see ftr_abi.hpp LearnLaunch). */
extern "C" __global__ void mlp_table_grad(const float* Y, int E, const double* lw, const double* lZ, const double* gC,
                                          const double* gS, const double* gsum, float* dY, double dmax, double smax) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (1u << E)) return;
  const float* y = Y + (size_t)g * 32;
  float* d = dY + (size_t)g * 32;
  const double S = gsum[0] > 0.0 ? gsum[0] : 1.0;
  double visits = 0.0;
  for (int e = 0; e < E; e++) visits += gC[(size_t)g * E + e];
  for (int c = 0; c < 32; c++) d[c] = 0.0f;
  for (int e = 0; e < E; e++) {
    if (!((g >> e) & 1u)) continue;
    const double q = lw[(size_t)g * E + e] / lZ[g];
    const double t = tanh((double)y[e] / dmax);
    d[e] = (float)(-(gC[(size_t)g * E + e] - q * visits) / S * (1.0 - t * t));
  }
  for (int e = 0; e < E; e++) {
    if (!((g >> e) & 1u) || g == (1u << e)) continue;
    const double t = tanh((double)y[E + e] / smax);
    d[E + e] = (float)(-gS[(size_t)g * E + e] / S * (1.0 - t * t));
  }
}
