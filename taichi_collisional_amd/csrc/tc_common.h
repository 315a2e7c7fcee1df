// Shared device helpers for the tcfmm HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcfmm {

constexpr int kWave = 64;

__host__ __device__ constexpr int nterm(int p) { return (p + 1) * (p + 1); }
__host__ __device__ constexpr int ixnm(int n, int m) { return n * n + n + m; }

// Body flag bits (packed from Body::issink / Body::issource, exafmm.h:44-45).
constexpr unsigned char kSink = 1, kSource = 2;

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
__device__ __forceinline__ cplx cscale(double s, cplx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }

// Work-group dispatch is round-robin over the 8 XCDs (linear block id mod 8),
// and each XCD has its own L2.  Neighbouring work items (target cells close
// in tree order) share most of their sources, so give every XCD one
// contiguous range of items: a bijection of [0, n) onto itself.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_contiguous(int bid, int n) {
  const int x = bid % kXcds, k = bid / kXcds;
  const int per = n / kXcds, rem = n % kXcds;
  return x * per + (x < rem ? x : rem) + k;
}

// 1/sqrt(x) for x > 0 in FP64 without a branch: the hardware estimate
// y0 = v_rsq_f64(x), then one second-order Newton-Householder step on
// f(y) = 1/y^2 - x:  e = 1 - x y0^2,  y = y0 + y0 e (1/2 + 3/8 e), which
// takes the estimate's relative error eps to ~(5/16) eps^3.  x = 0 gives a
// non-finite value that the caller discards.
__device__ __forceinline__ double rsqrt_fp64(double x) {
  const double y0 = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y0, y0, 1.0);
  return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// 1/x in FP64 without the division sequence: the hardware estimate
// v_rcp_f64 and two Newton steps y <- y + y (1 - x y), each squaring the
// relative error, to within an ulp or two of 1/x.  For finite nonzero x
// inside the estimate's range (|x| between about 1e-300 and 1e300).
__device__ __forceinline__ double rcp_fp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Wave-wide sum of a double, result in every lane.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Wave-wide argmax of (e, idx): larger e wins, ties -> smaller idx.
__device__ __forceinline__ void wave_argmax(double &e, int &idx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double e2 = __shfl_xor(e, o, kWave);
    int i2 = __shfl_xor(idx, o, kWave);
    if (e2 > e || (e2 == e && i2 < idx)) {
      e = e2;
      idx = i2;
    }
  }
}

// Exclusive prefix sum over the wave of a small int; total returned in *tot.
__device__ __forceinline__ int wave_excl_scan(int v, int *tot) {
  int lane = threadIdx.x & (kWave - 1);
  int x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  *tot = __shfl(x, kWave - 1, kWave);
  return x - v;
}

}  // namespace tcfmm
