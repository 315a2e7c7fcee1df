// Solid harmonics on the device (make_Gnm / make_Gnmdot,
// taichi_with_jerks/kernel_fast_lean.h:237-304).
//
// Y_n^m = r^n P_n^m(cos t) e^{i m phi} / (n+m)!  (no Condon-Shortley phase):
//   Y_0^0 = 1,  Y_n^n = Y_{n-1}^{n-1} (x + i y)/(2n),  Y_n^{n-1} = z Y_{n-1}^{n-1},
//   Y_n^m = [(2n-1) z Y_{n-1}^m - r^2 Y_{n-2}^m] / (n^2 - m^2),
//   Y_n^{-m} = (-1)^m conj(Y_n^m).
// Time derivative for a point moving with velocity v:
//   dY_n^m/dt = eta' Y_{n-1}^{m+1} + z' Y_{n-1}^m + xi' Y_{n-1}^{m-1},
//   eta' = (-v_x + i v_y)/2, z' = v_z, xi' = (v_x + i v_y)/2; Y_0^{-1} is taken
//   as 0 (the reference reads the slot before its array there).
#pragma once
#include "tc_common.h"

namespace tcfmm {

__host__ __device__ constexpr int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// All Y_n^m, m >= 0, n <= P into y[tri(n,m)] (per lane, fully unrolled).
template <int P>
__device__ __forceinline__ void harm_all(double dx, double dy, double dz, cplx *y) {
  double r2 = dx * dx + dy * dy + dz * dz;
  y[0] = {1.0, 0.0};
#pragma unroll
  for (int n = 1; n <= P; n++) {
    y[tri(n, n)] = cmul(y[tri(n - 1, n - 1)], {dx / (2 * n), dy / (2 * n)});
#pragma unroll
    for (int m = 0; m < n - 1; m++) {
      double q = (double)(n * n - m * m);
      double a = (2 * n - 1.0) / q * dz, b = r2 / q;
      y[tri(n, m)] = {a * y[tri(n - 1, m)].re - b * y[tri(n - 2, m)].re,
                      a * y[tri(n - 1, m)].im - b * y[tri(n - 2, m)].im};
    }
    y[tri(n, n - 1)] = cscale(dz, y[tri(n - 1, n - 1)]);
  }
}

// Recurrence coefficients for harm_row, index n*(P+1)+m (used for m < n-1):
// ra = (2n-1)/(n^2-m^2), rb = 1/(n^2-m^2).  Filled by the whole block; the
// caller synchronizes before use.
template <int P>
__device__ __forceinline__ void harm_coef_lds(double *ra, double *rb) {
  for (int t = threadIdx.x; t < (P + 1) * (P + 1); t += blockDim.x) {
    const int n = t / (P + 1), m = t - n * (P + 1);
    const double q = (double)(n * n - m * m);
    ra[t] = (m + 1 < n) ? (2 * n - 1.0) / q : 0.0;
    rb[t] = (m + 1 < n) ? 1.0 / q : 0.0;
  }
}

// Row n (runtime) of the solid harmonics, yn[m] = Y_n^m for m = 0..P (zero
// for m > n), from rows n-1 (y1) and n-2 (y2): the recurrences of harm_all,
// so orders can be streamed with three rows of registers.
template <int P>
__device__ __forceinline__ void harm_row(int n, double dx, double dy, double dz, double r2,
                                         const double *ra, const double *rb, const cplx *y1,
                                         const cplx *y2, cplx *yn) {
  const double tn = n > 0 ? 2.0 * n : 1.0;
  const cplx f = {dx / tn, dy / tn};
  const double *a_n = ra + n * (P + 1), *b_n = rb + n * (P + 1);
#pragma unroll
  for (int m = 0; m <= P; m++) {
    cplx v = {0.0, 0.0};
    if (n == 0) {
      if (m == 0) v = {1.0, 0.0};
    } else if (m + 1 < n) {
      const double a = a_n[m] * dz, b = r2 * b_n[m];
      v = {a * y1[m].re - b * y2[m].re, a * y1[m].im - b * y2[m].im};
    } else if (m + 1 == n) {
      v = cscale(dz, y1[m]);
    } else if (m == n && m >= 1) {
      v = cmul(y1[m >= 1 ? m - 1 : 0], f);
    }
    yn[m] = v;
  }
}

// dY_n^m/dt (m >= 0, compile-time; n runtime) from row n-1 (y1): harm_dot.
template <int P>
__device__ __forceinline__ cplx harm_dot_row(const cplx *y1, int n, int m, double vx, double vy,
                                             double vz) {
  cplx r = {0.0, 0.0};
  if (n == 0) return r;
  const cplx eta = {-vx * 0.5, vy * 0.5}, xi = {vx * 0.5, vy * 0.5};
  if (m + 2 <= n) r = cadd(r, cmul(y1[m + 1 <= P ? m + 1 : P], eta));
  if (m + 1 <= n) r = cadd(r, cscale(vz, y1[m]));
  if (m >= 1) {
    r = cadd(r, cmul(y1[m >= 1 ? m - 1 : 0], xi));
  } else if (n >= 2) {
    const cplx c1 = cconj(y1[1]);
    r = cadd(r, cmul({-c1.re, -c1.im}, xi));
  }
  return r;
}

// Row N (compile-time) of the solid harmonics from rows N-1 (y1) and N-2 (y2):
// the recurrences of harm_all, unrolled with no case tests.  yn[m] for m > N
// is left unset.
template <int N>
__device__ __forceinline__ void harm_row_c(double dx, double dy, double dz, double r2,
                                           const cplx *y1, const cplx *y2, cplx *yn) {
  if constexpr (N == 0) {
    yn[0] = {1.0, 0.0};
  } else {
#pragma unroll
    for (int m = 0; m < N - 1; m++) {
      const double q = (double)(N * N - m * m);
      const double a = (2 * N - 1.0) / q * dz, b = r2 / q;
      yn[m] = {a * y1[m].re - b * y2[m].re, a * y1[m].im - b * y2[m].im};
    }
    yn[N - 1] = cscale(dz, y1[N - 1]);
    yn[N] = cmul(y1[N - 1], {dx / (2 * N), dy / (2 * N)});
  }
}

// dY_N^m/dt (compile-time N, m) from row N-1 (y1): harm_dot.
template <int N, int m>
__device__ __forceinline__ cplx harm_dot_c(const cplx *y1, double vx, double vy, double vz) {
  cplx r = {0.0, 0.0};
  if constexpr (N > 0) {
    const cplx eta = {-vx * 0.5, vy * 0.5}, xi = {vx * 0.5, vy * 0.5};
    if constexpr (m <= N - 2) r = cadd(r, cmul(y1[m + 1], eta));
    if constexpr (m <= N - 1) r = cadd(r, cscale(vz, y1[m]));
    if constexpr (m >= 1) {
      r = cadd(r, cmul(y1[m - 1], xi));
    } else if constexpr (N >= 2) {
      const cplx c1 = cconj(y1[1]);
      r = cadd(r, cmul({-c1.re, -c1.im}, xi));
    }
  }
  return r;
}

// Y_k^l for any l (|l| <= k) from the m >= 0 table.
__device__ __forceinline__ cplx harm_at(const cplx *y, int k, int l) {
  if (l >= 0) return y[tri(k, l)];
  cplx c = cconj(y[tri(k, -l)]);
  return (l & 1) ? cplx{-c.re, -c.im} : c;
}

// dY_n^m/dt for m >= 0 (n >= 0) from the table of Y (orders <= n-1 needed).
__device__ __forceinline__ cplx harm_dot(const cplx *y, int n, int m, double vx, double vy,
                                         double vz) {
  cplx r = {0.0, 0.0};
  if (n == 0) return r;
  const cplx eta = {-vx * 0.5, vy * 0.5}, xi = {vx * 0.5, vy * 0.5};
  if (m <= n - 2) r = cadd(r, cmul(y[tri(n - 1, m + 1)], eta));
  if (m <= n - 1) r = cadd(r, cscale(vz, y[tri(n - 1, m)]));
  if (m >= 1)
    r = cadd(r, cmul(y[tri(n - 1, m - 1)], xi));
  else if (n >= 2)
    r = cadd(r, cmul(harm_at(y, n - 1, -1), xi));
  return r;
}

// Cooperative table of Y_n^m for all n <= P, -n <= m <= n into LDS
// (real-packed index order, complex values): the caller's column m (0..P)
// is computed by one thread; other values of m do nothing.
template <int P>
__device__ __forceinline__ void harm_table_col(int m, double dx, double dy, double dz, cplx *G) {
  if (m >= 0 && m <= P) {
    double r2 = dx * dx + dy * dy + dz * dz;
    cplx y = {1.0, 0.0};
    for (int k = 1; k <= m; k++) y = cmul(y, {dx / (2 * k), dy / (2 * k)});
    cplx y2 = y, y1 = y;  // Y_{n-2}^m, Y_{n-1}^m
    auto put = [&](int n, cplx v) {
      G[ixnm(n, m)] = v;
      if (m > 0) {
        cplx c = cconj(v);
        G[ixnm(n, -m)] = (m & 1) ? cplx{-c.re, -c.im} : c;
      }
    };
    put(m, y);
    if (m + 1 <= P) {
      y1 = cscale(dz, y);
      put(m + 1, y1);
      y2 = y;
      for (int n = m + 2; n <= P; n++) {
        double q = (double)(n * n - m * m);
        double a = (2 * n - 1.0) / q * dz, b = r2 / q;
        cplx yn = {a * y1.re - b * y2.re, a * y1.im - b * y2.im};
        put(n, yn);
        y2 = y1;
        y1 = yn;
      }
    }
  }
}

// Whole table by the block's first P+1 threads.
template <int P>
__device__ __forceinline__ void harm_table_lds(double dx, double dy, double dz, cplx *G) {
  harm_table_col<P>((int)threadIdx.x, dx, dy, dz, G);
}

// Complex form of a real-packed expansion (real_2_complex,
// kernel_fast_lean.h:164-179): c_n^m = (a[s+m], a[s-m]), c_n^-m = (-1)^m conj.
__device__ __forceinline__ cplx packed_to_c(const double *a, int n, int m) {
  int s = n * n + n;
  int am = m < 0 ? -m : m;
  if (am == 0) return {a[s], 0.0};
  cplx c = {a[s + am], a[s - am]};
  if (m > 0) return c;
  c = cconj(c);
  return (am & 1) ? cplx{-c.re, -c.im} : c;
}

}  // namespace tcfmm
