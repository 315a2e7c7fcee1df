// Smallest enclosing ball, computed by one wave or one block (device side of
// the cell geometry of P2M / M2M, taichi_with_jerks/kernel_fast_lean.h:733-763
// and :875-935, which use the Miniball of miniball.hpp).
//
// Algorithm (same as the CPU oracle, so both produce the same ball up to
// rounding; the minimum ball is unique): start from point 0; repeatedly find
// the point of largest excess |q-c|^2 - r^2 (threads scan subsets, then an
// argmax with ties to the lowest index); if it lies outside, the new ball is the smallest ball of
// (support set + q) with q on its boundary, found by testing the circumball of
// q with every subset of the support (<= 15 candidates, each of <= 4 points).
// The radius grows strictly, so the loop ends; the final ball encloses every
// point and is the minimum ball of its own support, hence of all points.
// Every thread runs the small solves on identical data, so the result is
// uniform across the block.  Contraction is off to match the CPU oracle.
#pragma once
#include "tc_common.h"

namespace tcfmm {

struct DBall {
  double c[3];
  double r2;
};

// Smallest ball with the k (1..4) points p[0..k-1] on its boundary: the
// circumcentre in their affine hull.  Loops have fixed bounds with guards so
// every array stays in registers; the arithmetic is that of the CPU oracle.
__device__ inline bool circumball_dev(const double (&p)[4][3], int k, DBall &b) {
#pragma clang fp contract(off)
  if (k == 1) {
    for (int d = 0; d < 3; d++) b.c[d] = p[0][d];
    b.r2 = 0;
    return true;
  }
  double v[3][3];
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int d = 0; d < 3; d++) v[i - 1][d] = i < k ? p[i][d] - p[0][d] : 0.0;
  const int q = k - 1;
  double A[3][3], rhs[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int j = 0; j < 3; j++) A[i][j] = v[i][0] * v[j][0] + v[i][1] * v[j][1] + v[i][2] * v[j][2];
    rhs[i] = 0.5 * A[i][i];
  }
  double lam[3] = {0, 0, 0};
  if (q == 1) {
    if (!(A[0][0] > 0)) return false;
    lam[0] = 0.5;
  } else if (q == 2) {
    double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    if (!(fabs(det) > 1e-13 * A[0][0] * A[1][1])) return false;
    lam[0] = (rhs[0] * A[1][1] - A[0][1] * rhs[1]) / det;
    lam[1] = (A[0][0] * rhs[1] - rhs[0] * A[1][0]) / det;
  } else {
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    if (!(fabs(det) > 1e-13 * A[0][0] * A[1][1] * A[2][2])) return false;
    double inv[3][3];
    inv[0][0] = c00;
    inv[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    inv[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    inv[1][0] = c01;
    inv[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    inv[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    inv[2][0] = c02;
    inv[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    inv[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
#pragma unroll
    for (int i = 0; i < 3; i++)
      lam[i] = (inv[i][0] * rhs[0] + inv[i][1] * rhs[1] + inv[i][2] * rhs[2]) / det;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double c = p[0][d];
#pragma unroll
    for (int i = 0; i < 3; i++)
      if (i < q) c += lam[i] * v[i][d];
    b.c[d] = c;
  }
  double r2 = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (i >= k) continue;
    double dd = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) dd += (p[i][d] - b.c[d]) * (p[i][d] - b.c[d]);
    r2 = fmax(r2, dd);
  }
  b.r2 = r2;
  return true;
}

// Points bq, then the support points selected by mask, in support order.
__device__ inline int gather_support(int mask, const double (&bq)[3], const double (&sup)[4][3],
                                     double (&p)[4][3]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    p[0][d] = bq[d];
    p[1][d] = p[2][d] = p[3][d] = 0.0;
  }
  int k = 1;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    if (mask & (1 << s)) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double v = sup[s][d];
        if (k == 1) p[1][d] = v;
        if (k == 2) p[2][d] = v;
        if (k == 3) p[3][d] = v;
      }
      k++;
    }
  }
  return k;
}

// Candidate `mask` of one refinement step: the ball through bq and the
// selected support points, valid when it also holds the other support points.
__device__ inline bool candidate_ball(int mask, int ns, const double (&bq)[3],
                                      const double (&sup)[4][3], DBall &cb) {
#pragma clang fp contract(off)
  if (mask >= (1 << ns) || __popc(mask) > 3) return false;
  double pts[4][3];
  const int k = gather_support(mask, bq, sup, pts);
  if (!circumball_dev(pts, k, cb)) return false;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    if (s >= ns || (mask & (1 << s))) continue;
    double dd = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) dd += (sup[s][d] - cb.c[d]) * (sup[s][d] - cb.c[d]);
    if (dd > cb.r2 * (1 + 1e-12)) return false;
  }
  return true;
}

constexpr double kBallTol = 1e-13;

// Smallest ball over the points one block of NW waves can enumerate.
//   get(i, q)   writes point i (uniform index; used for point 0 and the
//               selected point of each iteration),
//   scan(f)     calls f(i, q) for this thread's share of the points, in
//               increasing i.
// Per iteration the point of largest excess is found with ties going to the
// lowest index (thread, wave, then block argmax), so the sequence of support
// sets -- and hence the ball -- does not depend on NW.  s_e / s_i hold 2 x NW
// slots (double-buffered: one barrier per iteration).
template <int NW, class Get, class Scan>
__device__ DBall block_minball(Get get, Scan scan, double *s_e, int *s_i) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  DBall b;
  double sup[4][3];
  int ns = 1;
  get(0, sup[0]);
  for (int d = 0; d < 3; d++) b.c[d] = sup[0][d];
  b.r2 = 0;
  for (int iter = 0; iter < 10000; iter++) {
    double best = -1;
    int bidx = 0x7fffffff;
    scan([&](int i, const double *q) {
      double e = 0;
      for (int d = 0; d < 3; d++) e += (q[d] - b.c[d]) * (q[d] - b.c[d]);
      e -= b.r2;
      if (e > best) {
        best = e;
        bidx = i;
      }
    });
    wave_argmax(best, bidx);
    if (NW > 1) {
      const int par = iter & 1;
      if (lane == 0) {
        s_e[par * NW + wid] = best;
        s_i[par * NW + wid] = bidx;
      }
      __syncthreads();
      best = s_e[par * NW];
      bidx = s_i[par * NW];
      for (int w = 1; w < NW; w++) {
        const double e2 = s_e[par * NW + w];
        const int i2 = s_i[par * NW + w];
        if (e2 > best || (e2 == best && i2 < bidx)) {
          best = e2;
          bidx = i2;
        }
      }
    }
    if (best <= kBallTol * b.r2 || best <= 0) break;
    double bq[3];
    get(bidx, bq);
    // the <= 15 candidates (subsets of the support) one per lane; the valid
    // one of smallest radius wins, ties to the lowest mask (the order a
    // sequential scan over masks would keep)
    DBall cb;
    const bool ok = candidate_ball(lane, ns, bq, sup, cb);
    double key = ok ? cb.r2 : INFINITY;
    int bestmask = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double k2 = __shfl_xor(key, o, kWave);
      const int m2 = __shfl_xor(bestmask, o, kWave);
      if (k2 < key || (k2 == key && m2 < bestmask)) {
        key = k2;
        bestmask = m2;
      }
    }
    if (!(key < INFINITY)) break;  // numerically degenerate; keep the current ball
    DBall bestb;
    candidate_ball(bestmask, ns, bq, sup, bestb);
    double nsup[4][3];
    const int k = gather_support(bestmask, bq, sup, nsup);
#pragma unroll
    for (int s2 = 0; s2 < 4; s2++)
#pragma unroll
      for (int d = 0; d < 3; d++) sup[s2][d] = nsup[s2][d];
    ns = k;
    if (bestb.r2 <= b.r2) {
      b = bestb;
      break;
    }
    b = bestb;
  }
  return b;
}

// One wave, points 0..n-1 from get(), lane-strided scan.
template <class Get>
__device__ DBall wave_minball(int n, Get get) {
  const int lane = threadIdx.x & (kWave - 1);
  return block_minball<1>(
      get,
      [&](auto f) {
        for (int i = lane; i < n; i += kWave) {
          double q[3];
          get(i, q);
          f(i, q);
        }
      },
      nullptr, nullptr);
}

}  // namespace tcfmm
