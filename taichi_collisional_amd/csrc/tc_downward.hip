// Downward pass on the device (taichi_with_jerks/traverse_eager.h:272-292):
// L2L from each internal cell into its children, one launch per level from the
// root down, then L2P at every leaf, then the results are scattered back to
// the caller's body order.
//
//   L2L (kernel_fast_lean.h:1001-1058): for each child,
//       L_child[n,m] += sum_{k<=p-n} sum_{|l|<=k} L[n+k, m+l] conj Y_k^l(X - X_child)
//       two blocks of 256 threads per parent cell, four children each: 64
//       threads per child, each thread a strided share of the child's (output
//       term, L or Ldot) tasks; the four harmonic tables and the parent's
//       expansion sit in LDS (16 KB at p = 12).  The P2P kernel running
//       alongside holds 12 one-wave blocks per CU (154 KB of LDS, all but 8
//       VGPRs per SIMD lane), so a level launched while it runs waits for its
//       blocks to drain.  A task's sum runs over contiguous runs of both
//       tables, two terms at a time into two partial sums: the top levels are
//       a few blocks each, and their time (35-40 us) is that of one thread's
//       chain of LDS reads and adds.
//   L2P (kernel_fast_lean.h:1060-1128): two leaves per wave, 32 lanes each
//       (leaves average ~19 bodies: one leaf per wave left 70% of the lanes
//       idle); lane = body, Y = Y(X_c - x_b) and
//       dY/dt with velocity -v_b, streamed order by order (three harmonic
//       rows in registers); only n <= 1 of the local field is formed:
//       p += Re Phi_0^0, a -= (Re Phi_1^1, Im Phi_1^1, Re Phi_1^0),
//       j -= the same of Phidot = L conj dY + Ldot conj Y.
#include "tc_common.h"
#include "tc_harm.h"
#include "tc_kernels.h"

#ifndef TC_SHARD
#error "compiled once per order shard, with -DTC_SHARD=k (see TC_NSHARDS in tc_kernels.h)"
#endif

namespace tcfmm {
namespace TC_SHARD_NS(TC_SHARD) {

template <int P>
__global__ __launch_bounds__(256) void k_l2l(CellsDev c, const int *__restrict__ ids, int ls) {
  constexpr int NT = nterm(P);
  constexpr int NO = tri(P, P) + 1;
  constexpr int kGroup = 64;  // threads per child
  static_assert(P + 1 <= kGroup, "one thread per harmonic column");
  __shared__ cplx G[4][NT], cl[NT], cld[NT];
  const int tid = threadIdx.x;
  const int cell = blockIdx.x >> 1, half = blockIdx.x & 1;
  const int ci = ids ? ids[cell] : ls + cell;
  const int nch = c.nchild[ci] - 4 * half;  // children of this block's half
  if (nch <= 0) return;
  const int ch0 = c.child[ci] + 4 * half;
  const double *Lp = c.L + (size_t)ci * NT;
  const double *Ldp = c.Ld + (size_t)ci * NT;
  for (int t = tid; t < NT; t += blockDim.x) {
    int n = 0;
    while ((n + 1) * (n + 1) <= t) n++;
    const int m = t - n * n - n;
    cl[t] = packed_to_c(Lp, n, m);
    cld[t] = packed_to_c(Ldp, n, m);
  }
  const int k = tid / kGroup, u = tid - k * kGroup;
  const int ch = ch0 + k;
  if (k < nch)
    harm_table_col<P>(u, c.X[3 * ci] - c.X[3 * ch], c.X[3 * ci + 1] - c.X[3 * ch + 1],
                      c.X[3 * ci + 2] - c.X[3 * ch + 2], G[k]);
  __syncthreads();
  if (k >= nch) return;
  double *Lc = c.L + (size_t)ch * NT;
  double *Ldc = c.Ld + (size_t)ch * NT;
  const cplx *Gk = G[k];
  // tasks (output term, L or Ldot) of this child, strided over the group
  for (int j = u; j < 2 * NO; j += kGroup) {
    const int which = j >= NO, o = j - which * NO;
    int n = 0;
    while (tri(n + 1, 0) <= o) n++;
    const int m = o - tri(n, 0);
    const cplx *src = which ? cld : cl;
    cplx a = {0, 0}, a2 = {0, 0};
    for (int kk = 0; kk <= P - n; kk++) {
      // l = -kk..kk: src[ixnm(n + kk, m + l)] and Gk[ixnm(kk, l)] run contiguously
      const cplx *sr = src + ixnm(n + kk, m - kk), *gr = Gk + ixnm(kk, -kk);
      const int len = 2 * kk + 1;
      int t = 0;
      for (; t + 1 < len; t += 2) {
        a = cadd(a, cmul(sr[t], cconj(gr[t])));
        a2 = cadd(a2, cmul(sr[t + 1], cconj(gr[t + 1])));
      }
      if (t < len) a = cadd(a, cmul(sr[t], cconj(gr[t])));
    }
    a = cadd(a, a2);
    double *dst = which ? Ldc : Lc;
    const int s = n * n + n;
    dst[s + m] += a.re;
    if (m > 0) dst[s - m] += a.im;
  }
}

// Orders K..P of L2P for one body: row K of Y(X_c - x_b) (compile-time K, no
// case tests) and its time derivative, contracted with L / Ldot of orders K
// and K+1 for the n <= 1 part of the local field.  l = -m uses
// Y_k^{-m} = (-1)^m conj(Y_k^m), and the same of L, Ldot and dY/dt, so the
// terms of the outputs of order m = 0 that are read as real parts (Phi_0^0,
// Phi_1^0 and its time derivative) come in equal pairs: only l >= 0 is
// formed, the l > 0 sums (h00, h10, e10) counting twice at the end.
// Phi_1^1 and its derivative keep both l = +m and l = -m.  A scheduling
// barrier per order keeps the later orders' coefficient reads from being
// hoisted (which multiplies the live registers).
struct L2PAcc {
  double p00, h00, p10, h10, d10, e10;  // Re parts: l = 0 terms, l > 0 terms
  cplx p11, d11;
};

template <int P, int K>
__device__ __forceinline__ void l2p_orders(double dx, double dy, double dz, double r2, double vx,
                                           double vy, double vz, const cplx *cl, const cplx *cld,
                                           cplx *y1, cplx *y2, L2PAcc &a) {
  __builtin_amdgcn_sched_barrier(0);
  cplx yk[P + 1];
  harm_row_c<K>(dx, dy, dz, r2, y1, y2, yk);
  const cplx *L0 = cl + K * K + K;
  const cplx *L1 = cl + (K + 1) * (K + 1) + (K + 1);
  const cplx *D1 = cld + (K + 1) * (K + 1) + (K + 1);
  constexpr bool up = K < P;
  auto redot = [](cplx u, cplx v) { return u.re * v.re + u.im * v.im; };  // Re(u conj v)
#pragma unroll
  for (int m = 0; m <= K; m++) {
    cplx dy_m = {0.0, 0.0};
    if constexpr (K > 0) {
      const cplx eta = {-vx * 0.5, vy * 0.5}, xi = {vx * 0.5, vy * 0.5};
      if (m <= K - 2) dy_m = cadd(dy_m, cmul(y1[m + 1 <= P ? m + 1 : P], eta));
      if (m <= K - 1) dy_m = cadd(dy_m, cscale(vz, y1[m]));
      if (m >= 1) {
        dy_m = cadd(dy_m, cmul(y1[m >= 1 ? m - 1 : 0], xi));
      } else if (K >= 2) {
        const cplx c1 = cconj(y1[1]);
        dy_m = cadd(dy_m, cmul({-c1.re, -c1.im}, xi));
      }
    }
    const cplx y = yk[m];
    if (m == 0) {
      a.p00 += redot(L0[0], y);
      if (up) {
        a.p10 += redot(L1[0], y);
        a.d10 += redot(L1[0], dy_m) + redot(D1[0], y);
      }
    } else {
      a.h00 += redot(L0[m], y);
      if (up) {
        a.h10 += redot(L1[m], y);
        a.e10 += redot(L1[m], dy_m) + redot(D1[m], y);
      }
    }
    if (up) {
      const cplx g = cconj(y), gd = cconj(dy_m);
      a.p11 = cadd(a.p11, cmul(L1[m + 1], g));
      a.d11 = cadd(a.d11, cadd(cmul(L1[m + 1], gd), cmul(D1[m + 1], g)));
      if (m > 0) {
        const double sg = (m & 1) ? -1.0 : 1.0;
        const cplx g2 = cscale(sg, y), gd2 = cscale(sg, dy_m);
        a.p11 = cadd(a.p11, cmul(L1[1 - m], g2));
        a.d11 = cadd(a.d11, cadd(cmul(L1[1 - m], gd2), cmul(D1[1 - m], g2)));
      }
    }
  }
  if constexpr (K < P) {
#pragma unroll
    for (int m = 0; m <= K; m++) {
      y2[m] = y1[m];
      y1[m] = yk[m];
    }
    l2p_orders<P, K + 1>(dx, dy, dz, r2, vx, vy, vz, cl, cld, y1, y2, a);
  }
}

// Up to p = 12 the kernel is held to three waves per SIMD (168 VGPRs; at
// p = 12 that spills 14 of its 196, 60 bytes per lane, and still runs the
// downward pass 0.02-0.08 ms sooner per 2^20-body evaluation than two waves).
template <int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(P <= 12 ? 3 : 1))) void k_l2p(
    CellsDev c, BodiesDev b, ResultsDev r,
                                            const int *__restrict__ leaf_ids, int nleaves) {
  constexpr int NT = nterm(P);
  constexpr int kHalf = kWave / 2;  // lanes per leaf: two leaves per wave
  __shared__ cplx cl[2][NT], cld[2][NT];
  const int lane = threadIdx.x, h = lane / kHalf, u = lane - h * kHalf;
  const int li = 2 * (int)blockIdx.x + h;
  const int ci = li < nleaves ? leaf_ids[li] : -1;
  const bool on = ci >= 0 && c.has_sink[ci];
  if (on) {
    const double *Lp = c.L + (size_t)ci * NT;
    const double *Ldp = c.Ld + (size_t)ci * NT;
    for (int t = u; t < NT; t += kHalf) {
      int n = 0;
      while ((n + 1) * (n + 1) <= t) n++;
      int m = t - n * n - n;
      cl[h][t] = packed_to_c(Lp, n, m);
      cld[h][t] = packed_to_c(Ldp, n, m);
    }
  }
  __syncthreads();
  if (!on) return;
  const int b0 = c.body[ci], nb = c.nbody[ci];
  const double X0 = c.X[3 * ci], X1 = c.X[3 * ci + 1], X2 = c.X[3 * ci + 2];
  for (int base = 0; base < nb; base += kHalf) {
    const int i = b0 + base + u;
    if (base + u >= nb) continue;
    if (!(b.flags[i] & kSink)) continue;
    const double vx = -b.vx[i], vy = -b.vy[i], vz = -b.vz[i];
    const double dx = X0 - b.x[i], dy = X1 - b.y[i], dz = X2 - b.z[i];
    const double r2 = dx * dx + dy * dy + dz * dz;
    cplx y1[P + 1], y2[P + 1];
    L2PAcc a = {0, 0, 0, 0, 0, 0, {0, 0}, {0, 0}};
    l2p_orders<P, 0>(dx, dy, dz, r2, vx, vy, vz, cl[h], cld[h], y1, y2, a);
    r.pot[i] += a.p00 + 2.0 * a.h00;
    r.ax[i] -= a.p11.re;
    r.ay[i] -= a.p11.im;
    r.az[i] -= a.p10 + 2.0 * a.h10;
    r.jx[i] -= a.d11.re;
    r.jy[i] -= a.d11.im;
    r.jz[i] -= a.d10 + 2.0 * a.e10;
  }
}

template <int P>
static void l2l_go(CellsDev c, const int *ids, int ls, int ncell, hipStream_t s) {
  hipLaunchKernelGGL(k_l2l<P>, dim3(2 * ncell), dim3(256), 0, s, c, ids, ls);
}
template <int P>
static void l2p_go(CellsDev c, BodiesDev b, ResultsDev r, const int *ids, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_l2p<P>, dim3((n + 1) / 2), dim3(64), 0, s, c, b, r, ids, n);
}

int launch_l2l(int p, CellsDev c, int ls, int le, hipStream_t s) {
  if (le <= ls) return 0;
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      l2l_go<P>(c, nullptr, ls, le - ls, s);\
      return 0;\
    } else {\
      return -1;\
    }
    TC_FOR_EACH_P(TC_CASE)
#undef TC_CASE
    default:
      return -1;
  }
}

int launch_l2l_ids(int p, CellsDev c, const int *ids, int nids, hipStream_t s) {
  if (nids <= 0) return 0;
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      l2l_go<P>(c, ids, 0, nids, s);\
      return 0;\
    } else {\
      return -1;\
    }
    TC_FOR_EACH_P(TC_CASE)
#undef TC_CASE
    default:
      return -1;
  }
}

int launch_l2p(int p, CellsDev c, BodiesDev b, ResultsDev r, const int *leaf_ids, int nleaves,
               hipStream_t s) {
  if (nleaves <= 0) return 0;
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      l2p_go<P>(c, b, r, leaf_ids, nleaves, s);\
      return 0;\
    } else {\
      return -1;\
    }
    TC_FOR_EACH_P(TC_CASE)
#undef TC_CASE
    default:
      return -1;
  }
}

}  // namespace TC_SHARD_NS(TC_SHARD)
}  // namespace tcfmm
