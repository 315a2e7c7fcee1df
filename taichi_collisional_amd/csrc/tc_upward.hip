// Upward pass on the device: cell geometry (smallest enclosing balls), P2M and
// M2M for moments M and their time derivatives Mdot (taichi_with_jerks/
// kernel_fast_lean.h:701-999, driven post-order by traverse_eager.h:9-32).
//
//   P2M  one wave per leaf: the leaf's bodies are staged in LDS, their ball
//        (centre X, radius r + 1e-6) is found, then lane = body walks the
//        orders n = 0..P (compile-time, unrolled) with rolling harmonic rows;
//        per order the terms
//        q Y_n^m(x_b - X) and q dY_n^m/dt(v_b) are summed over the bodies
//        through an LDS transpose (lane j sums column j).  min_acc and
//        has_sink are reduced alongside.
//   M2M  one block of 4 waves per internal cell (one launch per level,
//        deepest first): the ball of the bodies of leaf children and of 64
//        design points on each grandchild sphere (points generated from LDS),
//        then for each child
//          M_parent[n,m] += sum_{k,l} M_child[n-k, m-l] Y_k^l(X_child - X_parent)
//        with the eight children's harmonics and moments in LDS and lanes
//        over (output term, child), the children's sums added across lanes
//        (M first, then Mdot); one child at a time when that does not fit.
#include "tc_common.h"
#include "tc_geom.h"
#include "tc_harm.h"
#include "tc_kernels.h"
#include "tdesign.h"

#ifndef TC_SHARD
#error "compiled once per order shard, with -DTC_SHARD=k (see TC_NSHARDS in tc_kernels.h)"
#endif

namespace tcfmm {
namespace TC_SHARD_NS(TC_SHARD) {

__host__ __device__ constexpr int pow2_floor(int v) {
  int p = 1;
  while (2 * p <= v) p *= 2;
  return p;
}

// Orders N..P of P2M for one lane's body (w = its mass, 0 when idle): row N of
// the harmonics from rows N-1, N-2 (compile-time N, so no case tests), then
// q Y_N^m and q dY_N^m/dt summed over the bodies through the LDS transpose
// (lane j sums column j; cnt bodies).
template <int P, int N>
__device__ __forceinline__ void p2m_orders(double dx, double dy, double dz, double r2, double vx,
                                           double vy, double vz, double w, cplx *y1, cplx *y2,
                                           double (*red)[2 * (P + 1) + 1], int lane, int cnt,
                                           double *Mc, double *Mdc) {
  cplx yn[P + 1];
  harm_row_c<N>(dx, dy, dz, r2, y1, y2, yn);
  constexpr int s = N * N + N;
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
#pragma unroll
    for (int m = 0; m <= N; m++) {
      cplx v;
      if (pass == 0) {
        v = yn[m];
      } else {
        v = {0.0, 0.0};
        if constexpr (N > 0) {
          const cplx eta = {-vx * 0.5, vy * 0.5}, xi = {vx * 0.5, vy * 0.5};
          if (m <= N - 2) v = cadd(v, cmul(y1[m + 1 <= P ? m + 1 : P], eta));
          if (m <= N - 1) v = cadd(v, cscale(vz, y1[m]));
          if (m >= 1) {
            v = cadd(v, cmul(y1[m >= 1 ? m - 1 : 0], xi));
          } else if (N >= 2) {
            const cplx c1 = cconj(y1[1]);
            v = cadd(v, cmul({-c1.re, -c1.im}, xi));
          }
        }
      }
      red[lane][2 * m] = w * v.re;
      red[lane][2 * m + 1] = w * v.im;
    }
    __syncthreads();
    // column c = 2m + (imaginary) summed by G lanes (G a power of two, C G <=
    // 64), each over every G-th body row, then a butterfly inside the group:
    // 64 / G dependent LDS reads per lane instead of 64
    constexpr int C = 2 * (N + 1);
    constexpr int G = pow2_floor(kWave / C);
    const int col = lane / G, part = lane % G;
    double acc = 0;
    if (col < C)
      for (int r = part; r < cnt; r += G) acc += red[r][col];
#pragma unroll
    for (int o = 1; o < G; o <<= 1) acc += __shfl_xor(acc, o, kWave);
    const int m = col >> 1, im = col & 1;
    if (col < C && part == 0 && !(m == 0 && im)) {
      double *dst = (pass == 0 ? Mc : Mdc) + (im ? s - m : s + m);
      *dst += acc;
    }
    __syncthreads();
  }
  if constexpr (N < P) {
#pragma unroll
    for (int m = 0; m <= N; m++) {
      y2[m] = y1[m];
      y1[m] = yn[m];
    }
    p2m_orders<P, N + 1>(dx, dy, dz, r2, vx, vy, vz, w, y1, y2, red, lane, cnt, Mc, Mdc);
  }
}

// kWhat: kUpGeom (ball, has_sink), kUpMin (min_acc), kUpMoments (M, Mdot
// about the ball centre already in c.X), or a combination.  Split, the theta
// walk starts on the geometry while the moments are formed on another stream,
// and with the error MAC geometry and moments run alongside the pre-pass that
// min_acc waits for.
template <int P, int kWhat>
__global__ __launch_bounds__(64) void k_p2m(CellsDev c, BodiesDev b,
                                            const int *__restrict__ leaf_ids) {
  constexpr int NT = nterm(P);
  constexpr int NR = 2 * (P + 1);  // real, imaginary part for m = 0..P
  constexpr int kStage = 128;
  __shared__ double s_q[3][kStage];
  __shared__ double red[kWave][NR + 1];
  const int lane = threadIdx.x;
  const int ci = leaf_ids[blockIdx.x];
  const int b0 = c.body[ci], nb = c.nbody[ci];
  if constexpr (kWhat == kUpMin) {
    // min_acc alone (the error MAC's, once the pre-pass has set acc_old)
    double mina = 1e30;
    for (int i = lane; i < nb; i += kWave) mina = fmin(b.acc_old[b0 + i], mina);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mina = fmin(mina, __shfl_xor(mina, o, kWave));
    if (lane == 0) c.min_acc[ci] = mina;
    return;
  }
  double X[3];
  if constexpr (!(kWhat & kUpGeom)) {
    for (int d = 0; d < 3; d++) X[d] = c.X[3 * ci + d];
  } else {
    const bool staged = nb <= kStage;
    if (staged)
      for (int i = lane; i < nb; i += kWave) {
        s_q[0][i] = b.x[b0 + i];
        s_q[1][i] = b.y[b0 + i];
        s_q[2][i] = b.z[b0 + i];
      }
    __syncthreads();
    double R;
    if (nb > 0) {
      DBall ball = wave_minball(nb, [&](int i, double *q) {
        if (staged) {
          q[0] = s_q[0][i];
          q[1] = s_q[1][i];
          q[2] = s_q[2][i];
        } else {
          q[0] = b.x[b0 + i];
          q[1] = b.y[b0 + i];
          q[2] = b.z[b0 + i];
        }
      });
      for (int d = 0; d < 3; d++) X[d] = ball.c[d];
      R = sqrt(ball.r2) + 1e-6;
    } else {
      for (int d = 0; d < 3; d++) X[d] = c.Xg[3 * ci + d];
      R = 1e-6;
    }
    double mina = 1e30;
    int sink = 0;
    for (int i = lane; i < nb; i += kWave) {
      if (kWhat & kUpMin) mina = fmin(b.acc_old[b0 + i], mina);
      sink |= (b.flags[b0 + i] & kSink) ? 1 : 0;
    }
  #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mina = fmin(mina, __shfl_xor(mina, o, kWave));
      sink |= __shfl_xor(sink, o, kWave);
    }
    if (lane == 0) {
      for (int d = 0; d < 3; d++) c.X[3 * ci + d] = X[d];
      c.R[ci] = R;
      if (kWhat & kUpMin) c.min_acc[ci] = mina;
      c.has_sink[ci] = (unsigned char)(c.has_sink[ci] | sink);
    }
  }
  if constexpr (!(kWhat & kUpMoments)) return;
  double *Mc = c.M + (size_t)ci * NT;
  double *Mdc = c.Md + (size_t)ci * NT;
  for (int base = 0; base < nb || base == 0; base += kWave) {
    const int i = b0 + base + lane;
    const int cnt = min(kWave, nb - base);
    double dx = 0, dy = 0, dz = 0, vx = 0, vy = 0, vz = 0, w = 0;
    if (lane < cnt) {
      dx = b.x[i] - X[0];
      dy = b.y[i] - X[1];
      dz = b.z[i] - X[2];
      vx = b.vx[i];
      vy = b.vy[i];
      vz = b.vz[i];
      w = (b.flags[i] & kSource) ? b.q[i] : 0.0;
    }
    const double r2 = dx * dx + dy * dy + dz * dz;
    cplx y1[P + 1], y2[P + 1];
    p2m_orders<P, 0>(dx, dy, dz, r2, vx, vy, vz, w, y1, y2, red, lane, cnt, Mc, Mdc);
    if (nb == 0) break;
  }
}

template <int P, int kWhat>
__global__ __launch_bounds__(256) void k_m2m(CellsDev c, BodiesDev b, const int *__restrict__ ids,
                                             int ls) {
#pragma clang fp contract(off)
  constexpr int NT = nterm(P);
  constexpr int NO = tri(P, P) + 1;
  constexpr int NW = 4;
  constexpr int kStage = 512;  // body points of leaf children staged in LDS
  // Eight children's harmonic tables and moments (complex) side by side, when
  // they fit one work-group's LDS; else one child at a time.
  constexpr bool kPar = 16 * 16 * NT <= 60 * 1024;
  constexpr int kBall = 192 + 64 * 4 + 3 * kStage;       // doubles, ball phase
  constexpr int kTrans = kPar ? 2 * 8 * 2 * NT : 3 * 2 * NT;  // doubles, translation
  __shared__ __attribute__((aligned(16))) double smem[kBall > kTrans ? kBall : kTrans];
  double *s_td = smem;                                            // t-design
  double(*s_g)[4] = reinterpret_cast<double(*)[4]>(smem + 192);  // grandchild (X, R)
  double(*s_p)[kStage] = reinterpret_cast<double(*)[kStage]>(smem + 192 + 256);
  __shared__ int s_pre[9], s_base[8], s_leaf[8], s_body[8], s_nstage;
  __shared__ double s_e[2 * NW];
  __shared__ int s_i[2 * NW];
  const int tid = threadIdx.x;
  const int ci = ids ? ids[blockIdx.x] : ls + blockIdx.x;
  const int nch = c.nchild[ci];
  if (nch == 0) return;
  const int ch0 = c.child[ci];
  if constexpr (kWhat == kUpMin) {
    if (tid == 0) {
      double mina = 1e30;
      for (int k = 0; k < nch; k++) mina = fmin(c.min_acc[ch0 + k], mina);
      c.min_acc[ci] = mina;
    }
    return;
  }
  double X[3];
  if constexpr (!(kWhat & kUpGeom)) {
    for (int d = 0; d < 3; d++) X[d] = c.X[3 * ci + d];
  } else {
    if (tid < 192) s_td[tid] = kTDesign[tid];
    // the children's records, one thread each (their loads overlap), then the
    // prefix over at most eight children from LDS
    __shared__ int s_nc2[8], s_nb[8], s_g0[8];
    if (tid < nch) {
      const int ch = ch0 + tid;
      s_nc2[tid] = c.nchild[ch];
      s_nb[tid] = c.nbody[ch];
      s_body[tid] = c.body[ch];
      s_g0[tid] = c.child[ch];
    }
    __syncthreads();
    if (tid == 0) {
      int pre = 0, nbs = 0, ng = 0;
      for (int k = 0; k < nch; k++) {
        const int nc2 = s_nc2[k];
        s_pre[k] = pre;
        s_leaf[k] = nc2 == 0;
        if (nc2 == 0) {
          s_base[k] = nbs;
          nbs += s_nb[k];
          pre += s_nb[k];
        } else {
          s_base[k] = ng;
          ng += nc2;
          pre += 64 * nc2;
        }
      }
      s_pre[nch] = pre;
      s_nstage = nbs;
    }
    __syncthreads();
    const bool staged = s_nstage <= kStage;
    // grandchildren's balls of the internal children: thread 8 k + j takes
    // grandchild j of child k
    if (tid < 64) {
      const int k = tid >> 3, j = tid & 7;
      if (k < nch && !s_leaf[k] && j < s_nc2[k]) {
        const int g = s_g0[k] + j;
        s_g[s_base[k] + j][0] = c.X[3 * g];
        s_g[s_base[k] + j][1] = c.X[3 * g + 1];
        s_g[s_base[k] + j][2] = c.X[3 * g + 2];
        s_g[s_base[k] + j][3] = c.R[g];
      }
    }
    // bodies of the leaf children, all of them strided over the block
    if (staged)
      for (int i = tid; i < s_nstage; i += blockDim.x) {
        int k = 0;  // the leaf child whose staging range holds slot i
        while (k < nch - 1 && !(s_leaf[k] && i >= s_base[k] && i < s_base[k] + s_nb[k])) k++;
        const int bi = s_body[k] + (i - s_base[k]);
        s_p[0][i] = b.x[bi];
        s_p[1][i] = b.y[bi];
        s_p[2][i] = b.z[bi];
      }
    __syncthreads();
    // point `local` of child k, in the order the reference enumerates them
    auto point = [&](int k, int local, double *q) {
      if (s_leaf[k]) {
        if (staged) {
          const int o = s_base[k] + local;
          q[0] = s_p[0][o];
          q[1] = s_p[1][o];
          q[2] = s_p[2][o];
        } else {
          const int bi = s_body[k] + local;
          q[0] = b.x[bi];
          q[1] = b.y[bi];
          q[2] = b.z[bi];
        }
      } else {
        const int g = s_base[k] + (local >> 6), t = local & 63;
        const double Rg = s_g[g][3];
        for (int d = 0; d < 3; d++) q[d] = s_g[g][d] + Rg * s_td[3 * t + d];
      }
    };
    DBall ball = block_minball<NW>(
        [&](int i, double *q) {
          int k = 0;
          while (k < nch - 1 && i >= s_pre[k + 1]) k++;
          point(k, i - s_pre[k], q);
        },
        [&](auto f) {
          for (int k = 0; k < nch; k++) {
            const int p0 = s_pre[k], cnt = s_pre[k + 1] - p0;
            for (int l = tid; l < cnt; l += NW * kWave) {
              double q[3];
              point(k, l, q);
              f(p0 + l, q);
            }
          }
        },
        s_e, s_i);
    for (int d = 0; d < 3; d++) X[d] = ball.c[d];
    if (tid == 0) {
      double mina = 1e30;
      int sink = 0;
      for (int k = 0; k < nch; k++) {
        if (kWhat & kUpMin) mina = fmin(c.min_acc[ch0 + k], mina);
        sink |= c.has_sink[ch0 + k];
      }
      for (int d = 0; d < 3; d++) c.X[3 * ci + d] = X[d];
      c.R[ci] = sqrt(ball.r2) + 1e-6;
      if (kWhat & kUpMin) c.min_acc[ci] = mina;
      c.has_sink[ci] = (unsigned char)(c.has_sink[ci] | sink);
    }
    __syncthreads();  // the ball's buffers are reused below
  }
  if constexpr (!(kWhat & kUpMoments)) return;
  if constexpr (kPar) {
    // translation, children in parallel: lane = (task group g, child k) with
    // k fastest, each lane a strided share of the output terms of child k;
    // the 8 children's partial sums of one term sit in 8 neighbouring lanes
    // and are added with a 3-step butterfly
    cplx *Gall = reinterpret_cast<cplx *>(smem), *Call = Gall + 8 * NT;
    const int k = tid & 7, g = tid >> 3;
    const bool haveK = k < nch;
    const int ch = ch0 + (haveK ? k : 0);
    if (haveK)
      harm_table_col<P>(g, c.X[3 * ch] - X[0], c.X[3 * ch + 1] - X[1], c.X[3 * ch + 2] - X[2],
                        Gall + k * NT);
    for (int pass = 0; pass < 2; pass++) {
      const double *Msrc = pass ? c.Md : c.M;
      for (int t = tid; t < 8 * NT; t += blockDim.x) {
        const int kc = t / NT, idx = t - kc * NT;
        if (kc < nch) {
          int nn = 0;
          while ((nn + 1) * (nn + 1) <= idx) nn++;
          Call[t] = packed_to_c(Msrc + (size_t)(ch0 + kc) * NT, nn, idx - nn * nn - nn);
        }
      }
      __syncthreads();
      double *dst = (pass ? c.Md : c.M) + (size_t)ci * NT;
      const cplx *src = Call + k * NT, *Gk = Gall + k * NT;
      for (int o = g; o < NO; o += 32) {
        int n = 0;
        while (tri(n + 1, 0) <= o && n < P) n++;
        const int m = o - tri(n, 0);
        cplx acc = {0, 0}, acc2 = {0, 0};
        if (haveK)
          for (int kk = 0; kk <= n; kk++) {
            const int lo = -kk > m - n + kk ? -kk : m - n + kk;
            const int hi = kk < m + n - kk ? kk : m + n - kk;
            // src[ixnm(n - kk, m - l)] runs down and Gk[ixnm(kk, l)] up with
            // l; two partial sums halve the chain of dependent adds
            const cplx *sr = src + ixnm(n - kk, m - lo), *gr = Gk + ixnm(kk, lo);
            const int len = hi - lo + 1;
            int t = 0;
            for (; t + 1 < len; t += 2) {
              acc = cadd(acc, cmul(sr[-t], gr[t]));
              acc2 = cadd(acc2, cmul(sr[-t - 1], gr[t + 1]));
            }
            if (t < len) acc = cadd(acc, cmul(sr[-t], gr[t]));
          }
        acc = cadd(acc, acc2);
#pragma unroll
        for (int x = 1; x < 8; x <<= 1) {
          acc.re += __shfl_xor(acc.re, x, kWave);
          acc.im += __shfl_xor(acc.im, x, kWave);
        }
        if (k == 0) {
          const int s = n * n + n;
          dst[s + m] += acc.re;
          if (m > 0) dst[s - m] += acc.im;
        }
      }
      __syncthreads();  // Call is refilled by the next pass
    }
  } else {
    // translation, one child at a time: tasks (output term, M or Mdot)
    // strided over the block, the child's harmonics and moments in LDS
    cplx *G = reinterpret_cast<cplx *>(smem), *cm = G + NT, *cmd = cm + NT;
    constexpr int NTASK = 2 * NO, NRT = (NTASK + 255) / 256;
    cplx a[NRT];
    int tn[NRT], tm[NRT];
#pragma unroll
    for (int r = 0; r < NRT; r++) {
      a[r] = {0, 0};
      const int j = tid + 256 * r, o = j % NO;
      int n = 0;
      while (tri(n + 1, 0) <= o && n < P) n++;
      tn[r] = n;
      tm[r] = o - tri(n, 0);
    }
    for (int kc = 0; kc < nch; kc++) {
      const int ch = ch0 + kc;
      __syncthreads();
      harm_table_lds<P>(c.X[3 * ch] - X[0], c.X[3 * ch + 1] - X[1], c.X[3 * ch + 2] - X[2], G);
      const double *Mch = c.M + (size_t)ch * NT;
      const double *Mdch = c.Md + (size_t)ch * NT;
      for (int t = tid; t < NT; t += blockDim.x) {
        int nn = 0;
        while ((nn + 1) * (nn + 1) <= t) nn++;
        const int mm = t - nn * nn - nn;
        cm[t] = packed_to_c(Mch, nn, mm);
        cmd[t] = packed_to_c(Mdch, nn, mm);
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < NRT; r++) {
        const int j = tid + 256 * r;
        if (j >= NTASK) continue;
        const cplx *src = j >= NO ? cmd : cm;
        const int n = tn[r], m = tm[r];
        cplx acc = a[r];
        for (int kk = 0; kk <= n; kk++) {
          const int lo = -kk > m - n + kk ? -kk : m - n + kk;
          const int hi = kk < m + n - kk ? kk : m + n - kk;
          for (int l = lo; l <= hi; l++)
            acc = cadd(acc, cmul(src[ixnm(n - kk, m - l)], G[ixnm(kk, l)]));
        }
        a[r] = acc;
      }
    }
#pragma unroll
    for (int r = 0; r < NRT; r++) {
      const int j = tid + 256 * r;
      if (j >= NTASK) continue;
      double *dst = (j >= NO ? c.Md : c.M) + (size_t)ci * NT;
      const int n = tn[r], m = tm[r], s = n * n + n;
      dst[s + m] += a[r].re;
      if (m > 0) dst[s - m] += a[r].im;
    }
  }
}

// the combinations the engine uses: all, geometry (+ min_acc), moments,
// geometry + moments without min_acc, min_acc alone
template <int P>
static void p2m_launch(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, int what,
                       hipStream_t s) {
  const dim3 g(nleaves), t(64);
  switch (what) {
    case kUpGeom | kUpMin:
      hipLaunchKernelGGL((k_p2m<P, kUpGeom | kUpMin>), g, t, 0, s, c, b, leaf_ids);
      break;
    case kUpMoments:
      hipLaunchKernelGGL((k_p2m<P, kUpMoments>), g, t, 0, s, c, b, leaf_ids);
      break;
    case kUpGeom | kUpMoments:
      hipLaunchKernelGGL((k_p2m<P, kUpGeom | kUpMoments>), g, t, 0, s, c, b, leaf_ids);
      break;
    case kUpMin:
      hipLaunchKernelGGL((k_p2m<P, kUpMin>), g, t, 0, s, c, b, leaf_ids);
      break;
    default:
      hipLaunchKernelGGL((k_p2m<P, kUpAll>), g, t, 0, s, c, b, leaf_ids);
  }
}

template <int P>
static void m2m_launch(CellsDev c, BodiesDev b, const int *ids, int ls, int ncell, int what,
                       hipStream_t s) {
  const dim3 g(ncell), t(256);
  switch (what) {
    case kUpGeom | kUpMin:
      hipLaunchKernelGGL((k_m2m<P, kUpGeom | kUpMin>), g, t, 0, s, c, b, ids, ls);
      break;
    case kUpMoments:
      hipLaunchKernelGGL((k_m2m<P, kUpMoments>), g, t, 0, s, c, b, ids, ls);
      break;
    case kUpGeom | kUpMoments:
      hipLaunchKernelGGL((k_m2m<P, kUpGeom | kUpMoments>), g, t, 0, s, c, b, ids, ls);
      break;
    case kUpMin:
      hipLaunchKernelGGL((k_m2m<P, kUpMin>), g, t, 0, s, c, b, ids, ls);
      break;
    default:
      hipLaunchKernelGGL((k_m2m<P, kUpAll>), g, t, 0, s, c, b, ids, ls);
  }
}

int launch_p2m(int p, CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, hipStream_t s,
               int what) {
  if (nleaves <= 0) return 0;
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      p2m_launch<P>(c, b, leaf_ids, nleaves, what, s);\
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

int launch_m2m(int p, CellsDev c, BodiesDev b, int ls, int le, hipStream_t s, int what) {
  if (le <= ls) return 0;
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      m2m_launch<P>(c, b, nullptr, ls, le - ls, what, s);\
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

int launch_m2m_ids(int p, CellsDev c, BodiesDev b, const int *ids, int nids, hipStream_t s) {
  if (nids <= 0) return 0;
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      m2m_launch<P>(c, b, ids, 0, nids, kUpAll, s);\
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
