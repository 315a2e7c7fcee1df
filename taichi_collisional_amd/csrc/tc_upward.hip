// Upward pass on the device: cell geometry (smallest enclosing balls), P2M and
// M2M for moments M and their time derivatives Mdot (taichi_with_jerks/
// kernel_fast_lean.h:701-999, driven post-order by traverse_eager.h:9-32).
//
//   P2M  one wave per leaf: the ball of the leaf's bodies (centre X, radius
//        r + 1e-6), then lane = body: M += q Y(x_b - X), Mdot += q dY/dt(v_b)
//        with the absolute body velocity; per order the contributions are
//        summed across the wave.  min_acc / has_sink are reduced alongside.
//   M2M  one wave per internal cell (one launch per level, deepest first):
//        the ball of the bodies of leaf children and of 64 design points on
//        each grandchild sphere of non-leaf children, then for each child
//        M_parent[n,m] += sum_{k,l} M_child[n-k, m-l] Y_k^l(X_child - X_parent)
//        with lanes over the output terms and the child's harmonics staged
//        in LDS.
#include "tc_common.h"
#include "tc_geom.h"
#include "tc_harm.h"
#include "tc_kernels.h"
#include "tdesign.h"

namespace tcfmm {

__constant__ double c_tdesign[192];

template <int P>
__global__ __launch_bounds__(64) void k_p2m(CellsDev c, BodiesDev b,
                                            const int *__restrict__ leaf_ids) {
  constexpr int NT = nterm(P);
  constexpr int NY = tri(P, P) + 1;
  const int lane = threadIdx.x;
  const int ci = leaf_ids[blockIdx.x];
  const int b0 = c.body[ci], nb = c.nbody[ci];
  double X[3], R;
  if (nb > 0) {
    DBall ball = wave_minball(nb, [&](int i, double *q) {
      q[0] = b.x[b0 + i];
      q[1] = b.y[b0 + i];
      q[2] = b.z[b0 + i];
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
    mina = fmin(b.acc_old[b0 + i], mina);
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
    c.min_acc[ci] = mina;
    c.has_sink[ci] = (unsigned char)sink;
  }
  double *Mc = c.M + (size_t)ci * NT;
  double *Mdc = c.Md + (size_t)ci * NT;
  for (int base = 0; base < nb || base == 0; base += kWave) {
    const int i = b0 + base + lane;
    const bool act = base + lane < nb;
    double dx = 0, dy = 0, dz = 0, vx = 0, vy = 0, vz = 0, w = 0;
    if (act) {
      dx = b.x[i] - X[0];
      dy = b.y[i] - X[1];
      dz = b.z[i] - X[2];
      vx = b.vx[i];
      vy = b.vy[i];
      vz = b.vz[i];
      w = (b.flags[i] & kSource) ? b.q[i] : 0.0;
    }
    cplx y[NY];
    harm_all<P>(dx, dy, dz, y);
#pragma unroll
    for (int n = 0; n <= P; n++) {
      const int s = n * n + n;
#pragma unroll
      for (int m = 0; m <= n; m++) {
        cplx g = y[tri(n, m)];
        cplx gd = harm_dot(y, n, m, vx, vy, vz);
        double a0 = wave_sum(w * g.re), a1 = (m > 0) ? wave_sum(w * g.im) : 0.0;
        double d0 = wave_sum(w * gd.re), d1 = (m > 0) ? wave_sum(w * gd.im) : 0.0;
        if (lane == 0) {
          if (base == 0) {
            Mc[s + m] = a0;
            Mdc[s + m] = d0;
            if (m > 0) {
              Mc[s - m] = a1;
              Mdc[s - m] = d1;
            }
          } else {
            Mc[s + m] += a0;
            Mdc[s + m] += d0;
            if (m > 0) {
              Mc[s - m] += a1;
              Mdc[s - m] += d1;
            }
          }
        }
      }
    }
    if (nb == 0) break;
  }
}

template <int P>
__global__ __launch_bounds__(64) void k_m2m(CellsDev c, BodiesDev b, int ls) {
#pragma clang fp contract(off)
  constexpr int NT = nterm(P);
  constexpr int NO = tri(P, P) + 1;
  constexpr int RO = (NO + kWave - 1) / kWave;
  __shared__ cplx G[NT], cm[NT], cmd[NT];
  const int lane = threadIdx.x;
  const int ci = ls + blockIdx.x;
  const int nch = c.nchild[ci];
  if (nch == 0) return;
  const int ch0 = c.child[ci];
  int pre[9];
  pre[0] = 0;
  double mina = 1e30;
  int sink = 0;
  for (int k = 0; k < 8; k++) {
    int cnt = 0;
    if (k < nch) {
      int ch = ch0 + k;
      cnt = c.nchild[ch] == 0 ? c.nbody[ch] : 64 * c.nchild[ch];
      mina = fmin(c.min_acc[ch], mina);
      sink |= c.has_sink[ch];
    }
    pre[k + 1] = pre[k] + cnt;
  }
  auto get = [&](int i, double *q) {
    int k = 0;
    while (k < nch - 1 && i >= pre[k + 1]) k++;
    int ch = ch0 + k;
    int local = i - pre[k];
    if (c.nchild[ch] == 0) {
      int bi = c.body[ch] + local;
      q[0] = b.x[bi];
      q[1] = b.y[bi];
      q[2] = b.z[bi];
    } else {
      int g = c.child[ch] + (local >> 6);
      int t = local & 63;
      double Rg = c.R[g];
      for (int d = 0; d < 3; d++) q[d] = c.X[3 * g + d] + Rg * c_tdesign[3 * t + d];
    }
  };
  DBall ball = wave_minball(pre[nch], get);
  double X[3] = {ball.c[0], ball.c[1], ball.c[2]};
  if (lane == 0) {
    for (int d = 0; d < 3; d++) c.X[3 * ci + d] = X[d];
    c.R[ci] = sqrt(ball.r2) + 1e-6;
    c.min_acc[ci] = mina;
    c.has_sink[ci] = (unsigned char)sink;
  }
  cplx acc[RO], accd[RO];
  int on[RO], om[RO];
#pragma unroll
  for (int r = 0; r < RO; r++) {
    acc[r] = {0, 0};
    accd[r] = {0, 0};
    int o = lane + kWave * r, n = 0;
    while (tri(n + 1, 0) <= o) n++;
    on[r] = n;
    om[r] = o - tri(n, 0);
  }
  for (int k = 0; k < nch; k++) {
    const int ch = ch0 + k;
    __syncthreads();
    harm_table_lds<P>(c.X[3 * ch] - X[0], c.X[3 * ch + 1] - X[1], c.X[3 * ch + 2] - X[2], G);
    const double *Mch = c.M + (size_t)ch * NT;
    const double *Mdch = c.Md + (size_t)ch * NT;
    for (int t = lane; t < NT; t += kWave) {
      int n = 0;
      while ((n + 1) * (n + 1) <= t) n++;
      int m = t - n * n - n;
      cm[t] = packed_to_c(Mch, n, m);
      cmd[t] = packed_to_c(Mdch, n, m);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RO; r++) {
      int o = lane + kWave * r;
      if (o >= NO) continue;
      const int n = on[r], m = om[r];
      cplx a = acc[r], ad = accd[r];
      for (int kk = 0; kk <= n; kk++) {
        int lo = -kk > m - n + kk ? -kk : m - n + kk;
        int hi = kk < m + n - kk ? kk : m + n - kk;
        for (int l = lo; l <= hi; l++) {
          cplx g = G[ixnm(kk, l)];
          int src = ixnm(n - kk, m - l);
          a = cadd(a, cmul(cm[src], g));
          ad = cadd(ad, cmul(cmd[src], g));
        }
      }
      acc[r] = a;
      accd[r] = ad;
    }
  }
  double *Mc = c.M + (size_t)ci * NT;
  double *Mdc = c.Md + (size_t)ci * NT;
#pragma unroll
  for (int r = 0; r < RO; r++) {
    int o = lane + kWave * r;
    if (o >= NO) continue;
    const int n = on[r], m = om[r], s = n * n + n;
    Mc[s + m] = acc[r].re;
    Mdc[s + m] = accd[r].re;
    if (m > 0) {
      Mc[s - m] = acc[r].im;
      Mdc[s - m] = accd[r].im;
    }
  }
}

static void upload_tdesign() {
  static bool done = false;
  if (done) return;
  hipMemcpyToSymbol(HIP_SYMBOL(c_tdesign), kTDesign, sizeof(double) * 192);
  done = true;
}

template <int P>
static void p2m_launch(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, hipStream_t s) {
  hipLaunchKernelGGL(k_p2m<P>, dim3(nleaves), dim3(64), 0, s, c, b, leaf_ids);
}

template <int P>
static void m2m_launch(CellsDev c, BodiesDev b, int ls, int le, hipStream_t s) {
  hipLaunchKernelGGL(k_m2m<P>, dim3(le - ls), dim3(64), 0, s, c, b, ls);
}

int launch_p2m(int p, CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, hipStream_t s) {
  if (nleaves <= 0) return 0;
  switch (p) {
#define TC_CASE(P)                            \
  case P:                                     \
    p2m_launch<P>(c, b, leaf_ids, nleaves, s); \
    return 0;
    TC_FOR_EACH_P(TC_CASE)
#undef TC_CASE
    default:
      return -1;
  }
}

int launch_m2m(int p, CellsDev c, BodiesDev b, int ls, int le, hipStream_t s) {
  upload_tdesign();
  if (le <= ls) return 0;
  switch (p) {
#define TC_CASE(P)                     \
  case P:                              \
    m2m_launch<P>(c, b, ls, le, s);    \
    return 0;
    TC_FOR_EACH_P(TC_CASE)
#undef TC_CASE
    default:
      return -1;
  }
}

}  // namespace tcfmm
