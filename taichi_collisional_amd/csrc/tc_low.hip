// The p = 2 pre-pass of the error-controlled MAC path and the multipole
// norms it feeds (taichi_with_jerks/integrator.h:326-399).
//
// The pre-pass estimates |a| of every body with a cheap FMM on the tree's own
// geometry (cell centre Xg, half-width Rg), kernel_fast_lean.h:1446-1780:
//   P2M_low   M = sum q Y(x_b - Xg)                       (one thread per leaf)
//   M2M_low   O(p^4) shift by Y(Xg_child - Xg)            (one thread per cell)
//   M2L_low   the imaginary parts of M are negated, then
//             L_n^m += sum_{k,q} M_k^q Theta_{n+k}^{m+q}(Xg_i - Xg_j)
//             (one thread per target cell over its list; no atomics)
//   P2P_low   FP32 pair sums, a += m dx / r^3 per source cell, the cell's sum
//             added in FP64 (one lane per target body)
//   L2L_low, L2P_low  shifts by conj Y(Xg_parent - Xg_child), then n <= 1 of
//             the local field: a -= (Re Phi_1^1, Im Phi_1^1, Re Phi_1^0)
// The list comes from the dual traversal with |dXg|^2 > 3 (Rg_i + Rg_j)^2
// (traverse_eager.h:172-212).  acc_old = |a| (integrator.h:350-356).
//
// Pn: Pn[c][n] = sqrt(sum_m M_n^m^2 (n+m)! (n-m)!) after the main upward
// pass (integrator.h:387-399), with the reference's factorial table.
#include "tc_common.h"
#include "tc_kernels.h"

namespace tcfmm {

namespace {

constexpr int LP = 2, LNT = 9;

__device__ __forceinline__ int lix(int n, int m) { return n * n + n + m; }

// make_Gnm at p = 2, all -n <= m <= n.
__device__ __forceinline__ void make_g2(double dx, double dy, double dz, cplx *G) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  G[0] = {1.0, 0.0};
  for (int n = 1; n <= LP; n++) {
    G[lix(n, n)] = cmul(G[lix(n - 1, n - 1)], {dx / (2 * n), dy / (2 * n)});
    for (int m = 0; m < n - 1; m++) {
      const double q = (double)(n * n - m * m);
      const double a = (2 * n - 1.) / q * dz, b = r2 / q;
      G[lix(n, m)] = {a * G[lix(n - 1, m)].re - b * G[lix(n - 2, m)].re,
                      a * G[lix(n - 1, m)].im - b * G[lix(n - 2, m)].im};
    }
    G[lix(n, n - 1)] = cscale(dz, G[lix(n - 1, n - 1)]);
    double s = -1;
    for (int m = 1; m <= n; m++) {
      G[lix(n, -m)] = cconj(cscale(s, G[lix(n, m)]));
      s = -s;
    }
  }
}

// make_Tnm at p = 2 (singular harmonics), all -n <= m <= n.
__device__ __forceinline__ void make_t2(double dx, double dy, double dz, cplx *T) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double ir2 = 1 / r2;
  T[0] = {sqrt(ir2), 0.0};
  for (int n = 1; n <= LP; n++) {
    T[lix(n, n)] = cmul(T[lix(n - 1, n - 1)], {dx * (2 * n - 1) * ir2, dy * (2 * n - 1) * ir2});
    for (int m = 0; m < n - 1; m++) {
      const double a = ir2 * (2 * n - 1.) * dz, b = ir2 * (double)((n - 1) * (n - 1) - m * m);
      T[lix(n, m)] = {a * T[lix(n - 1, m)].re - b * T[lix(n - 2, m)].re,
                      a * T[lix(n - 1, m)].im - b * T[lix(n - 2, m)].im};
    }
    T[lix(n, n - 1)] = cscale(ir2 * (2 * n - 1.) * dz, T[lix(n - 1, n - 1)]);
    double s = -1;
    for (int m = 1; m <= n; m++) {
      T[lix(n, -m)] = cconj(cscale(s, T[lix(n, m)]));
      s = -s;
    }
  }
}

// real_2_complex / complex_2_real at p = 2.
__device__ __forceinline__ void r2c2(const double *a, cplx *c) {
  for (int n = 0; n <= LP; n++) {
    const int s = lix(n, 0);
    c[s] = {a[s], 0.0};
    double f = -1;
    for (int m = 1; m <= n; m++) {
      c[s - m] = {a[s + m] * f, -a[s - m] * f};
      c[s + m] = {a[s + m], a[s - m]};
      f = -f;
    }
  }
}
__device__ __forceinline__ void c2r2(const cplx *c, double *a) {
  for (int n = 0; n <= LP; n++) {
    const int s = lix(n, 0);
    a[s] = c[s].re;
    for (int m = 1; m <= n; m++) {
      a[s - m] = c[s + m].im;
      a[s + m] = c[s + m].re;
    }
  }
}

}  // namespace

__global__ void k_low_p2m(CellsDev c, BodiesDev b, const int *__restrict__ leaf_ids, int nleaves,
                          double *__restrict__ lowM) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int ci = leaf_ids[k];
  const double X0 = c.Xg[3 * ci], X1 = c.Xg[3 * ci + 1], X2 = c.Xg[3 * ci + 2];
  cplx cm[LNT], G[LNT];
  for (int t = 0; t < LNT; t++) cm[t] = {0.0, 0.0};
  const int b0 = c.body[ci], nb = c.nbody[ci];
  for (int i = b0; i < b0 + nb; i++) {
    make_g2(b.x[i] - X0, b.y[i] - X1, b.z[i] - X2, G);
    const double q = b.q[i];
    for (int t = 0; t < LNT; t++) cm[t] = cadd(cm[t], cscale(q, G[t]));
  }
  double r[LNT];
  c2r2(cm, r);
  for (int t = 0; t < LNT; t++) lowM[LNT * ci + t] = r[t];
}

__global__ void k_low_m2m(CellsDev c, int ls, int le, double *__restrict__ lowM) {
  const int ci = ls + blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= le) return;
  const int nch = c.nchild[ci];
  if (nch == 0) return;
  cplx cm[LNT], cj[LNT], G[LNT];
  for (int t = 0; t < LNT; t++) cm[t] = {0.0, 0.0};
  for (int k = c.child[ci]; k < c.child[ci] + nch; k++) {
    r2c2(lowM + LNT * k, cj);
    make_g2(c.Xg[3 * k] - c.Xg[3 * ci], c.Xg[3 * k + 1] - c.Xg[3 * ci + 1],
            c.Xg[3 * k + 2] - c.Xg[3 * ci + 2], G);
    for (int n = 0; n <= LP; n++)
      for (int m = 0; m <= n; m++) {
        const int t = lix(n, m);
        for (int kk = 0; kk <= n; kk++)
          for (int l = max(-kk, m - n + kk); l <= min(kk, m + n - kk); l++)
            cm[t] = cadd(cm[t], cmul(cj[lix(n - kk, m - l)], G[lix(kk, l)]));
      }
  }
  double r[LNT];
  c2r2(cm, r);
  for (int t = 0; t < LNT; t++) lowM[LNT * ci + t] = r[t];
}

__global__ void k_low_m2l(CellsDev c, int nc, const int *__restrict__ off,
                          const int *__restrict__ cnt, const int *__restrict__ src,
                          const double *__restrict__ lowM, double *__restrict__ lowL) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const int e0 = off[i], ne = cnt[i];
  if (ne == 0) return;
  double acc[LNT];
  for (int t = 0; t < LNT; t++) acc[t] = 0.0;
  for (int e = e0; e < e0 + ne; e++) {
    const int j = src[e];
    double mm[LNT];
    for (int t = 0; t < LNT; t++) mm[t] = lowM[LNT * j + t];
    for (int n = 1; n <= LP; n++)
      for (int m = 1; m <= n; m++) mm[lix(n, -m)] *= -1;
    cplx cj[LNT], T[LNT], l[LNT];
    r2c2(mm, cj);
    make_t2(c.Xg[3 * i] - c.Xg[3 * j], c.Xg[3 * i + 1] - c.Xg[3 * j + 1],
            c.Xg[3 * i + 2] - c.Xg[3 * j + 2], T);
    for (int t = 0; t < LNT; t++) l[t] = {0.0, 0.0};
    for (int n = 0; n <= LP; n++)
      for (int m = 0; m <= n; m++)
        for (int k = 0; k <= LP - n; k++)
          for (int q = -k; q <= k; q++)
            l[lix(n, m)] = cadd(l[lix(n, m)], cmul(cj[lix(k, q)], T[lix(n + k, m + q)]));
    double r[LNT];
    c2r2(l, r);
    for (int t = 0; t < LNT; t++) acc[t] += r[t];
  }
  for (int t = 0; t < LNT; t++) lowL[LNT * i + t] += acc[t];
}

// The bodies as P2P_low reads them: (float) x, y, z, q, one 16-byte record each.
__global__ void k_low_stage(BodiesDev b, int n, float4 *__restrict__ P) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) P[i] = make_float4((float)b.x[i], (float)b.y[i], (float)b.z[i], (float)b.q[i]);
}

// One block of 64 per target leaf, lane = target body; the source bodies of
// each listed leaf come from the float records (every lane reads the same
// one: one load per source body instead of four and their conversions).
__global__ __launch_bounds__(64) void k_low_p2p(CellsDev c, BodiesDev b,
                                                const int *__restrict__ leaf_ids,
                                                const int *__restrict__ off,
                                                const int *__restrict__ cnt,
                                                const int *__restrict__ src,
                                                const float4 *__restrict__ P,
                                                double *__restrict__ F) {
  const int ci = leaf_ids[blockIdx.x];
  const int b0 = c.body[ci], nb = c.nbody[ci];
  const int e0 = off[ci], ne = cnt[ci];
  for (int base = 0; base < nb; base += kWave) {
    const int i = b0 + base + (int)threadIdx.x;
    const bool act = base + (int)threadIdx.x < nb;
    const int ii = act ? i : b0;
    const float xi = -(float)b.x[ii], yi = -(float)b.y[ii], zi = -(float)b.z[ii];
    double fx = 0, fy = 0, fz = 0;
    for (int e = e0; e < e0 + ne; e++) {
      const int j = src[e];
      const int j0 = c.body[j], nj = c.nbody[j];
      float ax = 0, ay = 0, az = 0;
      for (int y = j0; y < j0 + nj; y++) {
        const float4 q = P[y];
        const float mj = q.w;
        const float dx = xi + q.x, dy = yi + q.y, dz = zi + q.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 > 0)) continue;
        const float rr = sqrtf(r2);
        const float ir = 1 / rr;
        const float f = mj * ir * ir * ir;
        ax += dx * f;
        ay += dy * f;
        az += dz * f;
      }
      fx += ax;
      fy += ay;
      fz += az;
    }
    if (act) {
      F[3 * i] += fx;
      F[3 * i + 1] += fy;
      F[3 * i + 2] += fz;
    }
  }
}

__global__ void k_low_l2l(CellsDev c, int ls, int le, double *__restrict__ lowL) {
  const int ci = ls + blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= le) return;
  const int nch = c.nchild[ci];
  if (nch == 0) return;
  cplx cl[LNT], G[LNT];
  r2c2(lowL + LNT * ci, cl);
  for (int k = c.child[ci]; k < c.child[ci] + nch; k++) {
    make_g2(c.Xg[3 * ci] - c.Xg[3 * k], c.Xg[3 * ci + 1] - c.Xg[3 * k + 1],
            c.Xg[3 * ci + 2] - c.Xg[3 * k + 2], G);
    for (int t = 0; t < LNT; t++) G[t] = cconj(G[t]);
    cplx o[LNT];
    for (int t = 0; t < LNT; t++) o[t] = {0.0, 0.0};
    for (int n = 0; n <= LP; n++)
      for (int m = 0; m <= n; m++)
        for (int kk = 0; kk <= LP - n; kk++)
          for (int l = -kk; l <= kk; l++)
            o[lix(n, m)] = cadd(o[lix(n, m)], cmul(cl[lix(n + kk, m + l)], G[lix(kk, l)]));
    double r[LNT];
    c2r2(o, r);
    for (int t = 0; t < LNT; t++) lowL[LNT * k + t] += r[t];
  }
}

__global__ void k_low_l2p(CellsDev c, BodiesDev b, const int *__restrict__ leaf_ids, int nleaves,
                          const double *__restrict__ lowL, double *__restrict__ F) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int ci = leaf_ids[k];
  cplx cl[LNT], G[LNT];
  r2c2(lowL + LNT * ci, cl);
  const double X0 = c.Xg[3 * ci], X1 = c.Xg[3 * ci + 1], X2 = c.Xg[3 * ci + 2];
  const int b0 = c.body[ci], nb = c.nbody[ci];
  for (int i = b0; i < b0 + nb; i++) {
    make_g2(X0 - b.x[i], X1 - b.y[i], X2 - b.z[i], G);
    for (int t = 0; t < LNT; t++) G[t] = cconj(G[t]);
    cplx phi[4];
    for (int t = 0; t < 4; t++) phi[t] = {0.0, 0.0};
    for (int n = 0; n <= 1; n++)
      for (int m = 0; m <= n; m++)
        for (int kk = 0; kk <= LP - n; kk++)
          for (int l = -kk; l <= kk; l++)
            phi[lix(n, m)] = cadd(phi[lix(n, m)], cmul(cl[lix(n + kk, m + l)], G[lix(kk, l)]));
    F[3 * i] -= phi[3].re;
    F[3 * i + 1] -= phi[3].im;
    F[3 * i + 2] -= phi[2].re;
  }
}

__global__ void k_acc_old(BodiesDev b, int n, const double *__restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double fx = F[3 * i], fy = F[3 * i + 1], fz = F[3 * i + 2];
  b.acc_old[i] = sqrt(fx * fx + fy * fy + fz * fz);
}

__global__ void k_pn(const double *__restrict__ M, int nc, int p, FactTable ft,
                     double *__restrict__ Pn) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= nc) return;
  const int NT = (p + 1) * (p + 1);
  const double *Mc = M + (size_t)ci * NT;
  for (int n = 0; n <= p; n++) {
    double s = 0;
    for (int m = -n; m <= n; m++) {
      const double a = Mc[n * n + n + m];
      s += a * a * ft.fact[n + m] * ft.fact[n - m];
    }
    Pn[(size_t)ci * (p + 1) + n] = sqrt(s);
  }
}

__global__ void k_scatter_acc_old(BodiesDev b, int n, double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[b.order[k]] = b.acc_old[k];
}

// Master P2P_low (master kernel_fast_lean.h:1479-1556): the vector loop
// overwrites its register at every source, so each listed source leaf adds
// q / r^2 of its last body only (0 at r = 0), in FP64.  One lane per target.
__global__ __launch_bounds__(64) void k_low_p2p_master(CellsDev c, BodiesDev b,
                                                       const int *__restrict__ leaf_ids,
                                                       const int *__restrict__ off,
                                                       const int *__restrict__ cnt,
                                                       const int *__restrict__ src,
                                                       double *__restrict__ F) {
#pragma clang fp contract(off)
  const int ci = leaf_ids[blockIdx.x];
  const int b0 = c.body[ci], nb = c.nbody[ci];
  const int e0 = off[ci], ne = cnt[ci];
  for (int base = 0; base < nb; base += kWave) {
    const int i = b0 + base + (int)threadIdx.x;
    if (base + (int)threadIdx.x >= nb) break;
    const double xi = b.x[i], yi = b.y[i], zi = b.z[i];
    double near = 0;
    for (int e = e0; e < e0 + ne; e++) {
      const int j = src[e];
      const int y = c.body[j] + c.nbody[j] - 1;
      const double dx = xi - b.x[y], dy = yi - b.y[y], dz = zi - b.z[y];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > 0) near += b.q[y] / r2;
    }
    F[3 * i] += near;
  }
}

// Master L2P_low (:1609-1646): every body of the leaf gets the magnitude of
// the far-field gradient, sqrt(Re Phi_1^1^2 + Im Phi_1^1^2 + Re Phi_1^0^2).
__global__ void k_low_l2p_master(CellsDev c, BodiesDev b, const int *__restrict__ leaf_ids,
                                 int nleaves, const double *__restrict__ lowL,
                                 double *__restrict__ F) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int ci = leaf_ids[k];
  cplx cl[LNT], G[LNT];
  r2c2(lowL + LNT * ci, cl);
  const double X0 = c.Xg[3 * ci], X1 = c.Xg[3 * ci + 1], X2 = c.Xg[3 * ci + 2];
  const int b0 = c.body[ci], nb = c.nbody[ci];
  for (int i = b0; i < b0 + nb; i++) {
    make_g2(X0 - b.x[i], X1 - b.y[i], X2 - b.z[i], G);
    for (int t = 0; t < LNT; t++) G[t] = cconj(G[t]);
    cplx phi[4];
    for (int t = 0; t < 4; t++) phi[t] = {0.0, 0.0};
    for (int n = 0; n <= 1; n++)
      for (int m = 0; m <= n; m++)
        for (int kk = 0; kk <= LP - n; kk++)
          for (int l = -kk; l <= kk; l++)
            phi[lix(n, m)] = cadd(phi[lix(n, m)], cmul(cl[lix(n + kk, m + l)], G[lix(kk, l)]));
    F[3 * i + 1] +=
        sqrt(phi[3].re * phi[3].re + phi[3].im * phi[3].im + phi[2].re * phi[2].re);
  }
}

__global__ void k_acc_old_master(BodiesDev b, int n, const double *__restrict__ F,
                                 const double *__restrict__ ain) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a0 = ain ? ain[b.order[i]] : 0.0;
  b.acc_old[i] = a0 + (F[3 * i] + F[3 * i + 1]);
}

static inline dim3 blocks_for(int n, int t) { return dim3((n + t - 1) / t); }

int launch_low_p2p_master(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                          const int *p2p_off, const int *p2p_cnt, const int *p2p_src, double *F,
                          hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_low_p2p_master, dim3(nleaves), dim3(64), 0, s, c, b, leaf_ids, p2p_off,
                       p2p_cnt, p2p_src, F);
  return 0;
}

int launch_low_downward_master(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                               const std::vector<int> &level_start, double *lowL, double *F,
                               int n, const double *ain, hipStream_t s) {
  const int L = (int)level_start.size() - 1;
  for (int l = 0; l < L; l++) {
    const int ls = level_start[l], le = level_start[l + 1];
    if (le > ls)
      hipLaunchKernelGGL(k_low_l2l, blocks_for(le - ls, 128), dim3(128), 0, s, c, ls, le, lowL);
  }
  if (nleaves > 0)
    hipLaunchKernelGGL(k_low_l2p_master, blocks_for(nleaves, 128), dim3(128), 0, s, c, b,
                       leaf_ids, nleaves, lowL, F);
  if (n > 0)
    hipLaunchKernelGGL(k_acc_old_master, blocks_for(n, 256), dim3(256), 0, s, b, n, F, ain);
  return 0;
}

int launch_low_upward(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                      const std::vector<int> &level_start, double *lowM, hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_low_p2m, blocks_for(nleaves, 128), dim3(128), 0, s, c, b, leaf_ids,
                       nleaves, lowM);
  const int L = (int)level_start.size() - 1;
  for (int l = L - 1; l >= 0; l--) {
    const int ls = level_start[l], le = level_start[l + 1];
    if (le > ls)
      hipLaunchKernelGGL(k_low_m2m, blocks_for(le - ls, 128), dim3(128), 0, s, c, ls, le, lowM);
  }
  return 0;
}

int launch_low_horizontal(CellsDev c, BodiesDev b, int nc, const int *leaf_ids, int nleaves,
                          const int *m2l_off, const int *m2l_cnt, const int *m2l_src,
                          const int *p2p_off, const int *p2p_cnt, const int *p2p_src,
                          const double *lowM, double *lowL, double *F, int n, float4 *P,
                          hipStream_t s) {
  if (nc > 0)
    hipLaunchKernelGGL(k_low_m2l, blocks_for(nc, 128), dim3(128), 0, s, c, nc, m2l_off, m2l_cnt,
                       m2l_src, lowM, lowL);
  if (nleaves > 0) {
    hipLaunchKernelGGL(k_low_stage, dim3((n + 255) / 256), dim3(256), 0, s, b, n, P);
    hipLaunchKernelGGL(k_low_p2p, dim3(nleaves), dim3(64), 0, s, c, b, leaf_ids, p2p_off,
                       p2p_cnt, p2p_src, (const float4 *)P, F);
  }
  return 0;
}

int launch_low_downward(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                        const std::vector<int> &level_start, double *lowL, double *F, int n,
                        hipStream_t s) {
  const int L = (int)level_start.size() - 1;
  for (int l = 0; l < L; l++) {
    const int ls = level_start[l], le = level_start[l + 1];
    if (le > ls)
      hipLaunchKernelGGL(k_low_l2l, blocks_for(le - ls, 128), dim3(128), 0, s, c, ls, le, lowL);
  }
  if (nleaves > 0)
    hipLaunchKernelGGL(k_low_l2p, blocks_for(nleaves, 128), dim3(128), 0, s, c, b, leaf_ids,
                       nleaves, lowL, F);
  if (n > 0) hipLaunchKernelGGL(k_acc_old, blocks_for(n, 256), dim3(256), 0, s, b, n, F);
  return 0;
}

int launch_scatter_acc_old(BodiesDev b, int n, double *out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_scatter_acc_old, blocks_for(n, 256), dim3(256), 0, s, b, n, out);
  return 0;
}

int launch_pn(const double *M, int nc, int p, const FactTable &ft, double *Pn, hipStream_t s) {
  if (nc <= 0) return 0;
  hipLaunchKernelGGL(k_pn, blocks_for(nc, 128), dim3(128), 0, s, M, nc, p, ft, Pn);
  return 0;
}

}  // namespace tcfmm
