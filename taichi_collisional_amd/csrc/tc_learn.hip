// Learned per-cell expansion order and opening angle.
//
// A small MLP (11 inputs, three hidden layers of 64 tanh units, one linear
// output, FP32) predicts log10 of a leaf's rms relative force error from the
// leaf's features and a candidate (order q, opening angle theta).  It is
// fitted with an asymmetric (expectile) squared loss that weighs
// under-predicted errors more, so that its predictions lean to the safe side.  Before the
// walk, every leaf takes the cheapest candidate whose predicted error is below
// the target (or, if none is, the most accurate one); internal cells take the
// largest order and the smallest angle below them, so a leaf's ancestors are
// at least as accurate as the leaf.  The walk then accepts a pair (Ci, Cj)
// with Ci's own angle, and M2L truncates the pair at Ci's order.
//
// The network is trained on the GPU by the kernels below (forward, backward,
// weight gradients, Adam) on per-leaf errors that the FMM measures against
// the FP64 direct sum (k_leaf_errors).  The reference has no learned part;
// its error-controlled walk (taichi_with_jerks/traverse_eager.h:59-122) and
// the multipole norms it uses (integrator.h:387-399) are the analogues: the
// error estimate there is a formula, here it is fitted.
//
// Parameter layout (floats): W1[64][11] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[64] b4[1]; 9,153 values.  Weights are read at addresses that are
// the same on every lane, so they load through the scalar cache and enter
// the FMAs as scalar operands.
#include <hip/hip_runtime.h>

#include "tc_common.h"
#include "tc_kernels.h"

namespace tcfmm {

constexpr int kW1 = 0, kB1 = kW1 + kMlpH * kMlpIn, kW2 = kB1 + kMlpH, kB2 = kW2 + kMlpH * kMlpH,
              kW3 = kB2 + kMlpH, kB3 = kW3 + kMlpH * kMlpH, kW4 = kB3 + kMlpH, kB4 = kW4 + kMlpH;
static_assert(kB4 + 1 == kMlpParams, "parameter layout");

// One dense layer with tanh: out[j] = tanh(b[j] + sum_i W[j][i] in[i]).
template <int K>
__device__ __forceinline__ void dense_tanh(const float *__restrict__ W,
                                           const float *__restrict__ b, const float *in,
                                           float *out) {
#pragma unroll
  for (int j = 0; j < kMlpH; j++) {
    float a = b[j];
#pragma unroll
    for (int i = 0; i < K; i++) a = fmaf(W[j * K + i], in[i], a);
    out[j] = tanhf(a);
  }
}

__device__ __forceinline__ float mlp_eval(const float *__restrict__ P, const float *x) {
  float h[kMlpH], g[kMlpH];
  dense_tanh<kMlpIn>(P + kW1, P + kB1, x, h);
  dense_tanh<kMlpH>(P + kW2, P + kB2, h, g);
  dense_tanh<kMlpH>(P + kW3, P + kB3, g, h);
  float y = P[kB4];
#pragma unroll
  for (int j = 0; j < kMlpH; j++) y = fmaf(P[kW4 + j], h[j], y);
  return y;
}

// ---------------------------------------------------------------- training
// Forward over a batch: Y[b]; with H, the three hidden layers H[l][b][64].
__global__ __launch_bounds__(64) void k_mlp_fwd(const float *__restrict__ P,
                                                const float *__restrict__ X, int B,
                                                float *__restrict__ Y, float *__restrict__ H) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= B) return;
  float x[kMlpIn], h[kMlpH], g[kMlpH];
#pragma unroll
  for (int i = 0; i < kMlpIn; i++) x[i] = X[(size_t)b * kMlpIn + i];
  dense_tanh<kMlpIn>(P + kW1, P + kB1, x, h);
  dense_tanh<kMlpH>(P + kW2, P + kB2, h, g);
  if (H) {
#pragma unroll
    for (int j = 0; j < kMlpH; j++) {
      H[(size_t)b * kMlpH + j] = h[j];
      H[((size_t)B + b) * kMlpH + j] = g[j];
    }
  }
  dense_tanh<kMlpH>(P + kW3, P + kB3, g, h);
  float y = P[kB4];
#pragma unroll
  for (int j = 0; j < kMlpH; j++) y = fmaf(P[kW4 + j], h[j], y);
  if (H) {
#pragma unroll
    for (int j = 0; j < kMlpH; j++) H[(2 * (size_t)B + b) * kMlpH + j] = h[j];
  }
  Y[b] = y;
}

// Backward of the expectile loss (1/B) sum_b w_b (Y[b] - T[b])^2, w_b = tau
// where Y[b] < T[b] (an under-predicted error) and 1 - tau elsewhere, through
// the output and the three tanh layers: the layer deltas D[l][b][64]
// (dLoss/dpre-activation), dY[b], and the loss (added into *loss).
__global__ __launch_bounds__(64) void k_mlp_bwd(const float *__restrict__ P,
                                                const float *__restrict__ H,
                                                const float *__restrict__ Y,
                                                const float *__restrict__ T, int B, float tau,
                                                float *__restrict__ D, float *__restrict__ dY,
                                                float *__restrict__ loss) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  float l = 0.0f;
  if (b < B) {
    const float r = Y[b] - T[b];
    const float w = r < 0.0f ? tau : 1.0f - tau;
    l = w * r * r / (float)B;
    const float dy = 2.0f * w * r / (float)B;
    dY[b] = dy;
    float d3[kMlpH], d2[kMlpH];
#pragma unroll
    for (int j = 0; j < kMlpH; j++) {
      const float h = H[(2 * (size_t)B + b) * kMlpH + j];
      d3[j] = dy * P[kW4 + j] * (1.0f - h * h);
      D[(2 * (size_t)B + b) * kMlpH + j] = d3[j];
    }
#pragma unroll
    for (int i = 0; i < kMlpH; i++) {
      float a = 0.0f;
#pragma unroll
      for (int j = 0; j < kMlpH; j++) a = fmaf(P[kW3 + j * kMlpH + i], d3[j], a);
      const float h = H[((size_t)B + b) * kMlpH + i];
      d2[i] = a * (1.0f - h * h);
      D[((size_t)B + b) * kMlpH + i] = d2[i];
    }
#pragma unroll
    for (int i = 0; i < kMlpH; i++) {
      float a = 0.0f;
#pragma unroll
      for (int j = 0; j < kMlpH; j++) a = fmaf(P[kW2 + j * kMlpH + i], d2[j], a);
      const float h = H[(size_t)b * kMlpH + i];
      D[(size_t)b * kMlpH + i] = a * (1.0f - h * h);
    }
  }
  l = wave_sum_f(l);
  if (threadIdx.x == 0) atomicAdd(loss, l);
}

// Weight gradients G += sum_b (outer products of deltas and layer inputs).
// Thread j owns row j of W1, W2, W3 (and b1..b3, W4[j]); a block sums the
// samples [blockIdx.x * per, ...) and adds its rows into G once.
__global__ __launch_bounds__(64) void k_mlp_wgrad(const float *__restrict__ X,
                                                  const float *__restrict__ H,
                                                  const float *__restrict__ D,
                                                  const float *__restrict__ dY, int B, int per,
                                                  float *__restrict__ G) {
  const int j = threadIdx.x;
  const int b0 = blockIdx.x * per, b1 = min(B, b0 + per);
  float g1[kMlpIn], g2[kMlpH], g3[kMlpH];
#pragma unroll
  for (int i = 0; i < kMlpIn; i++) g1[i] = 0.0f;
#pragma unroll
  for (int i = 0; i < kMlpH; i++) g2[i] = g3[i] = 0.0f;
  float gb1 = 0, gb2 = 0, gb3 = 0, g4 = 0, gb4 = 0;
  for (int b = b0; b < b1; b++) {
    const float d1 = D[(size_t)b * kMlpH + j];
    const float d2 = D[((size_t)B + b) * kMlpH + j];
    const float d3 = D[(2 * (size_t)B + b) * kMlpH + j];
    const float dy = dY[b];
    const float *x = X + (size_t)b * kMlpIn;
    const float *h1 = H + (size_t)b * kMlpH;
    const float *h2 = H + ((size_t)B + b) * kMlpH;
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) g1[i] = fmaf(d1, x[i], g1[i]);
#pragma unroll
    for (int i = 0; i < kMlpH; i++) {
      g2[i] = fmaf(d2, h1[i], g2[i]);
      g3[i] = fmaf(d3, h2[i], g3[i]);
    }
    gb1 += d1;
    gb2 += d2;
    gb3 += d3;
    g4 = fmaf(dy, H[(2 * (size_t)B + b) * kMlpH + j], g4);
    gb4 += dy;
  }
#pragma unroll
  for (int i = 0; i < kMlpIn; i++) atomicAdd(&G[kW1 + j * kMlpIn + i], g1[i]);
#pragma unroll
  for (int i = 0; i < kMlpH; i++) {
    atomicAdd(&G[kW2 + j * kMlpH + i], g2[i]);
    atomicAdd(&G[kW3 + j * kMlpH + i], g3[i]);
  }
  atomicAdd(&G[kB1 + j], gb1);
  atomicAdd(&G[kB2 + j], gb2);
  atomicAdd(&G[kB3 + j], gb3);
  atomicAdd(&G[kW4 + j], g4);
  if (j == 0) atomicAdd(&G[kB4], gb4);
}

__global__ void k_fill_f32(float *__restrict__ p, int n, float v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam (Kingma & Ba 2015) on every parameter, step t >= 1.
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------- features
// Per-leaf features (kLeafFeatures floats): level / 16; log10(R / R_root);
// log2(bodies) / 6; log10(mass / mean body mass); log10 of the order-2 and
// order-4 multipole norms relative to M_0 R^n; log10(1 + M2L and P2P list
// lengths of the leaf in the probe walk); log10 of the leaf's distance from
// the root's centre over R_root (a leaf far out in a halo feels the whole
// system as a few large far-field terms).
__global__ void k_leaf_features(CellsDev c, const int *__restrict__ leaf_ids, int nleaves,
                                int NT, int P, const int *__restrict__ m2l_cnt,
                                const int *__restrict__ p2p_cnt, double mean_mass,
                                float *__restrict__ F) {
  const int l = blockIdx.x * 256 + threadIdx.x;
  if (l >= nleaves) return;
  const int ci = leaf_ids[l];
  const double R = c.R[ci], R0 = c.R[0];
  const double *M = c.M + (size_t)ci * NT;
  const double m0 = M[0];
  auto norm = [&](int n) {
    double s = 0.0;
    for (int m = -n; m <= n; m++) s += M[n * n + n + m] * M[n * n + n + m];
    return sqrt(s);
  };
  const double rn2 = R * R;
  const double q2 = m0 > 0 ? norm(2) / (m0 * rn2) : 0.0;
  const double q4 = (m0 > 0 && P >= 4) ? norm(4) / (m0 * rn2 * rn2) : 0.0;
  float *f = F + (size_t)l * kLeafFeatures;
  f[0] = (float)c.level[ci] / 16.0f;
  f[1] = (float)log10(fmax(R / R0, 1e-30));
  f[2] = (float)(log2((double)max(c.nbody[ci], 1)) / 6.0);
  f[3] = (float)log10(fmax(m0 / mean_mass, 1e-30));
  f[4] = (float)log10(q2 + 1e-6);
  f[5] = (float)log10(q4 + 1e-6);
  f[6] = (float)log10(1.0 + (double)m2l_cnt[ci]);
  f[7] = (float)log10(1.0 + (double)p2p_cnt[ci]);
  const double ex = c.X[3 * ci] - c.X[0], ey = c.X[3 * ci + 1] - c.X[1],
               ez = c.X[3 * ci + 2] - c.X[2];
  f[8] = (float)log10(sqrt(ex * ex + ey * ey + ez * ez) / R0 + 1e-3);
}

// Pairs per target cell of a walk's list (counts zeroed by the caller).
__global__ void k_count_targets(const int2 *__restrict__ pairs, int n, int *__restrict__ cnt) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) atomicAdd(&cnt[pairs[k].x], 1);
}

// ---------------------------------------------------------------- choice
// One block per leaf, one lane per candidate (q, theta): predicted log10
// error (the features clamped to their training range: no extrapolation),
// then the cheapest candidate at or below log_target (else the most
// accurate); the leaf's choice goes to q[leaf cell], theta[leaf cell].
__global__ __launch_bounds__(64) void k_leaf_select(const float *__restrict__ P,
                                                    const float *__restrict__ F,
                                                    const int *__restrict__ leaf_ids,
                                                    const OrderGrid g, float log_target,
                                                    int *__restrict__ q, double *__restrict__ theta,
                                                    float *__restrict__ pred) {
  const int l = blockIdx.x, c = threadIdx.x;
  const bool act = c < g.ncand;
  const int cc = act ? c : 0;
  float x[kMlpIn];
#pragma unroll
  for (int i = 0; i < kLeafFeatures; i++)
    x[i] = fminf(fmaxf(F[(size_t)l * kLeafFeatures + i], g.lo[i]), g.hi[i]);
  x[kLeafFeatures] = (float)g.q[cc] / 20.0f;
  x[kLeafFeatures + 1] = (float)g.theta[cc];
  const float y = mlp_eval(P, x);
  // key: feasible candidates by cost, then the rest by predicted error
  float key = !act ? 3e38f : (y <= log_target ? g.cost[cc] : 1e30f + y);
  int best = cc;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float k2 = __shfl_xor(key, o, kWave);
    const int b2 = __shfl_xor(best, o, kWave);
    if (k2 < key || (k2 == key && b2 < best)) {
      key = k2;
      best = b2;
    }
  }
  if (c == best) {
    const int ci = leaf_ids[l];
    q[ci] = g.q[best];
    theta[ci] = g.theta[best];
    if (pred) pred[l] = y;
  }
}

__global__ void k_leaf_uniform(const int *__restrict__ leaf_ids, int nleaves, int qv,
                               double tv, int *__restrict__ q, double *__restrict__ theta) {
  const int l = blockIdx.x * 256 + threadIdx.x;
  if (l >= nleaves) return;
  q[leaf_ids[l]] = qv;
  theta[leaf_ids[l]] = tv;
}

// Internal cells of one level: the largest order and smallest angle of their
// children (children are set first: levels run deepest first).
__global__ void k_propagate(CellsDev c, int ls, int le, int *__restrict__ q,
                            double *__restrict__ theta) {
  const int i = ls + blockIdx.x * 256 + threadIdx.x;
  if (i >= le || c.nchild[i] == 0) return;
  const int c0 = c.child[i], nc = c.nchild[i];
  int qm = 0;
  double tm = 1e30;
  for (int k = 0; k < nc; k++) {
    qm = max(qm, q[c0 + k]);
    tm = fmin(tm, theta[c0 + k]);
  }
  q[i] = qm;
  theta[i] = tm;
}

// ---------------------------------------------------------------- measurement
// log10 of each leaf's rms relative force error against reference
// accelerations (both in the caller's body order); -30 for a leaf without
// nonzero reference forces.
__global__ void k_leaf_errors(CellsDev c, BodiesDev b, const int *__restrict__ leaf_ids,
                              int nleaves, const double *__restrict__ acc,
                              const double *__restrict__ ref, float *__restrict__ out) {
  const int l = blockIdx.x * 256 + threadIdx.x;
  if (l >= nleaves) return;
  const int ci = leaf_ids[l];
  double s = 0.0;
  int cnt = 0;
  for (int k = c.body[ci]; k < c.body[ci] + c.nbody[ci]; k++) {
    const size_t i = (size_t)b.order[k];
    const double rx = ref[3 * i], ry = ref[3 * i + 1], rz = ref[3 * i + 2];
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (!(r2 > 0)) continue;
    const double ex = acc[3 * i] - rx, ey = acc[3 * i + 1] - ry, ez = acc[3 * i + 2] - rz;
    s += (ex * ex + ey * ey + ez * ez) / r2;
    cnt++;
  }
  out[l] = cnt ? (float)(0.5 * log10(s / cnt + 1e-300)) : -30.0f;
}

// Leaf index (into the leaf list) of every body, in the caller's order.
__global__ void k_leaf_of_body(CellsDev c, BodiesDev b, const int *__restrict__ leaf_ids,
                               int nleaves, int *__restrict__ out) {
  const int l = blockIdx.x;
  if (l >= nleaves) return;
  const int ci = leaf_ids[l];
  for (int k = c.body[ci] + (int)threadIdx.x; k < c.body[ci] + c.nbody[ci]; k += 64)
    out[b.order[k]] = l;
}

// ---------------------------------------------------------------- launches
static dim3 g256(int n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_mlp_fwd(const float *P, const float *X, int B, float *Y, float *H, hipStream_t s) {
  if (B > 0) hipLaunchKernelGGL(k_mlp_fwd, dim3((B + 63) / 64), dim3(64), 0, s, P, X, B, Y, H);
}

void launch_mlp_grad(const float *P, const float *X, const float *T, int B, float tau, float *Y,
                     float *H, float *D, float *dY, float *G, float *loss, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f32, g256(kMlpParams), dim3(256), 0, s, G, kMlpParams, 0.0f);
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, s, loss, 1, 0.0f);
  if (B <= 0) return;
  hipLaunchKernelGGL(k_mlp_fwd, dim3((B + 63) / 64), dim3(64), 0, s, P, X, B, Y, H);
  hipLaunchKernelGGL(k_mlp_bwd, dim3((B + 63) / 64), dim3(64), 0, s, P, H, Y, T, B, tau, D, dY,
                     loss);
  const int per = 256;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((B + per - 1) / per), dim3(64), 0, s, X, H, D, dY, B, per,
                     G);
}

void launch_adam(float *P, const float *G, float *m, float *v, float lr, int t, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, g256(kMlpParams), dim3(256), 0, s, P, G, m, v, kMlpParams, lr, b1,
                     b2, 1e-8f, c1, c2);
}

void launch_count_targets(const int2 *pairs, int n, int *cnt, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_count_targets, g256(n), dim3(256), 0, s, pairs, n, cnt);
}

void launch_leaf_features(CellsDev c, const int *leaf_ids, int nleaves, int p,
                          const int *m2l_cnt, const int *p2p_cnt, double mean_mass, float *F,
                          hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_leaf_features, g256(nleaves), dim3(256), 0, s, c, leaf_ids, nleaves,
                       nterm(p), p, m2l_cnt, p2p_cnt, mean_mass, F);
}

void launch_leaf_select(const float *P, const float *F, const int *leaf_ids, int nleaves,
                        const OrderGrid &g, float log_target, int *q, double *theta, float *pred,
                        hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_leaf_select, dim3(nleaves), dim3(64), 0, s, P, F, leaf_ids, g,
                       log_target, q, theta, pred);
}

void launch_leaf_uniform(const int *leaf_ids, int nleaves, int qv, double tv, int *q,
                         double *theta, hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_leaf_uniform, g256(nleaves), dim3(256), 0, s, leaf_ids, nleaves, qv, tv,
                       q, theta);
}

void launch_propagate(CellsDev c, int ls, int le, int *q, double *theta, hipStream_t s) {
  if (le > ls) hipLaunchKernelGGL(k_propagate, g256(le - ls), dim3(256), 0, s, c, ls, le, q, theta);
}

void launch_leaf_errors(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                        const double *acc, const double *ref, float *out, hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_leaf_errors, g256(nleaves), dim3(256), 0, s, c, b, leaf_ids, nleaves, acc,
                       ref, out);
}

void launch_leaf_of_body(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, int *out,
                         hipStream_t s) {
  if (nleaves > 0)
    hipLaunchKernelGGL(k_leaf_of_body, dim3(nleaves), dim3(64), 0, s, c, b, leaf_ids, nleaves,
                       out);
}

}  // namespace tcfmm
