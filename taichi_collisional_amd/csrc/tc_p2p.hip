// Near field on the device, FP64: P2P between leaf cells (taichi_with_jerks/
// kernel_fast_lean.h:499-699, the Vec8d double-precision branch) and an
// all-pairs direct sum used to measure the FMM error.
//
// Per pair (target i, source j), dx = x_j - x_i, dv = v_j - v_i:
//   pot += m/r,  a += m dx / r^3,  j += m dv / r^3 - 3 m (dx.dv) dx / r^5,
// pairs with r^2 == 0 contribute nothing (kernel_fast_lean.h:583-584).
//
// P2P kernel: one wave per target leaf.  With n_t target bodies the wave is
// split into G = floor(64/n_t) slices of n_t lanes; lane (slice s, target t)
// walks every G-th body of each source leaf in the target's list, and the
// slices are summed through LDS at the end.  Source bodies are read from the
// sorted SoA arrays (each source leaf is a contiguous range).
#include "tc_common.h"
#include "tc_kernels.h"

namespace tcfmm {

struct Acc7 {
  double p, ax, ay, az, jx, jy, jz;
};

__device__ __forceinline__ void pair_fp64(double xi, double yi, double zi, double vxi, double vyi,
                                          double vzi, double xj, double yj, double zj, double vxj,
                                          double vyj, double vzj, double mj, Acc7 &a) {
  double dx = xj - xi, dy = yj - yi, dz = zj - zi;
  double dvx = vxj - vxi, dvy = vyj - vyi, dvz = vzj - vzi;
  double r2 = dx * dx + dy * dy + dz * dz;
  double vr = 3.0 * (dx * dvx + dy * dvy + dz * dvz);
  double ri = r2 > 0.0 ? rsqrt(r2) : 0.0;
  double ri2 = ri * ri;
  double f = mj * ri;
  a.p += f;
  f *= ri2;
  a.ax += dx * f;
  a.ay += dy * f;
  a.az += dz * f;
  a.jx += dvx * f;
  a.jy += dvy * f;
  a.jz += dvz * f;
  f *= -ri2 * vr;
  a.jx += dx * f;
  a.jy += dy * f;
  a.jz += dz * f;
}

__global__ __launch_bounds__(64) void k_p2p(CellsDev c, BodiesDev b, ResultsDev r,
                                            const int *__restrict__ leaf_ids,
                                            const int *__restrict__ offs,
                                            const int *__restrict__ count,
                                            const int *__restrict__ src) {
  __shared__ double red[7][kWave];
  const int lane = threadIdx.x;
  const int ci = leaf_ids[blockIdx.x];
  if (!c.has_sink[ci]) return;
  const int b0 = c.body[ci], nb = c.nbody[ci];
  const int l0 = offs[ci], ln = count[ci];
  for (int tb = 0; tb < nb; tb += kWave) {
    const int nt = nb - tb < kWave ? nb - tb : kWave;
    const int G = kWave / nt;
    const int s = lane / nt, t = lane - s * nt;
    const bool act = s < G;
    const int ib = b0 + tb + t;
    double xi = b.x[ib], yi = b.y[ib], zi = b.z[ib];
    double vxi = b.vx[ib], vyi = b.vy[ib], vzi = b.vz[ib];
    Acc7 a = {0, 0, 0, 0, 0, 0, 0};
    if (act) {
      for (int e = 0; e < ln; e++) {
        const int j = src[l0 + e];
        const int jb0 = c.body[j], jn = c.nbody[j];
        for (int k = s; k < jn; k += G) {
          const int jb = jb0 + k;
          if (!(b.flags[jb] & kSource)) continue;
          pair_fp64(xi, yi, zi, vxi, vyi, vzi, b.x[jb], b.y[jb], b.z[jb], b.vx[jb], b.vy[jb],
                    b.vz[jb], b.q[jb], a);
        }
      }
    }
    __syncthreads();
    red[0][lane] = a.p;
    red[1][lane] = a.ax;
    red[2][lane] = a.ay;
    red[3][lane] = a.az;
    red[4][lane] = a.jx;
    red[5][lane] = a.jy;
    red[6][lane] = a.jz;
    __syncthreads();
    if (s == 0 && (b.flags[ib] & kSink)) {
      double o[7];
      for (int q = 0; q < 7; q++) {
        double v = 0;
        for (int g = 0; g < G; g++) v += red[q][g * nt + t];
        o[q] = v;
      }
      r.pot[ib] += o[0];
      r.ax[ib] += o[1];
      r.ay[ib] += o[2];
      r.az[ib] += o[3];
      r.jx[ib] += o[4];
      r.jy[ib] += o[5];
      r.jz[ib] += o[6];
    }
  }
}

int launch_p2p(CellsDev c, BodiesDev b, ResultsDev r, const int *leaf_ids, int nleaves,
               const int *offs, const int *count, const int *src, hipStream_t s) {
  if (nleaves <= 0) return 0;
  hipLaunchKernelGGL(k_p2p, dim3(nleaves), dim3(64), 0, s, c, b, r, leaf_ids, offs, count, src);
  return 0;
}

// All-pairs sum for a list of targets (sorted-order indices), FP64.
__global__ __launch_bounds__(256) void k_direct(BodiesDev b, int n, const int *__restrict__ tg,
                                                int nt, double *__restrict__ pot,
                                                double *__restrict__ acc,
                                                double *__restrict__ jerk) {
  __shared__ double sx[256], sy[256], sz[256], svx[256], svy[256], svz[256], sq[256];
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int i = k < nt ? tg[k] : 0;
  double xi = b.x[i], yi = b.y[i], zi = b.z[i], vxi = b.vx[i], vyi = b.vy[i], vzi = b.vz[i];
  Acc7 a = {0, 0, 0, 0, 0, 0, 0};
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    int j = base + threadIdx.x;
    if (j < n) {
      sx[threadIdx.x] = b.x[j];
      sy[threadIdx.x] = b.y[j];
      sz[threadIdx.x] = b.z[j];
      svx[threadIdx.x] = b.vx[j];
      svy[threadIdx.x] = b.vy[j];
      svz[threadIdx.x] = b.vz[j];
      sq[threadIdx.x] = (b.flags[j] & kSource) ? b.q[j] : 0.0;
    }
    __syncthreads();
    int lim = n - base < 256 ? n - base : 256;
    for (int q = 0; q < lim; q++)
      pair_fp64(xi, yi, zi, vxi, vyi, vzi, sx[q], sy[q], sz[q], svx[q], svy[q], svz[q], sq[q], a);
  }
  if (k < nt) {
    pot[k] = a.p;
    acc[3 * k] = a.ax;
    acc[3 * k + 1] = a.ay;
    acc[3 * k + 2] = a.az;
    jerk[3 * k] = a.jx;
    jerk[3 * k + 1] = a.jy;
    jerk[3 * k + 2] = a.jz;
  }
}

int launch_direct(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                  double *jerk, hipStream_t s) {
  if (nt <= 0) return 0;
  hipLaunchKernelGGL(k_direct, dim3((nt + 255) / 256), dim3(256), 0, s, b, n, targets, nt, pot,
                     acc, jerk);
  return 0;
}

}  // namespace tcfmm
