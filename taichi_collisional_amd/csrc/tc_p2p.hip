// Near field on the device, FP64: P2P between leaf cells (taichi_with_jerks/
// kernel_fast_lean.h:499-699, the Vec8d double-precision branch) and an
// all-pairs direct sum used to measure the FMM error.
//
// Per pair (target i, source j), dx = x_j - x_i, dv = v_j - v_i:
//   pot += m/r,  a += m dx / r^3,  j += m dv / r^3 - 3 m (dx.dv) dx / r^5,
// pairs with r^2 == 0 contribute nothing (kernel_fast_lean.h:583-584).
//
// P2P kernel: one wave per target group (<= 32 bodies of a leaf).  With n_t
// targets the wave is split into G = floor(64/n_t) slices of n_t lanes; the
// source leaves of the leaf's list are copied tile by tile into LDS, lane
// (slice s, target t) walks every G-th tile body, and the slices are summed
// through LDS at the end.
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
  const double rq = rsqrt_fp64(r2);
  double ri = r2 > 0.0 ? rq : 0.0;
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

// Source bodies per LDS tile.  The final slice reduction reuses the tile's
// memory, so LDS per wave is 64 * kTile bytes + 512 (occupancy is set by
// LDS, not registers).
#ifndef TC_P2P_TILE
#define TC_P2P_TILE 128
#endif
constexpr int kTile = TC_P2P_TILE;
static_assert(kTile >= 64 && 2 * kTile * 32 >= 7 * kWave * 8, "tile holds the reduction");

// One block (one wave) per work item: a target group (<= 32 consecutive bodies
// of one leaf) and a slice of at most kItemList entries of the leaf's P2P list
// (long lists -- a sparse halo leaf can touch most of the system -- are split
// so no single wave serializes them; partial sums are added atomically).
// The leaf's P2P list is a list of source body ranges (one per source leaf);
// up to 64 ranges are fetched at once (one per lane), a wave prefix sum over
// their lengths places as many whole ranges as fit into an LDS tile of kTile
// bodies (position+mass, velocity as double4), and every lane then copies tile
// slots in parallel.  Lane (slice s, target t) takes every G-th tile body,
// G = floor(64 / group size).  A single range longer than the tile (only when
// ncrit > kTile) is streamed through the tile in pieces.
__global__ __launch_bounds__(64) void k_p2p(CellsDev c, BodiesDev b, ResultsDev r,
                                            const int4 *__restrict__ groups,
                                            const int *__restrict__ offs,
                                            const int *__restrict__ count,
                                            const int2 *__restrict__ ranges,
                                            unsigned long long *__restrict__ npairs) {
  __shared__ double4 sbuf[2 * kTile];
  __shared__ int rstart[kWave], rexcl[kWave];
  double4 *sp = sbuf, *sv = sbuf + kTile;
  double(*red)[kWave] = reinterpret_cast<double(*)[kWave]>(sbuf);  // after the loop
  const int lane = threadIdx.x;
  // work item: target leaf, first target, list slice start, slice length | group size << 16
  // round-robin dispatch spreads the consecutive, heavy items of one large
  // list over all XCDs; a contiguous per-XCD range would pile them on one
  const int4 gw = groups[blockIdx.x];
  const int ci = gw.x;
  if (!c.has_sink[ci]) return;
  const int nt = gw.w >> 16;
  const int G = kWave / nt;
  const int s = lane / nt, t = lane - s * nt;
  const bool act = s < G;
  const int ib = c.body[ci] + gw.y + t;
  const double xi = b.x[ib], yi = b.y[ib], zi = b.z[ib];
  const double vxi = b.vx[ib], vyi = b.vy[ib], vzi = b.vz[ib];
  Acc7 a = {0, 0, 0, 0, 0, 0, 0};
  const int l0 = offs[ci] + gw.z, ln = gw.w & 0xffff;
  long long nsrc = 0;
  int e = 0, off = 0;
  while (e < ln) {
    const int2 rg = (e + lane < ln) ? ranges[l0 + e + lane] : make_int2(0, 0);
    int tot;
    const int excl = wave_excl_scan(rg.y, &tot);
    const bool fits = (e + lane < ln) && (excl + rg.y <= kTile);
    const unsigned long long fm = __ballot(fits);
    const int nfit = __popcll(fm);  // ranges 0..nfit-1 fit (prefix sums grow)
    int fill;
    if (nfit > 0) {
      rstart[lane] = rg.x;
      rexcl[lane] = excl;
      fill = __shfl(excl + rg.y, nfit - 1, kWave);
      __syncthreads();
      for (int q = lane; q < fill; q += kWave) {
        int lo = 0, hi = nfit - 1;  // last range with rexcl <= q
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (rexcl[mid] <= q)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int bq = rstart[lo] + q - rexcl[lo];
        const double mq = (b.flags[bq] & kSource) ? b.q[bq] : 0.0;
        sp[q] = make_double4(b.x[bq], b.y[bq], b.z[bq], mq);
        sv[q] = make_double4(b.vx[bq], b.vy[bq], b.vz[bq], 0.0);
      }
      e += nfit;
    } else {
      // first range longer than the tile: stream it piece by piece
      const int base = __shfl(rg.x, 0, kWave), len = __shfl(rg.y, 0, kWave);
      fill = min(kTile, len - off);
      for (int q = lane; q < fill; q += kWave) {
        const int bq = base + off + q;
        const double mq = (b.flags[bq] & kSource) ? b.q[bq] : 0.0;
        sp[q] = make_double4(b.x[bq], b.y[bq], b.z[bq], mq);
        sv[q] = make_double4(b.vx[bq], b.vy[bq], b.vz[bq], 0.0);
      }
      off += fill;
      if (off >= len) {
        off = 0;
        e += 1;
      }
    }
    nsrc += fill;
    __syncthreads();
    if (act) {
      // fill is the same on every lane: the first fill / G bodies of every
      // slice run as a loop with a uniform trip count (scalar control, no
      // per-pair exec-mask updates); the ragged last round follows
      const int nfull = __builtin_amdgcn_readfirstlane(fill / G);
      int k = s;
#pragma unroll 2
      for (int i = 0; i < nfull; i++, k += G) {
        const double4 P = sp[k], V = sv[k];
        pair_fp64(xi, yi, zi, vxi, vyi, vzi, P.x, P.y, P.z, V.x, V.y, V.z, P.w, a);
      }
      if (k < fill) {
        const double4 P = sp[k], V = sv[k];
        pair_fp64(xi, yi, zi, vxi, vyi, vzi, P.x, P.y, P.z, V.x, V.y, V.z, P.w, a);
      }
    }
    __syncthreads();
  }
  if (lane == 0) atomicAdd(npairs, (unsigned long long)(nsrc * nt));
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
    // several work items may cover one target group (long lists are split)
    atomicAdd(&r.pot[ib], o[0]);
    atomicAdd(&r.ax[ib], o[1]);
    atomicAdd(&r.ay[ib], o[2]);
    atomicAdd(&r.az[ib], o[3]);
    atomicAdd(&r.jx[ib], o[4]);
    atomicAdd(&r.jy[ib], o[5]);
    atomicAdd(&r.jz[ib], o[6]);
  }
}

// Source body range (first body, count) of every P2P list entry.
__global__ void k_p2p_ranges(CellsDev c, const int *__restrict__ src, int n,
                             int2 *__restrict__ ranges) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int j = src[k];
  ranges[k] = make_int2(c.body[j], c.nbody[j]);
}

int launch_p2p_ranges(CellsDev c, const int *src, int n, int2 *ranges, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_p2p_ranges, dim3((n + 255) / 256), dim3(256), 0, s, c, src, n, ranges);
  return 0;
}

int launch_p2p(CellsDev c, BodiesDev b, ResultsDev r, const int4 *groups, int ngroups,
               const int *offs, const int *count, const int2 *ranges,
               unsigned long long *npairs, hipStream_t s) {
  if (ngroups <= 0) return 0;
  hipLaunchKernelGGL(k_p2p, dim3(ngroups), dim3(64), 0, s, c, b, r, groups, offs, count, ranges,
                     npairs);
  return 0;
}

// All-pairs sum for a list of targets (sorted-order indices), FP64.  The
// sources are split into gridDim.y slabs so enough blocks run; each slab adds
// its partial sums into the (zeroed) outputs.
__global__ __launch_bounds__(256) void k_direct(BodiesDev b, int n, const int *__restrict__ tg,
                                                int nt, int slab, double *__restrict__ pot,
                                                double *__restrict__ acc,
                                                double *__restrict__ jerk) {
  __shared__ double4 sp[256], sv[256];
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int i = k < nt ? tg[k] : 0;
  const double xi = b.x[i], yi = b.y[i], zi = b.z[i], vxi = b.vx[i], vyi = b.vy[i],
               vzi = b.vz[i];
  Acc7 a = {0, 0, 0, 0, 0, 0, 0};
  const int j0 = blockIdx.y * slab, j1 = min(n, j0 + slab);
  for (int base = j0; base < j1; base += 256) {
    __syncthreads();
    int j = base + threadIdx.x;
    if (j < j1) {
      double mq = (b.flags[j] & kSource) ? b.q[j] : 0.0;
      sp[threadIdx.x] = make_double4(b.x[j], b.y[j], b.z[j], mq);
      sv[threadIdx.x] = make_double4(b.vx[j], b.vy[j], b.vz[j], 0.0);
    }
    __syncthreads();
    int lim = min(256, j1 - base);
    for (int q = 0; q < lim; q++) {
      const double4 P = sp[q], V = sv[q];
      pair_fp64(xi, yi, zi, vxi, vyi, vzi, P.x, P.y, P.z, V.x, V.y, V.z, P.w, a);
    }
  }
  if (k < nt) {
    atomicAdd(&pot[k], a.p);
    atomicAdd(&acc[3 * k], a.ax);
    atomicAdd(&acc[3 * k + 1], a.ay);
    atomicAdd(&acc[3 * k + 2], a.az);
    atomicAdd(&jerk[3 * k], a.jx);
    atomicAdd(&jerk[3 * k + 1], a.jy);
    atomicAdd(&jerk[3 * k + 2], a.jz);
  }
}

__global__ void k_zero_f64(double *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

void launch_zero_f64(double *p, size_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

int launch_direct(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                  double *jerk, hipStream_t s) {
  if (nt <= 0) return 0;
  hipMemsetAsync(pot, 0, sizeof(double) * nt, s);
  hipMemsetAsync(acc, 0, sizeof(double) * 3 * nt, s);
  hipMemsetAsync(jerk, 0, sizeof(double) * 3 * nt, s);
  return launch_direct_sum(b, n, targets, nt, pot, acc, jerk, s);
}

// The all-pairs sum alone: the outputs must already be zero.
int launch_direct_sum(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                      double *jerk, hipStream_t s) {
  if (nt <= 0) return 0;
  int bx = (nt + 255) / 256;
  int slabs = (2048 + bx - 1) / bx;  // aim for >= 2048 blocks
  int slab = (n + slabs - 1) / slabs;
  slab = ((slab + 255) / 256) * 256;
  slabs = (n + slab - 1) / slab;
  hipLaunchKernelGGL(k_direct, dim3(bx, slabs), dim3(256), 0, s, b, n, targets, nt, slab, pot,
                     acc, jerk);
  return 0;
}

}  // namespace tcfmm
