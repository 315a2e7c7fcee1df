// Master (non-jerk) variant of the force evaluation: what the reference's
// top-level sources run by default (master Makefile:4 leaves FMM undefined,
// so get_force_and_potential is direct(), integrator.h:557-559 and
// traverse_eager.h:349-370).  File and line citations below refer to those
// top-level sources.
//
// direct() visits every pair of leaves of the tree (directPass,
// traverse_eager.h:262-297) and runs P2P_simple (kernel_fast_lean.h:268-460)
// or, when the caller asks for time steps, P2P (:462-674) on it.  Summed over
// the leaf pairs that is, per sink body i:
//   p_i += sum_j q_j / r_ij,  F_i += sum_j q_j (x_j - x_i) / r_ij^3
// over sources j with r_ij > 0, and, with time steps,
//   timestep_i += sum over leaves J of (min_{j in J} tau_ij)^-4
// where tau_ij is the smaller of a free-fall and a fly-by estimate (see
// pair_tau).  The leaf structure matters only for the time steps: the minimum
// is taken per source leaf, and in the vector form the mass in the free-fall
// estimate accumulates along the leaf.
//
// Layout (default build): one thread per target body in tree (sorted) order, so each block's
// targets are neighbours; sources stream through LDS in tiles of kMT bodies in
// the same sorted order, with a flag byte per source marking "is a source" and
// "last body of its leaf".  Results are written at the caller's index.
#include <hip/hip_runtime.h>

#include "tc_common.h"
#include <algorithm>
#include <climits>

#include "tc_kernels.h"

namespace tcfmm {

constexpr int kMT = 256;

__global__ void k_fill_u8(unsigned char *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_iota(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// lend[last body of every leaf] = 1
__global__ void k_leaf_ends(const int *leaf_ids, int nleaves, const int *cbody,
                            const int *cnbody, unsigned char *lend) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int c = leaf_ids[k];
  if (cnbody[c] > 0) lend[cbody[c] + cnbody[c] - 1] = 1;
}

// Fold one (target, source) pair into the running minimum ts of the current
// source leaf.
//  kPairMass = false: the vector P2P of the master build (SIMD_P2P,
//    kernel_fast_lean.h:474-592): mi, reset to q_i at every source leaf,
//    accumulates every source met with r2 > 0 (mi += mj); r2 = 0 is replaced by
//    1e38; v2 carries +1e-20; both estimates are always taken.
//  kPairMass = true: the scalar P2P (:593-672): r2 = 0 skipped, pair mass
//    q_i + q_j, the fly-by estimate only for v2 > 0.
// tau_ff = dt/sqrt(2) r sqrt(r / M), tau_fb = dt r / |dv|, each divided by
// (1 - dtau/2) with its rate term dtau capped at 1.
// The vector form is written with reciprocal square roots and reciprocals
// instead of the divisions and square roots of the source expression
// (ir = 1/r, r sqrt(r / M) = r^2 / sqrt(r M), mi / (v2 r) = mi ir / v2): the
// same quantities to an ulp or two each, at a fraction of the instructions.
// ir is the target-source 1/r that the force uses as well (r2 = 0 pairs use
// the 1e38 stand-in, whose force is masked).
template <bool kPairMass>
__device__ __forceinline__ void pair_tau(double dx, double dy, double dz, double dvx, double dvy,
                                         double dvz, double r2, double ir, double qi, double qj,
                                         double dt, double &mi, double &ts) {
  const double dts2 = dt / M_SQRT2;
  if (!kPairMass) {
    const double v2 = dvx * dvx + dvy * dvy + dvz * dvz + 1e-20;
    const bool pos = r2 > 0;
    mi += pos ? qj : 0.0;
    const double rr = pos ? r2 : 1e38;
    const double vdr = (dx * dvx + dy * dvy + dz * dvz) * (ir * ir);
    const double r = rr * ir;
    double tau = dts2 * rr * rsqrt_fp64(r * mi);
    double dtau = 3.0 * tau * vdr / 2.0;
    dtau = dtau < 1.0 ? dtau : 1.0;
    tau *= rcp_fp64(1.0 - dtau / 2.0);
    ts = fmin(tau, ts);
    const double iv = rsqrt_fp64(v2);
    tau = dt * r * iv;
    dtau = tau * vdr * (1.0 + mi * ir * (iv * iv));
    dtau = dtau < 1.0 ? dtau : 1.0;
    tau *= rcp_fp64(1.0 - dtau / 2.0);
    ts = fmin(tau, ts);
  } else {
    if (!(r2 > 0)) return;
    const double r = sqrt(r2), v2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const double vdr = (dx * dvx + dy * dvy + dz * dvz) / r2, mm = qi + qj;
    double tau = dts2 * sqrt(r * r2 / mm);
    double dtau = 3.0 * tau * vdr / 2.0;
    if (dtau > 1.0) dtau = 1.0;
    tau /= (1.0 - dtau / 2.0);
    if (tau < ts) ts = tau;
    if (v2 > 0) {
      tau = dt * r / sqrt(v2);
      dtau = tau * vdr * (1.0 + mm / (v2 * r));
      if (dtau > 1.0) dtau = 1.0;
      tau /= (1.0 - dtau / 2.0);
      if (tau < ts) ts = tau;
    }
  }
}

// The smallest k >= j that starts a source leaf (k == 0, or lend[k - 1] set),
// or n.  The whole block searches kMT candidates per round (a leaf spans at
// most ncrit bodies, so a round or two); every thread returns the same value.
__device__ int next_leaf_start(const unsigned char *__restrict__ lend, int n, int j) {
  __shared__ int s_min;
  const int t = threadIdx.x;
  for (int base = j;; base += kMT) {
    if (t == 0) s_min = INT_MAX;
    __syncthreads();
    const int k = base + t;
    if (k >= n || k == 0 || lend[k - 1]) atomicMin(&s_min, k);
    __syncthreads();
    const int r = s_min;
    __syncthreads();  // everyone has read s_min before the next round resets it
    if (r != INT_MAX) return r;
  }
}

// Sources may be split into gridDim.y slabs of about `slab` bodies each, so
// that a small system still fills the GPU.  A slab starts and ends where a
// source leaf does (lend marks each leaf's last body): the running minimum
// of the time-step estimate belongs to one leaf, so a leaf is never cut.
// Each slab writes its partial sums, in sorted order, to part[slab][nc][n]
// (potential, acceleration and, with steps, the time-step sum: nc = 4 or 5);
// k_master_direct_sum adds the slabs in order.  With one slab the results
// are written directly.
template <bool kSteps, bool kPairMass>
__global__ __launch_bounds__(kMT) void k_master_direct(BodiesDev b, const unsigned char *lend,
                                                       int n, int slab, double dt, double *pot,
                                                       double *acc, double *tstep,
                                                       double *__restrict__ part) {
  __shared__ double sx[kMT], sy[kMT], sz[kMT], sq[kMT], svx[kMT], svy[kMT], svz[kMT];
  __shared__ unsigned char sf[kMT];
  const int t = threadIdx.x;
  int j0 = 0, j1 = n;
  if (gridDim.y > 1) {
    // both ends move forward to the next leaf start (or n), by the same rule,
    // so neighbouring slabs meet
    j0 = (int)min((long long)n, (long long)blockIdx.y * slab);
    j1 = (int)min((long long)n, (long long)blockIdx.y * slab + slab);
    if (kSteps) {
      j0 = next_leaf_start(lend, n, j0);
      j1 = next_leaf_start(lend, n, j1);
    }
  }
  const int i = blockIdx.x * kMT + t;
  const bool live = i < n;
  // lanes past n run the loop on a harmless stand-in body and write nothing
  const double xi = live ? b.x[i] : 0.0, yi = live ? b.y[i] : 0.0, zi = live ? b.z[i] : 0.0;
  const double vxi = live ? b.vx[i] : 0.0, vyi = live ? b.vy[i] : 0.0,
               vzi = live ? b.vz[i] : 0.0;
  const double qi = live ? b.q[i] : 1.0;
  double ax = 0, ay = 0, az = 0, pt = 0, tsum = 0, mi = qi, ts = 1e38;
  for (int base = j0; base < j1; base += kMT) {
    __syncthreads();
    const int j = base + t;
    if (j < j1) {
      sx[t] = b.x[j];
      sy[t] = b.y[j];
      sz[t] = b.z[j];
      sq[t] = b.q[j];
      if (kSteps) {
        svx[t] = b.vx[j];
        svy[t] = b.vy[j];
        svz[t] = b.vz[j];
      }
      sf[t] = ((b.flags[j] & kSource) ? 1 : 0) | (lend[j] ? 2 : 0);
    }
    __syncthreads();
    const int cnt = min(kMT, j1 - base);
    for (int k = 0; k < cnt; k++) {
      const double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const unsigned char f = sf[k];
      const double ir = rsqrt_fp64(r2 > 0 ? r2 : 1e38);
      if (kSteps)
        pair_tau<kPairMass>(dx, dy, dz, svx[k] - vxi, svy[k] - vyi, svz[k] - vzi, r2, ir, qi,
                            sq[k], dt, mi, ts);
      // P2P_simple / P2P force: sources only, r2 = 0 masked
      const bool use = (f & 1) && r2 > 0;
      const double w = use ? sq[k] * ir : 0.0;
      pt += w;
      const double w3 = w * ir * ir;
      ax += dx * w3;
      ay += dy * w3;
      az += dz * w3;
      if (kSteps && (f & 2)) {
        double t4 = ts * ts;
        t4 *= t4;
        tsum += 1.0 / t4;
        ts = 1e38;
        mi = qi;
      }
    }
  }
  if (!live) return;
  if (gridDim.y > 1) {
    constexpr int nc = kSteps ? 5 : 4;
    double *q = part + (size_t)blockIdx.y * nc * n + i;
    q[0] = pt;
    q[(size_t)n] = ax;
    q[2 * (size_t)n] = ay;
    q[3 * (size_t)n] = az;
    if (kSteps) q[4 * (size_t)n] = tsum;
    return;
  }
  const bool sink = b.flags[i] & kSink;
  const int o = b.order[i];
  pot[o] = sink ? pt : 0.0;
  acc[3 * o] = sink ? ax : 0.0;
  acc[3 * o + 1] = sink ? ay : 0.0;
  acc[3 * o + 2] = sink ? az : 0.0;
  if (kSteps) tstep[o] = sink ? tsum : 0.0;
}

// ---- the build with -DFMM ---------------------------------------------------
// Symmetric M2L_rotate (kernel_fast_lean.h:1034-1356) returns at once unless
// Ci < Cj (cell order) and one of the two cells holds a sink; it then adds to
// the locals of both cells.  As a list of one-sided translations: every
// accepted (i, j) with i < j and a sink in i or j gives (i, j) and (j, i).
__global__ void k_sym_count(const int2 *__restrict__ pairs, int npairs,
                            const unsigned char *__restrict__ has_sink, int *__restrict__ cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > npairs) return;
  if (k == npairs) {
    cnt[k] = 0;
    return;
  }
  const int2 pr = pairs[k];
  cnt[k] = (pr.x < pr.y && (has_sink[pr.x] || has_sink[pr.y])) ? 2 : 0;
}

__global__ void k_sym_write(const int2 *__restrict__ pairs, int npairs,
                            const unsigned char *__restrict__ has_sink,
                            const int *__restrict__ off, int2 *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const int2 pr = pairs[k];
  if (!(pr.x < pr.y && (has_sink[pr.x] || has_sink[pr.y]))) return;
  out[off[k]] = pr;
  out[off[k] + 1] = make_int2(pr.y, pr.x);
}

// P2P_simple / P2P (kernel_fast_lean.h:268-674) of one listed (target,
// source) cell pair per block: lanes take target bodies, every lane walks the
// source cell's bodies in order; results are added with atomics (a body's
// leaf and a tiny ancestor may both be listed targets).  Only pairs whose
// target holds a sink run (Ci->has_sink), only sink bodies receive.
template <bool kSteps, bool kPairMass>
__global__ __launch_bounds__(64) void k_master_p2p(CellsDev c, BodiesDev b,
                                                   const int2 *__restrict__ pairs, double dt,
                                                   ResultsDev r, double *__restrict__ ts_sorted) {
  const int2 pr = pairs[blockIdx.x];
  if (!c.has_sink[pr.x]) return;
  const int i0 = c.body[pr.x], ni = c.nbody[pr.x];
  const int j0 = c.body[pr.y], nj = c.nbody[pr.y];
  for (int base = 0; base < ni; base += kWave) {
    const int a = i0 + base + (int)threadIdx.x;
    if (base + (int)threadIdx.x >= ni) break;
    if (!(b.flags[a] & kSink)) continue;
    const double xi = b.x[a], yi = b.y[a], zi = b.z[a], qi = b.q[a];
    const double vxi = b.vx[a], vyi = b.vy[a], vzi = b.vz[a];
    double ax = 0, ay = 0, az = 0, pt = 0, mi = qi, ts = 1e38;
    for (int j = j0; j < j0 + nj; j++) {
      const double dx = b.x[j] - xi, dy = b.y[j] - yi, dz = b.z[j] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double ir = rsqrt_fp64(r2 > 0 ? r2 : 1e38);
      if (kSteps)
        pair_tau<kPairMass>(dx, dy, dz, b.vx[j] - vxi, b.vy[j] - vyi, b.vz[j] - vzi, r2, ir, qi,
                            b.q[j], dt, mi, ts);
      const bool use = (b.flags[j] & kSource) && r2 > 0;
      const double w = use ? b.q[j] * ir : 0.0;
      pt += w;
      const double w3 = w * ir * ir;
      ax += dx * w3;
      ay += dy * w3;
      az += dz * w3;
    }
    atomicAdd(r.pot + a, pt);
    atomicAdd(r.ax + a, ax);
    atomicAdd(r.ay + a, ay);
    atomicAdd(r.az + a, az);
    if (kSteps) {
      double t4 = ts * ts;
      t4 *= t4;
      atomicAdd(ts_sorted + a, 1.0 / t4);
    }
  }
}

__global__ void k_scatter_ts(BodiesDev b, const double *__restrict__ ts, int n,
                             double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[b.order[k]] = ts[k];
}

void launch_sym_count(const int2 *pairs, int npairs, const unsigned char *has_sink, int *cnt,
                      hipStream_t s) {
  k_sym_count<<<(npairs + 1 + 255) / 256, 256, 0, s>>>(pairs, npairs, has_sink, cnt);
}

void launch_sym_write(const int2 *pairs, int npairs, const unsigned char *has_sink,
                      const int *off, int2 *out, hipStream_t s) {
  if (npairs > 0) k_sym_write<<<(npairs + 255) / 256, 256, 0, s>>>(pairs, npairs, has_sink, off, out);
}

void launch_master_p2p(const CellsDev &c, const BodiesDev &b, const int2 *pairs, int npairs,
                       bool steps, bool pair_mass, double dt, ResultsDev r, double *ts_sorted,
                       hipStream_t s) {
  if (npairs <= 0) return;
  if (!steps)
    k_master_p2p<false, false><<<npairs, 64, 0, s>>>(c, b, pairs, dt, r, ts_sorted);
  else if (!pair_mass)
    k_master_p2p<true, false><<<npairs, 64, 0, s>>>(c, b, pairs, dt, r, ts_sorted);
  else
    k_master_p2p<true, true><<<npairs, 64, 0, s>>>(c, b, pairs, dt, r, ts_sorted);
}

void launch_scatter_ts(const BodiesDev &b, const double *ts_sorted, int n, double *out,
                       hipStream_t s) {
  if (n > 0) k_scatter_ts<<<(n + 255) / 256, 256, 0, s>>>(b, ts_sorted, n, out);
}

void launch_iota(int *p, int n, hipStream_t s) {
  if (n > 0) k_iota<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

void launch_zero_u8(unsigned char *p, int n, hipStream_t s) {
  if (n > 0) k_fill_u8<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

void launch_leaf_ends(const int *leaf_ids, int nleaves, const int *cbody, const int *cnbody,
                      int n, unsigned char *lend, hipStream_t s) {
  k_fill_u8<<<(n + 255) / 256, 256, 0, s>>>(lend, n);
  if (nleaves > 0)
    k_leaf_ends<<<(nleaves + 255) / 256, 256, 0, s>>>(leaf_ids, nleaves, cbody, cnbody, lend);
}

// The slabs' partial sums of k_master_direct, added in slab order, written
// to caller order for sinks (zero for the others).
template <bool kSteps>
__global__ void k_master_direct_sum(BodiesDev b, int n, int nslab, const double *__restrict__ part,
                                    double *pot, double *acc, double *tstep) {
  constexpr int nc = kSteps ? 5 : 4;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double v[nc];
#pragma unroll
  for (int k = 0; k < nc; k++) v[k] = 0;
  for (int q = 0; q < nslab; q++) {
    const double *r = part + (size_t)q * nc * n + i;
#pragma unroll
    for (int k = 0; k < nc; k++) v[k] += r[(size_t)k * n];
  }
  const bool sink = b.flags[i] & kSink;
  const int o = b.order[i];
  pot[o] = sink ? v[0] : 0.0;
  acc[3 * o] = sink ? v[1] : 0.0;
  acc[3 * o + 1] = sink ? v[2] : 0.0;
  acc[3 * o + 2] = sink ? v[3] : 0.0;
  if (kSteps) tstep[o] = sink ? v[nc - 1] : 0.0;
}

// The slabs' partial sums of the force-only direct sum, added in slab order,
// and the kick applied at once (the step loop's kicks): for every sink k of
// the union (bodies [0, ns) in the caller's order, body is[k] of the
// particle arrays) vel += F dt (gradient kicks: F dt 2/3, as ((F dt) 2) / 3)
// and, with set_pot, pot = p -- the arithmetic of a separate kick kernel on
// the summed results, without writing them out first.
__global__ void k_master_direct_sum_kick(BodiesDev b, int n, int nslab,
                                         const double *__restrict__ part,
                                         const int *__restrict__ is, int ns, double dt,
                                         bool gradient, bool set_pot, double *__restrict__ vel,
                                         double *__restrict__ pot) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int o = b.order[i];
  if (o >= ns) return;
  double v[4] = {0, 0, 0, 0};
  for (int q = 0; q < nslab; q++) {
    const double *r = part + (size_t)q * 4 * n + i;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] += r[(size_t)k * n];
  }
  const int j = is[o];
  for (int d = 0; d < 3; d++) {
    const double F = v[1 + d];
    vel[3 * j + d] += gradient ? F * dt * 2.0 / 3.0 : F * dt;
  }
  if (set_pot) pot[j] = v[0];
}

int launch_master_direct_kick(const BodiesDev &b, const unsigned char *lend, int n, bool pair_mass,
                              double dtp, int nslab, double *part, const int *is, int ns,
                              double dt, bool gradient, bool set_pot, double *vel, double *pot,
                              hipStream_t s) {
  if (nslab < 2 || !part) return -1;  // the caller kicks from the written results
  const int nb = (n + kMT - 1) / kMT;
  const int slab = (n + nslab - 1) / nslab;
  const dim3 grid(nb, nslab);
  (void)pair_mass;  // the forces alone do not depend on the step estimate's form
  k_master_direct<false, false><<<grid, kMT, 0, s>>>(b, lend, n, slab, dtp, nullptr, nullptr,
                                                      nullptr, part);
  k_master_direct_sum_kick<<<(n + 255) / 256, 256, 0, s>>>(b, n, nslab, part, is, ns, dt, gradient,
                                                           set_pot, vel, pot);
  return 0;
}

int master_direct_slabs(int n) {
  // about 1024 blocks (four per CU) when n is small, at least 64 sources a
  // slab, at most 32 slabs, so that the partial sums (32 x 5 x n doubles
  // written and read back) stay well below the force kernel's own time; one
  // slab (no partial sums) from 256 K bodies on
  const int nb = (n + kMT - 1) / kMT;
  const int want = (1024 + nb - 1) / nb;
  return std::max(1, std::min({want, (n + 63) / 64, 32}));
}

void launch_master_direct(const BodiesDev &b, const unsigned char *lend, int n, bool steps,
                          bool pair_mass, double dt, double *pot, double *acc, double *tstep,
                          int nslab, double *part, hipStream_t s) {
  const int nb = (n + kMT - 1) / kMT;
  if (nslab < 1 || !part) nslab = 1;
  const int slab = (n + nslab - 1) / nslab;
  const dim3 grid(nb, nslab);
  if (!steps)
    k_master_direct<false, false><<<grid, kMT, 0, s>>>(b, lend, n, slab, dt, pot, acc, tstep, part);
  else if (!pair_mass)
    k_master_direct<true, false><<<grid, kMT, 0, s>>>(b, lend, n, slab, dt, pot, acc, tstep, part);
  else
    k_master_direct<true, true><<<grid, kMT, 0, s>>>(b, lend, n, slab, dt, pot, acc, tstep, part);
  if (nslab > 1) {
    if (steps)
      k_master_direct_sum<true><<<(n + 255) / 256, 256, 0, s>>>(b, n, nslab, part, pot, acc, tstep);
    else
      k_master_direct_sum<false><<<(n + 255) / 256, 256, 0, s>>>(b, n, nslab, part, pot, acc, tstep);
  }
}

}  // namespace tcfmm
