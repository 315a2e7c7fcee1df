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

// Build-time shape of the direct sums (defaults measured at N = 4096 on
// MI355X): unroll of the force-only source loop (4: 19 us, 1: 22 us), targets
// per thread (2 gave fewer blocks and 10-30 % more time), the most source
// slabs of the force-only sum (64: slower with its larger partial sums), and
// the unroll of the time-step loop (2 and 4 changed nothing).
#ifndef TC_MD_UNROLL
#define TC_MD_UNROLL 4
#endif
#ifndef TC_MD_R
#define TC_MD_R 1
#endif
#ifndef TC_MD_MAXSLAB
#define TC_MD_MAXSLAB 32
#endif
#ifndef TC_MDS_UNROLL
#define TC_MDS_UNROLL 1
#endif

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

// The force alone (no time steps): one thread per target i < nt, sources
// [0, n).  Sources may be split into gridDim.y uniform slabs of `slab` bodies,
// so that a small system still fills the GPU; each slab then writes its
// partial sums, in sorted order, to part[slab][4][n] (potential,
// acceleration) and k_master_direct_sum / k_master_direct_sum_kick add the
// slabs in order.  With one slab the results are written directly.  The
// leaves do not matter here: P2P_simple sums every pair.
// The bodies of the force-only sum, read from the context's body arrays.
struct BodySource {
  BodiesDev b;
  __device__ void target(int i, double &x, double &y, double &z) const {
    x = b.x[i];
    y = b.y[i];
    z = b.z[i];
  }
  // P2P_simple: sources only (a non-source weighs nothing)
  __device__ double4 source(int j) const {
    return make_double4(b.x[j], b.y[j], b.z[j], (b.flags[j] & kSource) ? b.q[j] : 0.0);
  }
};

// The same bodies straight from the step loop's particle arrays: the union
// sinks ++ sources of a kick (body k < ns is particle is[k], a sink and, with
// sinks_are_sources, a source; body ns + k is particle js[k], a source), the
// positions shifted by acc * shift_fac (kick_gradient; 0 for plain kicks) as
// the union gather shifts them.
struct UnionSource {
  const int *is, *js;
  int ns;
  bool sinks_are_sources;
  const double *pos, *acc, *mass;
  double shift_fac;
  __device__ void at(int i, double &x, double &y, double &z) const {
    const double *p = pos + 3 * (size_t)i;
    if (shift_fac != 0.0) {
      const double *a = acc + 3 * (size_t)i;
      x = p[0] + a[0] * shift_fac;
      y = p[1] + a[1] * shift_fac;
      z = p[2] + a[2] * shift_fac;
    } else {
      x = p[0];
      y = p[1];
      z = p[2];
    }
  }
  __device__ void target(int k, double &x, double &y, double &z) const { at(is[k], x, y, z); }
  __device__ double4 source(int j) const {
    const bool sink = j < ns;
    const int i = sink ? is[j] : js[j - ns];
    double x, y, z;
    at(i, x, y, z);
    return make_double4(x, y, z, (!sink || sinks_are_sources) ? mass[i] : 0.0);
  }
};

// One thread per target i < nt (kR targets a thread), sources [0, n).
// Sources may be split into gridDim.y uniform slabs of `slab` bodies, so that
// a small system still fills the GPU; each slab then writes its partial sums,
// in target order, to part[slab][4][n] (potential, acceleration) and
// k_master_direct_sum / k_master_direct_sum_kick add the slabs in order.
// With one slab (BodySource only) the results are written directly.  The
// leaves do not matter here: P2P_simple sums every pair.
template <int kR, class Src>
__global__ __launch_bounds__(kMT) void k_master_direct(Src src, BodiesDev b, int n, int nt,
                                                       int slab, double *pot, double *acc,
                                                       double *__restrict__ part) {
  // a source as (x, y, z, q): two 16-byte LDS reads per source for kR targets
  __shared__ double4 ss[kMT];
  const int t = threadIdx.x;
  int j0 = 0, j1 = n;
  if (gridDim.y > 1) {
    j0 = (int)min((long long)n, (long long)blockIdx.y * slab);
    j1 = (int)min((long long)n, (long long)blockIdx.y * slab + slab);
  }
  // targets blockIdx.x * kR * kMT + r * kMT + t; lanes past nt run the loop
  // on a harmless stand-in body and write nothing
  double xi[kR], yi[kR], zi[kR], ax[kR], ay[kR], az[kR], pt[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) {
    const int i = (blockIdx.x * kR + r) * kMT + t;
    xi[r] = yi[r] = zi[r] = 0.0;
    if (i < nt) src.target(i, xi[r], yi[r], zi[r]);
    ax[r] = ay[r] = az[r] = pt[r] = 0.0;
  }
  for (int base = j0; base < j1; base += kMT) {
    __syncthreads();
    const int j = base + t;
    if (j < j1) ss[t] = src.source(j);
    __syncthreads();
    const int cnt = min(kMT, j1 - base);
#pragma unroll TC_MD_UNROLL
    for (int k = 0; k < cnt; k++) {
      const double4 sk = ss[k];
#pragma unroll
      for (int r = 0; r < kR; r++) {
        const double dx = sk.x - xi[r], dy = sk.y - yi[r], dz = sk.z - zi[r];
        const double r2 = dx * dx + dy * dy + dz * dz;
        // r2 = 0 masked
        const double ir = rsqrt_fp64(r2 > 0 ? r2 : 1e38);
        const double w = r2 > 0 ? sk.w * ir : 0.0;
        pt[r] += w;
        const double w3 = w * ir * ir;
        ax[r] += dx * w3;
        ay[r] += dy * w3;
        az[r] += dz * w3;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kR; r++) {
    const int i = (blockIdx.x * kR + r) * kMT + t;
    if (i >= nt) continue;
    if (gridDim.y > 1) {
      double *q = part + (size_t)blockIdx.y * 4 * n + i;
      q[0] = pt[r];
      q[(size_t)n] = ax[r];
      q[2 * (size_t)n] = ay[r];
      q[3 * (size_t)n] = az[r];
      continue;
    }
    const bool sink = b.flags[i] & kSink;
    const int o = b.order[i];
    pot[o] = sink ? pt[r] : 0.0;
    acc[3 * o] = sink ? ax[r] : 0.0;
    acc[3 * o + 1] = sink ? ay[r] : 0.0;
    acc[3 * o + 2] = sink ? az[r] : 0.0;
  }
}

// ---- the direct sum with time steps, in slabs that may cut leaves -----------
// With the time-step estimate the sum runs per source leaf: a running
// minimum, and (vector form) a target mass mi that grows along the leaf.  A
// slab that must not cut leaves leaves most of the GPU idle on a few thousand
// bodies with ncrit = 300.  Here, when the tree is a one-launch small tree
// (launch_small_tree gives every body its leaf's last index lidx and the mass
// of the leaf's bodies before it, pmx) and no two bodies are at zero distance
// (flag[0] == 0: no r^2 = 0 pair other than a body with itself), slabs are
// uniform: a slab starting inside a leaf starts mi where the leaf's running
// sum would be there -- q_i + pmx[j0], less q_i itself when the target sits
// in that leaf before j0 -- and each slab reports, per target, its leaf
// pieces: the first and the last (leaf id = last index, running minimum) and
// the sum over the leaves it holds whole; k_master_direct_merge joins the
// pieces of a leaf across slabs (minimum), then adds min^-4 per leaf, slab by
// slab.  Otherwise (the level-built tree, or coincident bodies) the slabs are
// leaf-aligned as in k_master_direct and the same format and merge serve.
// part: [slab][9][n] doubles: potential, acceleration (3), the whole leaves'
// sum, first piece minimum, last piece minimum, first and last piece ids.
constexpr int kStepCols = 9;

template <bool kPairMass>
__global__ __launch_bounds__(kMT) void k_master_direct_steps(BodiesDev b,
                                                             const unsigned char *lend,
                                                             const int *__restrict__ lidx,
                                                             const double *__restrict__ pmx,
                                                             const int *__restrict__ flag, int n,
                                                             int slab, double dt,
                                                             double *__restrict__ part) {
  __shared__ double sx[kMT], sy[kMT], sz[kMT], sq[kMT], svx[kMT], svy[kMT], svz[kMT];
  __shared__ unsigned char sf[kMT];
  const int t = threadIdx.x;
  // uniform slabs: the small tree's leaf data, and (vector form) no zero
  // distances between distinct bodies
  const bool uniform = lidx && pmx && (kPairMass || (flag && flag[0] == 0));
  int j0 = (int)min((long long)n, (long long)blockIdx.y * slab);
  int j1 = (int)min((long long)n, (long long)blockIdx.y * slab + slab);
  if (!uniform) {
    j0 = next_leaf_start(lend, n, j0);
    j1 = next_leaf_start(lend, n, j1);
  }
  const int i = blockIdx.x * kMT + t;
  const bool live = i < n;
  const double xi = live ? b.x[i] : 0.0, yi = live ? b.y[i] : 0.0, zi = live ? b.z[i] : 0.0;
  const double vxi = live ? b.vx[i] : 0.0, vyi = live ? b.vy[i] : 0.0,
               vzi = live ? b.vz[i] : 0.0;
  const double qi = live ? b.q[i] : 1.0;
  double ax = 0, ay = 0, az = 0, pt = 0, ts = 1e38, mi = qi;
  if (uniform && !kPairMass && j0 < j1 && j0 > 0 && !lend[j0 - 1]) {
    // the slab starts inside a leaf: its running mass up to j0
    const bool own = live && lidx[i] == lidx[j0] && i < j0;
    mi = own ? pmx[j0] : qi + pmx[j0];
  }
  // pieces: first (fid, fmin), whole leaves (whole), the one open (cid)
  double fmin_ = 0, whole = 0, lmin = 0;
  int fid = -1, lid = -1;
  for (int base = j0; base < j1; base += kMT) {
    __syncthreads();
    const int j = base + t;
    if (j < j1) {
      sx[t] = b.x[j];
      sy[t] = b.y[j];
      sz[t] = b.z[j];
      sq[t] = b.q[j];
      svx[t] = b.vx[j];
      svy[t] = b.vy[j];
      svz[t] = b.vz[j];
      sf[t] = ((b.flags[j] & kSource) ? 1 : 0) | (lend[j] ? 2 : 0);
    }
    __syncthreads();
    const int cnt = min(kMT, j1 - base);
#pragma unroll TC_MDS_UNROLL
    for (int k = 0; k < cnt; k++) {
      const double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const unsigned char f = sf[k];
      const double ir = rsqrt_fp64(r2 > 0 ? r2 : 1e38);
      pair_tau<kPairMass>(dx, dy, dz, svx[k] - vxi, svy[k] - vyi, svz[k] - vzi, r2, ir, qi,
                          sq[k], dt, mi, ts);
      const bool use = (f & 1) && r2 > 0;
      const double w = use ? sq[k] * ir : 0.0;
      pt += w;
      const double w3 = w * ir * ir;
      ax += dx * w3;
      ay += dy * w3;
      az += dz * w3;
      const bool end = (f & 2) != 0;
      if (end || base + k == j1 - 1) {
        // a piece closes: at the end of its leaf, or at the end of the slab
        const int id = end ? base + k : lidx[base + k];
        if (fid < 0) {
          fid = id;
          fmin_ = ts;
        } else {
          if (lid >= 0) {
            double t4 = lmin * lmin;
            t4 *= t4;
            whole += 1.0 / t4;
          }
          lid = id;
          lmin = ts;
        }
        if (end) {
          ts = 1e38;
          mi = qi;
        }
      }
    }
  }
  if (!live) return;
  if (lid < 0) {  // one piece (or none, for an empty slab)
    lid = fid;
    lmin = fmin_;
  }
  double *o = part + (size_t)blockIdx.y * kStepCols * n + i;
  o[0] = pt;
  o[(size_t)n] = ax;
  o[2 * (size_t)n] = ay;
  o[3 * (size_t)n] = az;
  o[4 * (size_t)n] = whole;
  o[5 * (size_t)n] = fmin_;
  o[6 * (size_t)n] = lmin;
  o[7 * (size_t)n] = (double)fid;
  o[8 * (size_t)n] = (double)lid;
}

// A run of consecutive slabs, per target: the force sums, and its source
// leaf pieces -- the first (id fid = the leaf's last body, running minimum
// fmin), the sum of min^-4 over the leaves strictly between (whole), and the
// last (lid, lmin); fid == lid for a single piece, fid = -1 for no source.
struct StepRun {
  double v[4], whole, fmin, lmin;
  int fid, lid;
};

__device__ __forceinline__ double inv4(double m) {
  double t4 = m * m;
  t4 *= t4;
  return 1.0 / t4;
}

// a then b: a leaf cut between the two (a's last piece and b's first have
// one id) joins by the minimum; a piece that ends up between the new first
// and last adds its min^-4 to whole
__device__ __forceinline__ StepRun join_runs(const StepRun &a, const StepRun &b) {
  StepRun r;
#pragma unroll
  for (int k = 0; k < 4; k++) r.v[k] = a.v[k] + b.v[k];
  if (a.fid < 0 || b.fid < 0) {
    const StepRun &c = a.fid < 0 ? b : a;
    r.whole = c.whole;
    r.fmin = c.fmin;
    r.lmin = c.lmin;
    r.fid = c.fid;
    r.lid = c.lid;
    return r;
  }
  const bool sa = a.fid == a.lid, sb = b.fid == b.lid;
  r.fid = a.fid;
  r.fmin = a.fmin;
  r.lid = b.lid;
  r.lmin = b.lmin;
  r.whole = a.whole + b.whole;
  if (a.lid == b.fid) {
    const double m = fmin(a.lmin, b.fmin);
    if (sa) r.fmin = m;
    if (sb) r.lmin = m;
    if (!sa && !sb) r.whole += inv4(m);
  } else {
    if (!sa) r.whole += inv4(a.lmin);
    if (!sb) r.whole += inv4(b.fmin);
  }
  return r;
}

// Per target, the slabs' runs joined in slab order: kMergeGroups groups of
// threads take consecutive ranges of slabs for the same 64 targets, then the
// first group joins the groups' runs; the time-step sum is the first piece's
// min^-4, whole, and the last piece's (for more than one piece).  Results go
// to the caller's order for sinks.
constexpr int kMergeGroups = 8;
__global__ __launch_bounds__(64 * kMergeGroups) void k_master_direct_merge(
    BodiesDev b, int n, int nslab, const double *__restrict__ part, double *pot, double *acc,
    double *tstep) {
  __shared__ StepRun runs[kMergeGroups][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int per = (nslab + kMergeGroups - 1) / kMergeGroups;
  const int q0 = min(nslab, g * per), q1 = min(nslab, q0 + per);
  StepRun acc_run;
#pragma unroll
  for (int k = 0; k < 4; k++) acc_run.v[k] = 0.0;
  acc_run.whole = acc_run.fmin = acc_run.lmin = 0.0;
  acc_run.fid = acc_run.lid = -1;
  if (i < n)
    for (int q = q0; q < q1; q++) {
      const double *r = part + (size_t)q * kStepCols * n + i;
      StepRun c;
#pragma unroll
      for (int k = 0; k < 4; k++) c.v[k] = r[(size_t)k * n];
      c.whole = r[4 * (size_t)n];
      c.fmin = r[5 * (size_t)n];
      c.lmin = r[6 * (size_t)n];
      c.fid = (int)r[7 * (size_t)n];
      c.lid = (int)r[8 * (size_t)n];
      acc_run = join_runs(acc_run, c);
    }
  runs[g][lane] = acc_run;
  __syncthreads();
  if (g != 0 || i >= n) return;
  for (int u = 1; u < kMergeGroups; u++) acc_run = join_runs(acc_run, runs[u][lane]);
  double tsum = 0.0;
  if (acc_run.fid >= 0) {
    tsum = inv4(acc_run.fmin) + acc_run.whole;
    if (acc_run.lid != acc_run.fid) tsum += inv4(acc_run.lmin);
  }
  const bool sink = b.flags[i] & kSink;
  const int o = b.order[i];
  pot[o] = sink ? acc_run.v[0] : 0.0;
  acc[3 * o] = sink ? acc_run.v[1] : 0.0;
  acc[3 * o + 1] = sink ? acc_run.v[2] : 0.0;
  acc[3 * o + 2] = sink ? acc_run.v[3] : 0.0;
  tstep[o] = sink ? tsum : 0.0;
}

int master_direct_steps_slabs(int n, bool uniform) {
  if (!uniform) {
    // slabs aligned to leaves: at least ncrit-sized, so fewer of them
    const int nb = (n + kMT - 1) / kMT;
    return std::max(1, std::min({(1024 + nb - 1) / nb, (n + 63) / 64, 32}));
  }
  // uniform slabs of >= 64 sources, about 1024 blocks
  const int nb = (n + kMT - 1) / kMT;
  const int want = (1024 + nb - 1) / nb;
  return std::max(1, std::min({want, (n + 63) / 64, 64}));
}

void launch_master_direct_steps(const BodiesDev &b, const unsigned char *lend, const int *lidx,
                                const double *pmx, const int *flag, int n, bool pair_mass,
                                double dt, double *pot, double *acc, double *tstep, int nslab,
                                double *part, hipStream_t s) {
  const int nb = (n + kMT - 1) / kMT;
  if (nslab < 1) nslab = 1;
  const int slab = (n + nslab - 1) / nslab;
  const dim3 grid(nb, nslab);
  if (!pair_mass)
    k_master_direct_steps<false><<<grid, kMT, 0, s>>>(b, lend, lidx, pmx, flag, n, slab, dt,
                                                       part);
  else
    k_master_direct_steps<true><<<grid, kMT, 0, s>>>(b, lend, lidx, pmx, flag, n, slab, dt,
                                                      part);
  k_master_direct_merge<<<(n + 63) / 64, 64 * kMergeGroups, 0, s>>>(b, n, nslab, part, pot, acc,
                                                                     tstep);
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

// The slabs' partial sums of k_master_direct for target i, added in slab
// order by kMergeGroups groups of threads (consecutive ranges of slabs for
// the same 64 targets), then the groups' sums in group order by the first
// group; true for the first group's live lanes, which hold the totals.
__device__ __forceinline__ bool sum_slabs(int i, int nlive, int n, int nslab,
                                          const double *__restrict__ part, double v[4]) {
  __shared__ double gs[kMergeGroups][4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int per = (nslab + kMergeGroups - 1) / kMergeGroups;
  const int q0 = min(nslab, g * per), q1 = min(nslab, q0 + per);
#pragma unroll
  for (int c = 0; c < 4; c++) v[c] = 0.0;
  if (i < nlive)
    for (int q = q0; q < q1; q++) {
      const double *r = part + (size_t)q * 4 * n + i;
#pragma unroll
      for (int c = 0; c < 4; c++) v[c] += r[(size_t)c * n];
    }
#pragma unroll
  for (int c = 0; c < 4; c++) gs[g][c][lane] = v[c];
  __syncthreads();
  if (g != 0 || i >= nlive) return false;
  for (int u = 1; u < kMergeGroups; u++)
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] += gs[u][c][lane];
  return true;
}

// written to caller order for sinks (zero for the others)
__global__ __launch_bounds__(64 * kMergeGroups) void k_master_direct_sum(
    BodiesDev b, int n, int nslab, const double *__restrict__ part, double *pot, double *acc) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  double v[4];
  if (!sum_slabs(i, n, n, nslab, part, v)) return;
  const bool sink = b.flags[i] & kSink;
  const int o = b.order[i];
  pot[o] = sink ? v[0] : 0.0;
  acc[3 * o] = sink ? v[1] : 0.0;
  acc[3 * o + 1] = sink ? v[2] : 0.0;
  acc[3 * o + 2] = sink ? v[3] : 0.0;
}

// The same sums with the kick applied at once (the step loop's kicks): for
// every sink k of the union (bodies [0, ns), stored in the caller's order,
// body is[k] of the particle arrays) vel += F dt (gradient kicks: F dt 2/3,
// as ((F dt) 2) / 3) and, with set_pot, pot = p -- the arithmetic of a
// separate kick kernel on the summed results, without writing them out
// first.
__global__ __launch_bounds__(64 * kMergeGroups) void k_master_direct_sum_kick(
    int n, int nslab, const double *__restrict__ part, const int *__restrict__ is, int ns,
    double dt, bool gradient, bool set_pot, double *__restrict__ vel, double *__restrict__ pot) {
  const int k = blockIdx.x * 64 + (threadIdx.x & 63);
  double v[4];
  if (!sum_slabs(k, ns, n, nslab, part, v)) return;
  const int j = is[k];
  for (int d = 0; d < 3; d++) {
    const double F = v[1 + d];
    vel[3 * j + d] += gradient ? F * dt * 2.0 / 3.0 : F * dt;
  }
  if (set_pot) pot[j] = v[0];
}

int master_direct_slabs(int nt, int n) {
  // about 1024 blocks (four per CU) over the nt targets, at least 64 sources
  // a slab, at most 32 slabs, so that the partial sums (32 x 4 x n doubles
  // written and read back) stay well below the force kernel's own time; one
  // slab (no partial sums) from 256 K targets on
  const int nb = (std::max(nt, 1) + kMT - 1) / kMT;
  const int want = (1024 + nb - 1) / nb;
  return std::max(1, std::min({want, (n + 63) / 64, TC_MD_MAXSLAB}));
}

int launch_master_direct_kick(const int *is, int ns, const int *js, int nj,
                              bool sinks_are_sources, const double *pos, const double *acc_in,
                              double shift_fac, const double *mass, int nslab, double *part,
                              double dt, bool gradient, bool set_pot, double *vel, double *pot,
                              hipStream_t s) {
  if (nslab < 2 || !part) return -1;  // the caller kicks from the written results
  if (ns < 1) return 0;
  const int n = ns + nj;
  const int slab = (n + nslab - 1) / nslab;
  const dim3 grid((ns + kMT * TC_MD_R - 1) / (kMT * TC_MD_R), nslab);
  const UnionSource src{is, js, ns, sinks_are_sources, pos, acc_in, mass, shift_fac};
  k_master_direct<TC_MD_R><<<grid, kMT, 0, s>>>(src, BodiesDev{}, n, ns, slab, nullptr, nullptr,
                                                part);
  k_master_direct_sum_kick<<<(ns + 63) / 64, 64 * kMergeGroups, 0, s>>>(
      n, nslab, part, is, ns, dt, gradient, set_pot, vel, pot);
  return 0;
}

void launch_master_direct(const BodiesDev &b, int n, double *pot, double *acc, int nslab,
                          double *part, hipStream_t s) {
  const int nb = (n + kMT * TC_MD_R - 1) / (kMT * TC_MD_R);
  if (nslab < 1 || !part) nslab = 1;
  const int slab = (n + nslab - 1) / nslab;
  const dim3 grid(nb, nslab);
  k_master_direct<TC_MD_R><<<grid, kMT, 0, s>>>(BodySource{b}, b, n, n, slab, pot, acc, part);
  if (nslab > 1)
    k_master_direct_sum<<<(n + 63) / 64, 64 * kMergeGroups, 0, s>>>(b, n, nslab, part, pot, acc);
}

}  // namespace tcfmm
