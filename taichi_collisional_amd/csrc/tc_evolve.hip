// The master variant's step loop on the device: do_evolve with the
// fourth-order force-gradient scheme the master Makefile builds (GRADIENT:
// evolve_split_omelyan + kick_gradient, master integrator.h:720-812, 313-376,
// with kick_naive :182-306, kick :378-457, kick_cpu :65-98, split :592-640,
// drift with compensated summation :642-655, COMPSUM :42).  File and line
// citations refer to the reference's top-level (master) sources.
//
// The particles stay resident in HBM as arrays (positions, their compensated
// error terms, velocities, masses, time steps, accelerations, potentials,
// acc_old).  The control flow -- the recursion over time-step levels and the
// sequence of kicks and drifts -- runs on the host; every kick is one force
// evaluation of the master variant (tc_engine.hip) on the union of its sinks
// and sources, gathered by index into the context's staging arrays.
//
// split() in the reference swaps particle records to put the fast ones first;
// here a subsystem is an index list and split is a stable partition of it,
// so the caller's arrays keep their order (the reference tracks identity by
// particle id).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "tc_engine.h"

namespace tcfmm {

namespace {

// union = sinks ++ sources: positions (shifted by acc * shift_fac, the
// kick_gradient position of integrator.h:322-341; 0 for plain kicks),
// velocities, masses and sink/source flags.
__global__ void k_gather_union(const int *__restrict__ is, int ns, const int *__restrict__ js,
                               int nj, bool sinks_are_sources, const double *__restrict__ pos,
                               const double *__restrict__ vel, const double *__restrict__ mass,
                               const double *__restrict__ acc, double shift_fac,
                               double *__restrict__ x, double *__restrict__ v,
                               double *__restrict__ m, unsigned char *__restrict__ fl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns + nj) return;
  const bool sink = k < ns;
  const int i = sink ? is[k] : js[k - ns];
  for (int d = 0; d < 3; d++) {
    x[3 * k + d] = shift_fac != 0.0 ? pos[3 * i + d] + acc[3 * i + d] * shift_fac : pos[3 * i + d];
    v[3 * k + d] = vel[3 * i + d];
  }
  m[k] = mass[i];
  fl[k] = sink ? (kSink | (sinks_are_sources ? kSource : 0)) : kSource;
}

// The same union straight into the context's body arrays (structure of
// arrays, order = identity, no leaf ends), for the all-pairs force without a
// tree: one launch instead of a gather into staging and a second gather.
__global__ void k_gather_union_bodies(const int *__restrict__ is, int ns,
                                      const int *__restrict__ js, int nj, bool sinks_are_sources,
                                      const double *__restrict__ pos,
                                      const double *__restrict__ vel,
                                      const double *__restrict__ mass,
                                      const double *__restrict__ acc, double shift_fac,
                                      BodiesDev b, unsigned char *__restrict__ lend) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns + nj) return;
  const bool sink = k < ns;
  const int i = sink ? is[k] : js[k - ns];
  double x[3];
  for (int d = 0; d < 3; d++)
    x[d] = shift_fac != 0.0 ? pos[3 * i + d] + acc[3 * i + d] * shift_fac : pos[3 * i + d];
  b.x[k] = x[0];
  b.y[k] = x[1];
  b.z[k] = x[2];
  b.vx[k] = vel[3 * i];
  b.vy[k] = vel[3 * i + 1];
  b.vz[k] = vel[3 * i + 2];
  b.q[k] = mass[i];
  b.flags[k] = sink ? (kSink | (sinks_are_sources ? kSource : 0)) : kSource;
  b.order[k] = k;
  b.acc_old[k] = 0.0;
  lend[k] = 0;
}

// vel[sinks] += F * dt * fac (fac: 1 for kicks, 2/3 for gradient kicks, applied
// as ((F dt) 2) / 3 like integrator.h:350-353); pot[sinks] = p when set_pot.
// F = p = null: a force and potential of zero (no source), the same
// arithmetic on 0.0.
__global__ void k_apply_kick(const int *__restrict__ is, int ns, const double *__restrict__ F,
                             const double *__restrict__ p, double dt, bool gradient,
                             bool set_pot, double *__restrict__ vel, double *__restrict__ pot) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int i = is[k];
  for (int d = 0; d < 3; d++) {
    const double f = F ? F[3 * k + d] : 0.0;
    const double dv = gradient ? f * dt * 2.0 / 3.0 : f * dt;
    vel[3 * i + d] += dv;
  }
  if (set_pot) pot[i] = p ? p[k] : 0.0;
}

// Two kicks of zero force in a row on the same sinks (their sources empty):
// vel += 0 * dt0 (* 2/3 for a gradient kick), then the same with dt1, and
// pot = 0 with set_pot -- what two k_apply_kick launches with F = p = null
// do, in one launch.
__global__ void k_zero_kicks(const int *__restrict__ is, int ns, double dt0, bool gradient0,
                             double dt1, bool gradient1, bool set_pot, double *__restrict__ vel,
                             double *__restrict__ pot) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int i = is[k];
  const double f = 0.0;
  for (int d = 0; d < 3; d++) {
    double v = vel[3 * i + d];
    v += gradient0 ? f * dt0 * 2.0 / 3.0 : f * dt0;
    v += gradient1 ? f * dt1 * 2.0 / 3.0 : f * dt1;
    vel[3 * i + d] = v;
  }
  if (set_pot) pot[i] = 0.0;
}

// kick_naive's bookkeeping (integrator.h:250-300) for the subsystem's bodies:
// pot, acc, acc_old copied; the time step 1/sqrt(sqrt(sum)) into tmp for the
// minimum; vel += F * dt with dt = 0 is skipped (it changes nothing finite).
__global__ void k_naive_results(const int *__restrict__ is, int ns, const double *__restrict__ F,
                                const double *__restrict__ p, const double *__restrict__ tsum,
                                const double *__restrict__ ao, double *__restrict__ acc,
                                double *__restrict__ pot, double *__restrict__ acc_old,
                                double *__restrict__ tmp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int i = is[k];
  pot[i] = p[k];
  for (int d = 0; d < 3; d++) acc[3 * i + d] = F[3 * k + d];
  acc_old[i] = ao ? ao[k] : 0.0;
  tmp[k] = 1.0 / sqrt(sqrt(tsum[k]));
}

// k_naive_results, k_min and k_set_timestep in one workgroup, for
// subsystems of at most kNaiveSmall bodies: the same writes and the same
// minimum (fmin is exact and ignores NaN in any order).
constexpr int kNaiveSmall = 16384;
__global__ __launch_bounds__(1024) void k_naive_small(const int *__restrict__ is, int ns,
                                                      const double *__restrict__ F,
                                                      const double *__restrict__ p,
                                                      const double *__restrict__ tsum,
                                                      const double *__restrict__ ao,
                                                      double *__restrict__ acc,
                                                      double *__restrict__ pot,
                                                      double *__restrict__ acc_old,
                                                      double *__restrict__ ts) {
  __shared__ double red[16];
  double v = INFINITY;
  for (int k = threadIdx.x; k < ns; k += 1024) {
    const int i = is[k];
    pot[i] = p[k];
    for (int d = 0; d < 3; d++) acc[3 * i + d] = F[3 * k + d];
    acc_old[i] = ao ? ao[k] : 0.0;
    v = fmin(v, 1.0 / sqrt(sqrt(tsum[k])));
  }
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < 16; w++) r = fmin(r, red[w]);
  for (int k = threadIdx.x; k < ns; k += 1024) ts[is[k]] = r;
}

// k_naive_small followed by the split of the same subsystem (omelyan calls
// split right after kick_naive): every body of s now has the minimum step r,
// so the stable partition by r < dt is s itself, all fast or all slow;
// nfast[0] = ns or 0 and nfast[1] = *sticky go to pinned host memory as
// k_split_small writes them.
__global__ __launch_bounds__(1024) void k_naive_split_small(
    const int *__restrict__ is, int ns, const double *__restrict__ F, const double *__restrict__ p,
    const double *__restrict__ tsum, const double *__restrict__ ao, double *__restrict__ acc,
    double *__restrict__ pot, double *__restrict__ acc_old, double *__restrict__ ts, double dt,
    int *__restrict__ out, int *__restrict__ nfast, const int *sticky) {
  __shared__ double red[16];
  double v = INFINITY;
  for (int k = threadIdx.x; k < ns; k += 1024) {
    const int i = is[k];
    pot[i] = p[k];
    for (int d = 0; d < 3; d++) acc[3 * i + d] = F[3 * k + d];
    acc_old[i] = ao ? ao[k] : 0.0;
    v = fmin(v, 1.0 / sqrt(sqrt(tsum[k])));
  }
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < 16; w++) r = fmin(r, red[w]);
  for (int k = threadIdx.x; k < ns; k += 1024) {
    const int i = is[k];
    ts[i] = r;
    out[k] = i;
  }
  if (threadIdx.x == 0) {
    nfast[0] = r < dt ? ns : 0;
    nfast[1] = sticky ? *sticky : 0;
  }
}

__global__ void k_set_timestep(const int *__restrict__ is, int ns, const double *__restrict__ mn,
                               double *__restrict__ ts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ns) ts[is[k]] = *mn;
}

// One block: minimum of tmp[0..n) into *out (fmin ignores NaN as the
// reference's `if(min > t)` test does).
__global__ __launch_bounds__(1024) void k_min(const double *__restrict__ tmp, int n,
                                              double *__restrict__ out) {
  __shared__ double red[16];
  double v = INFINITY;
  for (int k = threadIdx.x; k < n; k += 1024) v = fmin(v, tmp[k]);
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 16; w++) r = fmin(r, red[w]);
    *out = r;
  }
}

// COMPSUM(s, e, ds): a = s; e += ds; s = a + e; e += a - s (integrator.h:42)
__global__ void k_drift(const int *__restrict__ is, int ns, double dt,
                        const double *__restrict__ vel, double *__restrict__ pos,
                        double *__restrict__ pos_e) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int i = is[k];
  for (int d = 0; d < 3; d++) {
    const double a = pos[3 * i + d];
    double e = pos_e[3 * i + d] + dt * vel[3 * i + d];
    const double s = a + e;
    e += a - s;
    pos[3 * i + d] = s;
    pos_e[3 * i + d] = e;
  }
}

__global__ void k_split_flags(const int *__restrict__ is, int ns, const double *__restrict__ ts,
                              double dt, int *__restrict__ fast) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= ns) fast[k] = k < ns ? (ts[is[k]] < dt ? 1 : 0) : 0;
}

// stable partition: fast ones first (their order kept), then the slow ones
__global__ void k_split_scatter(const int *__restrict__ is, int ns, const int *__restrict__ fast,
                                const int *__restrict__ off, int nfast, int *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int pos = fast[k] ? off[k] : nfast + (k - off[k]);
  out[pos] = is[k];
}

// split in one launch for a subsystem of at most 4096 bodies (one
// workgroup): the flags, their exclusive scan, the count of fast bodies
// (*nfast) and the same stable partition as k_split_flags + scan +
// k_split_scatter.
constexpr int kSplitSmall = 4096;
// nfast[1] = *sticky (the step loop's small-tree status, 0 without one);
// nfast is the context's pinned host buffer, written straight from the
// kernel (the host reads it once the stream has passed the kernel).
__global__ __launch_bounds__(1024) void k_split_small(const int *__restrict__ is, int ns,
                                                      const double *__restrict__ ts, double dt,
                                                      int *__restrict__ out, int *__restrict__ nfast,
                                                      const int *sticky) {
  __shared__ int wsum[16];
  constexpr int PER = kSplitSmall / 1024;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  int f[PER], idx[PER], cnt = 0;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int k = t * PER + q;
    idx[q] = k < ns ? is[k] : 0;
    f[q] = (k < ns && ts[idx[q]] < dt) ? 1 : 0;
    cnt += f[q];
  }
  int wt;
  const int ex = wave_excl_scan(cnt, &wt);
  if (lane == 0) wsum[w] = wt;
  __syncthreads();
  int before = 0, total = 0;
  for (int k = 0; k < 16; k++) {
    if (k < w) before += wsum[k];
    total += wsum[k];
  }
  int run = before + ex;  // fast bodies before this thread's first entry
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int k = t * PER + q;
    if (k < ns) out[f[q] ? run : total + (k - run)] = idx[q];
    run += f[q];
  }
  if (t == 0) {
    nfast[0] = total;
    nfast[1] = sticky ? *sticky : 0;
  }
}

// time-step sums of zero: kick_naive's 1/sqrt(sqrt(0)) gives every body +inf
__global__ void k_no_step_sums(double *__restrict__ out, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = 0.0;
}

__global__ void k_iota(int *__restrict__ out, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = k;
}

inline dim3 grid_for(int n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

// A subsystem: a range of one index buffer (device).
struct Sub {
  const int *idx = nullptr;
  int n = 0;
};

// A kick to apply to the sinks of a force call where the force kernels can
// (the all-pairs sum in slabs): vel[idx[k]] += F_k dt (2/3 of it for a
// gradient kick), pot[idx[k]] = p_k with set_pot; done tells the caller.
struct KickSpec {
  double dt;
  bool gradient, set_pot;
  bool done = false;
};

// The scratch buffers live in the context (grown on demand, released by
// tc_destroy), so repeated calls allocate nothing once they have grown.
class Evolver {
 public:
  Evolver(tc_context *ctx, const tc_particles &p, int mode, double dt_param)
      : ctx_(ctx), p_(p), mode_(mode & (TC_MASTER_FMM | TC_MASTER_PAIR_MASS)),
        jerk_fmm_((mode & TC_EVOLVE_JERK_FMM) != 0), dtp_(dt_param),
        lev_(ctx->ev_lev), all_(ctx->ev_all), ux_(ctx->ev_ux), uv_(ctx->ev_uv), um_(ctx->ev_um),
        ufl_(ctx->ev_ufl), upot_(ctx->ev_upot), uacc_(ctx->ev_uacc), uts_(ctx->ev_uts),
        uao_(ctx->ev_uao), tmp_(ctx->ev_tmp), flags_(ctx->ev_flags), offs_(ctx->ev_offs),
        scan_tmp_(ctx->ev_scan_tmp) {}

  int run(double *simtime, double dt);

 private:
  tc_context *ctx_;
  tc_particles p_;
  int mode_;
  bool jerk_fmm_;  // forces from the jerk build's evaluation (no time steps)
  double dtp_;
  double simtime_ = 0;
  // index buffers: one per recursion level holds that level's subsystem
  // partitioned into [fast | slow]
  std::vector<DBuf> &lev_;
  DBuf &all_, &ux_, &uv_, &um_, &ufl_, &upot_, &uacc_, &uts_, &uao_, &tmp_, &flags_, &offs_,
      &scan_tmp_;

  int ensure_union(int nu);
  int small_tree_error() {
    ctx_->err = "tree deeper than 64 levels (more than ncrit coincident bodies?)";
    return -1;
  }
  // exact: drop TC_MASTER_FMM, so that the sum is the all-pairs one
  int force(Sub sinks, Sub sources, bool sinks_are_sources, double shift_fac, bool steps,
            bool with_acc_old, bool exact = false, KickSpec *kick = nullptr);
  int kick(Sub sinks, Sub sources, double dt, bool sinks_are_fast);
  int kick_cpu(Sub s1, Sub s2, double dt);
  int kick_gradient(Sub sinks, Sub sources, double dt, bool sinks_are_fast);
  // kick_naive, and with sf the split of the same subsystem that follows it
  struct SplitFuse {
    int level;
    double dt;
    Sub slow, fast;
    bool done = false;
  };
  int kick_naive(Sub s, SplitFuse *sf = nullptr);
  int drift(Sub s, double dt);
  int split(int level, Sub s, double dt, Sub &slow, Sub &fast);
  int omelyan(int level, Sub s, double stime, double etime, double dt, bool calc_timestep);
};

int Evolver::ensure_union(int nu) {
  tc_context *ctx = ctx_;
  const size_t n = (size_t)std::max(nu, 1);
  TC_CHECK(ux_.ensure(24 * n));
  TC_CHECK(uv_.ensure(24 * n));
  TC_CHECK(um_.ensure(8 * n));
  TC_CHECK(ufl_.ensure(n));
  TC_CHECK(upot_.ensure(8 * n));
  TC_CHECK(uacc_.ensure(24 * n));
  TC_CHECK(uts_.ensure(8 * n));
  TC_CHECK(uao_.ensure(8 * n));
  TC_CHECK(tmp_.ensure(8 * n + 8));
  return 0;
}

// One get_force_and_potential on sinks ++ sources (kick_naive / kick /
// kick_gradient set the bodies up this way, integrator.h:190-245, 318-351,
// 383-424).  Results for the sinks are the first sinks.n entries of upot_ /
// uacc_ (and uts_ with steps).
int Evolver::force(Sub sinks, Sub sources, bool sinks_are_sources, double shift_fac, bool steps,
                   bool with_acc_old, bool exact, KickSpec *kick) {
  tc_context *ctx = ctx_;
  const int nu = sinks.n + sources.n;
  if (nu == 0) return 0;
  if (int e = ensure_union(nu)) return e;
  hipStream_t s = ctx->stream;
  if (!sinks_are_sources && sources.n == 0 && !steps) {
    // no body exerts a force (kick(fast, slow) with an empty slow set): the
    // reference's evaluation returns zeros; skip it (a kick of zero in one
    // launch)
    if (kick) {
      hipLaunchKernelGGL(k_apply_kick, grid_for(sinks.n), dim3(256), 0, s, sinks.idx, sinks.n,
                         nullptr, nullptr, kick->dt, kick->gradient, kick->set_pot, p_.vel,
                         p_.pot);
      kick->done = true;
      return 0;
    }
    TC_CHECK(hipMemsetAsync(upot_.p, 0, 8 * (size_t)sinks.n, s));
    TC_CHECK(hipMemsetAsync(uacc_.p, 0, 24 * (size_t)sinks.n, s));
    return 0;
  }
  const int mode = exact ? (mode_ & ~TC_MASTER_FMM) : mode_;
  if (!steps && !(mode & TC_MASTER_FMM) && !(jerk_fmm_ && !exact)) {
    // the default build's all-pairs force: gathered straight into the body
    // arrays and queued with no wait (the loop reads nothing back from it)
    if (int e = master_direct_prepare(ctx, nu, sinks.n)) return e;
    if (kick && master_direct_queue_kick(ctx, sinks.idx, sinks.n, sources.idx, sources.n,
                                         sinks_are_sources, p_.pos, p_.acc, shift_fac, p_.mass,
                                         kick->dt, kick->gradient, kick->set_pot, p_.vel,
                                         p_.pot)) {
      kick->done = true;
      return 0;
    }
    hipLaunchKernelGGL(k_gather_union_bodies, grid_for(nu), dim3(256), 0, s, sinks.idx, sinks.n,
                       sources.idx, sources.n, sinks_are_sources, p_.pos, p_.vel, p_.mass,
                       p_.acc, shift_fac, bodies_of(ctx), ctx->mlend.as<unsigned char>());
    master_direct_queue(ctx, nu, upot_.as<double>(), uacc_.as<double>());
    return 0;
  }
  if (steps && !(mode & TC_MASTER_FMM) && !jerk_fmm_ && sources.n == 0 && sinks_are_sources &&
      shift_fac == 0.0 && nu <= std::min(ctx->small_tree_max, kSmallTree)) {
    // kick_naive's time-step sums on a small subsystem: the one-launch tree
    // reads the particles through the index list (no union gather)
    return master_eval_device(ctx, nu, p_.pos, p_.vel, p_.mass, nullptr,
                              mode | TC_MASTER_STEPS, dtp_, upot_.as<double>(),
                              uacc_.as<double>(), uts_.as<double>(), nullptr, sinks.idx);
  }
  hipLaunchKernelGGL(k_gather_union, grid_for(nu), dim3(256), 0, s, sinks.idx, sinks.n,
                     sources.idx, sources.n, sinks_are_sources, p_.pos, p_.vel, p_.mass, p_.acc,
                     shift_fac, ux_.as<double>(), uv_.as<double>(), um_.as<double>(),
                     ufl_.as<unsigned char>());
  if (jerk_fmm_ && !exact) {
    // the jerk build's get_force_and_potential on the union (no time step:
    // +inf, so that split never makes a fast level)
    if (steps) {
      hipLaunchKernelGGL(k_no_step_sums, grid_for(nu), dim3(256), 0, s, uts_.as<double>(), nu);
    }
    return eval_soa_device(ctx, nu, ux_.as<double>(), uv_.as<double>(), um_.as<double>(),
                           ufl_.as<unsigned char>(), upot_.as<double>(), uacc_.as<double>(),
                           nullptr);
  }
  double *ao = nullptr;
  if (with_acc_old && (mode & TC_MASTER_FMM)) {
    // kick_naive builds fresh bodies: the pre-pass starts from 0
    ao = uao_.as<double>();
    TC_CHECK(hipMemsetAsync(ao, 0, 8 * (size_t)nu, s));
  }
  return master_eval_device(ctx, nu, ux_.as<double>(), uv_.as<double>(), um_.as<double>(),
                            ufl_.as<unsigned char>(), mode | (steps ? TC_MASTER_STEPS : 0), dtp_,
                            upot_.as<double>(), uacc_.as<double>(),
                            steps ? uts_.as<double>() : nullptr, ao);
}

// kick (integrator.h:378-457): vel[sinks] += F dt, pot[sinks] = p
int Evolver::kick(Sub sinks, Sub sources, double dt, bool sinks_are_fast) {
  if (sinks.n == 0) return 0;
  KickSpec ks{dt, false, true};
  if (int e = force(sinks, sources, !sinks_are_fast, 0.0, false, false, false, &ks)) return e;
  if (ks.done) return 0;
  hipLaunchKernelGGL(k_apply_kick, grid_for(sinks.n), dim3(256), 0, ctx_->stream, sinks.idx,
                     sinks.n, uacc_.as<double>(), upot_.as<double>(), dt, false, true, p_.vel,
                     p_.pot);
  return 0;
}

// kick_cpu (integrator.h:65-98): the all-pairs kick of s1 by s2 the
// reference uses for subsystems of <= 128 bodies; the potential is not set.
// It is an exact sum in every build, so the force call drops TC_MASTER_FMM
// (s1.n + s2.n may reach 256 bodies, above ncrit).
int Evolver::kick_cpu(Sub s1, Sub s2, double dt) {
  if (s1.n == 0) return 0;
  // s2 may contain s1 (join(slow, fast)); r = 0 pairs are skipped either way,
  // so s1 as sinks that are not sources plus s2 as sources is the same sum
  KickSpec ks{dt, false, false};
  if (int e = force(s1, s2, false, 0.0, false, false, true, &ks)) return e;
  if (ks.done) return 0;
  hipLaunchKernelGGL(k_apply_kick, grid_for(s1.n), dim3(256), 0, ctx_->stream, s1.idx, s1.n,
                     uacc_.as<double>(), upot_.as<double>(), dt, false, false, p_.vel, p_.pot);
  return 0;
}

// kick_gradient (integrator.h:313-376): positions shifted by acc dt^2/24,
// vel[sinks] += F dt 2/3.  (The reference also sets pot from an array it
// never fills; that is not reproduced.)
int Evolver::kick_gradient(Sub sinks, Sub sources, double dt, bool sinks_are_fast) {
  if (sinks.n == 0) return 0;
  const double fac = dt * dt / 24.0;
  KickSpec ks{dt, true, false};
  if (int e = force(sinks, sources, !sinks_are_fast, fac, false, false, false, &ks)) return e;
  if (ks.done) return 0;
  hipLaunchKernelGGL(k_apply_kick, grid_for(sinks.n), dim3(256), 0, ctx_->stream, sinks.idx,
                     sinks.n, uacc_.as<double>(), upot_.as<double>(), dt, true, false, p_.vel,
                     p_.pot);
  return 0;
}

// kick_naive(clevel, s, zerosys, zerosys, 0, true) (integrator.h:182-306)
// with GRADIENT: every body of s gets the minimum over s of the time step.
int Evolver::kick_naive(Sub s, SplitFuse *sf) {
  tc_context *ctx = ctx_;
  if (s.n == 0) return 0;
  if (int e = force(s, Sub{}, true, 0.0, true, true)) return e;
  hipStream_t st = ctx->stream;
  double *tmp = tmp_.as<double>();
  const double *ao = (mode_ & TC_MASTER_FMM) ? uao_.as<double>() : nullptr;
  if (sf && s.n <= kSplitSmall && s.n <= kNaiveSmall) {
    // with the split of s that follows, in one launch
    if ((int)lev_.size() <= sf->level) lev_.resize(sf->level + 1);
    DBuf &out = lev_[sf->level];
    TC_CHECK(out.ensure(sizeof(int) * (size_t)std::max(s.n, 1)));
    hipLaunchKernelGGL(k_naive_split_small, dim3(1), dim3(1024), 0, st, s.idx, s.n,
                       uacc_.as<double>(), upot_.as<double>(), uts_.as<double>(), ao, p_.acc,
                       p_.pot, p_.acc_old, p_.timestep, std::fabs(sf->dt), out.as<int>(),
                       ctx->h_pin_dev + 12, ctx->small_sticky.as<int>());
    TC_CHECK(spin_wait(ctx, st));
    int nfast = ctx->h_pin[12];
    if (ctx->h_pin[13] == 2) return small_tree_error();
    if (nfast == 1) nfast = 0;
    sf->fast = Sub{out.as<int>(), nfast};
    sf->slow = Sub{out.as<int>() + nfast, s.n - nfast};
    sf->done = true;
    return 0;
  }
  if (s.n <= kNaiveSmall) {
    hipLaunchKernelGGL(k_naive_small, dim3(1), dim3(1024), 0, st, s.idx, s.n, uacc_.as<double>(),
                       upot_.as<double>(), uts_.as<double>(), ao, p_.acc, p_.pot, p_.acc_old,
                       p_.timestep);
    return 0;
  }
  hipLaunchKernelGGL(k_naive_results, grid_for(s.n), dim3(256), 0, st, s.idx, s.n,
                     uacc_.as<double>(), upot_.as<double>(), uts_.as<double>(), ao, p_.acc,
                     p_.pot, p_.acc_old, tmp);
  hipLaunchKernelGGL(k_min, dim3(1), dim3(1024), 0, st, tmp, s.n, tmp + s.n);
  hipLaunchKernelGGL(k_set_timestep, grid_for(s.n), dim3(256), 0, st, s.idx, s.n, tmp + s.n,
                     p_.timestep);
  return 0;
}

int Evolver::drift(Sub s, double dt) {
  if (s.n == 0) return 0;
  hipLaunchKernelGGL(k_drift, grid_for(s.n), dim3(256), 0, ctx_->stream, s.idx, s.n, dt, p_.vel,
                     p_.pos, p_.pos_e);
  return 0;
}

// split (integrator.h:592-640): fast = time step < |dt|, first; a single fast
// body is not split off.
int Evolver::split(int level, Sub s, double dt, Sub &slow, Sub &fast) {
  tc_context *ctx = ctx_;
  hipStream_t st = ctx->stream;
  if ((int)lev_.size() <= level) lev_.resize(level + 1);
  DBuf &out = lev_[level];
  TC_CHECK(out.ensure(sizeof(int) * (size_t)std::max(s.n, 1)));
  TC_CHECK(flags_.ensure(sizeof(int) * ((size_t)s.n + 1)));
  TC_CHECK(offs_.ensure(sizeof(int) * ((size_t)s.n + 1)));
  TC_CHECK(scan_tmp_.ensure(scan_temp_bytes(s.n + 1) + 256));
  int nfast = 0;
  if (s.n <= kSplitSmall) {
    hipLaunchKernelGGL(k_split_small, dim3(1), dim3(1024), 0, st, s.idx, s.n, p_.timestep,
                       std::fabs(dt), out.as<int>(), ctx->h_pin_dev + 12,
                       ctx->small_sticky.as<int>());
    TC_CHECK(spin_wait(ctx, st));
    nfast = ctx->h_pin[12];
    if (ctx->h_pin[13] == 2) return small_tree_error();
  } else {
    hipLaunchKernelGGL(k_split_flags, grid_for(s.n + 1), dim3(256), 0, st, s.idx, s.n,
                       p_.timestep, std::fabs(dt), flags_.as<int>());
    if (launch_scan(flags_.as<int>(), offs_.as<int>(), s.n + 1, scan_tmp_.p, scan_tmp_.bytes,
                    st)) {
      ctx->err = "scan failed";
      return -4;
    }
    TC_CHECK(hipMemcpyAsync(ctx->h_pin + 12, offs_.as<int>() + s.n, sizeof(int),
                            hipMemcpyDeviceToHost, st));
    TC_CHECK(spin_wait(ctx, st));
    nfast = ctx->h_pin[12];
    hipLaunchKernelGGL(k_split_scatter, grid_for(s.n), dim3(256), 0, st, s.idx, s.n,
                       flags_.as<int>(), offs_.as<int>(), nfast, out.as<int>());
  }
  if (nfast == 1) nfast = 0;
  fast = Sub{out.as<int>(), nfast};
  slow = Sub{out.as<int>() + nfast, s.n - nfast};
  return 0;
}

// evolve_split_omelyan (integrator.h:720-812)
int Evolver::omelyan(int level, Sub s, double stime, double etime, double dt,
                     bool calc_timestep) {
  tc_context *ctx = ctx_;
  if (level >= 64) {  // MAXLEVEL (integrator.h:8)
    ctx->err = "time step too small: more than 64 levels";
    return -2;
  }
  Sub slow, fast;
  SplitFuse sf{level, dt};
  if (calc_timestep)
    if (int e = kick_naive(s, &sf)) return e;
  if (sf.done) {
    slow = sf.slow;
    fast = sf.fast;
  } else if (int e = split(level, s, dt, slow, fast)) {
    return e;
  }
  if (fast.n == 0) simtime_ += dt;
  auto outer_kicks = [&]() -> int {
    if (s.n > 128) {
      if (int e = kick(slow, fast, dt / 6, false)) return e;
      if (fast.n > 0)
        if (int e = kick(fast, slow, dt / 6, true)) return e;
    } else if (slow.n > 0) {
      // join(slow, fast) is s itself
      if (int e = kick_cpu(slow, s, dt / 6)) return e;
      if (fast.n > 0)
        if (int e = kick_cpu(fast, slow, dt / 6)) return e;
    }
    return 0;
  };
  if (int e = outer_kicks()) return e;
  if (fast.n > 0)
    if (int e = omelyan(level + 1, fast, stime, stime + dt / 2, dt / 2, false)) return e;
  if (int e = drift(slow, dt / 2)) return e;
  if (s.n > 0) {
    const Sub none;
    if (int e = kick(slow, none, 0.0, false)) return e;
    if (int e = kick_gradient(slow, none, dt, false)) return e;
    if (fast.n == 0 || slow.n == 0) {
      // one of the two sets is empty: of the four kicks between them, the
      // two with the other set as sinks do nothing, and the other two (the
      // pot-setting kick with dt 0, then the gradient kick) have no source,
      // so they add zero forces -- applied in one launch
      const Sub z = fast.n == 0 ? slow : fast;
      hipLaunchKernelGGL(k_zero_kicks, grid_for(z.n), dim3(256), 0, ctx_->stream, z.idx, z.n, 0.0,
                         false, dt, true, true, p_.vel, p_.pot);
    } else {
      if (int e = kick(slow, fast, 0.0, true)) return e;
      if (int e = kick(fast, slow, 0.0, true)) return e;
      if (int e = kick_gradient(slow, fast, dt, true)) return e;
      if (int e = kick_gradient(fast, slow, dt, true)) return e;
    }
  }
  if (int e = drift(slow, dt / 2)) return e;
  if (fast.n > 0)
    if (int e = omelyan(level + 1, fast, stime + dt / 2, etime, dt / 2, true)) return e;
  return outer_kicks();
}

// do_evolve (integrator.h:864-884) with GRADIENT
int Evolver::run(double *simtime, double dt) {
  tc_context *ctx = ctx_;
  simtime_ = *simtime;
  if (dt == 0) return 0;
  const int n = p_.n;
  if (n > ctx->ev_all_n) {
    TC_CHECK(all_.ensure(sizeof(int) * (size_t)n));
    hipLaunchKernelGGL(k_iota, grid_for(n), dim3(256), 0, ctx->stream, all_.as<int>(), n);
    ctx->ev_all_n = n;
  }
  const Sub all{all_.as<int>(), n};
  // the time-step sums' one-launch trees report a tree too deep through
  // small_sticky, read with every split of up to kSplitSmall bodies and at
  // the end, instead of a wait after each tree
  TC_CHECK(ctx->small_sticky.ensure(sizeof(int)));
  TC_CHECK(hipMemsetAsync(ctx->small_sticky.p, 0, sizeof(int), ctx->stream));
  struct Defer {
    tc_context *c;
    ~Defer() { c->defer_small_check = false; }
  } defer{ctx};
  ctx->defer_small_check = true;
  if (int e = kick_naive(all)) return e;
  if (int e = omelyan(0, all, simtime_, simtime_ + dt, dt, true)) return e;
  TC_CHECK(hipMemcpyAsync(ctx->h_pin + 13, ctx->small_sticky.p, sizeof(int),
                          hipMemcpyDeviceToHost, ctx->stream));
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  TC_CHECK(hipGetLastError());
  if (ctx->h_pin[13] == 2) return small_tree_error();
  *simtime = simtime_;
  return 0;
}

}  // namespace tcfmm

extern "C" int tc_master_evolve(tc_context *ctx, const tc_particles *p, double *simtime,
                                double dt, int mode, double dt_param) {
  if (!ctx) return -1;
  if (!p || !simtime || p->n <= 0 || !p->pos || !p->pos_e || !p->vel || !p->mass ||
      !p->timestep || !p->acc || !p->pot || !p->acc_old) {
    ctx->err = "tc_master_evolve: every particle array is required";
    return -1;
  }
  if (!(dt_param > 0) || !std::isfinite(dt)) {
    ctx->err = "tc_master_evolve: dt_param > 0 and a finite dt are required";
    return -1;
  }
  if (mode & ~(TC_MASTER_FMM | TC_MASTER_PAIR_MASS | TC_EVOLVE_JERK_FMM)) {
    ctx->err = "tc_master_evolve: unknown mode bits (TC_MASTER_FMM, TC_MASTER_PAIR_MASS, "
               "TC_EVOLVE_JERK_FMM)";
    return -1;
  }
  if ((mode & TC_EVOLVE_JERK_FMM) && (mode & (TC_MASTER_FMM | TC_MASTER_PAIR_MASS))) {
    ctx->err = "tc_master_evolve: TC_EVOLVE_JERK_FMM excludes the master force modes";
    return -1;
  }
  tcfmm::Evolver ev(ctx, *p, mode, dt_param);
  return ev.run(simtime, dt);
}

extern "C" int tc_master_evolve_particles(tc_context *ctx, tc_particle *part, int n,
                                          double *simtime, double dt, int mode,
                                          double dt_param) {
  if (!ctx) return -1;
  if (!part || n <= 0 || !simtime) {
    ctx->err = "tc_master_evolve_particles: particles and simtime are required";
    return -1;
  }
  using tcfmm::DBuf;
  const size_t N = (size_t)n;
  // device arrays: pos, pos_e, vel, acc [3n]; mass, timestep, pot, acc_old [n]
  DBuf buf;
  TC_CHECK(buf.ensure(sizeof(double) * 16 * N));
  std::vector<double> h(16 * N);
  double *hp = h.data(), *he = hp + 3 * N, *hv = he + 3 * N, *ha = hv + 3 * N;
  double *hm = ha + 3 * N, *ht = hm + N, *hpot = ht + N, *hao = hpot + N;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      hp[3 * i + d] = part[i].pos[d];
      he[3 * i + d] = part[i].pos_e[d];
      hv[3 * i + d] = part[i].vel[d];
      ha[3 * i + d] = part[i].acc[d];
    }
    hm[i] = part[i].mass;
    ht[i] = part[i].timestep;
    hpot[i] = part[i].pot;
    hao[i] = part[i].acc_old;
  }
  hipStream_t s = ctx->stream;
  TC_CHECK(hipMemcpyAsync(buf.p, h.data(), sizeof(double) * 16 * N, hipMemcpyHostToDevice, s));
  double *d = buf.as<double>();
  tc_particles p;
  p.n = n;
  p.pos = d;
  p.pos_e = d + 3 * N;
  p.vel = d + 6 * N;
  p.acc = d + 9 * N;
  p.mass = d + 12 * N;
  p.timestep = d + 13 * N;
  p.pot = d + 14 * N;
  p.acc_old = d + 15 * N;
  const double t0 = *simtime;
  int e = tc_master_evolve(ctx, &p, simtime, dt, mode, dt_param);
  if (e == 0) {
    TC_CHECK(hipMemcpyAsync(h.data(), buf.p, sizeof(double) * 16 * N, hipMemcpyDeviceToHost, s));
    TC_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) {
      for (int d3 = 0; d3 < 3; d3++) {
        part[i].pos[d3] = hp[3 * i + d3];
        part[i].pos_e[d3] = he[3 * i + d3];
        part[i].vel[d3] = hv[3 * i + d3];
        part[i].acc[d3] = ha[3 * i + d3];
      }
      part[i].timestep = ht[i];
      part[i].pot = hpot[i];
      part[i].acc_old = hao[i];
      if (dt != 0) part[i].postime = t0 + dt;
    }
  }
  buf.release();
  return e;
}
