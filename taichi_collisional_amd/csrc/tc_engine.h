// Host-side state of a tcfmm context, shared by the engine (tc_engine.hip)
// and the per-kernel entry points (tc_batch.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tcfmm.h"
#include "tc_common.h"
#include "tc_kernels.h"

namespace tcfmm {

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  // grow to at least b bytes; keep `keep` bytes of old content
  hipError_t ensure(size_t b, size_t keep = 0, hipStream_t s = nullptr) {
    if (b <= bytes) return hipSuccess;
    // a quarter of headroom on the first allocation and 1.5x growth after it:
    // sizes that vary from call to call (a new root shift, a new tree) settle
    // after a call or two instead of reallocating (and synchronizing) again
    size_t nb = std::max(b + b / 4, bytes + bytes / 2);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return e;
    if (keep && p) {
      e = hipMemcpyAsync(q, p, std::min(keep, bytes), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
      hipStreamSynchronize(s);
    }
    if (p) hipFree(p);
    p = q;
    bytes = nb;
    return hipSuccess;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
};

// Page-locked host memory for the arrays the engine moves every evaluation
// (level counts, child counts, leaf ids, P2P work items): copies from or to
// it are asynchronous and skip the driver's staging.
template <class T>
struct HBuf {
  T *p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t m) {
    if (m <= n) return hipSuccess;
    size_t nn = std::max(m + m / 4, n + n / 2);  // headroom as DBuf::ensure
    T *q = nullptr;
    hipError_t e = hipHostMalloc((void **)&q, sizeof(T) * nn, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    if (p) hipHostFree(p);
    p = q;
    n = nn;
    return hipSuccess;
  }
  void release() {
    if (p) hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Lcg48 {  // srand48 / drand48: X' = (0x5DEECE66D X + 0xB) mod 2^48
  uint64_t X = 0x1234ABCD330EULL;
  void seed(long s) { X = (((uint64_t)(uint32_t)s) << 16) | 0x330E; }
  double next() {
    X = (0x5DEECE66DULL * X + 0xB) & ((1ULL << 48) - 1);
    return std::ldexp((double)X, -48);
  }
};

}  // namespace tcfmm

using namespace tcfmm;

struct tc_context {
  tc_config cfg;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  tc_stats st{};
  Lcg48 rng;
  bool fixed_shift = false;
  int p2p_mode = 0;           // TC_P2P_* (tc_set_precision)
  double close_radius = 0.0;  // mixed P2P: pairs closer than this in FP64
  // P2P launch shape (tc_set_p2p_shape / tc_tune_p2p): source bodies per LDS
  // tile, list entries per work item, the staging cost (in pair rounds) of
  // the target-group split, and targets per lane (FP64 form)
  int p2p_tile = 128;
  int p2p_list = 128;
  double p2p_stage = 8.0;
  int p2p_per_lane = 1;
  P2PSplit p2p_split;
  bool p2p_ready = false;  // the last evaluation's near-field inputs are in place
  double shift[3] = {0, 0, 0};
  // bodies
  int cap_n = 0;
  DBuf bx, by, bz, bvx, bvy, bvz, bq, bflags, border, bacc;
  DBuf rpot, rax, ray, raz, rjx, rjy, rjz;
  DBuf key_hi, key_a, key_b, val_a, val_b, sort_tmp, part, tp, dshift;
  // cells
  int cap_c = 0;
  DBuf clevel, cnchild, cchild, cbody, cnbody, cXg, cRg, cX, cR, cmin, csink;
  DBuf cM, cMd, cL, cLd;
  DBuf split_offs, split_nch, split_scan, scan_tmp;
  std::vector<int> level_start;
  DBuf leaf_ids;
  HBuf<int> h_leaves, h_nbd, h_pcnt, h_tree;
  DBuf tree_state;      // TreeState of the level-by-level build
  int tree_levels = 0;  // levels of the last tree (batch size of the next build)
  HBuf<int4> h_items;
  size_t n_items = 0;
  DBuf p2p_items, p2p_ranges, npairs;
  int nleaves = 0;
  // traversal / lists
  DBuf front_a, front_b, m2l, p2p, tcnt, tblk, tscan, tstate, tcls;
  int trav_steps[4] = {0, 0, 0, 0};  // steps of the last walk per MAC mode (batch size)
  DBuf gsort1[4], gsort2[4];  // list grouping scratch per stream: keys in, out, sources, temp
  int *h_pin = nullptr;  // small pinned host buffer for per-step counts (256 ints)
  int *h_pin_dev = nullptr;  // the same memory as kernels address it
  DBuf m2l_cnt, m2l_off, m2l_cur, m2l_src, m2l_coff, m2l_chunks;
  DBuf p2p_cnt, p2p_off, p2p_cur, p2p_src;
  // error-controlled MAC: p = 2 pre-pass, multipole norms, acc_old output
  DBuf lowM, lowL, lowF, cPn, accout, lowP;  // lowP: float records for P2P_low
  bool has_acc_old = false;  // the last evaluation ran the pre-pass
  int last_n = 0;
  // direct
  DBuf dtg, dpot, dacc, djerk, dx3, dv3, dm, dfl;
  // host Body records: page-locked staging and device result buffers
  HBuf<double> h_stage;
  DBuf hop, hoa, hoj;
  // master variant: last-body-of-leaf flags, time-step sums (caller order,
  // tree order), symmetric M2L list and its scan
  DBuf mlend, mts, mts_sorted, msym_cnt, msym_off, msym_tmp, m2l_sym;
  // master time-step sums up to small_tree_max bodies build their tree in one
  // launch (launch_small_tree; 0: always the multi-kernel build)
  int small_tree_max = kSmallTree;
  DBuf small_ci, small_cg, small_out, small_lidx, small_pmx;
  // the last one-launch tree's sorted order and its n (-1: none)
  DBuf small_order;
  int small_order_n = -1;
  // the step loop's calls do not wait for the small tree's status: a tree
  // too deep sets small_sticky (one int on the device), which the loop reads
  // with its next split
  bool defer_small_check = false;
  DBuf small_sticky;
  DBuf mdpart;  // partial sums of the source slabs of the master direct sum
  // [0..8] phase boundaries; [9], [10] start and end of P2P on near_stream;
  // [11] the tree's leaf ids and body counts are in h_leaves, h_nbd; [12],
  // [13] end of the upward geometry (stream) and moments (near_stream); [14]
  // the walk's lists written (stream), [15] the P2P list grouped and its
  // lengths in h_pcnt (near_stream)
  hipEvent_t ev[16];
  hipEvent_t ev_wait = nullptr;  // spin_wait's event (no timing)
  // second stream (default priority): P2P runs on it while the downward pass
  // runs on `stream`; its far-field results go to a second result set
  hipStream_t near_stream = nullptr;
  DBuf fpot, fax, fay, faz, fjx, fjy, fjz;
  // the step loop (tc_evolve.hip): one index buffer per time-step level, the
  // identity list, the gathered union of sinks and sources and its results,
  // and the split's flags and scan.  Kept between calls: released by
  // tc_destroy.
  std::vector<DBuf> ev_lev;
  DBuf ev_all, ev_ux, ev_uv, ev_um, ev_ufl, ev_upot, ev_uacc, ev_uts, ev_uao, ev_tmp,
      ev_flags, ev_offs, ev_scan_tmp;
  int ev_all_n = 0;  // entries of the identity list written so far
  // per-cell order and angle (tc_learn.hip): adaptive 0 = off (every pair at
  // cfg.p, cfg.theta), 1 = every leaf at (ad_q, ad_theta) through the
  // per-cell path, 2 = the learned choice.  The probe walk at probe_theta
  // gives the list-length features.
  // forest of independent clusters (tc_eval_clusters): ncl roots, cluster k
  // = bodies [start[k], start[k+1]) of the input; ncl = 1 otherwise
  int ncl = 1;
  DBuf cl_start, cl_id;
  int adaptive = 0, ad_q = 0;
  double ad_theta = 0.0, probe_theta = 0.7;
  float log_target = -7.0f;
  OrderGrid grid{};
  bool have_model = false;
  DBuf mlp_params, cq, ctheta, probe_m2l, probe_p2p, leaf_feat, leaf_pred;
  DBuf chunk_part, chunk_hist, chunk_sort[5];  // chunk_sort: keys, keys, idx, idx, temp
  long long m2l_by_order[kMaxP + 1] = {};  // M2L pairs per truncation order
  DBuf mlp_Y, mlp_H, mlp_D, mlp_dY;         // tc_mlp_grad scratch
  int mlp_cap = 0;
};

#define TC_CHECK(expr)                                                   \
  do {                                                                   \
    hipError_t _e = (expr);                                              \
    if (_e != hipSuccess) {                                              \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);      \
      return -3;                                                         \
    }                                                                    \
  } while (0)

// The master variant's get_force_and_potential on device arrays (x, v [n][3],
// m, flags): pot, acc [n][3], tstep (TC_MASTER_STEPS), acc_old (in/out, the
// TC_MASTER_FMM pre-pass).  tc_engine.hip; used by the step loop too.
// With ix (the step loop's time-step sums), body i is record ix[i] of x, v,
// m, flags; only for the one-launch small tree (error otherwise).
int master_eval_device(tc_context *ctx, int n, const double *x, const double *v,
                       const double *m, const unsigned char *flags, int mode, double dt,
                       double *pot, double *acc, double *tstep, double *acc_old,
                       const int *ix = nullptr);
// The master default build's all-pairs force without time steps, queued on
// the context's stream with no wait: the bodies must already be in the
// context's body arrays (order = identity) with the leaf-end flags zeroed.
// master_direct_prepare sizes those arrays for n bodies, of which the first
// ns are the sinks, first.
int master_direct_prepare(tc_context *ctx, int n, int ns);
// The same sum for the step loop's kicks, read straight from the particle
// arrays (the union of sinks is and sources js, no gather into the body
// arrays), with the kick applied in the sum (false: nothing launched, the
// caller gathers, queues master_direct_queue and its own kick).
bool master_direct_queue_kick(tc_context *ctx, const int *is, int ns, const int *js, int nj,
                              bool sinks_are_sources, const double *pos, const double *acc,
                              double shift_fac, const double *mass, double dt, bool gradient,
                              bool set_pot, double *vel, double *pot);
void master_direct_queue(tc_context *ctx, int n, double *pot, double *acc);

// The jerk build's get_force_and_potential on device arrays (tc_eval_soa);
// used by the step loop's TC_EVOLVE_JERK_FMM mode.
int eval_soa_device(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                    const unsigned char *flags, double *pot, double *acc, double *jerk);

// Wait for the work queued on s by polling an event (spinning) instead of
// hipStreamSynchronize, for the short waits inside an evaluation where the
// host needs a few counts to go on: a blocking wait may sleep until an
// interrupt, and that wake-up has been measured at hundreds of microseconds
// on some runs, per wait.
inline hipError_t spin_event(hipEvent_t ev) {
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e != hipErrorNotReady) return e;
    __builtin_ia32_pause();
  }
}

inline hipError_t spin_wait(tc_context *ctx, hipStream_t s) {
  hipError_t e = hipEventRecord(ctx->ev_wait, s);
  if (e != hipSuccess) return e;
  return spin_event(ctx->ev_wait);
}

inline BodiesDev bodies_of(tc_context *ctx) {
  return {ctx->bx.as<double>(),   ctx->by.as<double>(),  ctx->bz.as<double>(),
          ctx->bvx.as<double>(),  ctx->bvy.as<double>(), ctx->bvz.as<double>(),
          ctx->bq.as<double>(),   ctx->bflags.as<unsigned char>(), ctx->border.as<int>(),
          ctx->bacc.as<double>()};
}

inline CellsDev cells_of(tc_context *ctx) {
  return {ctx->clevel.as<int>(), ctx->cnchild.as<int>(), ctx->cchild.as<int>(),
          ctx->cbody.as<int>(),  ctx->cnbody.as<int>(),  ctx->cXg.as<double>(),
          ctx->cRg.as<double>(), ctx->cX.as<double>(),   ctx->cR.as<double>(),
          ctx->cmin.as<double>(), ctx->csink.as<unsigned char>(), ctx->cM.as<double>(),
          ctx->cMd.as<double>(), ctx->cL.as<double>(),   ctx->cLd.as<double>()};
}

inline ResultsDev results_of(tc_context *ctx) {
  return {ctx->rpot.as<double>(), ctx->rax.as<double>(), ctx->ray.as<double>(),
          ctx->raz.as<double>(),  ctx->rjx.as<double>(), ctx->rjy.as<double>(),
          ctx->rjz.as<double>()};
}

