// Host engine and C ABI of tcfmm (include/tcfmm.h).
//
// One evaluation = get_force_and_potential (taichi_with_jerks/integrator.h:
// 308-432) on device-resident bodies: bounds -> keys -> sort -> level-by-level
// octree -> upward pass (ball geometry, P2M, M2M per level) -> breadth-first
// dual traversal into M2L / P2P lists -> M2L -> P2P -> L2L per level -> L2P ->
// results back in the caller's order.  All kernels run on the context's
// stream; the host only reads level sizes, list sizes and the leaf set
// between phases.  Buffers grow on demand and are kept between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "tc_engine.h"

using namespace tcfmm;

static int ensure_bodies(tc_context *ctx, int n) {
  if (n <= ctx->cap_n) return 0;
  size_t d = sizeof(double) * (size_t)n;
  for (DBuf *b : {&ctx->bx, &ctx->by, &ctx->bz, &ctx->bvx, &ctx->bvy, &ctx->bvz, &ctx->bq,
                  &ctx->bacc, &ctx->rpot, &ctx->rax, &ctx->ray, &ctx->raz, &ctx->rjx, &ctx->rjy,
                  &ctx->rjz, &ctx->fpot, &ctx->fax, &ctx->fay, &ctx->faz, &ctx->fjx, &ctx->fjy,
                  &ctx->fjz})
    TC_CHECK(b->ensure(d));
  TC_CHECK(ctx->lowF.ensure(sizeof(double) * 3 * (size_t)n));
  TC_CHECK(ctx->bflags.ensure(n));
  TC_CHECK(ctx->border.ensure(sizeof(int) * (size_t)n));
  TC_CHECK(ctx->key_hi.ensure(8 * (size_t)n));
  TC_CHECK(ctx->key_a.ensure(8 * (size_t)n));
  TC_CHECK(ctx->key_b.ensure(8 * (size_t)n));
  TC_CHECK(ctx->val_a.ensure(4 * (size_t)n));
  TC_CHECK(ctx->val_b.ensure(4 * (size_t)n));
  TC_CHECK(ctx->sort_tmp.ensure(sort_temp_bytes(n) + 256));
  TC_CHECK(ctx->part.ensure(sizeof(double) * 6 * 1024));
  TC_CHECK(ctx->tp.ensure(sizeof(TreeParams)));
  TC_CHECK(ctx->dshift.ensure(sizeof(double) * 3));
  ctx->cap_n = n;
  return 0;
}

static int ensure_cells(tc_context *ctx, int nc, int keep) {
  if (nc <= ctx->cap_c) return 0;
  int cap = std::max(nc + nc / 4, ctx->cap_c + ctx->cap_c / 2);
  size_t NT = (size_t)nterm(ctx->cfg.p);
  size_t ki = sizeof(int) * (size_t)keep, kd = sizeof(double) * (size_t)keep;
  hipStream_t s = ctx->stream;
  TC_CHECK(ctx->clevel.ensure(sizeof(int) * cap, ki, s));
  TC_CHECK(ctx->cnchild.ensure(sizeof(int) * cap, ki, s));
  TC_CHECK(ctx->cchild.ensure(sizeof(int) * cap, ki, s));
  TC_CHECK(ctx->cbody.ensure(sizeof(int) * cap, ki, s));
  TC_CHECK(ctx->cnbody.ensure(sizeof(int) * cap, ki, s));
  TC_CHECK(ctx->cXg.ensure(sizeof(double) * 3 * cap, 3 * kd, s));
  TC_CHECK(ctx->cRg.ensure(sizeof(double) * cap, kd, s));
  TC_CHECK(ctx->cX.ensure(sizeof(double) * 3 * cap, 3 * kd, s));
  TC_CHECK(ctx->cR.ensure(sizeof(double) * cap, kd, s));
  TC_CHECK(ctx->cmin.ensure(sizeof(double) * cap, kd, s));
  TC_CHECK(ctx->csink.ensure(cap, keep, s));
  for (DBuf *b : {&ctx->cM, &ctx->cMd, &ctx->cL, &ctx->cLd})
    TC_CHECK(b->ensure(sizeof(double) * NT * cap));
  TC_CHECK(ctx->lowM.ensure(sizeof(double) * 9 * cap));
  TC_CHECK(ctx->lowL.ensure(sizeof(double) * 9 * cap));
  TC_CHECK(ctx->cPn.ensure(sizeof(double) * (ctx->cfg.p + 1) * cap));
  for (DBuf *b : {&ctx->m2l_cnt, &ctx->m2l_off, &ctx->m2l_cur, &ctx->m2l_coff, &ctx->p2p_cnt,
                  &ctx->p2p_off, &ctx->p2p_cur})
    TC_CHECK(b->ensure(sizeof(int) * (cap + 1)));
  TC_CHECK(ctx->scan_tmp.ensure(scan_temp_bytes(cap + 1) + 256));
  TC_CHECK(ctx->leaf_ids.ensure(sizeof(int) * cap, ki, s));  // leaves so far <= keep
  ctx->cap_c = cap;
  return 0;
}

// Tree levels are queued in batches without host reads (TreeState on the
// device holds the level range and counts); after a batch one small copy says
// whether the last level had no children.  A level whose children do not fit
// the cell arrays ends the build with a flag and what it needed: the arrays
// grow and the build starts again from the root (a new workload's first
// evaluation, or a tree that a new root shift makes larger).
static int build_tree(tc_context *ctx, int n) {
  hipStream_t s = ctx->stream;
  BodiesDev b = bodies_of(ctx);
  const int ncrit = ctx->cfg.ncrit;
  // no wait for the sort queued before: the cell arrays are not in its use,
  // and a reallocation's hipFree synchronizes by itself
  if (int e = ensure_cells(ctx, std::max(1024, 4 * n / std::max(ncrit, 1) + 64 + ctx->ncl), 0))
    return e;
  TC_CHECK(ctx->tree_state.ensure(sizeof(TreeState)));
  TC_CHECK(ctx->h_tree.ensure(sizeof(TreeState) / sizeof(int)));
  TreeState *hst = reinterpret_cast<TreeState *>(ctx->h_tree.p);
  for (int attempt = 0;; attempt++) {
    if (attempt > 16) {
      ctx->err = "tree build: cell arrays keep overflowing";
      return -3;
    }
    const int cap = ctx->cap_c;
    TC_CHECK(ctx->split_offs.ensure(sizeof(int) * 8 * (size_t)cap));
    TC_CHECK(ctx->split_nch.ensure(sizeof(int) * 2 * (size_t)cap));
    TC_CHECK(ctx->split_scan.ensure(sizeof(int) * 2 * (size_t)cap));
    std::memset(hst, 0, sizeof(TreeState));
    hst->ls = 0;
    hst->le = ctx->ncl;  // one root, or one per cluster of a forest
    TC_CHECK(hipMemcpyAsync(ctx->tree_state.p, hst, sizeof(TreeState), hipMemcpyHostToDevice, s));
    launch_root(cells_of(ctx), n, ctx->tp.as<TreeParams>(), s,
                ctx->ncl > 1 ? ctx->cl_start.as<int>() : nullptr, ctx->ncl);
    int queued = 0, batch = std::max(4, ctx->tree_levels + 1);
    for (;;) {
      for (int i = 0; i < batch; i++, queued++)
        launch_tree_level(cells_of(ctx), b, ctx->tree_state.as<TreeState>(), ncrit, cap,
                          ctx->split_offs.as<int>(), ctx->split_nch.as<int>(),
                          ctx->split_scan.as<int>(), ctx->tp.as<TreeParams>(),
                          ctx->leaf_ids.as<int>(), s);
      TC_CHECK(hipMemcpyAsync(hst, ctx->tree_state.p, sizeof(TreeState), hipMemcpyDeviceToHost,
                              s));
      TC_CHECK(spin_wait(ctx, s));
      if (hst->done || hst->overflow) break;
      if (queued > 4 * kMaxTreeLevels) {
        ctx->err = "tree build does not end";
        return -3;
      }
      batch = 4;
    }
    if (hst->overflow == 2) {
      ctx->err = "tree deeper than 64 levels (more than ncrit coincident bodies?)";
      return -1;
    }
    if (!hst->overflow) break;
    if (int e = ensure_cells(ctx, hst->need, 0)) return e;
  }
  ctx->tree_levels = hst->nlev;
  ctx->level_start.assign(hst->level_start, hst->level_start + hst->nlev + 1);
  const int depth = hst->nlev - 1, nleaf = hst->nleaf;
  // level_start holds the start of every level and, last, the total cell count
  int nc = ctx->level_start.back();
  ctx->st.ncells = nc;
  ctx->st.depth = depth;
  ctx->st.nlevels = (int)ctx->level_start.size() - 1;
  ctx->nleaves = nleaf;
  ctx->st.nleaves = nleaf;
  // leaf ids and body counts for the host-built P2P work items: copied now,
  // read after ev[11] (build_p2p_items, while the GPU runs M2L)
  TC_CHECK(ctx->h_nbd.ensure(nc));
  TC_CHECK(ctx->h_leaves.ensure(std::max(nleaf, 1)));
  TC_CHECK(hipMemcpyAsync(ctx->h_leaves.p, ctx->leaf_ids.p, sizeof(int) * nleaf,
                          hipMemcpyDeviceToHost, s));
  TC_CHECK(hipMemcpyAsync(ctx->h_nbd.p, ctx->cnbody.p, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
  TC_CHECK(hipEventRecord(ctx->ev[11], s));
  return 0;
}

// Upward pass, in one of three arrangements:
//   kUpwardWhole   all of it on `stream`;
//   kUpwardTheta   the geometry (balls, min_acc, has_sink: all the theta walk
//                  reads) on `stream`, then the moments on near_stream after
//                  it, while the walk and the lists run;
//   kUpwardErrMac  geometry and moments on near_stream, after the tree,
//                  while `stream` runs the pre-pass (they do not read
//                  acc_old); upward_min() then forms min_acc on `stream`.
// In the last two ev[13] marks the end of the near_stream part, which
// `stream` waits for before it reads M.
enum { kUpwardWhole, kUpwardTheta, kUpwardErrMac };

static int upward(tc_context *ctx, int how = kUpwardWhole) {
  hipStream_t s = ctx->stream, s2 = ctx->near_stream;
  int p = ctx->cfg.p;
  int nc = ctx->st.ncells;
  size_t NT = (size_t)nterm(p);
  hipStream_t sz = s;  // where the moments are cleared
  if (how == kUpwardErrMac) {
    TC_CHECK(hipEventRecord(ctx->ev[12], s));  // the tree
    TC_CHECK(hipStreamWaitEvent(s2, ctx->ev[12], 0));
    sz = s2;
  }
  TC_CHECK(hipMemsetAsync(ctx->cM.p, 0, sizeof(double) * NT * nc, sz));
  TC_CHECK(hipMemsetAsync(ctx->cMd.p, 0, sizeof(double) * NT * nc, sz));
  TC_CHECK(hipMemsetAsync(ctx->cL.p, 0, sizeof(double) * NT * nc, sz));
  TC_CHECK(hipMemsetAsync(ctx->cLd.p, 0, sizeof(double) * NT * nc, sz));
  TC_CHECK(hipMemsetAsync(ctx->csink.p, 0, nc, sz));
  CellsDev c = cells_of(ctx);
  BodiesDev b = bodies_of(ctx);
  const int L = (int)ctx->level_start.size() - 1;
  auto pass = [&](int what, hipStream_t st) {
    if (launch_p2m(p, c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, st, what)) return -2;
    for (int l = L - 1; l >= 0; l--)
      launch_m2m(p, c, b, ctx->level_start[l], ctx->level_start[l + 1], st, what);
    return 0;
  };
  int e = 0;
  if (how == kUpwardWhole) {
    e = pass(kUpAll, s);
  } else if (how == kUpwardTheta) {
    e = pass(kUpGeom | kUpMin, s);
    if (!e) {
      TC_CHECK(hipEventRecord(ctx->ev[12], s));
      TC_CHECK(hipStreamWaitEvent(s2, ctx->ev[12], 0));
      e = pass(kUpMoments, s2);
    }
  } else {
    e = pass(kUpGeom | kUpMoments, s2);
  }
  if (e) {
    ctx->err = "expansion order not compiled";
    return -2;
  }
  if (how != kUpwardWhole) TC_CHECK(hipEventRecord(ctx->ev[13], s2));
  return 0;
}

// min_acc of every cell from acc_old (the error MAC's), on `stream`, after the
// near_stream part of a kUpwardErrMac upward pass.
static int upward_min(tc_context *ctx) {
  hipStream_t s = ctx->stream;
  TC_CHECK(hipStreamWaitEvent(s, ctx->ev[13], 0));
  CellsDev c = cells_of(ctx);
  BodiesDev b = bodies_of(ctx);
  const int p = ctx->cfg.p, L = (int)ctx->level_start.size() - 1;
  launch_p2m(p, c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, s, kUpMin);
  for (int l = L - 1; l >= 0; l--)
    launch_m2m(p, c, b, ctx->level_start[l], ctx->level_start[l + 1], s, kUpMin);
  return 0;
}

// Breadth-first dual traversal from (root, root).  Steps are queued in
// batches without reading counts back (TravState holds them on the device);
// after a batch one small copy says whether the frontier is empty.  A step
// whose output would not fit its buffer ends the walk with a flag and what it
// needed: the buffers then grow and the walk starts again (in practice only
// on the first evaluation of a workload).
__global__ void k_root_pairs(int2 *__restrict__ f, int ncl) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < ncl) f[k] = make_int2(k, k);
}

static int traverse(tc_context *ctx, int n, const MacArgs &mac) {
  hipStream_t s = ctx->stream;
  TravState *hst = reinterpret_cast<TravState *>(ctx->h_pin + 40);
  static_assert(sizeof(TravState) <= 12 * sizeof(int), "pinned slots 40..51 (52, 53: P2P pairs)");
  TC_CHECK(ctx->tstate.ensure(sizeof(TravState)));
  TC_CHECK(ctx->tblk.ensure(sizeof(int) * 6 * kTravBlocks));
  TC_CHECK(ctx->front_a.ensure(sizeof(int2) * (size_t)std::max(1 << 16, 4 * n)));
  TC_CHECK(ctx->front_b.ensure(ctx->front_a.bytes));
  TC_CHECK(ctx->tcls.ensure(ctx->front_a.bytes / sizeof(int2)));
  TC_CHECK(ctx->m2l.ensure(sizeof(int2) * (size_t)std::max(1 << 16, 24 * n)));
  TC_CHECK(ctx->p2p.ensure(sizeof(int2) * (size_t)std::max(1 << 16, 8 * n)));
  int &known = ctx->trav_steps[mac.mode & 3];
  for (int attempt = 0;; attempt++) {
    if (attempt > 16) {
      ctx->err = "traversal: list buffers keep overflowing";
      return -3;
    }
    auto cap = [](const DBuf &d) {
      return (long long)std::min<size_t>(d.bytes / sizeof(int2), 0x7fffffff);
    };
    const long long cm = cap(ctx->m2l), cp = cap(ctx->p2p),
                    cf = std::min({cap(ctx->front_a), cap(ctx->front_b),
                                   (long long)std::min<size_t>(ctx->tcls.bytes, 0x7fffffff)});
    std::memset(hst, 0, sizeof(TravState));
    hst->nin[0] = ctx->ncl;  // the root pair (0, 0), or (k, k) for every cluster
    TC_CHECK(hipMemcpyAsync(ctx->tstate.p, hst, sizeof(TravState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_root_pairs, dim3((ctx->ncl + 255) / 256), dim3(256), 0, s,
                       ctx->front_a.as<int2>(), ctx->ncl);
    int k = 0, batch = std::max(8, known + 1);
    for (;;) {
      for (int i = 0; i < batch; i++, k++) {
        const bool odd = k & 1;
        TravArgs a;
        a.fin = (odd ? ctx->front_b : ctx->front_a).as<int2>();
        a.fout = (odd ? ctx->front_a : ctx->front_b).as<int2>();
        a.m2l = ctx->m2l.as<int2>();
        a.p2p = ctx->p2p.as<int2>();
        a.cls = ctx->tcls.as<unsigned char>();
        a.mac = mac;
        launch_trav_step(cells_of(ctx), a, ctx->tstate.as<TravState>(), (int)odd,
                         ctx->tblk.as<int>(), cm, cp, cf, s);
      }
      TC_CHECK(hipMemcpyAsync(hst, ctx->tstate.p, sizeof(TravState), hipMemcpyDeviceToHost, s));
      TC_CHECK(spin_wait(ctx, s));
      if (hst->overflow || hst->nin[k & 1] == 0) break;
      if (k > 100000) {
        ctx->err = "traversal does not end";
        return -3;
      }
      batch = 4;
    }
    if (!hst->overflow) break;
    // grow what overflowed (a quarter more than the failing step needed) and walk again
    const size_t gm = (size_t)hst->need_m2l, gp = (size_t)hst->need_p2p,
                 gf = (size_t)hst->need_front;
    if ((long long)gm > cm) TC_CHECK(ctx->m2l.ensure(sizeof(int2) * (gm + gm / 4)));
    if ((long long)gp > cp) TC_CHECK(ctx->p2p.ensure(sizeof(int2) * (gp + gp / 4)));
    if ((long long)gf > cf) {
      TC_CHECK(ctx->front_a.ensure(sizeof(int2) * (gf + gf / 4)));
      TC_CHECK(ctx->front_b.ensure(ctx->front_a.bytes));
      TC_CHECK(ctx->tcls.ensure(ctx->front_a.bytes / sizeof(int2)));
    }
  }
  known = hst->steps;
  ctx->st.n_m2l = hst->nm2l;
  ctx->st.n_p2p_cells = hst->np2p;
  return 0;
}

// group (target, source) pairs by target: stable radix sort + run bounds
// (second: on near_stream with its own sort scratch, so that two lists are
// grouped at once)
static int build_csr(tc_context *ctx, const int2 *pairs, int npairs, DBuf &cnt, DBuf &off,
                     DBuf &src, bool second = false) {
  hipStream_t s = second ? ctx->near_stream : ctx->stream;
  DBuf *sc = second ? ctx->gsort2 : ctx->gsort1;  // target keys in/out, sources in, temp
  int nc = ctx->st.ncells;
  TC_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int) * (nc + 1), s));
  TC_CHECK(hipMemsetAsync(off.p, 0, sizeof(int) * (nc + 1), s));
  size_t m = (size_t)std::max(npairs, 1);
  TC_CHECK(src.ensure(sizeof(int) * m));
  TC_CHECK(sc[0].ensure(sizeof(int) * m));
  TC_CHECK(sc[1].ensure(sizeof(int) * m));
  TC_CHECK(sc[2].ensure(sizeof(int) * m));
  TC_CHECK(sc[3].ensure(pair_sort_temp_bytes(npairs) + 256));
  int bits = 1;
  while ((1 << bits) < nc) bits++;
  if (launch_group_pairs(pairs, npairs, bits, nc, sc[0].as<int>(), sc[1].as<int>(),
                         sc[2].as<int>(), src.as<int>(), off.as<int>(), cnt.as<int>(), sc[3].p,
                         sc[3].bytes, s))
    return -4;
  return 0;
}

__global__ void k_chunk_counts(const int *__restrict__ cnt, int n, int *__restrict__ out) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i <= n) out[i] = i < n ? (cnt[i] + 63) / 64 : 0;
}

static int build_lists(tc_context *ctx, int &nchunks) {
  hipStream_t s = ctx->stream;
  int nc = ctx->st.ncells;
  // the P2P list on near_stream meanwhile (after the walk that wrote it), with
  // the per-leaf list lengths for the host-built work items; ev[15] marks both
  TC_CHECK(hipEventRecord(ctx->ev[14], s));
  TC_CHECK(hipStreamWaitEvent(ctx->near_stream, ctx->ev[14], 0));
  if (int e = build_csr(ctx, ctx->p2p.as<int2>(), (int)ctx->st.n_p2p_cells, ctx->p2p_cnt,
                        ctx->p2p_off, ctx->p2p_src, true))
    return e;
  TC_CHECK(ctx->h_pcnt.ensure(nc));
  TC_CHECK(hipMemcpyAsync(ctx->h_pcnt.p, ctx->p2p_cnt.p, sizeof(int) * nc, hipMemcpyDeviceToHost,
                          ctx->near_stream));
  TC_CHECK(hipEventRecord(ctx->ev[15], ctx->near_stream));
  if (int e = build_csr(ctx, ctx->m2l.as<int2>(), (int)ctx->st.n_m2l, ctx->m2l_cnt, ctx->m2l_off,
                        ctx->m2l_src))
    return e;
  // chunks of <= 64 sources per target for M2L
  hipLaunchKernelGGL(k_chunk_counts, dim3((nc + 1 + 255) / 256), dim3(256), 0, s,
                     ctx->m2l_cnt.as<int>(), nc, ctx->m2l_cur.as<int>());
  if (launch_scan(ctx->m2l_cur.as<int>(), ctx->m2l_coff.as<int>(), nc + 1, ctx->scan_tmp.p,
                  ctx->scan_tmp.bytes, s))
    return -4;
  TC_CHECK(hipMemcpyAsync(ctx->h_pin + 10, ctx->m2l_coff.as<int>() + nc, sizeof(int),
                          hipMemcpyDeviceToHost, s));
  TC_CHECK(spin_wait(ctx, s));
  nchunks = ctx->h_pin[10];
  TC_CHECK(ctx->m2l_chunks.ensure(sizeof(int4) * (size_t)std::max(nchunks, 1)));
  launch_make_chunks(ctx->m2l_off.as<int>(), ctx->m2l_cnt.as<int>(), ctx->m2l_coff.as<int>(), nc,
                     ctx->m2l_chunks.as<int4>(), s);
  return 0;
}

// P2P work items: every target group x slices of <= p2p_list list entries.
// Each leaf is cut into target groups of <= 64 bodies by the lane model of
// the kernel (P2PSplit).  The per-leaf list lengths were copied to pinned
// memory before the M2L launch (and the stream synchronized after that
// copy), so this host loop runs while the GPU is busy with M2L; the upload is
// queued behind it.
static int build_p2p_items(tc_context *ctx) {
  const int kList = ctx->p2p_list;
  TC_CHECK(spin_event(ctx->ev[11]));  // leaf ids and body counts (build_tree)
  TC_CHECK(spin_event(ctx->ev[15]));  // list lengths (build_lists)
  const int *pcnt = ctx->h_pcnt.p, *leaves = ctx->h_leaves.p, *nbd = ctx->h_nbd.p;
  P2PSplit &sp = ctx->p2p_split;
  sp.configure(ctx->p2p_stage, (ctx->p2p_tile * 7) / 8,
               p2p_targets_per_lane(ctx->p2p_mode, ctx->p2p_per_lane));
  int nbmax = 1;
  for (int l = 0; l < ctx->nleaves; l++) nbmax = std::max(nbmax, nbd[leaves[l]]);
  sp.grow(nbmax);
  size_t nitem = 0;
  for (int l = 0; l < ctx->nleaves; l++) {
    const int i = leaves[l];
    nitem += (size_t)sp.count(nbd[i]) * ((pcnt[i] + kList - 1) / kList);
  }
  TC_CHECK(ctx->h_items.ensure(std::max<size_t>(nitem, 1)));
  int4 *items = ctx->h_items.p;
  size_t k = 0;
  for (int l = 0; l < ctx->nleaves; l++) {
    const int i = leaves[l], nb = nbd[i], cnt = pcnt[i];
    const int ng = sp.count(nb);
    const int *gs = sp.sizes(nb);
    int o = 0;
    for (int g = 0; g < ng; g++) {
      for (int e0 = 0; e0 < cnt; e0 += kList)
        items[k++] = make_int4(i, o, e0, std::min(kList, cnt - e0) | (gs[g] << 16));
      o += gs[g];
    }
  }
  ctx->n_items = nitem;
  TC_CHECK(ctx->p2p_items.ensure(sizeof(int4) * std::max<size_t>(nitem, 1)));
  TC_CHECK(hipMemcpyAsync(ctx->p2p_items.p, items, sizeof(int4) * nitem, hipMemcpyHostToDevice,
                          ctx->stream));
  return 0;
}

// The reference's factorial_table (kernel_fast_lean.h:124-131): products of
// doubles up to index 2p-1; with the quirk index 2p stays 0.
static FactTable fact_table(const tc_config &cfg) {
  FactTable f;
  for (double &v : f.fact) v = 0.0;
  f.fact[0] = 1.0;
  const int top = cfg.quirk_factorial ? 2 * cfg.p : 2 * cfg.p + 1;
  for (int i = 1; i < top; i++) f.fact[i] = (double)i * f.fact[i - 1];
  return f;
}

static MacArgs mac_args(tc_context *ctx, int mode) {
  MacArgs a;
  std::memset(&a, 0, sizeof(a));
  a.mode = mode;
  a.p = ctx->cfg.p;
  a.theta = ctx->cfg.theta;
  a.facc = ctx->cfg.force_accuracy;
  a.Pn = ctx->cPn.as<double>();
  const FactTable f = fact_table(ctx->cfg);
  for (int k = 0; k <= a.p; k++) a.comb[k] = f.fact[a.p] / f.fact[k] / f.fact[a.p - k];
  return a;
}

// p = 2 pre-pass of the error-MAC path (integrator.h:326-356): acc_old of
// every body (sorted order) from a cheap FMM on the tree geometry.
static int low_pass(tc_context *ctx, int n) {
  hipStream_t s = ctx->stream;
  const int nc = ctx->st.ncells;
  CellsDev c = cells_of(ctx);
  BodiesDev b = bodies_of(ctx);
  TC_CHECK(hipMemsetAsync(ctx->lowM.p, 0, sizeof(double) * 9 * nc, s));
  TC_CHECK(hipMemsetAsync(ctx->lowL.p, 0, sizeof(double) * 9 * nc, s));
  TC_CHECK(hipMemsetAsync(ctx->lowF.p, 0, sizeof(double) * 3 * (size_t)n, s));
  TC_CHECK(ctx->lowP.ensure(sizeof(float4) * (size_t)n));
  launch_low_upward(c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->level_start,
                    ctx->lowM.as<double>(), s);
  if (int e = traverse(ctx, n, mac_args(ctx, kMacLow))) return e;
  ctx->st.n_m2l_low = ctx->st.n_m2l;
  ctx->st.n_p2p_low = ctx->st.n_p2p_cells;
  if (int e = build_csr(ctx, ctx->m2l.as<int2>(), (int)ctx->st.n_m2l, ctx->m2l_cnt, ctx->m2l_off,
                        ctx->m2l_src))
    return e;
  if (int e = build_csr(ctx, ctx->p2p.as<int2>(), (int)ctx->st.n_p2p_cells, ctx->p2p_cnt,
                        ctx->p2p_off, ctx->p2p_src))
    return e;
  launch_low_horizontal(c, b, nc, ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->m2l_off.as<int>(),
                        ctx->m2l_cnt.as<int>(), ctx->m2l_src.as<int>(), ctx->p2p_off.as<int>(),
                        ctx->p2p_cnt.as<int>(), ctx->p2p_src.as<int>(), ctx->lowM.as<double>(),
                        ctx->lowL.as<double>(), ctx->lowF.as<double>(), n,
                        ctx->lowP.as<float4>(), s);
  launch_low_downward(c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->level_start,
                      ctx->lowL.as<double>(), ctx->lowF.as<double>(), n, s);
  return 0;
}

// Split [0, n) into contiguous slices over a few host threads.  Packing and
// unpacking the 152-byte Body records is bound by host memory bandwidth,
// which one core does not reach.
template <class F>
static void host_parallel(int n, F f) {
  const int kMinPerThread = 1 << 15;
  int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
  nt = std::max(1, std::min(nt, n / kMinPerThread));
  if (nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  for (int t = 1; t < nt; t++) {
    const int lo = (int)((int64_t)n * t / nt), hi = (int)((int64_t)n * (t + 1) / nt);
    th.emplace_back([=, &f] { f(lo, hi); });
  }
  f(0, (int)((int64_t)n / nt));
  for (auto &t : th) t.join();
}

// acc_old of the last evaluation into out[n] in the caller's body order.
static int copy_acc_old(tc_context *ctx, int n, double *out) {
  if (!ctx->has_acc_old || n != ctx->last_n) {
    ctx->err = "acc_old is only computed by an error-MAC evaluation (TC_MAC_ERROR)";
    return -2;
  }
  launch_scatter_acc_old(bodies_of(ctx), n, out, ctx->stream);
  return 0;
}

// Bounds, keys, sort, gather into tree order, cells (buildTree,
// build_tree.h:239-487); the root shift is whatever ctx->dshift holds.
static int sort_and_build(tc_context *ctx, int n, const double *x, const double *v,
                          const double *m, const unsigned char *flags) {
  hipStream_t s = ctx->stream;
  const bool forest = ctx->ncl > 1;
  if (forest) {
    launch_bounds_clusters(x, ctx->ncl, ctx->cl_start.as<int>(), ctx->tp.as<TreeParams>(),
                           ctx->dshift.as<double>(), s);
    launch_cluster_ids(ctx->cl_start.as<int>(), ctx->ncl, n, ctx->cl_id.as<int>(), s);
  } else {
    launch_bounds(x, n, ctx->part.as<double>(), 1024, ctx->tp.as<TreeParams>(),
                  ctx->dshift.as<double>(), s);
  }
  launch_morton(x, n, ctx->tp.as<TreeParams>(), ctx->key_hi.as<uint64_t>(),
                ctx->key_a.as<uint64_t>(), ctx->val_a.as<int>(), s,
                forest ? ctx->cl_id.as<int>() : nullptr);
  if (launch_sort(ctx->key_a.as<uint64_t>(), ctx->key_b.as<uint64_t>(), ctx->val_a.as<int>(),
                  ctx->val_b.as<int>(), ctx->key_hi.as<uint64_t>(), n, ctx->sort_tmp.p,
                  ctx->sort_tmp.bytes, s) ||
      (forest && launch_sort_clusters(ctx->key_a.as<uint64_t>(), ctx->key_b.as<uint64_t>(),
                                      ctx->val_a.as<int>(), ctx->val_b.as<int>(),
                                      ctx->cl_id.as<int>(), ctx->ncl, n, ctx->sort_tmp.p,
                                      ctx->sort_tmp.bytes, s))) {
    ctx->err = "radix sort failed";
    return -4;
  }
  launch_gather(x, v, m, flags, ctx->val_a.as<int>(), n, bodies_of(ctx), s);
  return build_tree(ctx, n);
}

// Per-cell orders and angles (tc_learn.hip): a probe walk at probe_theta
// gives each leaf's M2L and P2P list lengths; with the moments done, the
// leaves' features are formed and every leaf takes its (q, theta) -- the
// learned choice, or the uniform one -- and the internal cells take the
// largest order and smallest angle below them.
__global__ void k_mass_sum(const double *__restrict__ m, int n, double *__restrict__ out) {
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += m[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

static int choose_cell_orders(tc_context *ctx, int n, const double *m) {
  hipStream_t s = ctx->stream;
  const int nc = ctx->st.ncells, nl = ctx->nleaves;
  TC_CHECK(ctx->cq.ensure(sizeof(int) * (size_t)nc));
  TC_CHECK(ctx->ctheta.ensure(sizeof(double) * (size_t)nc));
  TC_CHECK(ctx->probe_m2l.ensure(sizeof(int) * (size_t)nc));
  TC_CHECK(ctx->probe_p2p.ensure(sizeof(int) * (size_t)nc));
  TC_CHECK(ctx->leaf_feat.ensure(sizeof(float) * kLeafFeatures * (size_t)std::max(nl, 1)));
  TC_CHECK(ctx->leaf_pred.ensure(sizeof(float) * (size_t)std::max(nl, 1)));
  MacArgs pa = mac_args(ctx, kMacTheta);
  pa.theta = ctx->probe_theta;
  if (int e = traverse(ctx, n, pa)) return e;
  TC_CHECK(hipMemsetAsync(ctx->probe_m2l.p, 0, sizeof(int) * (size_t)nc, s));
  TC_CHECK(hipMemsetAsync(ctx->probe_p2p.p, 0, sizeof(int) * (size_t)nc, s));
  launch_count_targets(ctx->m2l.as<int2>(), (int)ctx->st.n_m2l, ctx->probe_m2l.as<int>(), s);
  launch_count_targets(ctx->p2p.as<int2>(), (int)ctx->st.n_p2p_cells, ctx->probe_p2p.as<int>(),
                       s);
  // mean body mass (sum over the gathered bodies)
  double *msum = reinterpret_cast<double *>(ctx->h_pin + 120);
  TC_CHECK(ctx->accout.ensure(sizeof(double)));
  TC_CHECK(hipMemsetAsync(ctx->accout.p, 0, sizeof(double), s));
  hipLaunchKernelGGL(k_mass_sum, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, s,
                     bodies_of(ctx).q, n, ctx->accout.as<double>());
  TC_CHECK(hipMemcpyAsync(msum, ctx->accout.p, sizeof(double), hipMemcpyDeviceToHost, s));
  TC_CHECK(hipStreamWaitEvent(s, ctx->ev[13], 0));  // the moments
  TC_CHECK(spin_wait(ctx, s));
  const double mean_mass = *msum / n;
  (void)m;
  launch_leaf_features(cells_of(ctx), ctx->leaf_ids.as<int>(), nl, ctx->cfg.p,
                       ctx->probe_m2l.as<int>(), ctx->probe_p2p.as<int>(), mean_mass,
                       ctx->leaf_feat.as<float>(), s);
  if (ctx->adaptive == 2) {
    launch_leaf_select(ctx->mlp_params.as<float>(), ctx->leaf_feat.as<float>(),
                       ctx->leaf_ids.as<int>(), nl, ctx->grid, ctx->log_target,
                       ctx->cq.as<int>(), ctx->ctheta.as<double>(), ctx->leaf_pred.as<float>(),
                       s);
  } else {
    launch_leaf_uniform(ctx->leaf_ids.as<int>(), nl, ctx->ad_q, ctx->ad_theta, ctx->cq.as<int>(),
                        ctx->ctheta.as<double>(), s);
  }
  const int L = (int)ctx->level_start.size() - 1;
  for (int l = L - 1; l >= 0; l--)
    launch_propagate(cells_of(ctx), ctx->level_start[l], ctx->level_start[l + 1],
                     ctx->cq.as<int>(), ctx->ctheta.as<double>(), s);
  return 0;
}

// M2L chunks (grouped by target) partitioned by their target's order, and
// one launch per order present; M2L pair counts per order for the flop count.
// Per-order chunk and pair counts: a block histogram in LDS, then one global
// atomic per order per block (most chunks share a few orders, so direct
// global atomics would all hit the same two words).
__global__ __launch_bounds__(256) void k_chunk_hist(const int4 *__restrict__ ch, int nch,
                                                    const int *__restrict__ q,
                                                    int *__restrict__ hist) {
  constexpr int H = kMaxP + 1;
  __shared__ int sh[2 * H];
  for (int t = threadIdx.x; t < 2 * H; t += 256) sh[t] = 0;
  __syncthreads();
  for (int k = blockIdx.x * 256 + threadIdx.x; k < nch; k += gridDim.x * 256) {
    const int4 c = ch[k];
    const int o = q[c.x];
    atomicAdd(&sh[o], 1);
    atomicAdd(&sh[H + o], c.z);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 2 * H; t += 256)
    if (sh[t]) atomicAdd(&hist[t], sh[t]);
}

static int m2l_by_order(tc_context *ctx, M2LArgs ma, int nchunks) {
  hipStream_t s = ctx->stream;
  const int H = kMaxP + 1;
  TC_CHECK(ctx->chunk_hist.ensure(sizeof(int) * 3 * H));
  TC_CHECK(ctx->chunk_part.ensure(sizeof(int4) * (size_t)std::max(nchunks, 1)));
  int *hist = ctx->chunk_hist.as<int>();
  TC_CHECK(hipMemsetAsync(hist, 0, sizeof(int) * 3 * H, s));
  if (nchunks > 0)
    hipLaunchKernelGGL(k_chunk_hist, dim3(std::min(1024, (nchunks + 255) / 256)), dim3(256), 0,
                       s, ma.chunks, nchunks, ctx->cq.as<int>(), hist);
  int *h = ctx->h_pin + 64;  // [64, 64 + 2H): chunk and pair counts per order
  TC_CHECK(hipMemcpyAsync(h, hist, sizeof(int) * 2 * H, hipMemcpyDeviceToHost, s));
  TC_CHECK(spin_wait(ctx, s));
  int off[kMaxP + 2];
  off[0] = 0;
  for (int q = 0; q < H; q++) off[q + 1] = off[q] + h[q];
  for (int q = 0; q < H; q++) ctx->m2l_by_order[q] = h[H + q];
  // a stable sort by order: each order's chunks keep the list's target order
  const size_t nb = sizeof(int) * (size_t)std::max(nchunks, 1);
  for (int k = 0; k < 4; k++) TC_CHECK(ctx->chunk_sort[k].ensure(nb));
  TC_CHECK(ctx->chunk_sort[4].ensure(pair_sort_temp_bytes(std::max(nchunks, 1)) + 256));
  if (launch_sort_chunks_by_order(ma.chunks, nchunks, ctx->cq.as<int>(),
                                  ctx->chunk_sort[0].as<int>(), ctx->chunk_sort[1].as<int>(),
                                  ctx->chunk_sort[2].as<int>(), ctx->chunk_sort[3].as<int>(),
                                  ctx->chunk_part.as<int4>(), ctx->chunk_sort[4].p,
                                  ctx->chunk_sort[4].bytes, s)) {
    ctx->err = "chunk sort failed";
    return -4;
  }
  const int P = ctx->cfg.p;
  for (int q = 0; q <= P; q++) {
    if (off[q + 1] == off[q]) continue;
    if (q < 2) {
      ctx->err = "per-cell order below 2";
      return -2;
    }
    M2LArgs mq = ma;
    mq.chunks = ctx->chunk_part.as<int4>() + off[q];
    mq.nchunks = off[q + 1] - off[q];
    mq.stride = nterm(P);
    mq.quirk = ma.quirk && q == P;
    if (launch_m2l(q, mq, s)) {
      ctx->err = "expansion order not compiled";
      return -2;
    }
  }
  if (off[H] != off[P + 1]) {
    ctx->err = "per-cell order above the context's order";
    return -2;
  }
  return 0;
}

// The all-pairs branch's sums (caller order) into the outputs; bodies that
// are not sinks get zeros (any output may be null).
__global__ void k_direct_out(const unsigned char *__restrict__ flags, int n,
                             const double *__restrict__ p, const double *__restrict__ a,
                             const double *__restrict__ j, double *__restrict__ pot,
                             double *__restrict__ acc, double *__restrict__ jerk) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const bool sink = !flags || (flags[i] & kSink);
  if (pot) pot[i] = sink ? p[i] : 0.0;
  for (int d = 0; d < 3; d++) {
    if (acc) acc[3 * (size_t)i + d] = sink ? a[3 * (size_t)i + d] : 0.0;
    if (jerk) jerk[3 * (size_t)i + d] = sink ? j[3 * (size_t)i + d] : 0.0;
  }
}

static int eval_device(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                       const unsigned char *flags, double *pot, double *acc, double *jerk) {
  hipStream_t s = ctx->stream;
  const tc_config &cfg = ctx->cfg;
  std::memset(&ctx->st, 0, sizeof(ctx->st));
  std::memset(ctx->m2l_by_order, 0, sizeof(ctx->m2l_by_order));
  ctx->has_acc_old = false;
  ctx->p2p_ready = false;
  ctx->last_n = n;
  if (n <= 0) {
    ctx->err = "n must be positive";
    return -1;
  }
  if (int e = ensure_bodies(ctx, n)) return e;
  // root-shift values: three drand48() per evaluation, as getBounds draws them
  // (a forest: three per cluster, in cluster order)
  double u[3];
  if (ctx->ncl > 1) {
    std::vector<double> us(3 * (size_t)ctx->ncl);
    for (int k = 0; k < ctx->ncl; k++) {
      for (int d = 0; d < 3; d++)
        u[d] = ctx->fixed_shift ? ctx->shift[d] : ctx->rng.next() - 0.5;
      std::memcpy(&us[3 * (size_t)k], u, sizeof(u));
    }
    TC_CHECK(hipMemcpy(ctx->dshift.p, us.data(), sizeof(double) * us.size(),
                       hipMemcpyHostToDevice));
  } else {
    if (ctx->fixed_shift) {
      std::memcpy(u, ctx->shift, sizeof(u));
    } else {
      for (int d = 0; d < 3; d++) u[d] = ctx->rng.next() - 0.5;
    }
    // through pinned memory: a copy from pageable memory is staged and holds the host
    std::memcpy(ctx->h_pin + 24, u, sizeof(u));
    TC_CHECK(hipMemcpyAsync(ctx->dshift.p, ctx->h_pin + 24, sizeof(u), hipMemcpyHostToDevice, s));
  }
  std::memcpy(ctx->shift, u, sizeof(u));
  TC_CHECK(hipEventRecord(ctx->ev[0], s));

  if (cfg.direct_below > 0 && n < cfg.direct_below && ctx->ncl == 1) {
    // all-pairs sum in the caller's order (direct(), traverse_eager.h:295-304)
    BodiesDev b = bodies_of(ctx);
    TC_CHECK(ctx->dtg.ensure(sizeof(int) * n));
    launch_iota(ctx->dtg.as<int>(), n, s);
    launch_gather(x, v, m, flags, ctx->dtg.as<int>(), n, b, s);
    TC_CHECK(ctx->dpot.ensure(sizeof(double) * n));
    TC_CHECK(ctx->dacc.ensure(sizeof(double) * 3 * n));
    TC_CHECK(ctx->djerk.ensure(sizeof(double) * 3 * n));
    launch_direct(b, n, ctx->dtg.as<int>(), n, ctx->dpot.as<double>(), ctx->dacc.as<double>(),
                  ctx->djerk.as<double>(), s);
    // to the outputs; non-sinks receive nothing
    hipLaunchKernelGGL(k_direct_out, dim3((n + 255) / 256), dim3(256), 0, s, flags, n,
                       ctx->dpot.as<double>(), ctx->dacc.as<double>(), ctx->djerk.as<double>(),
                       pot, acc, jerk);
    TC_CHECK(hipStreamSynchronize(s));
    TC_CHECK(hipGetLastError());
    return 0;
  }

  // ---- tree
  if (int e = sort_and_build(ctx, n, x, v, m, flags)) return e;
  const bool err_mac = cfg.mac == TC_MAC_ERROR;
  if (err_mac) {
    // the high-order geometry and moments do not read acc_old: they run on
    // the second stream alongside the pre-pass
    if (int e = upward(ctx, kUpwardErrMac)) return e;
    TC_CHECK(hipEventRecord(ctx->ev[8], s));
    if (int e = low_pass(ctx, n)) return e;
    ctx->has_acc_old = true;
  }
  TC_CHECK(hipEventRecord(ctx->ev[1], s));
  // ---- upward: with the theta MAC the walk needs only the geometry, so the
  // moments are formed on the second stream while the walk and the lists run
  if (err_mac) {
    if (int e = upward_min(ctx)) return e;
    launch_pn(ctx->cM.as<double>(), ctx->st.ncells, cfg.p, fact_table(cfg), ctx->cPn.as<double>(),
              s);
  } else {
    if (int e = upward(ctx, kUpwardTheta)) return e;
  }
  TC_CHECK(hipEventRecord(ctx->ev[2], s));
  // ---- traversal and lists
  const bool per_cell = !err_mac && ctx->adaptive != 0;
  if (per_cell) {
    if (int e = choose_cell_orders(ctx, n, m)) return e;
    MacArgs ma = mac_args(ctx, kMacTheta);
    ma.theta_cell = ctx->ctheta.as<double>();
    if (int e = traverse(ctx, n, ma)) return e;
  } else {
    if (int e = traverse(ctx, n, mac_args(ctx, err_mac ? kMacError : kMacTheta))) return e;
  }
  TC_CHECK(hipEventRecord(ctx->ev[3], s));
  int nchunks = 0;
  if (int e = build_lists(ctx, nchunks)) return e;
  if (!err_mac) TC_CHECK(hipStreamWaitEvent(s, ctx->ev[13], 0));  // the moments
  TC_CHECK(hipEventRecord(ctx->ev[4], s));
  // ---- far field
  M2LArgs ma;
  ma.chunks = ctx->m2l_chunks.as<int4>();
  ma.nchunks = nchunks;
  ma.src = ctx->m2l_src.as<int>();
  ma.cX = ctx->cX.as<double>();
  ma.has_sink = ctx->csink.as<unsigned char>();
  ma.M = ctx->cM.as<double>();
  ma.Md = ctx->cMd.as<double>();
  ma.L = ctx->cL.as<double>();
  ma.Ld = ctx->cLd.as<double>();
  ma.quirk = cfg.quirk_factorial;
  if (per_cell) {
    if (int e = m2l_by_order(ctx, ma, nchunks)) return e;
  } else {
    if (launch_m2l(cfg.p, ma, s)) {
      ctx->err = "expansion order not compiled";
      return -2;
    }
    ctx->m2l_by_order[cfg.p] = ctx->st.n_m2l;
  }
  // the near field's set-up on the second stream, beside M2L: it reads only
  // the tree and the P2P list (already queued on that stream)
  hipStream_t s2 = ctx->near_stream;
  ResultsDev r = results_of(ctx);
  for (double *q : {r.pot, r.ax, r.ay, r.az, r.jx, r.jy, r.jz})
    TC_CHECK(hipMemsetAsync(q, 0, sizeof(double) * n, s2));
  TC_CHECK(ctx->p2p_ranges.ensure(sizeof(int2) * (size_t)std::max<long long>(ctx->st.n_p2p_cells, 1)));
  TC_CHECK(ctx->npairs.ensure(sizeof(unsigned long long)));
  TC_CHECK(hipMemsetAsync(ctx->npairs.p, 0, sizeof(unsigned long long), s2));
  launch_p2p_ranges(cells_of(ctx), ctx->p2p_src.as<int>(), (int)ctx->st.n_p2p_cells,
                    ctx->p2p_ranges.as<int2>(), s2);
  // the GPU runs M2L while the host builds the P2P work items
  if (int e = build_p2p_items(ctx)) return e;
  TC_CHECK(hipEventRecord(ctx->ev[5], s));
  // ---- near field on the second stream, after M2L (and the work-item upload
  // queued behind it), alongside the downward pass
  TC_CHECK(hipStreamWaitEvent(s2, ctx->ev[5], 0));
  TC_CHECK(hipEventRecord(ctx->ev[9], s2));
  launch_p2p(cells_of(ctx), bodies_of(ctx), r, ctx->p2p_items.as<int4>(),
             (int)ctx->n_items, ctx->p2p_off.as<int>(), ctx->p2p_cnt.as<int>(),
             ctx->p2p_ranges.as<int2>(), ctx->npairs.as<unsigned long long>(), s2,
             ctx->p2p_mode, ctx->close_radius, ctx->p2p_tile,
             p2p_targets_per_lane(ctx->p2p_mode, ctx->p2p_per_lane));
  TC_CHECK(hipEventRecord(ctx->ev[10], s2));
  ctx->p2p_ready = true;
  // ---- downward into the far-field result set
  ResultsDev rf = {ctx->fpot.as<double>(), ctx->fax.as<double>(), ctx->fay.as<double>(),
                   ctx->faz.as<double>(), ctx->fjx.as<double>(), ctx->fjy.as<double>(),
                   ctx->fjz.as<double>()};
  for (double *q : {rf.pot, rf.ax, rf.ay, rf.az, rf.jx, rf.jy, rf.jz})
    TC_CHECK(hipMemsetAsync(q, 0, sizeof(double) * n, s));
  int L = (int)ctx->level_start.size() - 1;
  for (int l = 0; l < L; l++)
    launch_l2l(cfg.p, cells_of(ctx), ctx->level_start[l], ctx->level_start[l + 1], s);
  launch_l2p(cfg.p, cells_of(ctx), bodies_of(ctx), rf, ctx->leaf_ids.as<int>(), ctx->nleaves, s);
  TC_CHECK(hipEventRecord(ctx->ev[6], s));
  TC_CHECK(hipStreamWaitEvent(s, ctx->ev[10], 0));
  launch_scatter_results_sum(bodies_of(ctx), r, rf, n, pot, acc, jerk, s);
  TC_CHECK(hipEventRecord(ctx->ev[7], s));
  // body pairs evaluated by P2P (statistics, counted inside the P2P kernel)
  unsigned long long *pairs = reinterpret_cast<unsigned long long *>(ctx->h_pin + 52);
  TC_CHECK(hipMemcpyAsync(pairs, ctx->npairs.p, sizeof(*pairs), hipMemcpyDeviceToHost, s));
  TC_CHECK(spin_wait(ctx, s));
  ctx->st.n_p2p_pairs = (long long)*pairs;
  float ms[8];
  for (int k = 0; k < 5; k++) hipEventElapsedTime(&ms[k], ctx->ev[k], ctx->ev[k + 1]);
  hipEventElapsedTime(&ms[5], ctx->ev[9], ctx->ev[10]);  // P2P (second stream)
  hipEventElapsedTime(&ms[6], ctx->ev[5], ctx->ev[6]);   // downward (alongside P2P)
  hipEventElapsedTime(&ms[7], ctx->ev[0], ctx->ev[7]);
  if (err_mac) {
    // tree phase ends where the pre-pass starts
    float pre = 0;
    hipEventElapsedTime(&ms[0], ctx->ev[0], ctx->ev[8]);
    hipEventElapsedTime(&pre, ctx->ev[8], ctx->ev[1]);
    ctx->st.ms_prepass = pre;
  }
  ctx->st.ms_tree = ms[0];
  ctx->st.ms_upward = ms[1];
  ctx->st.ms_traverse = ms[2];
  ctx->st.ms_lists = ms[3];
  ctx->st.ms_m2l = ms[4];
  ctx->st.ms_p2p = ms[5];
  ctx->st.ms_downward = ms[6];
  ctx->st.ms_total = ms[7];
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("kernel launch: ") + hipGetErrorString(e);
    return -3;
  }
  return 0;
}

// The master build with -DFMM above ncrit bodies (master integrator.h:466-556):
// the master p = 2 pre-pass (acc_old accumulates onto the caller's values),
// upward pass, Pn, the symmetric error-controlled walk, symmetric M2L,
// P2P_simple / P2P over the near list, downward pass.  Jerk-free outputs; the
// M2L and downward kernels also carry Mdot/Ldot, whose results are dropped.
static int master_fmm_device(tc_context *ctx, int n, bool steps, bool pair_mass, double dt,
                             double *pot, double *acc, double *tstep, double *acc_old) {
  hipStream_t s = ctx->stream;
  const tc_config &cfg = ctx->cfg;
  const int nc = ctx->st.ncells;
  CellsDev c = cells_of(ctx);
  BodiesDev b = bodies_of(ctx);
  // ---- pre-pass
  TC_CHECK(hipEventRecord(ctx->ev[8], s));
  TC_CHECK(hipMemsetAsync(ctx->lowM.p, 0, sizeof(double) * 9 * nc, s));
  TC_CHECK(hipMemsetAsync(ctx->lowL.p, 0, sizeof(double) * 9 * nc, s));
  TC_CHECK(hipMemsetAsync(ctx->lowF.p, 0, sizeof(double) * 3 * (size_t)n, s));
  launch_low_upward(c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->level_start,
                    ctx->lowM.as<double>(), s);
  if (int e = traverse(ctx, n, mac_args(ctx, kMacLow))) return e;
  ctx->st.n_m2l_low = ctx->st.n_m2l;
  ctx->st.n_p2p_low = ctx->st.n_p2p_cells;
  if (int e = build_csr(ctx, ctx->m2l.as<int2>(), (int)ctx->st.n_m2l, ctx->m2l_cnt, ctx->m2l_off,
                        ctx->m2l_src))
    return e;
  if (int e = build_csr(ctx, ctx->p2p.as<int2>(), (int)ctx->st.n_p2p_cells, ctx->p2p_cnt,
                        ctx->p2p_off, ctx->p2p_src))
    return e;
  launch_low_horizontal(c, b, nc, ctx->leaf_ids.as<int>(), 0, ctx->m2l_off.as<int>(),
                        ctx->m2l_cnt.as<int>(), ctx->m2l_src.as<int>(), ctx->p2p_off.as<int>(),
                        ctx->p2p_cnt.as<int>(), ctx->p2p_src.as<int>(), ctx->lowM.as<double>(),
                        ctx->lowL.as<double>(), ctx->lowF.as<double>(), n, nullptr, s);
  launch_low_p2p_master(c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->p2p_off.as<int>(),
                        ctx->p2p_cnt.as<int>(), ctx->p2p_src.as<int>(), ctx->lowF.as<double>(), s);
  launch_low_downward_master(c, b, ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->level_start,
                             ctx->lowL.as<double>(), ctx->lowF.as<double>(), n, acc_old, s);
  ctx->has_acc_old = true;
  TC_CHECK(hipEventRecord(ctx->ev[1], s));
  // ---- upward, norms
  if (int e = upward(ctx)) return e;
  launch_pn(ctx->cM.as<double>(), nc, cfg.p, fact_table(cfg), ctx->cPn.as<double>(), s);
  TC_CHECK(hipEventRecord(ctx->ev[2], s));
  // ---- symmetric walk
  MacArgs mac = mac_args(ctx, kMacMaster);
  mac.steps = steps;
  if (int e = traverse(ctx, n, mac)) return e;
  const long long accepted = ctx->st.n_m2l;
  if (accepted > 0) {
    const int na = (int)accepted;
    TC_CHECK(ctx->msym_cnt.ensure(sizeof(int) * ((size_t)na + 1)));
    TC_CHECK(ctx->msym_off.ensure(sizeof(int) * ((size_t)na + 1)));
    TC_CHECK(ctx->msym_tmp.ensure(scan_temp_bytes(na + 1) + 256));
    launch_sym_count(ctx->m2l.as<int2>(), na, ctx->csink.as<unsigned char>(),
                     ctx->msym_cnt.as<int>(), s);
    if (launch_scan(ctx->msym_cnt.as<int>(), ctx->msym_off.as<int>(), na + 1, ctx->msym_tmp.p,
                    ctx->msym_tmp.bytes, s)) {
      ctx->err = "scan failed";
      return -4;
    }
    TC_CHECK(hipMemcpyAsync(ctx->h_pin + 11, ctx->msym_off.as<int>() + na, sizeof(int),
                            hipMemcpyDeviceToHost, s));
    TC_CHECK(spin_wait(ctx, s));
    const int nsym = ctx->h_pin[11];
    TC_CHECK(ctx->m2l_sym.ensure(sizeof(int2) * (size_t)std::max(nsym, 1)));
    launch_sym_write(ctx->m2l.as<int2>(), na, ctx->csink.as<unsigned char>(),
                     ctx->msym_off.as<int>(), ctx->m2l_sym.as<int2>(), s);
    // the lists below read ctx->m2l: hand them the symmetric list
    std::swap(ctx->m2l, ctx->m2l_sym);
    ctx->st.n_m2l = nsym;
  }
  TC_CHECK(hipEventRecord(ctx->ev[3], s));
  int nchunks = 0;
  if (int e = build_lists(ctx, nchunks)) return e;
  TC_CHECK(hipStreamWaitEvent(s, ctx->ev[15], 0));  // the P2P list's grouping (near_stream)
  TC_CHECK(hipEventRecord(ctx->ev[4], s));
  // ---- far field
  if (nchunks > 0) {
    M2LArgs ma;
    ma.chunks = ctx->m2l_chunks.as<int4>();
    ma.nchunks = nchunks;
    ma.src = ctx->m2l_src.as<int>();
    ma.cX = ctx->cX.as<double>();
    ma.has_sink = ctx->csink.as<unsigned char>();
    ma.M = ctx->cM.as<double>();
    ma.Md = ctx->cMd.as<double>();
    ma.L = ctx->cL.as<double>();
    ma.Ld = ctx->cLd.as<double>();
    ma.quirk = cfg.quirk_factorial;
    if (launch_m2l(cfg.p, ma, s)) {
      ctx->err = "expansion order not compiled";
      return -2;
    }
  }
  TC_CHECK(hipEventRecord(ctx->ev[5], s));
  // ---- near field
  ResultsDev r = results_of(ctx);
  for (double *q : {r.pot, r.ax, r.ay, r.az, r.jx, r.jy, r.jz})
    TC_CHECK(hipMemsetAsync(q, 0, sizeof(double) * n, s));
  TC_CHECK(ctx->mts_sorted.ensure(sizeof(double) * (size_t)n));
  TC_CHECK(hipMemsetAsync(ctx->mts_sorted.p, 0, sizeof(double) * n, s));
  launch_master_p2p(c, b, ctx->p2p.as<int2>(), (int)ctx->st.n_p2p_cells, steps, pair_mass, dt, r,
                    ctx->mts_sorted.as<double>(), s);
  TC_CHECK(hipEventRecord(ctx->ev[6], s));
  // ---- downward
  int L = (int)ctx->level_start.size() - 1;
  for (int l = 0; l < L; l++)
    launch_l2l(cfg.p, c, ctx->level_start[l], ctx->level_start[l + 1], s);
  launch_l2p(cfg.p, c, b, r, ctx->leaf_ids.as<int>(), ctx->nleaves, s);
  launch_scatter_results(b, r, n, pot, acc, nullptr, s);
  if (steps) launch_scatter_ts(b, ctx->mts_sorted.as<double>(), n, tstep, s);
  if (acc_old) launch_scatter_acc_old(b, n, acc_old, s);
  TC_CHECK(hipEventRecord(ctx->ev[7], s));
  TC_CHECK(spin_wait(ctx, s));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("kernel launch: ") + hipGetErrorString(e);
    return -3;
  }
  float ms[8], pre = 0;
  for (int k = 1; k < 7; k++) hipEventElapsedTime(&ms[k], ctx->ev[k], ctx->ev[k + 1]);
  hipEventElapsedTime(&ms[0], ctx->ev[0], ctx->ev[8]);
  hipEventElapsedTime(&pre, ctx->ev[8], ctx->ev[1]);
  hipEventElapsedTime(&ms[7], ctx->ev[0], ctx->ev[7]);
  ctx->st.ms_prepass = pre;
  ctx->st.ms_tree = ms[0];
  ctx->st.ms_upward = ms[1];
  ctx->st.ms_traverse = ms[2];
  ctx->st.ms_lists = ms[3];
  ctx->st.ms_m2l = ms[4];
  ctx->st.ms_p2p = ms[5];
  ctx->st.ms_downward = ms[6];
  ctx->st.ms_total = ms[7];
  long long pairs = 0;
  (void)pairs;
  ctx->st.n_m2l = accepted;
  return 0;
}

// ---- master (non-jerk) variant -------------------------------------------
// get_force_and_potential of the reference's top-level sources (master
// integrator.h:459-562).  The tree is the master one: getBounds there has no
// root shift (master build_tree.h), so the shift is zero here.
int master_eval_device(tc_context *ctx, int n, const double *x, const double *v,
                       const double *m, const unsigned char *flags, int mode, double dt,
                       double *pot, double *acc, double *tstep, double *acc_old,
                       const int *ix) {
  hipStream_t s = ctx->stream;
  std::memset(&ctx->st, 0, sizeof(ctx->st));
  std::memset(ctx->m2l_by_order, 0, sizeof(ctx->m2l_by_order));
  ctx->has_acc_old = false;
  ctx->last_n = n;
  const bool steps = mode & TC_MASTER_STEPS, pair_mass = mode & TC_MASTER_PAIR_MASS;
  if (n <= 0 || !x || !v || !m || !pot || !acc) {
    ctx->err = "n > 0 and x, v, m, pot, acc are required";
    return -1;
  }
  if (mode & ~(TC_MASTER_STEPS | TC_MASTER_FMM | TC_MASTER_PAIR_MASS)) {
    ctx->err = "unknown mode bits";
    return -1;
  }
  if (steps && (!tstep || !(dt > 0))) {
    ctx->err = "TC_MASTER_STEPS needs a timestep array and dt_param > 0";
    return -1;
  }
  if (int e = ensure_bodies(ctx, n)) return e;
  TC_CHECK(hipEventRecord(ctx->ev[0], s));
  // the -DFMM build runs direct() below ncrit bodies (integrator.h:462-465)
  const bool fmm = (mode & TC_MASTER_FMM) && n >= ctx->cfg.ncrit;
  const bool small = steps && !fmm && n <= std::min(ctx->small_tree_max, kSmallTree);
  if (ix && !small) {
    ctx->err = "an index list needs the one-launch tree (time steps, n <= its limit)";
    return -1;
  }
  if (steps || fmm) {
    // the master tree has no root shift (the one-launch tree takes none)
    const double zero[3] = {0, 0, 0};
    std::memcpy(ctx->shift, zero, sizeof(zero));
    if (!small) {
      std::memcpy(ctx->h_pin + 16, zero, sizeof(zero));
      TC_CHECK(hipMemcpyAsync(ctx->dshift.p, ctx->h_pin + 16, sizeof(zero),
                              hipMemcpyHostToDevice, s));
    }
  }
  TC_CHECK(ctx->mlend.ensure((size_t)n));
  if (!steps && !fmm) {
    // direct() without time steps: the sums over leaf pairs are the all-pairs
    // sums, and the leaves do not enter them, so no tree is built
    TC_CHECK(ctx->dtg.ensure(sizeof(int) * (size_t)n));
    launch_iota(ctx->dtg.as<int>(), n, s);
    launch_gather(x, v, m, flags, ctx->dtg.as<int>(), n, bodies_of(ctx), s);
    launch_zero_u8(ctx->mlend.as<unsigned char>(), n, s);
  } else if (small) {
    // the time-step sums of a small system: its whole tree in one launch
    TC_CHECK(ctx->small_ci.ensure(sizeof(int) * 4 * (size_t)n));
    TC_CHECK(ctx->small_cg.ensure(sizeof(double) * 8 * (size_t)n));
    TC_CHECK(ctx->small_out.ensure(sizeof(int) * 4));
    TC_CHECK(ctx->small_lidx.ensure(sizeof(int) * (size_t)n));
    TC_CHECK(ctx->small_pmx.ensure(sizeof(double) * (size_t)n));
    if (ctx->small_order.bytes < sizeof(unsigned short) * kSmallTree) {
      TC_CHECK(ctx->small_order.ensure(sizeof(unsigned short) * kSmallTree));
      ctx->small_order_n = -1;
    }
    launch_small_tree(x, v, m, flags, n, ctx->cfg.ncrit, bodies_of(ctx),
                      ctx->mlend.as<unsigned char>(), ctx->small_lidx.as<int>(),
                      ctx->small_pmx.as<double>(), ctx->small_ci.as<int>(),
                      ctx->small_cg.as<double>(), ctx->small_out.as<int>(),
                      ctx->defer_small_check ? ctx->small_sticky.as<int>() : nullptr, ix,
                      ctx->small_order.as<unsigned short>(), ctx->small_order_n == n, s);
    ctx->small_order_n = n;
  } else {
    if (int e = sort_and_build(ctx, n, x, v, m, flags)) return e;
    if (fmm) return master_fmm_device(ctx, n, steps, pair_mass, dt, pot, acc, tstep, acc_old);
    launch_leaf_ends(ctx->leaf_ids.as<int>(), ctx->nleaves, ctx->cbody.as<int>(),
                     ctx->cnbody.as<int>(), n, ctx->mlend.as<unsigned char>(), s);
  }
  if (steps) {
    // leaf-cutting slabs where the one-launch tree gives the leaf data
    const int nslab = master_direct_steps_slabs(n, small);
    TC_CHECK(ctx->mdpart.ensure(sizeof(double) * 9 * (size_t)nslab * n));
    TC_CHECK(hipEventRecord(ctx->ev[1], s));
    launch_master_direct_steps(bodies_of(ctx), ctx->mlend.as<unsigned char>(),
                               small ? ctx->small_lidx.as<int>() : nullptr,
                               small ? ctx->small_pmx.as<double>() : nullptr,
                               small ? ctx->small_out.as<int>() + 2 : nullptr, n, pair_mass, dt,
                               pot, acc, tstep, nslab, ctx->mdpart.as<double>(), s);
  } else {
    const int nslab = master_direct_slabs(n, n);
    if (nslab > 1) TC_CHECK(ctx->mdpart.ensure(sizeof(double) * 4 * (size_t)nslab * n));
    TC_CHECK(hipEventRecord(ctx->ev[1], s));
    launch_master_direct(bodies_of(ctx), n, pot, acc, nslab, ctx->mdpart.as<double>(), s);
  }
  TC_CHECK(hipEventRecord(ctx->ev[2], s));
  if (small && ctx->defer_small_check) {
    // the step loop goes on without a wait; a tree too deep shows in
    // small_sticky at its next split
    ctx->st.n_p2p_pairs = (long long)n * n;
    return 0;
  }
  if (small)
    TC_CHECK(hipMemcpyAsync(ctx->h_pin + 136, ctx->small_out.p, sizeof(int) * 2,
                            hipMemcpyDeviceToHost, s));
  TC_CHECK(spin_wait(ctx, s));
  if (small) {
    if (ctx->h_pin[137] == 2) {
      ctx->err = "tree deeper than 64 levels (more than ncrit coincident bodies?)";
      return -1;
    }
    ctx->nleaves = ctx->h_pin[136];
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->err = std::string("kernel launch: ") + hipGetErrorString(e);
    return -3;
  }
  float ms[3];
  hipEventElapsedTime(&ms[0], ctx->ev[0], ctx->ev[1]);
  hipEventElapsedTime(&ms[1], ctx->ev[1], ctx->ev[2]);
  hipEventElapsedTime(&ms[2], ctx->ev[0], ctx->ev[2]);
  ctx->st.ms_tree = ms[0];
  ctx->st.ms_p2p = ms[1];
  ctx->st.ms_total = ms[2];
  ctx->st.n_p2p_cells = (!steps && !fmm) ? 1 : (long long)ctx->nleaves * ctx->nleaves;
  ctx->st.n_p2p_pairs = (long long)n * n;
  return 0;
}

// ------------------------------------------------------------------ C ABI ---
int eval_soa_device(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                    const unsigned char *flags, double *pot, double *acc, double *jerk) {
  return eval_device(ctx, n, x, v, m, flags, pot, acc, jerk);
}

int master_direct_prepare(tc_context *ctx, int n, int ns) {
  if (int e = ensure_bodies(ctx, n)) return e;
  TC_CHECK(ctx->mlend.ensure((size_t)n));
  const int nslab = std::max(master_direct_slabs(n, n), master_direct_slabs(ns, n));
  if (nslab > 1) TC_CHECK(ctx->mdpart.ensure(sizeof(double) * 4 * (size_t)nslab * n));
  return 0;
}

bool master_direct_queue_kick(tc_context *ctx, const int *is, int ns, const int *js, int nj,
                              bool sinks_are_sources, const double *pos, const double *acc,
                              double shift_fac, const double *mass, double dt, bool gradient,
                              bool set_pot, double *vel, double *pot) {
  const int n = ns + nj;
  return launch_master_direct_kick(is, ns, js, nj, sinks_are_sources, pos, acc, shift_fac, mass,
                                   master_direct_slabs(ns, n), ctx->mdpart.as<double>(), dt,
                                   gradient, set_pot, vel, pot, ctx->stream) == 0;
}

void master_direct_queue(tc_context *ctx, int n, double *pot, double *acc) {
  launch_master_direct(bodies_of(ctx), n, pot, acc, master_direct_slabs(n, n),
                       ctx->mdpart.as<double>(), ctx->stream);
}

extern "C" {

int tc_supported_orders(int *orders, int maxn) {
  int all[] = {
#define TC_LIST(P) P,
      TC_FOR_EACH_P(TC_LIST)
#undef TC_LIST
  };
  int n = (int)(sizeof(all) / sizeof(all[0]));
  for (int i = 0; i < n && i < maxn; i++) orders[i] = all[i];
  return n;
}

static bool order_supported(int p) {
  int o[64];
  int n = tc_supported_orders(o, 64);
  for (int i = 0; i < n; i++)
    if (o[i] == p) return true;
  return false;
}

int tc_create(const tc_config *cfg, void *stream, tc_context **out) {
  if (!cfg || !out) return -1;
  if (!order_supported(cfg->p) || cfg->ncrit < 1 || !(cfg->theta > 0)) return -1;
  if (cfg->mac != TC_MAC_THETA && cfg->mac != TC_MAC_ERROR) return -1;
  if (cfg->mac == TC_MAC_ERROR && !(cfg->force_accuracy > 0)) return -1;
  tc_context *ctx = new tc_context();
  ctx->cfg = *cfg;
  if (stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreate(&ctx->stream) != hipSuccess) {
      delete ctx;
      return -3;
    }
    ctx->own_stream = true;
  }
  for (auto &e : ctx->ev) hipEventCreate(&e);
  hipEventCreateWithFlags(&ctx->ev_wait, hipEventDisableTiming);
  if (hipStreamCreateWithFlags(&ctx->near_stream, hipStreamNonBlocking) != hipSuccess) {
    ctx->near_stream = nullptr;
    tc_destroy(ctx);
    return -3;
  }
  if (hipHostMalloc((void **)&ctx->h_pin, 256 * sizeof(int)) != hipSuccess ||
      hipHostGetDevicePointer((void **)&ctx->h_pin_dev, ctx->h_pin, 0) != hipSuccess) {
    tc_destroy(ctx);
    return -3;
  }
  // the no-allocation entry point's fixed scratch: its address on this device
  unsigned char *na = nullptr;
  if (noalloc_scratch(&na) != hipSuccess) {
    tc_destroy(ctx);
    return -3;
  }
  *out = ctx;
  return 0;
}

int tc_destroy(tc_context *ctx) {
  if (!ctx) return -1;
  hipStreamSynchronize(ctx->stream);
  DBuf *all[] = {&ctx->bx, &ctx->by, &ctx->bz, &ctx->bvx, &ctx->bvy, &ctx->bvz, &ctx->bq,
                 &ctx->bflags, &ctx->border, &ctx->bacc, &ctx->rpot, &ctx->rax, &ctx->ray,
                 &ctx->raz, &ctx->rjx, &ctx->rjy, &ctx->rjz, &ctx->key_hi, &ctx->key_a,
                 &ctx->key_b, &ctx->val_a, &ctx->val_b, &ctx->sort_tmp, &ctx->part, &ctx->tp,
                 &ctx->dshift, &ctx->clevel, &ctx->cnchild, &ctx->cchild, &ctx->cbody,
                 &ctx->cnbody, &ctx->cXg, &ctx->cRg, &ctx->cX, &ctx->cR, &ctx->cmin, &ctx->csink,
                 &ctx->cM, &ctx->cMd, &ctx->cL, &ctx->cLd, &ctx->split_offs, &ctx->split_nch,
                 &ctx->split_scan, &ctx->scan_tmp, &ctx->leaf_ids, &ctx->front_a, &ctx->front_b,
                 &ctx->m2l, &ctx->p2p, &ctx->tcnt, &ctx->tblk, &ctx->tscan, &ctx->tstate, &ctx->tcls, &ctx->tree_state,
                 &ctx->gsort1[0], &ctx->gsort1[1], &ctx->gsort1[2], &ctx->gsort1[3], &ctx->gsort2[0], &ctx->gsort2[1], &ctx->gsort2[2], &ctx->gsort2[3], &ctx->m2l_cnt, &ctx->m2l_off, &ctx->m2l_cur,
                 &ctx->m2l_src, &ctx->m2l_coff, &ctx->m2l_chunks, &ctx->p2p_cnt, &ctx->p2p_off,
                 &ctx->p2p_cur, &ctx->p2p_src, &ctx->dtg, &ctx->dpot, &ctx->dacc, &ctx->djerk,
                 &ctx->dx3, &ctx->dv3, &ctx->dm, &ctx->dfl, &ctx->hop, &ctx->hoa, &ctx->hoj, &ctx->mlend, &ctx->mts, &ctx->mts_sorted,
                 &ctx->msym_cnt, &ctx->msym_off, &ctx->msym_tmp, &ctx->m2l_sym, &ctx->mdpart,
                 &ctx->fpot, &ctx->fax, &ctx->fay, &ctx->faz, &ctx->fjx, &ctx->fjy, &ctx->fjz,
                 &ctx->p2p_ranges, &ctx->npairs, &ctx->p2p_items, &ctx->lowM, &ctx->lowL,
                 &ctx->lowF, &ctx->cPn, &ctx->accout, &ctx->lowP, &ctx->ev_all, &ctx->ev_ux,
                 &ctx->ev_uv, &ctx->ev_um, &ctx->ev_ufl, &ctx->ev_upot, &ctx->ev_uacc,
                 &ctx->ev_uts, &ctx->ev_uao, &ctx->ev_tmp, &ctx->ev_flags, &ctx->ev_offs,
                 &ctx->ev_scan_tmp, &ctx->mlp_params, &ctx->cq, &ctx->ctheta, &ctx->probe_m2l,
                 &ctx->probe_p2p, &ctx->leaf_feat, &ctx->leaf_pred, &ctx->chunk_part,
                 &ctx->chunk_hist, &ctx->mlp_Y, &ctx->mlp_H, &ctx->mlp_D, &ctx->mlp_dY,
                 &ctx->chunk_sort[0], &ctx->chunk_sort[1], &ctx->chunk_sort[2],
                 &ctx->chunk_sort[3], &ctx->chunk_sort[4], &ctx->small_ci, &ctx->small_cg,
                 &ctx->small_out, &ctx->small_lidx, &ctx->small_pmx, &ctx->small_sticky,
                 &ctx->small_order};
  for (DBuf *b : all) b->release();
  for (DBuf &b : ctx->ev_lev) b.release();
  ctx->h_leaves.release();
  ctx->h_nbd.release();
  ctx->h_tree.release();
  ctx->h_pcnt.release();
  ctx->h_items.release();
  ctx->h_stage.release();
  for (auto &e : ctx->ev) hipEventDestroy(e);
  hipEventDestroy(ctx->ev_wait);
  if (ctx->h_pin) hipHostFree(ctx->h_pin);
  if (ctx->near_stream) {
    hipStreamSynchronize(ctx->near_stream);
    hipStreamDestroy(ctx->near_stream);
  }
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int tc_set_seed(tc_context *ctx, long seed) {
  if (!ctx) return -1;
  ctx->rng.seed(seed);
  ctx->fixed_shift = false;
  return 0;
}

int tc_set_shift(tc_context *ctx, const double frac[3]) {
  if (!ctx || !frac) return -1;
  for (int d = 0; d < 3; d++) {
    if (!(frac[d] >= -0.5 && frac[d] < 0.5)) return -1;
    ctx->shift[d] = frac[d];
  }
  ctx->fixed_shift = true;
  return 0;
}

int tc_set_precision(tc_context *ctx, int p2p_mode, double close_radius) {
  if (!ctx) return -1;
  if (p2p_mode != TC_P2P_FP64 && p2p_mode != TC_P2P_MIXED && p2p_mode != TC_P2P_MIXED_HILO) {
    ctx->err = "unknown P2P precision mode";
    return -1;
  }
  if (p2p_mode != TC_P2P_FP64 && !(close_radius >= 0 && close_radius < 1e18)) {
    ctx->err = "close_radius must be >= 0";
    return -1;
  }
  ctx->p2p_mode = p2p_mode;
  ctx->close_radius = p2p_mode == TC_P2P_FP64 ? 0.0 : close_radius;
  return 0;
}

int tc_set_p2p_shape(tc_context *ctx, int tile, int list_slice, double stage_rounds,
                     int targets_per_lane) {
  if (!ctx) return -1;
  if (!p2p_tile_supported(tile)) {
    ctx->err = "P2P tile must be 64, 128 or 192 bodies";
    return -1;
  }
  if (list_slice < 1 || list_slice > 0xffff) {
    ctx->err = "P2P list slice must be in 1..65535";
    return -1;
  }
  if (!(stage_rounds >= 0.0 && stage_rounds <= 1e6)) {
    ctx->err = "P2P staging cost must be in [0, 1e6]";
    return -1;
  }
  if (targets_per_lane != 1 && targets_per_lane != 2) {
    ctx->err = "P2P targets per lane must be 1 or 2";
    return -1;
  }
  ctx->p2p_tile = tile;
  ctx->p2p_list = list_slice;
  ctx->p2p_stage = stage_rounds;
  ctx->p2p_per_lane = targets_per_lane;
  return 0;
}

int tc_set_small_tree(tc_context *ctx, int max_bodies) {
  if (!ctx) return -1;
  if (max_bodies < 0) {
    ctx->err = "max_bodies must be >= 0";
    return -1;
  }
  ctx->small_tree_max = std::min(max_bodies, kSmallTree);
  return 0;
}

int tc_get_p2p_shape(tc_context *ctx, int *tile, int *list_slice, double *stage_rounds,
                     int *targets_per_lane) {
  if (!ctx) return -1;
  if (tile) *tile = ctx->p2p_tile;
  if (list_slice) *list_slice = ctx->p2p_list;
  if (stage_rounds) *stage_rounds = ctx->p2p_stage;
  if (targets_per_lane) *targets_per_lane = ctx->p2p_per_lane;
  return 0;
}

// Re-run the last evaluation's near field under every launch shape of a small
// grid (targets per lane x tile x list slice x staging cost, in that nesting
// order), each `reps` times, timed with events on the near-field stream; keep
// the fastest.  The lists, ranges and host copies of the list lengths are
// those of the last evaluation; the launches add into the context's own
// near-field result set (zeroed by the next evaluation), never into the
// caller's arrays.  The mixed-precision forms run one target per lane, so
// their grid's two halves time the same kernels.
int tc_tune_p2p(tc_context *ctx, int reps, double *best_ms, double *default_ms,
                double *grid_ms) {
  if (!ctx) return -1;
  if (!ctx->p2p_ready) {
    ctx->err = "tc_tune_p2p needs a completed tree evaluation on this context";
    return -1;
  }
  if (reps < 1 || reps > 100) {
    ctx->err = "reps must be in 1..100";
    return -1;
  }
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  TC_CHECK(hipStreamSynchronize(ctx->near_stream));
  const int per_lanes[] = {1, 2};
  const int tiles[] = {64, 128, 192};
  const int lists[] = {64, 128, 256};
  const double stages[] = {0.0, 2.0, 4.0, 8.0};
  const int t0 = ctx->p2p_tile, l0 = ctx->p2p_list, r0 = ctx->p2p_per_lane;
  const double g0 = ctx->p2p_stage;
  hipStream_t s2 = ctx->near_stream;
  auto run = [&](int per_lane, int tile, int list, double stage, float *ms) -> int {
    ctx->p2p_per_lane = per_lane;
    ctx->p2p_tile = tile;
    ctx->p2p_list = list;
    ctx->p2p_stage = stage;
    if (int e = build_p2p_items(ctx)) return e;
    TC_CHECK(hipStreamSynchronize(ctx->stream));  // the item upload
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
      TC_CHECK(hipEventRecord(ctx->ev[9], s2));
      launch_p2p(cells_of(ctx), bodies_of(ctx), results_of(ctx), ctx->p2p_items.as<int4>(),
                 (int)ctx->n_items, ctx->p2p_off.as<int>(), ctx->p2p_cnt.as<int>(),
                 ctx->p2p_ranges.as<int2>(), ctx->npairs.as<unsigned long long>(), s2,
                 ctx->p2p_mode, ctx->close_radius, tile,
                 p2p_targets_per_lane(ctx->p2p_mode, per_lane));
      TC_CHECK(hipEventRecord(ctx->ev[10], s2));
      TC_CHECK(hipEventSynchronize(ctx->ev[10]));
      float t = 0;
      TC_CHECK(hipEventElapsedTime(&t, ctx->ev[9], ctx->ev[10]));
      best = std::min(best, t);
    }
    *ms = best;
    return 0;
  };
  float tdef = 0;
  if (int e = run(r0, t0, l0, g0, &tdef)) return e;
  float tbest = tdef;
  int br = r0, bt = t0, bl = l0;
  double bg = g0;
  int k = 0;
  for (int per_lane : per_lanes)
    for (int tile : tiles)
      for (int list : lists)
        for (double stage : stages) {
          float t = 0;
          if (int e = run(per_lane, tile, list, stage, &t)) return e;
          if (grid_ms) grid_ms[k] = t;
          k++;
          if (t < tbest) {
            tbest = t;
            br = per_lane;
            bt = tile;
            bl = list;
            bg = stage;
          }
        }
  ctx->p2p_per_lane = br;
  ctx->p2p_tile = bt;
  ctx->p2p_list = bl;
  ctx->p2p_stage = bg;
  if (best_ms) *best_ms = tbest;
  if (default_ms) *default_ms = tdef;
  return 0;
}

int tc_set_adaptive(tc_context *ctx, int mode, int q, double theta, double probe_theta) {
  if (!ctx) return -1;
  if (mode < 0 || mode > 2 || !(probe_theta > 0 && probe_theta < 2)) {
    ctx->err = "tc_set_adaptive: mode 0, 1 or 2 and 0 < probe_theta < 2";
    return -1;
  }
  if (mode == 1 && (q < 2 || q > ctx->cfg.p || !(theta > 0 && theta < 2))) {
    ctx->err = "tc_set_adaptive: uniform order must be in [2, p] and 0 < theta < 2";
    return -1;
  }
  if (mode == 2 && !ctx->have_model) {
    ctx->err = "tc_set_adaptive: no order model (tc_set_order_model)";
    return -1;
  }
  if (mode != 0 && ctx->cfg.mac != TC_MAC_THETA) {
    ctx->err = "tc_set_adaptive: per-cell orders need the theta MAC";
    return -1;
  }
  ctx->adaptive = mode;
  ctx->ad_q = q;
  ctx->ad_theta = theta;
  ctx->probe_theta = probe_theta;
  return 0;
}

int tc_set_order_model(tc_context *ctx, const float *params, int nparams, const int *q,
                       const float *theta, const float *cost, int ncand, const float *feat_lo,
                       const float *feat_hi, double target_rms) {
  if (!ctx) return -1;
  if (!params || nparams != kMlpParams || !q || !theta || !cost || ncand < 1 ||
      ncand > kMaxCand || !(target_rms > 0) || !feat_lo || !feat_hi) {
    ctx->err = "tc_set_order_model: bad arguments";
    return -1;
  }
  OrderGrid g{};
  g.ncand = ncand;
  for (int i = 0; i < kLeafFeatures; i++) {
    if (!(feat_lo[i] <= feat_hi[i])) {
      ctx->err = "tc_set_order_model: feature range with lo > hi";
      return -1;
    }
    g.lo[i] = feat_lo[i];
    g.hi[i] = feat_hi[i];
  }
  for (int c = 0; c < ncand; c++) {
    if (q[c] < 2 || q[c] > ctx->cfg.p || !(theta[c] > 0 && theta[c] < 2)) {
      ctx->err = "tc_set_order_model: candidate order outside [2, p] or angle outside (0, 2)";
      return -1;
    }
    g.q[c] = q[c];
    g.theta[c] = theta[c];
    g.cost[c] = cost[c];
  }
  TC_CHECK(ctx->mlp_params.ensure(sizeof(float) * kMlpParams));
  TC_CHECK(hipMemcpy(ctx->mlp_params.p, params, sizeof(float) * kMlpParams,
                     hipMemcpyHostToDevice));
  ctx->grid = g;
  ctx->log_target = (float)std::log10(target_rms);
  ctx->have_model = true;
  return 0;
}

int tc_leaf_count(const tc_context *ctx) { return ctx ? ctx->nleaves : -1; }

int tc_leaf_data(tc_context *ctx, float *features, int *leaf_of_body, float *pred) {
  if (!ctx) return -1;
  if (ctx->adaptive == 0 || ctx->last_n <= 0) {
    ctx->err = "tc_leaf_data: the last evaluation did not choose per-cell orders";
    return -2;
  }
  hipStream_t s = ctx->stream;
  const int nl = ctx->nleaves;
  if (features)
    TC_CHECK(hipMemcpyAsync(features, ctx->leaf_feat.p, sizeof(float) * kLeafFeatures * nl,
                            hipMemcpyDeviceToDevice, s));
  if (pred && ctx->adaptive == 2)
    TC_CHECK(hipMemcpyAsync(pred, ctx->leaf_pred.p, sizeof(float) * nl, hipMemcpyDeviceToDevice,
                            s));
  if (leaf_of_body)
    launch_leaf_of_body(cells_of(ctx), bodies_of(ctx), ctx->leaf_ids.as<int>(), nl, leaf_of_body,
                        s);
  TC_CHECK(hipStreamSynchronize(s));
  return 0;
}

int tc_leaf_errors(tc_context *ctx, const double *acc, const double *acc_ref, float *out) {
  if (!ctx || !acc || !acc_ref || !out) return -1;
  if (ctx->last_n <= 0 || ctx->nleaves <= 0) {
    ctx->err = "tc_leaf_errors: no tree (evaluate first)";
    return -2;
  }
  launch_leaf_errors(cells_of(ctx), bodies_of(ctx), ctx->leaf_ids.as<int>(), ctx->nleaves, acc,
                     acc_ref, out, ctx->stream);
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tc_m2l_order_counts(const tc_context *ctx, long long *counts, int maxp) {
  if (!ctx || !counts) return -1;
  for (int q = 0; q <= maxp; q++) counts[q] = q <= kMaxP ? ctx->m2l_by_order[q] : 0;
  return kMaxP;
}

int tc_mlp_param_count(void) { return kMlpParams; }
int tc_mlp_input_count(void) { return kMlpIn; }

int tc_mlp_forward(tc_context *ctx, const float *params, const float *X, int B, float *Y) {
  if (!ctx || !params || (B > 0 && (!X || !Y)) || B < 0) return -1;
  launch_mlp_fwd(params, X, B, Y, nullptr, ctx->stream);
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tc_mlp_grad(tc_context *ctx, const float *params, const float *X, const float *T, int B,
                float tau, float *G, float *loss) {
  if (!ctx || !params || !G || !loss || B < 0 || (B > 0 && (!X || !T)) || !(tau > 0) ||
      !(tau < 1))
    return -1;
  if (B > ctx->mlp_cap) {
    const size_t b = (size_t)B;
    TC_CHECK(ctx->mlp_Y.ensure(sizeof(float) * b));
    TC_CHECK(ctx->mlp_dY.ensure(sizeof(float) * b));
    TC_CHECK(ctx->mlp_H.ensure(sizeof(float) * 3 * kMlpH * b));
    TC_CHECK(ctx->mlp_D.ensure(sizeof(float) * 3 * kMlpH * b));
    ctx->mlp_cap = B;
  }
  launch_mlp_grad(params, X, T, B, tau, ctx->mlp_Y.as<float>(), ctx->mlp_H.as<float>(),
                  ctx->mlp_D.as<float>(), ctx->mlp_dY.as<float>(), G, loss, ctx->stream);
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tc_mlp_adam(tc_context *ctx, float *params, const float *G, float *m, float *v, float lr,
                int step) {
  if (!ctx || !params || !G || !m || !v || step < 1 || !(lr > 0)) return -1;
  launch_adam(params, G, m, v, lr, step, ctx->stream);
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tc_last_shift(const tc_context *ctx, double frac[3]) {
  if (!ctx || !frac) return -1;
  for (int d = 0; d < 3; d++) frac[d] = ctx->shift[d];
  return 0;
}

int tc_eval_soa(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                const unsigned char *flags, double *pot, double *acc, double *jerk) {
  if (!ctx) return -1;
  if (!x || !v || !m) {
    ctx->err = "x, v and m are required";
    return -1;
  }
  return eval_device(ctx, n, x, v, m, flags, pot, acc, jerk);
}

int tc_eval_clusters(tc_context *ctx, int ncl, const int *starts, const double *x,
                     const double *v, const double *m, const unsigned char *flags, double *pot,
                     double *acc, double *jerk) {
  if (!ctx) return -1;
  if (ncl < 1 || !starts || !x || !v || !m || starts[0] != 0) {
    ctx->err = "tc_eval_clusters: ncl >= 1, starts[0] == 0, x, v and m required";
    return -1;
  }
  for (int k = 0; k < ncl; k++)
    if (starts[k + 1] <= starts[k]) {
      ctx->err = "tc_eval_clusters: every cluster needs at least one body (starts increasing)";
      return -1;
    }
  if (ctx->adaptive != 0) {
    ctx->err = "tc_eval_clusters: per-cell orders are not available for clusters";
    return -1;
  }
  const int n = starts[ncl];
  if (ncl == 1) return eval_device(ctx, n, x, v, m, flags, pot, acc, jerk);
  TC_CHECK(ctx->cl_start.ensure(sizeof(int) * (size_t)(ncl + 1)));
  TC_CHECK(ctx->cl_id.ensure(sizeof(int) * (size_t)n));
  TC_CHECK(ctx->tp.ensure(sizeof(TreeParams) * (size_t)ncl));
  TC_CHECK(ctx->dshift.ensure(sizeof(double) * 3 * (size_t)ncl));
  TC_CHECK(hipMemcpyAsync(ctx->cl_start.p, starts, sizeof(int) * (ncl + 1),
                          hipMemcpyHostToDevice, ctx->stream));
  ctx->ncl = ncl;
  const int e = eval_device(ctx, n, x, v, m, flags, pot, acc, jerk);
  ctx->ncl = 1;
  return e;
}

int tc_eval_bodies(tc_context *ctx, tc_body *bodies, int n) {
  if (!ctx) return -1;
  if (!bodies || n <= 0) {
    ctx->err = "bodies must be non-empty";
    return -1;
  }
  hipStream_t s = ctx->stream;
  const size_t N = (size_t)n;
  // page-locked staging: inputs x[3n] v[3n] m[n] flags[n bytes]; outputs
  // pot[n] acc[3n] jerk[3n] acc_old[n] reuse the same block afterwards
  TC_CHECK(ctx->h_stage.ensure(8 * N + (N + 7) / 8));
  double *hx = ctx->h_stage.p, *hv = hx + 3 * N, *hm = hv + 3 * N;
  unsigned char *hf = (unsigned char *)(hm + N);
  host_parallel(n, [&](int lo, int hi) {
    for (int i = lo; i < hi; i++) {
      const tc_body &b = bodies[i];
      hx[3 * i] = b.X[0];
      hx[3 * i + 1] = b.X[1];
      hx[3 * i + 2] = b.X[2];
      hv[3 * i] = b.V[0];
      hv[3 * i + 1] = b.V[1];
      hv[3 * i + 2] = b.V[2];
      hm[i] = b.q;
      hf[i] = (b.issink ? kSink : 0) | (b.issource ? kSource : 0);
    }
  });
  TC_CHECK(ctx->dx3.ensure(24 * N));
  TC_CHECK(ctx->dv3.ensure(24 * N));
  TC_CHECK(ctx->dm.ensure(8 * N));
  TC_CHECK(ctx->dfl.ensure(N));
  TC_CHECK(ctx->hop.ensure(8 * N));
  TC_CHECK(ctx->hoa.ensure(24 * N));
  TC_CHECK(ctx->hoj.ensure(24 * N));
  TC_CHECK(hipMemcpyAsync(ctx->dx3.p, hx, 24 * N, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(ctx->dv3.p, hv, 24 * N, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(ctx->dm.p, hm, 8 * N, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(ctx->dfl.p, hf, N, hipMemcpyHostToDevice, s));
  // the all-pairs branch of eval_device uses dpot/dacc/djerk as its own
  // scratch, so the results land in separate buffers
  double *op = ctx->hop.as<double>(), *oa = ctx->hoa.as<double>(), *oj = ctx->hoj.as<double>();
  int e = eval_device(ctx, n, ctx->dx3.as<double>(), ctx->dv3.as<double>(), ctx->dm.as<double>(),
                      ctx->dfl.as<unsigned char>(), op, oa, oj);
  if (e != 0) return e;
  double *rp = ctx->h_stage.p, *ra = rp + N, *rj = ra + 3 * N, *ro = rj + 3 * N;
  TC_CHECK(hipMemcpyAsync(rp, op, 8 * N, hipMemcpyDeviceToHost, s));
  TC_CHECK(hipMemcpyAsync(ra, oa, 24 * N, hipMemcpyDeviceToHost, s));
  TC_CHECK(hipMemcpyAsync(rj, oj, 24 * N, hipMemcpyDeviceToHost, s));
  // acc_old is set by the pre-pass of the error-MAC path (integrator.h:350-356)
  const bool want_acc = ctx->has_acc_old;
  if (want_acc) {
    // op has been copied out above (same stream), so it can hold acc_old
    if ((e = copy_acc_old(ctx, n, op)) != 0) return e;
    TC_CHECK(hipMemcpyAsync(ro, op, 8 * N, hipMemcpyDeviceToHost, s));
  }
  TC_CHECK(hipStreamSynchronize(s));
  host_parallel(n, [&](int lo, int hi) {
    for (int i = lo; i < hi; i++) {
      tc_body &b = bodies[i];
      b.p = rp[i];
      b.F[0] = ra[3 * i];
      b.F[1] = ra[3 * i + 1];
      b.F[2] = ra[3 * i + 2];
      b.J[0] = rj[3 * i];
      b.J[1] = rj[3 * i + 1];
      b.J[2] = rj[3 * i + 2];
      if (want_acc) b.acc_old = ro[i];
    }
  });
  return 0;
}

int tc_master_eval_soa(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                       const unsigned char *flags, int mode, double dt_param, double *pot,
                       double *acc, double *timestep, double *acc_old) {
  if (!ctx) return -1;
  return master_eval_device(ctx, n, x, v, m, flags, mode, dt_param, pot, acc, timestep, acc_old);
}

int tc_master_eval_bodies(tc_context *ctx, tc_master_body *bodies, int n, int mode,
                          double dt_param) {
  if (!ctx) return -1;
  if (!bodies || n <= 0) {
    ctx->err = "bodies must be non-empty";
    return -1;
  }
  hipStream_t s = ctx->stream;
  const size_t N = (size_t)n;
  const bool steps = mode & TC_MASTER_STEPS;
  // the -DFMM build's pre-pass adds to Body.acc_old (above ncrit bodies)
  const bool fmm = (mode & TC_MASTER_FMM) && n >= ctx->cfg.ncrit;
  // page-locked staging: x[3n] v[3n] m[n] acc_old[n] flags[n bytes]; the
  // outputs pot[n] acc[3n] timestep[n] acc_old[n] reuse it afterwards
  TC_CHECK(ctx->h_stage.ensure(9 * N));
  double *hx = ctx->h_stage.p, *hv = hx + 3 * N, *hm = hv + 3 * N, *ha = hm + N;
  unsigned char *hf = (unsigned char *)(ha + N);
  host_parallel(n, [&](int lo, int hi) {
    for (int i = lo; i < hi; i++) {
      const tc_master_body &b = bodies[i];
      for (int d = 0; d < 3; d++) {
        hx[3 * i + d] = b.X[d];
        hv[3 * i + d] = b.V[d];
      }
      hm[i] = b.q;
      ha[i] = b.acc_old;
      hf[i] = (b.issink ? kSink : 0) | (b.issource ? kSource : 0);
    }
  });
  TC_CHECK(ctx->dx3.ensure(24 * N));
  TC_CHECK(ctx->dv3.ensure(24 * N));
  TC_CHECK(ctx->dm.ensure(8 * N));
  TC_CHECK(ctx->dfl.ensure(N));
  TC_CHECK(ctx->hop.ensure(8 * N));
  TC_CHECK(ctx->hoa.ensure(24 * N));
  TC_CHECK(ctx->hoj.ensure(8 * N));
  TC_CHECK(ctx->mts.ensure(8 * N));
  TC_CHECK(hipMemcpyAsync(ctx->dx3.p, hx, 24 * N, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(ctx->dv3.p, hv, 24 * N, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(ctx->dm.p, hm, 8 * N, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(ctx->dfl.p, hf, N, hipMemcpyHostToDevice, s));
  double *op = ctx->hop.as<double>(), *oa = ctx->hoa.as<double>(), *ot = ctx->mts.as<double>();
  double *oo = ctx->hoj.as<double>();
  if (fmm) TC_CHECK(hipMemcpyAsync(oo, ha, 8 * N, hipMemcpyHostToDevice, s));
  int e = master_eval_device(ctx, n, ctx->dx3.as<double>(), ctx->dv3.as<double>(),
                             ctx->dm.as<double>(), ctx->dfl.as<unsigned char>(), mode, dt_param,
                             op, oa, ot, fmm ? oo : nullptr);
  if (e != 0) return e;
  double *rp = ctx->h_stage.p, *ra = rp + N, *rt = ra + 3 * N, *ro = rt + N;
  TC_CHECK(hipMemcpyAsync(rp, op, 8 * N, hipMemcpyDeviceToHost, s));
  TC_CHECK(hipMemcpyAsync(ra, oa, 24 * N, hipMemcpyDeviceToHost, s));
  if (steps) TC_CHECK(hipMemcpyAsync(rt, ot, 8 * N, hipMemcpyDeviceToHost, s));
  if (fmm) TC_CHECK(hipMemcpyAsync(ro, oo, 8 * N, hipMemcpyDeviceToHost, s));
  TC_CHECK(hipStreamSynchronize(s));
  host_parallel(n, [&](int lo, int hi) {
    for (int i = lo; i < hi; i++) {
      tc_master_body &b = bodies[i];
      b.p = rp[i];
      for (int d = 0; d < 3; d++) b.F[d] = ra[3 * i + d];
      if (steps) b.timestep = rt[i];
      if (fmm) b.acc_old = ro[i];
    }
  });
  return 0;
}

int tc_direct_soa(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                  const unsigned char *flags, int nt, const int *targets, double *pot,
                  double *acc, double *jerk) {
  if (!ctx) return -1;
  if (n <= 0 || nt < 0 || !x || !v || !m || (nt > 0 && !targets)) {
    ctx->err = "bad arguments";
    return -1;
  }
  if (int e = ensure_bodies(ctx, n)) return e;
  hipStream_t s = ctx->stream;
  TC_CHECK(ctx->dtg.ensure(sizeof(int) * (size_t)n));
  launch_iota(ctx->dtg.as<int>(), n, s);
  launch_gather(x, v, m, flags, ctx->dtg.as<int>(), n, bodies_of(ctx), s);
  launch_direct(bodies_of(ctx), n, targets, nt, pot, acc, jerk, s);
  TC_CHECK(hipStreamSynchronize(s));
  TC_CHECK(hipGetLastError());
  return 0;
}

int tc_get_acc_old(tc_context *ctx, double *acc_old) {
  if (!ctx) return -1;
  if (!acc_old) {
    ctx->err = "acc_old must be a device pointer";
    return -1;
  }
  if (int e = copy_acc_old(ctx, ctx->last_n, acc_old)) return e;
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tc_get_stats(const tc_context *ctx, tc_stats *out) {
  if (!ctx || !out) return -1;
  *out = ctx->st;
  return 0;
}

const char *tc_last_error(const tc_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"
