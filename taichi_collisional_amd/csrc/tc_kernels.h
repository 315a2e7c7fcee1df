// Kernel launch interfaces shared by the tcfmm host engine and the HIP
// translation units.  Device layout (all device pointers, sorted body order):
//   bodies  SoA x,y,z,vx,vy,vz,q (double), flags (kSink|kSource), order (orig idx)
//   cells   SoA level, nchild, child (first child), body (first body), nbody,
//           Xg[3], Rg (tree geometry), X[3], R (bounding-ball geometry),
//           has_sink, min_acc; coefficients M, Md, L, Ld: [cell][NT] doubles
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace tcfmm {

constexpr int kMaxP = 24;
constexpr int kMaxNT = (kMaxP + 1) * (kMaxP + 1);

// Expansion orders with compiled kernels.
#define TC_FOR_EACH_P(X)                                                                     \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) \
      X(19) X(20) X(21) X(22) X(23) X(24)

// Order shards.  tc_m2l.hip, tc_upward.hip and tc_downward.hip instantiate
// their kernels for every order; each is compiled TC_NSHARDS times, with
// -DTC_SHARD=k, and copy k holds the orders p with p % TC_NSHARDS == k inside
// namespace tcfmm::sh<k> (the copies build in parallel).  tc_dispatch.hip
// routes the launch_* calls of those units (declared below) to the copy that
// holds p.
#ifndef TC_NSHARDS
#define TC_NSHARDS 3
#endif
#define TC_SHARD_NS_(k) sh##k
#define TC_SHARD_NS(k) TC_SHARD_NS_(k)
constexpr bool in_shard(int p, int k) { return p % TC_NSHARDS == k; }

template <int... Ns>
struct IntSeq {};
template <int N, int... Ns>
struct MakeIntSeqImpl : MakeIntSeqImpl<N - 1, N - 1, Ns...> {};
template <int... Ns>
struct MakeIntSeqImpl<0, Ns...> {
  using type = IntSeq<Ns...>;
};
template <int N>
using MakeIntSeq = typename MakeIntSeqImpl<N>::type;

struct BodiesDev {
  double *x, *y, *z, *vx, *vy, *vz, *q;
  unsigned char *flags;
  int *order;
  double *acc_old;
};

struct CellsDev {
  int *level, *nchild, *child, *body, *nbody;
  double *Xg, *Rg, *X, *R, *min_acc;
  unsigned char *has_sink;
  double *M, *Md, *L, *Ld;
};

struct ResultsDev {
  double *pot, *ax, *ay, *az, *jx, *jy, *jz;
};

struct TreeParams {
  double X0[3], R0;
  double mn[3], mx[3];
};

// ---- tree (tc_tree.hip)
int launch_bounds(const double *x3, int n, double *partial, int nblocks_max, TreeParams *d_tp,
                  const double *shift_frac_dev, hipStream_t s);
// cid: cluster of every body (d_tp indexed by it), or null for one tree
int launch_morton(const double *x3, int n, const TreeParams *d_tp, uint64_t *key_hi,
                  uint64_t *key_lo, int *idx, hipStream_t s, const int *cid = nullptr);
// forest of clusters (bodies grouped by cluster, ncl + 1 starts)
int launch_bounds_clusters(const double *x3, int ncl, const int *start, TreeParams *d_tp,
                           const double *shift_frac_dev, hipStream_t s);
int launch_cluster_ids(const int *start, int ncl, int n, int *cid, hipStream_t s);
int launch_sort_clusters(uint64_t *key_a, uint64_t *key_b, int *val_a, int *val_b,
                         const int *cid, int ncl, int n, void *temp, size_t temp_bytes,
                         hipStream_t s);
size_t sort_temp_bytes(int n);
int launch_sort(uint64_t *key_a, uint64_t *key_b, int *val_a, int *val_b, const uint64_t *key_hi,
                int n, void *temp, size_t temp_bytes, hipStream_t s);
int launch_gather(const double *x3, const double *v3, const double *q,
                  const unsigned char *flags_in, const int *perm, int n, BodiesDev b,
                  hipStream_t s);
int launch_root(CellsDev c, int n, const TreeParams *d_tp, hipStream_t s,
                const int *start = nullptr, int ncl = 1);
// Device-side state of a tree build (ints): the level being split [ls, le),
// leaves and levels so far, the level being placed (cls, cle, its first leaf
// slot cbase, go), done, overflow (1: the cell arrays are too small, need
// cells; 2: deeper than kMaxTreeLevels) and level_start[0..nlev].
constexpr int kMaxTreeLevels = 64;
struct TreeState {
  int ls, le, nleaf, nlev;
  int cls, cle, cbase, go;
  int done, overflow, need, pad;
  int level_start[kMaxTreeLevels + 4];
};
// One level: split (offs, nch), scans (scan), children and leaves.  Arrays of
// per-level cells hold cap entries (nch and scan: 2 x cap).
int launch_tree_level(CellsDev c, BodiesDev b, TreeState *st, int ncrit, int cap, int *offs,
                      int *nch, int *scan, const TreeParams *d_tp, int *leaf_ids, hipStream_t s);
// The whole tree of at most kSmallTree bodies in one launch (one workgroup),
// for the master time-step sums: bodies gathered into tree order, lend set on
// the last body of every leaf; out[0] = leaves, out[1] = 0 or 2 (too deep),
// out[2] = 1 when two bodies are (next to) coincident; lidx / pmx (or null):
// each body's leaf's last index and the mass before it in its leaf; sticky
// (or null) is set to 2 with out[1] = 2 and otherwise left as it is; with
// ix, body i is record ix[i] of x3, v3, q, flags.  order (n ushorts, or
// null) receives the bodies' sorted order by key prefix and index; with
// use_order it holds the previous call's (any permutation of 0..n-1), which
// is kept when still sorted instead of running the sort.
// ci: 4 n ints, cg: 8 n doubles of scratch.
constexpr int kSmallTree = 4096;
int launch_small_tree(const double *x3, const double *v3, const double *q,
                      const unsigned char *flags, int n, int ncrit, BodiesDev b,
                      unsigned char *lend, int *lidx, double *pmx, int *ci, double *cg, int *out,
                      int *sticky, const int *ix, unsigned short *order, bool use_order,
                      hipStream_t s);
size_t scan_temp_bytes(int n);
int launch_scan(const int *in, int *out, int n, void *temp, size_t temp_bytes, hipStream_t s);

// ---- traversal (tc_tree.hip)
// Multipole acceptance: kMacTheta  |dX|^2 theta^2 > (Ri+Rj)^2 on the balls
// (traverse_eager.h:125-156); kMacLow  |dXg|^2 > 3 (Rgi+Rgj)^2 on the tree
// geometry (:172-212); kMacError  the error estimate of horizontalPassHigh
// (:59-122) from Pn of the source and min_acc of the target; kMacMaster  the
// symmetric estimate of the master sources' horizontalPassHigh (master
// traverse_eager.h:61-149): both directions below 1, accepted pairs of tiny
// cells to P2P, and with `steps` accepted far pairs dropped.
enum { kMacTheta = 0, kMacLow = 1, kMacError = 2, kMacMaster = 3 };
struct MacArgs {
  int mode, p;
  int steps;                // kMacMaster: the get_steps walk
  double theta, facc;
  const double *Pn;         // [cell][p+1], kMacError
  const double *theta_cell; // kMacTheta: the target cell's own angle (null: theta)
  double comb[kMaxP + 1];   // C(p,k), kMacError
};
struct FactTable {
  double fact[2 * kMaxP + 1];  // the reference's factorial_table (quirk applied)
};
struct TravArgs {
  const int2 *fin;  // this step's frontier
  int2 *fout;       // the next step's
  int2 *m2l, *p2p;  // the lists, appended to
  unsigned char *cls;  // per frontier pair: kind << 4 | child pairs (pass 1 -> pass 2)
  MacArgs mac;
};
// Device-side state of a walk (ints): frontier sizes nin[step parity], list
// lengths so far, this step's list offsets, whether pass 2 writes, steps
// taken, and on overflow what the failing step needed.
struct TravState {
  int nin[2];
  int nm2l, np2p, base_m2l, base_p2p;
  int write, overflow, steps;
  int need_m2l, need_p2p, need_front;
};
constexpr int kTravBlocks = 2048;  // blocks of a walk step; blk holds 6 x kTravBlocks ints
// One step of the walk (count, offsets, write) on the frontier of parity par;
// caps are the capacities of the lists and of the next frontier (pairs).
int launch_trav_step(CellsDev c, const TravArgs &a, TravState *st, int par, int *blk,
                     long long cap_m2l, long long cap_p2p, long long cap_front, hipStream_t s);
size_t pair_sort_temp_bytes(int n);
int launch_group_pairs(const int2 *pairs, int n, int key_bits, int ncells, int *tg_a, int *tg_b,
                       int *src_a, int *src, int *start, int *count, void *temp,
                       size_t temp_bytes, hipStream_t s);
int launch_make_chunks(const int *offs, const int *count, const int *chunk_offs, int ncells,
                       int4 *chunks, hipStream_t s);
// chunks stably sorted by q[target] (orders < 32); scratch key/idx arrays of n ints
int launch_sort_chunks_by_order(const int4 *chunks, int n, const int *q, int *key_a, int *key_b,
                                int *idx_a, int *idx_b, int4 *out, void *temp, size_t temp_bytes,
                                hipStream_t s);

// ---- upward (tc_upward.hip)
// what: kUpGeom (ball X, R, has_sink), kUpMin (min_acc from acc_old),
// kUpMoments (M, Mdot about the X already there); all by default.  Launched
// combinations: all, kUpGeom | kUpMin, kUpMoments, kUpGeom | kUpMoments, kUpMin.
enum { kUpGeom = 1, kUpMoments = 2, kUpMin = 4, kUpAll = 7 };
int launch_p2m(int p, CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, hipStream_t s,
               int what = kUpAll);
int launch_m2m(int p, CellsDev c, BodiesDev b, int ls, int le, hipStream_t s,
               int what = kUpAll);
int launch_m2m_ids(int p, CellsDev c, BodiesDev b, const int *ids, int nids, hipStream_t s);

// ---- M2L (tc_m2l.hip)
struct M2LArgs {
  const int4 *chunks;
  int nchunks;
  const int *src;
  const double *cX;
  const unsigned char *has_sink;
  const double *M, *Md;
  double *L, *Ld;
  int quirk;       // drop M_p^{+-p} (the reference's factorial_table[2p] == 0)
  int stride = 0;  // doubles per cell row of M, Md, L, Ld; 0: (p+1)^2
};
// p: the truncation order of the listed pairs (<= the order of the rows)
int launch_m2l(int p, const M2LArgs &a, hipStream_t s);

// ---- P2P (tc_p2p.hip)
// mode: 0 FP64 (the reference's DOUBLE_P2P form); 1 mixed: FP32 pairs on
// offsets from the target group's first body, FP64 for pairs closer than
// close_radius; 2 mixed with the offsets held as float hi + lo pairs
// tile: source bodies per LDS tile, one of 64, 128, 192 (p2p_tile_supported);
// per_lane: targets per lane, 1 or 2 (2 for the FP64 form only, see
// p2p_targets_per_lane); the groups' sizes must follow the same value
int launch_p2p(CellsDev c, BodiesDev b, ResultsDev r, const int4 *groups, int ngroups,
               const int *offs, const int *count, const int2 *ranges,
               unsigned long long *npairs, hipStream_t s, int mode = 0,
               double close_radius = 0.0, int tile = 128, int per_lane = 1);
bool p2p_tile_supported(int tile);
// targets per lane the P2P kernel of this precision mode runs for a request
int p2p_targets_per_lane(int mode, int requested);

// How the P2P kernel's target groups cut a leaf of nb bodies (host side):
// sizes(nb) lists the group sizes (<= 64 per_lane each, summing to nb) of
// least cost under the kernel's lane model; the table grows on demand.
class P2PSplit {
 public:
  void configure(double stage_rounds, int fill, int per_lane);
  void grow(int nmax);
  int count(int nb) const { return first_[nb + 1] - first_[nb]; }
  const int *sizes(int nb) const { return sizes_.data() + first_[nb]; }

 private:
  double stage_ = -1.0;
  int fill_ = -1, per_lane_ = 1;
  std::vector<int> first_{0}, sizes_, choice_{0};
  std::vector<double> cost_{0.0};
};
int launch_p2p_ranges(CellsDev c, const int *src, int n, int2 *ranges, hipStream_t s);
int launch_direct(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                  double *jerk, hipStream_t s);
int launch_direct_sum(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                      double *jerk, hipStream_t s);
void launch_zero_f64(double *p, size_t n, hipStream_t s);

// ---- p = 2 pre-pass and multipole norms (tc_low.hip)
int launch_low_upward(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                      const std::vector<int> &level_start, double *lowM, hipStream_t s);
int launch_low_horizontal(CellsDev c, BodiesDev b, int nc, const int *leaf_ids, int nleaves,
                          const int *m2l_off, const int *m2l_cnt, const int *m2l_src,
                          const int *p2p_off, const int *p2p_cnt, const int *p2p_src,
                          const double *lowM, double *lowL, double *F, int n, float4 *P,
                          hipStream_t s);  // P: n float records for P2P_low (nleaves > 0)
int launch_low_downward(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                        const std::vector<int> &level_start, double *lowL, double *F, int n,
                        hipStream_t s);
// master pre-pass variants (master kernel_fast_lean.h:1479-1556, 1609-1646):
// F[3i] collects the near terms, F[3i+1] the far-field magnitude; acc_old =
// ain (caller order, may be null) + both
int launch_low_p2p_master(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                          const int *p2p_off, const int *p2p_cnt, const int *p2p_src, double *F,
                          hipStream_t s);
int launch_low_downward_master(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                               const std::vector<int> &level_start, double *lowL, double *F,
                               int n, const double *ain, hipStream_t s);
int launch_pn(const double *M, int nc, int p, const FactTable &ft, double *Pn, hipStream_t s);
int launch_scatter_acc_old(BodiesDev b, int n, double *out, hipStream_t s);

// ---- downward (tc_downward.hip)
int launch_l2l(int p, CellsDev c, int ls, int le, hipStream_t s);
int launch_l2l_ids(int p, CellsDev c, const int *ids, int nids, hipStream_t s);
int launch_l2p(int p, CellsDev c, BodiesDev b, ResultsDev r, const int *leaf_ids, int nleaves,
               hipStream_t s);
// master (non-jerk) variant, tc_master.hip
void launch_iota(int *p, int n, hipStream_t s);
void launch_zero_u8(unsigned char *p, int n, hipStream_t s);
// ---- learned per-cell order and opening angle (tc_learn.hip)
constexpr int kLeafFeatures = 9, kMlpIn = kLeafFeatures + 2, kMlpH = 64;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + 2 * (kMlpH * kMlpH + kMlpH) + kMlpH + 1;
constexpr int kMaxCand = 64;
// Candidate (order, angle) pairs of the choice and their relative costs.
struct OrderGrid {
  int ncand;
  int q[kMaxCand];
  float theta[kMaxCand];
  float cost[kMaxCand];
  // each leaf feature clamped to the range seen in training before the MLP
  float lo[kLeafFeatures], hi[kLeafFeatures];
};
void launch_mlp_fwd(const float *P, const float *X, int B, float *Y, float *H, hipStream_t s);
// loss = mean w (Y - T)^2 with w = tau where Y < T and 1 - tau elsewhere
// (tau = 0.5: half the squared error) and its gradient G (zeroed first);
// Y [B], H and D [3][B][64], dY [B] are scratch
void launch_mlp_grad(const float *P, const float *X, const float *T, int B, float tau, float *Y,
                     float *H, float *D, float *dY, float *G, float *loss, hipStream_t s);
void launch_adam(float *P, const float *G, float *m, float *v, float lr, int t, hipStream_t s);
void launch_count_targets(const int2 *pairs, int n, int *cnt, hipStream_t s);
void launch_leaf_features(CellsDev c, const int *leaf_ids, int nleaves, int p,
                          const int *m2l_cnt, const int *p2p_cnt, double mean_mass, float *F,
                          hipStream_t s);
void launch_leaf_select(const float *P, const float *F, const int *leaf_ids, int nleaves,
                        const OrderGrid &g, float log_target, int *q, double *theta, float *pred,
                        hipStream_t s);
void launch_leaf_uniform(const int *leaf_ids, int nleaves, int qv, double tv, int *q,
                         double *theta, hipStream_t s);
void launch_propagate(CellsDev c, int ls, int le, int *q, double *theta, hipStream_t s);
void launch_leaf_errors(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,
                        const double *acc, const double *ref, float *out, hipStream_t s);
void launch_leaf_of_body(CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, int *out,
                         hipStream_t s);

// ---- fixed scratch of tc_direct_soa_noalloc (tc_noalloc.hip)
hipError_t noalloc_scratch(unsigned char **out);
int noalloc_max_bodies();
void launch_leaf_ends(const int *leaf_ids, int nleaves, const int *cbody, const int *cnbody,
                      int n, unsigned char *lend, hipStream_t s);
// The force-only direct sum of the step loop's kicks, read straight from the
// particle arrays, with the kick applied in the slab-sum kernel: the bodies
// are the union of the sinks is[0..ns) (sources too with sinks_are_sources)
// and the sources js[0..nj), positions pos + acc_in * shift_fac (pos alone
// for shift_fac = 0); vel[is[k]] += F_k dt (2/3 of it for gradient kicks),
// pot[is[k]] = p_k with set_pot.  part: 4 * nslab * (ns + nj) doubles.
// Returns -1 (nothing launched) when nslab < 2.
int launch_master_direct_kick(const int *is, int ns, const int *js, int nj,
                              bool sinks_are_sources, const double *pos, const double *acc_in,
                              double shift_fac, const double *mass, int nslab, double *part,
                              double dt, bool gradient, bool set_pot, double *vel, double *pot,
                              hipStream_t s);
// The direct sum with the time-step estimate, in slabs that may cut source
// leaves when the one-launch small tree gave lidx (each body's leaf's last
// index) and pmx (the mass of the leaf's bodies before it), and (vector form)
// flag[0] == 0 (no zero distance between distinct bodies); otherwise in
// leaf-aligned slabs.  part: 9 * nslab * n doubles.  Writes pot, acc, tstep
// in the caller's order (zero for non-sinks).
int master_direct_steps_slabs(int n, bool uniform);
void launch_master_direct_steps(const BodiesDev &b, const unsigned char *lend, const int *lidx,
                                const double *pmx, const int *flag, int n, bool pair_mass,
                                double dt, double *pot, double *acc, double *tstep, int nslab,
                                double *part, hipStream_t s);
// Source slabs of the force-only direct sum for nt targets among n bodies
// (1: no partial sums).
int master_direct_slabs(int nt, int n);
// The force-only direct sum (P2P_simple over every pair); part: 4 * nslab * n
// doubles of partial sums when nslab > 1.
void launch_master_direct(const BodiesDev &b, int n, double *pot, double *acc, int nslab,
                          double *part, hipStream_t s);
// symmetric M2L list: for every accepted (i, j) with i < j and a sink in
// either cell, (i, j) and (j, i); cnt/off hold npairs + 1 ints
void launch_sym_count(const int2 *pairs, int npairs, const unsigned char *has_sink, int *cnt,
                      hipStream_t s);
void launch_sym_write(const int2 *pairs, int npairs, const unsigned char *has_sink,
                      const int *off, int2 *out, hipStream_t s);
// P2P_simple / P2P of the master sources over a (target, source) cell list
void launch_master_p2p(const CellsDev &c, const BodiesDev &b, const int2 *pairs, int npairs,
                       bool steps, bool pair_mass, double dt, ResultsDev r, double *ts_sorted,
                       hipStream_t s);
void launch_scatter_ts(const BodiesDev &b, const double *ts_sorted, int n, double *out,
                       hipStream_t s);

int launch_scatter_results(BodiesDev b, ResultsDev r, int n, double *pot, double *acc3,
                           double *jerk3, hipStream_t s);
int launch_scatter_results_sum(BodiesDev b, ResultsDev r, ResultsDev f, int n, double *pot,
                               double *acc3, double *jerk3, hipStream_t s);

}  // namespace tcfmm
