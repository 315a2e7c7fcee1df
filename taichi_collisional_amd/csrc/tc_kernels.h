// Kernel launch interfaces shared by the tcfmm host engine and the HIP
// translation units.  Device layout (all device pointers, sorted body order):
//   bodies  SoA x,y,z,vx,vy,vz,q (double), flags (kSink|kSource), order (orig idx)
//   cells   SoA level, nchild, child (first child), body (first body), nbody,
//           Xg[3], Rg (tree geometry), X[3], R (bounding-ball geometry),
//           has_sink, min_acc; coefficients M, Md, L, Ld: [cell][NT] doubles
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcfmm {

constexpr int kMaxP = 20;
constexpr int kMaxNT = (kMaxP + 1) * (kMaxP + 1);

// Expansion orders with compiled kernels.
#define TC_FOR_EACH_P(X) X(2) X(4) X(6) X(8) X(10) X(12) X(14) X(16)

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
int launch_morton(const double *x3, int n, const TreeParams *d_tp, uint64_t *key_hi,
                  uint64_t *key_lo, int *idx, hipStream_t s);
size_t sort_temp_bytes(int n);
int launch_sort(uint64_t *key_a, uint64_t *key_b, int *val_a, int *val_b, const uint64_t *key_hi,
                int n, void *temp, size_t temp_bytes, hipStream_t s);
int launch_gather(const double *x3, const double *v3, const double *q,
                  const unsigned char *flags_in, const int *perm, int n, BodiesDev b,
                  hipStream_t s);
int launch_root(CellsDev c, int n, const TreeParams *d_tp, hipStream_t s);
int launch_split(CellsDev c, BodiesDev b, int ls, int le, int ncrit, int *offs, int *nch,
                 hipStream_t s);
int launch_children(CellsDev c, int ls, int le, const int *offs, const int *nch_scan,
                    const TreeParams *d_tp, hipStream_t s);
size_t scan_temp_bytes(int n);
int launch_scan(const int *in, int *out, int n, void *temp, size_t temp_bytes, hipStream_t s);

// ---- traversal (tc_tree.hip)
struct TravArgs {
  const int2 *fin;
  int nin;
  int2 *fout;
  int *cnt;  // [1] M2L pairs so far, [2] P2P pairs so far (device copy)
  int2 *m2l, *p2p;
  double theta;
};
int launch_trav_count(CellsDev c, const int2 *fin, int nin, double theta, int *bm2l, int *bp2p,
                      int *bsplit, hipStream_t s);
int launch_trav_write(CellsDev c, const TravArgs &a, const int *om2l, const int *op2p,
                      const int *osplit, hipStream_t s);
size_t pair_sort_temp_bytes(int n);
int launch_group_pairs(const int2 *pairs, int n, int key_bits, int ncells, int *tg_a, int *tg_b,
                       int *src_a, int *src, int *start, int *count, void *temp,
                       size_t temp_bytes, hipStream_t s);
int launch_make_chunks(const int *offs, const int *count, const int *chunk_offs, int ncells,
                       int4 *chunks, hipStream_t s);

// ---- upward (tc_upward.hip)
int launch_p2m(int p, CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, hipStream_t s);
int launch_m2m(int p, CellsDev c, BodiesDev b, int ls, int le, hipStream_t s);

// ---- M2L (tc_m2l.hip)
struct M2LArgs {
  const int4 *chunks;
  int nchunks;
  const int *src;
  const double *cX;
  const unsigned char *has_sink;
  const double *M, *Md;
  double *L, *Ld;
  int quirk;
};
int launch_m2l(int p, const M2LArgs &a, hipStream_t s);

// ---- P2P (tc_p2p.hip)
int launch_p2p(CellsDev c, BodiesDev b, ResultsDev r, const int4 *groups, int ngroups,
               const int *offs, const int *count, const int2 *ranges,
               unsigned long long *npairs, hipStream_t s);
int launch_p2p_ranges(CellsDev c, const int *src, int n, int2 *ranges, hipStream_t s);
int launch_direct(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                  double *jerk, hipStream_t s);

// ---- downward (tc_downward.hip)
int launch_l2l(int p, CellsDev c, int ls, int le, hipStream_t s);
int launch_l2p(int p, CellsDev c, BodiesDev b, ResultsDev r, const int *leaf_ids, int nleaves,
               hipStream_t s);
int launch_scatter_results(BodiesDev b, ResultsDev r, int n, double *pot, double *acc3,
                           double *jerk3, hipStream_t s);

}  // namespace tcfmm
