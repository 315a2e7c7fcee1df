// Device octree build and dual-tree traversal (FP64 geometry), gfx950.
//
// Tree: replaces buildTree (taichi_with_jerks/build_tree.h:454-487).
//   bounds      getBounds (build_tree.h:239-276): min/max reduction, centre,
//               root shift = (drand48()-0.5)*boxsize/25 per axis with the three
//               uniform values supplied by the host, R0 *= 1.00001
//   keys        28 bits per axis, floor((x-(X0-R0))/(2 R0) * 2^28)
//               (build_tree.h:459-470); interleaved z>y>x per level, which is
//               the order of Body::operator< (exafmm.h:31-40)
//   sort        stable LSD radix sort (low 21 key bits, then high 63)
//   cells       level-synchronous top-down split of every cell with more than
//               ncrit bodies into its non-empty octants, octant boundaries
//               found exactly as buildCellsSorted does (build_tree.h:360-451):
//               a count for <= 32 bodies, else a binary search on x >= centre
//               (which always keeps the first body in the first range)
// Traversal: replaces horizontalPass (traverse_eager.h:125-156): the
//   recursion from (root, root) is run breadth-first over a frontier of cell
//   pairs; each pair is accepted for M2L by the theta-MAC, sent to P2P when
//   both are leaves, or split (Ci if Cj is a leaf or Ci is larger, else Cj).
//   Each step runs in two passes (count per block, scan, write at the scanned
//   offsets), so the lists are written without atomics and in a fixed order;
//   they are then grouped by target cell with a stable radix sort, and the
//   M2L list is cut into chunks of <= 64 sources.
#include <hipcub/hipcub.hpp>

#include "tc_common.h"
#include "tc_kernels.h"

namespace tcfmm {

// ---------------------------------------------------------------- bounds ----
__global__ __launch_bounds__(256) void k_bounds_partial(const double *__restrict__ x3, int n,
                                                        double *__restrict__ part) {
  __shared__ double smn[3][256], smx[3][256];
  int t = threadIdx.x;
  double mn[3], mx[3];
  for (int d = 0; d < 3; d++) mn[d] = mx[d] = x3[d];
  for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256)
    for (int d = 0; d < 3; d++) {
      double v = x3[3 * (size_t)i + d];
      mn[d] = fmin(v, mn[d]);
      mx[d] = fmax(v, mx[d]);
    }
  for (int d = 0; d < 3; d++) smn[d][t] = mn[d], smx[d][t] = mx[d];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int d = 0; d < 3; d++) {
        smn[d][t] = fmin(smn[d][t], smn[d][t + o]);
        smx[d][t] = fmax(smx[d][t], smx[d][t + o]);
      }
    __syncthreads();
  }
  if (t == 0)
    for (int d = 0; d < 3; d++) {
      part[6 * blockIdx.x + d] = smn[d][0];
      part[6 * blockIdx.x + 3 + d] = smx[d][0];
    }
}

// One wave: lane-strided min / max over the block partials, then a wave
// reduction (min and max do not depend on the order); lane 0 finishes.
__global__ __launch_bounds__(64) void k_bounds_final(const double *__restrict__ part, int nb,
                                                     const double *__restrict__ shift_frac,
                                                     TreeParams *tp) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x;
  double mn[3], mx[3];
  for (int d = 0; d < 3; d++) mn[d] = part[d], mx[d] = part[3 + d];
  for (int b = lane; b < nb; b += kWave)
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(part[6 * b + d], mn[d]);
      mx[d] = fmax(part[6 * b + 3 + d], mx[d]);
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(mn[d], __shfl_xor(mn[d], o, kWave));
      mx[d] = fmax(mx[d], __shfl_xor(mx[d], o, kWave));
    }
  if (lane != 0) return;
  double X0[3];
  for (int d = 0; d < 3; d++) X0[d] = (mx[d] + mn[d]) / 2;
  double box = -1;
  for (int d = 0; d < 3; d++)
    if (mx[d] - mn[d] > box) box = mx[d] - mn[d];
  for (int d = 0; d < 3; d++) {
    double sh = shift_frac[d] * box / 25;
    X0[d] += sh;
  }
  double R0 = 0;
  for (int d = 0; d < 3; d++) {
    R0 = fmax(X0[d] - mn[d], R0);
    R0 = fmax(mx[d] - X0[d], R0);
  }
  R0 *= 1.00001;
  for (int d = 0; d < 3; d++) {
    tp->X0[d] = X0[d];
    tp->mn[d] = mn[d];
    tp->mx[d] = mx[d];
  }
  tp->R0 = R0;
}

int launch_bounds(const double *x3, int n, double *partial, int nblocks_max, TreeParams *d_tp,
                  const double *shift_frac_dev, hipStream_t s) {
  int nb = (n + 255) / 256;
  if (nb > nblocks_max) nb = nblocks_max;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(256), 0, s, x3, n, partial);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, partial, nb, shift_frac_dev, d_tp);
  return 0;
}

// Forest of independent clusters (bodies grouped by cluster, cluster k =
// [start[k], start[k+1])): one block per cluster reduces its bounds and forms
// its TreeParams as k_bounds_final does for a single root, with the cluster's
// own shift values shift_frac[3k..3k+2].
__global__ __launch_bounds__(256) void k_bounds_clusters(const double *__restrict__ x3,
                                                         const int *__restrict__ start,
                                                         const double *__restrict__ shift_frac,
                                                         TreeParams *tp) {
#pragma clang fp contract(off)
  __shared__ double smn[3][256], smx[3][256];
  const int k = blockIdx.x, t = threadIdx.x;
  const int b0 = start[k], b1 = start[k + 1];
  double mn[3], mx[3];
  for (int d = 0; d < 3; d++) mn[d] = mx[d] = x3[3 * (size_t)b0 + d];
  for (int i = b0 + t; i < b1; i += 256)
    for (int d = 0; d < 3; d++) {
      const double v = x3[3 * (size_t)i + d];
      mn[d] = fmin(v, mn[d]);
      mx[d] = fmax(v, mx[d]);
    }
  for (int d = 0; d < 3; d++) smn[d][t] = mn[d], smx[d][t] = mx[d];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int d = 0; d < 3; d++) {
        smn[d][t] = fmin(smn[d][t], smn[d][t + o]);
        smx[d][t] = fmax(smx[d][t], smx[d][t + o]);
      }
    __syncthreads();
  }
  if (t != 0) return;
  double X0[3], box = -1, R0 = 0;
  for (int d = 0; d < 3; d++) {
    mn[d] = smn[d][0];
    mx[d] = smx[d][0];
    X0[d] = (mx[d] + mn[d]) / 2;
    if (mx[d] - mn[d] > box) box = mx[d] - mn[d];
  }
  for (int d = 0; d < 3; d++) X0[d] += shift_frac[3 * k + d] * box / 25;
  for (int d = 0; d < 3; d++) {
    R0 = fmax(X0[d] - mn[d], R0);
    R0 = fmax(mx[d] - X0[d], R0);
  }
  R0 *= 1.00001;
  for (int d = 0; d < 3; d++) {
    tp[k].X0[d] = X0[d];
    tp[k].mn[d] = mn[d];
    tp[k].mx[d] = mx[d];
  }
  tp[k].R0 = R0;
}

int launch_bounds_clusters(const double *x3, int ncl, const int *start, TreeParams *d_tp,
                           const double *shift_frac_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds_clusters, dim3(ncl), dim3(256), 0, s, x3, start, shift_frac_dev,
                     d_tp);
  return 0;
}

// Cluster of every body (input order) from the cluster starts.
__global__ void k_cluster_ids(const int *__restrict__ start, int ncl, int n,
                              int *__restrict__ cid) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = ncl - 1;  // last cluster with start <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  cid[i] = lo;
}

int launch_cluster_ids(const int *start, int ncl, int n, int *cid, hipStream_t s) {
  hipLaunchKernelGGL(k_cluster_ids, dim3((n + 255) / 256), dim3(256), 0, s, start, ncl, n, cid);
  return 0;
}

// ---------------------------------------------------------------- keys ------
__device__ __forceinline__ uint64_t spread3(uint64_t v) {  // 21 bits -> every 3rd bit
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

__global__ __launch_bounds__(256) void k_morton(const double *__restrict__ x3, int n,
                                                const TreeParams *__restrict__ tps,
                                                const int *__restrict__ cid,
                                                uint64_t *__restrict__ key_hi,
                                                uint64_t *__restrict__ key_lo,
                                                int *__restrict__ idx) {
#pragma clang fp contract(off)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const TreeParams *tp = tps + (cid ? cid[i] : 0);
  const double maxint = (double)(1u << 28);
  unsigned int M[3];
  for (int d = 0; d < 3; d++) {
    double f = (x3[3 * (size_t)i + d] - (tp->X0[d] - tp->R0)) / (2.0 * tp->R0);
    M[d] = (unsigned int)floor(f * maxint);
  }
  // hi: bits 27..7 of each axis, lo: bits 6..0; within a level z > y > x
  uint64_t hi = (spread3(M[2] >> 7) << 2) | (spread3(M[1] >> 7) << 1) | spread3(M[0] >> 7);
  uint64_t lo = (spread3(M[2] & 0x7f) << 2) | (spread3(M[1] & 0x7f) << 1) | spread3(M[0] & 0x7f);
  key_hi[i] = hi;
  key_lo[i] = lo;
  idx[i] = i;
}

int launch_morton(const double *x3, int n, const TreeParams *d_tp, uint64_t *key_hi,
                  uint64_t *key_lo, int *idx, hipStream_t s, const int *cid) {
  hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, s, x3, n, d_tp, cid, key_hi,
                     key_lo, idx);
  return 0;
}

size_t sort_temp_bytes(int n) {
  size_t b = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                     (int *)nullptr, (int *)nullptr, n, 0, 64);
  return b;
}

__global__ void k_gather_keys(const uint64_t *__restrict__ key_hi, const int *__restrict__ perm,
                              int n, uint64_t *__restrict__ out) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = key_hi[perm[i]];
}

// Forest: a last stable pass on the cluster id of the permuted bodies, so the
// order is (cluster, key); each cluster keeps its input range.
__global__ void k_gather_cid(const int *__restrict__ cid, const int *__restrict__ perm, int n,
                             uint64_t *__restrict__ out) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (uint64_t)cid[perm[i]];
}

int launch_sort_clusters(uint64_t *key_a, uint64_t *key_b, int *val_a, int *val_b,
                         const int *cid, int ncl, int n, void *temp, size_t temp_bytes,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_gather_cid, dim3((n + 255) / 256), dim3(256), 0, s, cid, val_a, n, key_a);
  int bits = 1;
  while ((1 << bits) < ncl) bits++;
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, key_a, key_b, val_a, val_b, n, 0, bits, s) !=
      hipSuccess)
    return -1;
  return hipMemcpyAsync(val_a, val_b, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s) ==
                 hipSuccess
             ? 0
             : -1;
}

// key_a holds the low keys on entry; on return val_a holds the permutation.
int launch_sort(uint64_t *key_a, uint64_t *key_b, int *val_a, int *val_b, const uint64_t *key_hi,
                int n, void *temp, size_t temp_bytes, hipStream_t s) {
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, key_a, key_b, val_a, val_b, n, 0, 21, s) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_gather_keys, dim3((n + 255) / 256), dim3(256), 0, s, key_hi, val_b, n,
                     key_a);
  tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, key_a, key_b, val_b, val_a, n, 0, 63, s) !=
      hipSuccess)
    return -1;
  return 0;
}

__global__ __launch_bounds__(256) void k_gather(const double *__restrict__ x3,
                                                const double *__restrict__ v3,
                                                const double *__restrict__ q,
                                                const unsigned char *__restrict__ flags_in,
                                                const int *__restrict__ perm, int n, BodiesDev b) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int i = perm[k];
  b.x[k] = x3[3 * (size_t)i];
  b.y[k] = x3[3 * (size_t)i + 1];
  b.z[k] = x3[3 * (size_t)i + 2];
  b.vx[k] = v3[3 * (size_t)i];
  b.vy[k] = v3[3 * (size_t)i + 1];
  b.vz[k] = v3[3 * (size_t)i + 2];
  b.q[k] = q[i];
  b.flags[k] = flags_in ? flags_in[i] : (unsigned char)(kSink | kSource);
  b.order[k] = i;
  b.acc_old[k] = 0.0;
}

int launch_gather(const double *x3, const double *v3, const double *q,
                  const unsigned char *flags_in, const int *perm, int n, BodiesDev b,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, x3, v3, q, flags_in, perm,
                     n, b);
  return 0;
}

// ---------------------------------------------------------------- cells -----
// Root cells: one (all n bodies), or one per cluster of a forest (cluster k
// = [start[k], start[k+1]) with its own TreeParams).
__global__ void k_root(CellsDev c, int n, const TreeParams *__restrict__ tp,
                       const int *__restrict__ start, int ncl) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= ncl) return;
  c.level[k] = 0;
  c.nchild[k] = 0;
  c.child[k] = -1;
  c.body[k] = start ? start[k] : 0;
  c.nbody[k] = start ? start[k + 1] - start[k] : n;
  for (int d = 0; d < 3; d++) c.Xg[3 * k + d] = tp[k].X0[d];
  c.Rg[k] = tp[k].R0;
}

int launch_root(CellsDev c, int n, const TreeParams *d_tp, hipStream_t s, const int *start,
                int ncl) {
  hipLaunchKernelGGL(k_root, dim3((ncl + 63) / 64), dim3(64), 0, s, c, n, d_tp, start, ncl);
  return 0;
}

__device__ __forceinline__ int octant_of(const BodiesDev &b, int i, double X0, double X1,
                                         double X2) {
  return (b.x[i] >= X0) + ((b.y[i] >= X1) << 1) + ((b.z[i] >= X2) << 2);
}

// One level of the tree build, queued without host reads: the level's range
// and the running counts live in TreeState.  Arrays of the level's cells
// (offs: 8 octant starts per cell; nch: child counts, and at nch + cap the
// leaf flags; scan: their exclusive scans, the same layout) hold up to cap
// cells.
//
// Split: eight lanes per cell, lane j for octant j; above 32 bodies lane j > 0
// runs the binary search for the start of octants >= j (the search never
// tests the cell's first body, as the reference's does not, so its start is
// l + 1); the seven searches run side by side instead of one after another
// (the top levels' cells hold up to all N bodies, and their splits are chains
// of dependent loads).  Up to 32 bodies lane j counts octant j's bodies.
__global__ __launch_bounds__(256) void k_split(CellsDev c, BodiesDev b,
                                               const TreeState *__restrict__ st, int ncrit,
                                               int cap, int *__restrict__ offs,
                                               int *__restrict__ nch) {
  if (st->done || st->overflow) return;
  const int ls = st->ls, nl = st->le - ls;
  const int j = threadIdx.x & 7, lane = threadIdx.x & (kWave - 1);
  for (int k = (blockIdx.x * 256 + (int)threadIdx.x) >> 3; k < nl; k += gridDim.x * 32) {
    const int ci = ls + k;
    const int nb = c.nbody[ci];
    if (nb <= ncrit) {
      if (j == 0) {
        nch[cap + k] = 1;
        nch[k] = 0;
        c.nchild[ci] = 0;
        c.child[ci] = -1;
      }
      continue;
    }
    const int begin = c.body[ci], end = begin + nb;
    const double X0 = c.Xg[3 * ci], X1 = c.Xg[3 * ci + 1], X2 = c.Xg[3 * ci + 2];
    const bool small = nb <= 32;
    int o = begin, sz = 0;
    if (small) {
      for (int i = begin; i < end; i++) sz += octant_of(b, i, X0, X1, X2) == j;
    } else if (j > 0) {
      int l = begin, r = end;
      while (l < r - 1) {
        const int mid = (l + r) / 2;
        if (octant_of(b, mid, X0, X1, X2) >= j)
          r = mid;
        else
          l = mid;
      }
      o = l + 1;
    }
    // within the group of eight: starts from counts, or counts from starts
    int incl = sz;
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
      const int t = __shfl(incl, lane - d, kWave);
      if (j >= d) incl += t;
    }
    const int onext = __shfl(o, lane + 1, kWave);
    if (small)
      o = begin + incl - sz;
    else
      sz = (j < 7 ? onext : end) - o;
    offs[8 * k + j] = o;
    const unsigned long long nz = __ballot(sz != 0);
    const int cnt = __popcll((nz >> (lane & ~7)) & 0xffull);
    if (j == 0) {
      nch[cap + k] = 0;
      nch[k] = cnt;
      c.nchild[ci] = cnt;
    }
  }
}

// Exclusive scans of the level's child counts and leaf flags (one block), and
// the state: the level the children kernel places (cls, cle, cbase, go), the
// next level, the leaves and levels so far.  A level whose children would not
// fit the cell arrays sets overflow = 1 and what it needed; a level past
// kMaxTreeLevels sets overflow = 2.
__global__ __launch_bounds__(1024) void k_level_scan(const int *__restrict__ nch,
                                                     int *__restrict__ scan, TreeState *st,
                                                     int cap) {
  __shared__ int wsum[2][16];
  __shared__ int carry[2];
  if (st->done || st->overflow) {
    if (threadIdx.x == 0) st->go = 0;
    return;
  }
  const int ls = st->ls, le = st->le, nl = le - ls;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  if (t < 2) carry[t] = 0;
  __syncthreads();
  // both arrays in one pass over tiles of 16 per thread (a 24k-cell level is
  // two tiles): each tile is a few global-load and barrier round trips
  constexpr int PER = 16, TILE = 1024 * PER;
  for (int base = 0; base < nl; base += TILE) {
    int v0[PER], v1[PER], s0 = 0, s1 = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int i = base + t * PER + q;
      v0[q] = i < nl ? nch[i] : 0;
      v1[q] = i < nl ? nch[(size_t)cap + i] : 0;
      s0 += v0[q];
      s1 += v1[q];
    }
    int wt0, wt1;
    const int ex0 = wave_excl_scan(s0, &wt0);
    const int ex1 = wave_excl_scan(s1, &wt1);
    if (lane == 0) {
      wsum[0][w] = wt0;
      wsum[1][w] = wt1;
    }
    __syncthreads();
    int r0 = carry[0], r1 = carry[1], all0 = 0, all1 = 0;
    for (int k = 0; k < 16; k++) {
      if (k < w) {
        r0 += wsum[0][k];
        r1 += wsum[1][k];
      }
      all0 += wsum[0][k];
      all1 += wsum[1][k];
    }
    r0 += ex0;
    r1 += ex1;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int i = base + t * PER + q;
      if (i < nl) {
        scan[i] = r0;
        scan[(size_t)cap + i] = r1;
      }
      r0 += v0[q];
      r1 += v1[q];
    }
    __syncthreads();
    if (t == 0) {
      carry[0] += all0;
      carry[1] += all1;
    }
    __syncthreads();
  }
  if (t == 0) {
    const int total = carry[0], leaves = carry[1];
    st->cls = ls;
    st->cle = le;
    st->cbase = st->nleaf;
    if ((long long)le + total > cap) {
      st->overflow = 1;
      st->need = (int)((long long)le + total < 0x7fffffff ? le + total : 0x7fffffff);
      st->go = 0;
    } else if (st->nlev >= kMaxTreeLevels) {
      st->overflow = 2;
      st->go = 0;
    } else {
      st->go = 1;
      st->nleaf += leaves;
      st->nlev += 1;
      st->level_start[st->nlev] = le;
      if (total == 0) {
        st->done = 1;
      } else {
        st->ls = le;
        st->le = le + total;
      }
    }
  }
}

// The children of the level the scan kernel set up (cls, cle), and its leaves
// appended to leaf_ids from cbase, in cell order.
__global__ __launch_bounds__(256) void k_children(CellsDev c, const TreeState *__restrict__ st,
                                                  int cap, const int *__restrict__ offs,
                                                  const int *__restrict__ scan,
                                                  const TreeParams *__restrict__ tp,
                                                  int *__restrict__ leaf_ids) {
#pragma clang fp contract(off)
  if (!st->go) return;
  const int ls = st->cls, le = st->cle, leaf_base = st->cbase, nl = le - ls;
  for (int k = blockIdx.x * 256 + (int)threadIdx.x; k < nl; k += gridDim.x * 256) {
    const int ci = ls + k;
    const int nc = c.nchild[ci];
    if (nc == 0) {
      leaf_ids[leaf_base + scan[cap + k]] = ci;
      continue;
    }
    const int first = le + scan[k];
    c.child[ci] = first;
    const int level = c.level[ci];
    const int end = c.body[ci] + c.nbody[ci];
    // R0 2^-(level+1), from the parent's radius (exact: a power-of-two scaling),
    // so that the roots of a forest may each have their own R0
    const double r = ldexp(c.Rg[ci], -1);
    (void)tp;
    int jc = 0;
    for (int i = 0; i < 8; i++) {
      const int o = offs[8 * k + i];
      const int e = (i < 7) ? offs[8 * k + i + 1] : end;
      const int size = e - o;
      if (!size) continue;
      const int cc = first + jc;
      for (int d = 0; d < 3; d++)
        c.Xg[3 * cc + d] = c.Xg[3 * ci + d] + r * (((i >> d) & 1) * 2 - 1);
      c.Rg[cc] = r;
      c.level[cc] = level + 1;
      c.body[cc] = o;
      c.nbody[cc] = size;
      c.nchild[cc] = 0;
      c.child[cc] = -1;
      jc++;
    }
  }
}

int launch_tree_level(CellsDev c, BodiesDev b, TreeState *st, int ncrit, int cap, int *offs,
                      int *nch, int *scan, const TreeParams *d_tp, int *leaf_ids, hipStream_t s) {
  // grids cover the widest levels in a few sweeps; narrow levels leave most
  // blocks idle for a few microseconds
  hipLaunchKernelGGL(k_split, dim3(512), dim3(256), 0, s, c, b, (const TreeState *)st, ncrit, cap,
                     offs, nch);
  hipLaunchKernelGGL(k_level_scan, dim3(1), dim3(1024), 0, s, (const int *)nch, scan, st, cap);
  hipLaunchKernelGGL(k_children, dim3(256), dim3(256), 0, s, c, (const TreeState *)st, cap,
                     (const int *)offs, (const int *)scan, d_tp, leaf_ids);
  return 0;
}

// ---------------------------------------------------------- small trees ----
// The master direct sum with the time-step estimate needs the leaves of the
// tree of its bodies (the minimum of the estimate runs per source leaf), and
// the step loop asks for it dozens of times per step on a few thousand
// bodies, where the multi-kernel build above is some forty launches and a
// host wait.  For n <= kSmallTree one workgroup builds the same tree in one
// launch, with the same arithmetic:
//  * bounds without a root shift (k_bounds_final with zero shift values);
//  * 28-bit-per-axis keys (k_morton); the order of the stable radix sort on
//    (high, low) key is that of (high key, low key, input index): a bitonic
//    sort in registers of one 64-bit value per body, the high key's top 52
//    bits (its cell at level 17) above the input index, then each run of
//    equal 52-bit prefixes (bodies closer than 2^-17 of the box; rare)
//    reordered by (high key, low key, index);
//  * the gather (k_gather);
//  * levels split as k_split does (eight lanes per cell: a count per octant up
//    to 32 bodies, else the search for the start of octants >= j that never
//    tests the cell's first body), children placed as k_children does, with
//    the positions read from LDS.
// It marks the last body of every leaf in lend (cleared first) and counts
// the leaves: out[0] = leaves, out[1] = 0, or 2 when the tree would need more
// than kMaxTreeLevels levels (the multi-kernel build's error).  With lidx /
// pmx it also gives every body its leaf's last index and the mass of the
// leaf's bodies before it, and out[2] = 1 when two bodies sit closer than
// 1e-100 in every coordinate (checked for every pair within a run of equal
// prefixes and for neighbours in tree order): then a pair other than a body
// with itself may have r^2 = 0, which the time-step sums skip.  Equal
// positions always share a key; distinct ones give r^2 = 0 only when every
// coordinate difference is below 1e-162, which needs coordinates below about
// 1e-146.  Cells live in global scratch: two level buffers of n cells (a
// level never holds more cells than bodies), begin / count in ci, centre and
// radius in cg.
__device__ __forceinline__ void small_key_of(const double *p, const double *X0, double R0,
                                             uint64_t &hi, uint64_t &lo) {
#pragma clang fp contract(off)
  const double maxint = (double)(1u << 28);
  unsigned int M[3];
  for (int d = 0; d < 3; d++) {
    const double f = (p[d] - (X0[d] - R0)) / (2.0 * R0);
    M[d] = (unsigned int)floor(f * maxint);
  }
  hi = (spread3(M[2] >> 7) << 2) | (spread3(M[1] >> 7) << 1) | spread3(M[0] >> 7);
  lo = (spread3(M[2] & 0x7f) << 2) | (spread3(M[1] & 0x7f) << 1) | spread3(M[0] & 0x7f);
}

__device__ __forceinline__ void small_key(const double *x3, int i, const double *X0, double R0,
                                          uint64_t &hi, uint64_t &lo) {
  const double p[3] = {x3[3 * (size_t)i], x3[3 * (size_t)i + 1], x3[3 * (size_t)i + 2]};
  small_key_of(p, X0, R0, hi, lo);
}

// One compare-exchange step of the bitonic network for element e holding k
// against its partner e ^ stride holding pk: in blocks of `size` that sort
// ascending ((e & size) == 0) the lower index keeps the smaller value,
// descending blocks the reverse.
__device__ __forceinline__ void bitonic_take(int e, int size, int stride, uint64_t pk,
                                             uint64_t &k) {
  const bool up = (e & size) == 0, lower = (e & stride) == 0;
  if ((pk < k) == (lower == up)) k = pk;
}

// the same for two elements a < c of one thread
__device__ __forceinline__ void bitonic_pair(int ea, int size, uint64_t &ka, uint64_t &kc) {
  const bool up = (ea & size) == 0;
  if ((ka > kc) == up) {
    const uint64_t tk = ka;
    ka = kc;
    kc = tk;
  }
}

__global__ __launch_bounds__(1024) void k_small_tree(const double *__restrict__ x3,
                                                     const double *__restrict__ v3,
                                                     const double *__restrict__ q,
                                                     const unsigned char *__restrict__ flags_in,
                                                     int n, int ncrit, BodiesDev b,
                                                     unsigned char *__restrict__ lend,
                                                     int *__restrict__ lidx,
                                                     double *__restrict__ pmx,
                                                     int *__restrict__ ci, double *__restrict__ cg,
                                                     int *__restrict__ out, int *sticky,
                                                     const int *__restrict__ ix,
                                                     unsigned short *order, int use_order) {
#pragma clang fp contract(off)
  __shared__ uint64_t skey[kSmallTree];
  __shared__ unsigned short sidx[kSmallTree];
  // positions in tree order for the level splits (136 KB of LDS in all)
  __shared__ double spx[kSmallTree], spy[kSmallTree], spz[kSmallTree];
  __shared__ double wred[16][6];
  __shared__ double sX0[3], sR0;
  __shared__ int ncur, nnext, nleaf, status, near;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  // body i is record ix[i] of x3, v3, q, flags_in (record i without ix)
  auto rec = [&](int i) { return ix ? ix[i] : i; };
  // ---- bounds (min / max do not depend on the order); thread t holds
  // elements 4t .. 4t+3 from here to the end of the sort: bodies 4t .. 4t+3,
  // or with `order` (the previous call's sorted order of the n bodies) the
  // bodies at those places of it
  double px[4][3];
  int bid[4];
  double mn[3], mx[3];
  const int r0 = rec(0);
  for (int d = 0; d < 3; d++) mn[d] = mx[d] = x3[3 * (size_t)r0 + d];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int e = 4 * t + j;
    bid[j] = e < n ? (use_order ? (int)order[e] : e) : 0;
    const int i = bid[j];
    const int ri = e < n ? rec(i) : r0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      px[j][d] = x3[3 * (size_t)ri + d];
      mn[d] = fmin(px[j][d], mn[d]);
      mx[d] = fmax(px[j][d], mx[d]);
    }
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(mn[d], __shfl_xor(mn[d], o, kWave));
      mx[d] = fmax(mx[d], __shfl_xor(mx[d], o, kWave));
    }
  if (lane == 0)
    for (int d = 0; d < 3; d++) {
      wred[w][d] = mn[d];
      wred[w][3 + d] = mx[d];
    }
  __syncthreads();
  if (t == 0) {
    for (int k = 1; k < 16; k++)
      for (int d = 0; d < 3; d++) {
        mn[d] = fmin(wred[k][d], mn[d]);
        mx[d] = fmax(wred[k][3 + d], mx[d]);
      }
    double X0[3], box = -1;
    for (int d = 0; d < 3; d++) X0[d] = (mx[d] + mn[d]) / 2;
    for (int d = 0; d < 3; d++)
      if (mx[d] - mn[d] > box) box = mx[d] - mn[d];
    for (int d = 0; d < 3; d++) X0[d] += 0.0 * box / 25;  // the master tree: no shift
    double R0 = 0;
    for (int d = 0; d < 3; d++) {
      R0 = fmax(X0[d] - mn[d], R0);
      R0 = fmax(mx[d] - X0[d], R0);
    }
    R0 *= 1.00001;
    for (int d = 0; d < 3; d++) sX0[d] = X0[d];
    sR0 = R0;
    ncur = 1;
    nnext = 0;
    nleaf = 0;
    status = 0;
    near = 0;
  }
  __syncthreads();
  const double X0[3] = {sX0[0], sX0[1], sX0[2]};
  const double R0 = sR0;
  // ---- keys and the bitonic sort, in registers, on one 64-bit value per
  // body: the high key's top 52 bits (the cell at level 17) above the
  // 12-bit input index.  The partner e ^ stride sits in the same thread
  // (stride < 4), in another lane of the wave (stride < 256: a lane
  // shuffle) or in another wave (through LDS); padding (~0) sorts last.
  int n2 = 4;
  while (n2 < n) n2 <<= 1;
  uint64_t kk[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int e = 4 * t + j;
    uint64_t hi = 0, lo;
    if (e < n) small_key_of(px[j], X0, R0, hi, lo);
    kk[j] = e < n ? ((hi >> 11) << 12) | (uint64_t)bid[j] : ~0ull;
  }
  // The values are distinct, so if the previous order still has them
  // strictly increasing it is the sorted order and the network is skipped
  // (bodies move little between the step loop's calls).
  bool sorted = false;
  if (use_order) {
    // every neighbouring pair (e, e + 1) with e + 1 < n in increasing order
    const int e0 = 4 * t;
    bool up = true;
#pragma unroll
    for (int j = 0; j < 3; j++) up = up && (e0 + j + 1 >= n || kk[j] < kk[j + 1]);
    const uint64_t nxt = __shfl_down(kk[0], 1, kWave);
    if (lane == 0) skey[w] = kk[0];  // each wave's first value
    __syncthreads();
    const uint64_t nk = lane < kWave - 1 ? nxt : (w < 15 ? skey[w + 1] : ~0ull);
    up = up && (e0 + 4 >= n || kk[3] < nk);
    sorted = __syncthreads_and(up ? 1 : 0) != 0;
  }
  if (use_order && !sorted) {
    // Nearly sorted: an element larger than everything before it and
    // smaller than everything after it is already at its sorted place; the
    // others fill the gaps between such elements, each gap with exactly the
    // values that belong there, so sorting every gap finishes the sort.
    // Gaps longer than 64 fall back to the sorting network.
    const int e0 = 4 * t;
    uint64_t tmax = 0, tmin = ~0ull;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (e0 + j < n) tmax = kk[j] > tmax ? kk[j] : tmax;
      if (e0 + j < n) tmin = kk[j] < tmin ? kk[j] : tmin;
    }
    // inclusive wave scans: max from the front, min from the back
    uint64_t pmax = tmax, smin = tmin;
    for (int d = 1; d < kWave; d <<= 1) {
      const uint64_t a = __shfl_up(pmax, d, kWave), b = __shfl_down(smin, d, kWave);
      if (lane >= d) pmax = a > pmax ? a : pmax;
      if (lane + d < kWave) smin = b < smin ? b : smin;
    }
    __syncthreads();  // the check's skey reads are done
    if (lane == kWave - 1) skey[w] = pmax;   // wave totals
    if (lane == 0) skey[16 + w] = smin;
    __syncthreads();
    uint64_t before = 0, after = ~0ull;  // over earlier / later waves
    for (int u = 0; u < w; u++) before = skey[u] > before ? skey[u] : before;
    for (int u = w + 1; u < 16; u++) after = skey[16 + u] < after ? skey[16 + u] : after;
    // exclusive over this thread's elements' predecessors / successors
    uint64_t pm = __shfl_up(pmax, 1, kWave), sm = __shfl_down(smin, 1, kWave);
    pm = lane > 0 ? (pm > before ? pm : before) : before;
    sm = lane < kWave - 1 ? (sm < after ? sm : after) : after;
    bool clean[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      uint64_t sj = sm;
      for (int q = j + 1; q < 4; q++)
        if (e0 + q < n) sj = kk[q] < sj ? kk[q] : sj;
      const int e = e0 + j;
      clean[j] = e >= n || ((e == 0 || pm < kk[j]) && (e == n - 1 || kk[j] < sj));
      if (e < n) pm = kk[j] > pm ? kk[j] : pm;
    }
    __syncthreads();  // the totals are read
#pragma unroll
    for (int j = 0; j < 4; j++) {
      skey[e0 + j] = kk[j];
      sidx[e0 + j] = clean[j] ? 0 : 1;
    }
    __syncthreads();
    bool big = false;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int a = e0 + j;
      if (clean[j] || (a > 0 && sidx[a - 1])) continue;  // not the start of a gap
      int b = a;
      while (b + 1 < n && sidx[b + 1] && b - a < 64) b++;
      if (b - a >= 64) {
        big = true;
        continue;
      }
      for (int u = a + 1; u <= b; u++) {  // insertion sort of the gap
        const uint64_t v = skey[u];
        int c = u - 1;
        while (c >= a && skey[c] > v) {
          skey[c + 1] = skey[c];
          c--;
        }
        skey[c + 1] = v;
      }
    }
    if (__syncthreads_or(big ? 1 : 0) == 0) {
      sorted = true;
#pragma unroll
      for (int j = 0; j < 4; j++) kk[j] = skey[e0 + j];
    }
  }
  for (int size = 2; size <= (sorted ? 1 : n2); size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 4 * kWave) {
        __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int j = 0; j < 4; j++) skey[4 * t + j] = kk[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int e = 4 * t + j;
          bitonic_take(e, size, stride, skey[e ^ stride], kk[j]);
        }
      } else if (stride >= 4) {
        const int m = stride >> 2;
#pragma unroll
        for (int j = 0; j < 4; j++)
          bitonic_take(4 * t + j, size, stride, __shfl_xor(kk[j], m, kWave), kk[j]);
      } else if (stride == 2) {
        bitonic_pair(4 * t, size, kk[0], kk[2]);
        bitonic_pair(4 * t + 1, size, kk[1], kk[3]);
      } else {
        bitonic_pair(4 * t, size, kk[0], kk[1]);
        bitonic_pair(4 * t + 2, size, kk[2], kk[3]);
      }
    }
  __syncthreads();
  // skey: the 52-bit prefix, sidx: the input index; the order for the next
  // call (before the runs below reorder within equal prefixes)
#pragma unroll
  for (int j = 0; j < 4; j++) {
    skey[4 * t + j] = kk[j] >> 12;
    sidx[4 * t + j] = (unsigned short)(kk[j] & 0xfff);
    if (order && 4 * t + j < n) order[4 * t + j] = (unsigned short)(kk[j] & 0xfff);
  }
  __syncthreads();
  // runs of equal prefixes (bodies in one cell of 2^-17 of the box; rare):
  // by (high key, low key, index), insertion sort per run
  for (int k = t; k < n - 1; k += 1024) {
    if (skey[k] != skey[k + 1] || (k > 0 && skey[k - 1] == skey[k])) continue;
    int e = k + 1;
    while (e + 1 < n && skey[e + 1] == skey[k]) e++;
    for (int a = k + 1; a <= e; a++) {
      const unsigned short ia = sidx[a];
      uint64_t ha, la;
      small_key(x3, rec(ia), X0, R0, ha, la);
      int c = a - 1;
      while (c >= k) {
        uint64_t hc, lc;
        small_key(x3, rec(sidx[c]), X0, R0, hc, lc);
        if (hc < ha || (hc == ha && (lc < la || (lc == la && sidx[c] < ia)))) break;
        sidx[c + 1] = sidx[c];
        c--;
      }
      sidx[c + 1] = ia;
    }
  }
  __syncthreads();
  // ---- the gather into tree order (every load of a thread's four bodies
  // issued before its stores)
  {
    int src[4];
    double gx[4][3], gv[4][3], gq[4];
    unsigned char gf[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = t + 1024 * r;
      src[r] = k < n ? sidx[k] : 0;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = rec(src[r]);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        gx[r][d] = x3[3 * (size_t)i + d];
        gv[r][d] = v3[3 * (size_t)i + d];
      }
      gq[r] = q[i];
      gf[r] = flags_in ? flags_in[i] : (unsigned char)(kSink | kSource);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = t + 1024 * r;
      if (k >= n) continue;
      spx[k] = gx[r][0];
      spy[k] = gx[r][1];
      spz[k] = gx[r][2];
      b.x[k] = gx[r][0];
      b.y[k] = gx[r][1];
      b.z[k] = gx[r][2];
      b.vx[k] = gv[r][0];
      b.vy[k] = gv[r][1];
      b.vz[k] = gv[r][2];
      b.q[k] = gq[r];
      b.flags[k] = gf[r];
      b.order[k] = src[r];
      b.acc_old[k] = 0.0;
      lend[k] = 0;
    }
  }
  __syncthreads();
  // bodies at (next to) zero distance: neighbours in tree order, and every
  // pair within a run of equal prefixes
  {
    auto near_zero = [&](int a, int c) {
      return fabs(spx[a] - spx[c]) < 1e-100 && fabs(spy[a] - spy[c]) < 1e-100 &&
             fabs(spz[a] - spz[c]) < 1e-100;
    };
    bool nz = false;
    for (int k = t; k < n - 1; k += 1024) {
      nz = nz || near_zero(k, k + 1);
      if (skey[k] != skey[k + 1] || (k > 0 && skey[k - 1] == skey[k])) continue;
      int e = k + 1;
      while (e + 1 < n && skey[e + 1] == skey[k]) e++;
      for (int a = k; a <= e; a++)
        for (int c = a + 1; c <= e; c++) nz = nz || near_zero(a, c);
    }
    if (nz) near = 1;
  }
  // root
  if (t == 0) {
    ci[0] = 0;
    ci[1] = n;
    for (int d = 0; d < 3; d++) cg[d] = X0[d];
    cg[3] = R0;
  }
  __syncthreads();
  // ---- levels
  int *cur_i = ci, *nxt_i = ci + 2 * (size_t)n;
  double *cur_g = cg, *nxt_g = cg + 4 * (size_t)n;
  const int j = t & 7;
  for (int level = 0;; level++) {
    const int nc = ncur;
    if (nc == 0) break;
    if (level >= kMaxTreeLevels) {
      if (t == 0) status = 2;
      break;
    }
    for (int base = 0; base < nc; base += 128) {
      const int k = base + (t >> 3);
      const bool on = k < nc;
      int begin = 0, nb = 0;
      double c0 = 0, c1 = 0, c2 = 0, r = 0;
      if (on) {
        begin = cur_i[2 * k];
        nb = cur_i[2 * k + 1];
        c0 = cur_g[4 * k];
        c1 = cur_g[4 * k + 1];
        c2 = cur_g[4 * k + 2];
        r = cur_g[4 * k + 3];
      }
      const bool split = on && nb > ncrit;
      if (on && !split) {
        if (j == 0) {
          lend[begin + nb - 1] = 1;
          atomicAdd(&nleaf, 1);
        }
        if (lidx)
          for (int k2 = begin + j; k2 < begin + nb; k2 += 8) lidx[k2] = begin + nb - 1;
      }
      const int end = begin + nb;
      const bool small = nb <= 32;
      int o = begin, sz = 0;
      if (split) {
        if (small) {
          for (int i = begin; i < end; i++) {
            const int oc = (spx[i] >= c0) + ((spy[i] >= c1) << 1) + ((spz[i] >= c2) << 2);
            sz += oc == j;
          }
        } else if (j > 0) {
          int l = begin, rr = end;
          while (l < rr - 1) {
            const int mid = (l + rr) / 2;
            const int oc = (spx[mid] >= c0) + ((spy[mid] >= c1) << 1) + ((spz[mid] >= c2) << 2);
            if (oc >= j)
              rr = mid;
            else
              l = mid;
          }
          o = l + 1;
        }
      }
      // within the group of eight (every lane of a group shares `split`)
      int incl = sz;
#pragma unroll
      for (int d = 1; d < 8; d <<= 1) {
        const int tt = __shfl(incl, lane - d, kWave);
        if (j >= d) incl += tt;
      }
      const int onext = __shfl(o, lane + 1, kWave);
      if (split) {
        if (small)
          o = begin + incl - sz;
        else
          sz = (j < 7 ? onext : end) - o;
        if (sz > 0) {
          const int slot = atomicAdd(&nnext, 1);
          const double rc = ldexp(r, -1);
          nxt_i[2 * slot] = o;
          nxt_i[2 * slot + 1] = sz;
          nxt_g[4 * slot] = c0 + rc * (((j >> 0) & 1) * 2 - 1);
          nxt_g[4 * slot + 1] = c1 + rc * (((j >> 1) & 1) * 2 - 1);
          nxt_g[4 * slot + 2] = c2 + rc * (((j >> 2) & 1) * 2 - 1);
          nxt_g[4 * slot + 3] = rc;
        }
      }
    }
    __syncthreads();
    if (t == 0) {
      ncur = nnext;
      nnext = 0;
    }
    int *ti = cur_i;
    cur_i = nxt_i;
    nxt_i = ti;
    double *tg = cur_g;
    cur_g = nxt_g;
    nxt_g = tg;
    __syncthreads();
  }
  __syncthreads();
  if (pmx) {
    // the mass before each body in its leaf: a scan over tree order that
    // restarts after every leaf end; thread t holds bodies 4t .. 4t+3, its
    // carry (the mass after the chunk's last leaf end, or the whole chunk's
    // mass with a flag for "no end") joins the others' by a wave scan, then
    // the sixteen wave totals
    double qv[4], cs = 0.0;
    bool ev[4], cf = false;
    for (int e = 0; e < 4; e++) {
      const int k = 4 * t + e;
      qv[e] = k < n ? b.q[k] : 0.0;
      ev[e] = k < n && lend[k];
      cs = ev[e] ? 0.0 : cs + qv[e];
      cf = cf || ev[e];
    }
    // inclusive scan of (cf, cs): (fa, sa) then (fb, sb) -> (fa|fb, fb ? sb : sa+sb)
    for (int d = 1; d < kWave; d <<= 1) {
      const double us = __shfl_up(cs, d, kWave);
      const int uf = __shfl_up((int)cf, d, kWave);
      if (lane >= d) {
        cs = cf ? cs : us + cs;
        cf = cf || uf;
      }
    }
    if (lane == kWave - 1) {
      wred[w][4] = cs;
      wred[w][5] = cf ? 1.0 : 0.0;
    }
    __syncthreads();
    // exclusive carry into this thread: earlier waves, then earlier lanes
    double carry = 0.0;
    for (int u = 0; u < w; u++) carry = wred[u][5] != 0.0 ? wred[u][4] : carry + wred[u][4];
    const double ls = __shfl_up(cs, 1, kWave);
    const int lf = __shfl_up((int)cf, 1, kWave);
    if (lane > 0) carry = lf ? ls : carry + ls;
    for (int e = 0; e < 4; e++) {
      const int k = 4 * t + e;
      if (k < n) pmx[k] = carry;
      carry = ev[e] ? 0.0 : carry + qv[e];
    }
  }
  if (t == 0) {
    out[0] = nleaf;
    out[1] = status;
    out[2] = near;
    if (sticky && status == 2) *sticky = 2;
  }
}

int launch_small_tree(const double *x3, const double *v3, const double *q,
                      const unsigned char *flags, int n, int ncrit, BodiesDev b,
                      unsigned char *lend, int *lidx, double *pmx, int *ci, double *cg, int *out,
                      int *sticky, const int *ix, unsigned short *order, bool use_order,
                      hipStream_t s) {
  if (n < 1 || n > kSmallTree) return -1;
  hipLaunchKernelGGL(k_small_tree, dim3(1), dim3(1024), 0, s, x3, v3, q, flags, n, ncrit, b,
                     lend, lidx, pmx, ci, cg, out, sticky, ix, order, use_order ? 1 : 0);
  return 0;
}

size_t scan_temp_bytes(int n) {
  size_t b = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int *)nullptr, (int *)nullptr, n);
  return b;
}

int launch_scan(const int *in, int *out, int n, void *temp, size_t temp_bytes, hipStream_t s) {
  size_t tb = temp_bytes;
  return hipcub::DeviceScan::ExclusiveSum(temp, tb, in, out, n, s) == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ traversal -----
// One breadth-first step, in two passes so that no shared counter is hit by
// every wave: pass 1 classifies each frontier pair, keeps the class and writes
// per-block totals (M2L, P2P, child pairs); the totals are scanned over
// blocks; pass 2 writes every output at its scanned position, so the lists
// come out in frontier order (deterministic).
// The walks' squared distance norm(dX) (kernel_fast_lean.h:113-116) and one
// error-MAC term (traverse_eager.h:80-88, master :82-90) rounded as the
// reference's x86-64 build evaluates them: GCC -O3 contracts the sum of
// squares into fma(z, z, fma(x, x, y*y)) and the sums into
// E = fma(Pn * power, comb, E).  The p = 2 pre-pass compares the distance
// with 3(Ri+Rj)^2 on grid-aligned cells, where diagonal neighbours tie in
// exact arithmetic and this rounding decides them.
__device__ __forceinline__ double mac_norm(double dx, double dy, double dz) {
  return __builtin_fma(dz, dz, __builtin_fma(dx, dx, dy * dy));
}
__device__ __forceinline__ double mac_term(double E, double pn, double pw, double comb) {
  return __builtin_fma(pn * pw, comb, E);
}

__device__ __forceinline__ int trav_classify(const CellsDev &c, const MacArgs &a, int i, int j,
                                             int *nsplit) {
#pragma clang fp contract(off)
  const int nci = c.nchild[i], ncj = c.nchild[j];
  double Ri, Rj;
  bool accept;
  if (a.mode == kMacLow) {
    const double dx = c.Xg[3 * i] - c.Xg[3 * j];
    const double dy = c.Xg[3 * i + 1] - c.Xg[3 * j + 1];
    const double dz = c.Xg[3 * i + 2] - c.Xg[3 * j + 2];
    const double R2 = mac_norm(dx, dy, dz);
    Ri = c.Rg[i];
    Rj = c.Rg[j];
    accept = R2 > (Ri + Rj) * (Ri + Rj) * 3;
  } else {
    const double dx = c.X[3 * i] - c.X[3 * j];
    const double dy = c.X[3 * i + 1] - c.X[3 * j + 1];
    const double dz = c.X[3 * i + 2] - c.X[3 * j + 2];
    Ri = c.R[i];
    Rj = c.R[j];
    if (a.mode == kMacTheta) {
      const double th = a.theta_cell ? a.theta_cell[i] : a.theta;
      const double R2 = mac_norm(dx, dy, dz) * th * th;
      accept = R2 > (Ri + Rj) * (Ri + Rj);
    } else if (a.mode == kMacMaster) {
      const double R2 = mac_norm(dx, dy, dz);
      const double s = Ri + Rj;
      double f1 = 10, f2 = 10;
      if (R2 > s * s) {
        const int P = a.p;
        const double R = sqrt(R2);
        const double *Pi = a.Pn + (size_t)i * (P + 1), *Pj = a.Pn + (size_t)j * (P + 1);
        double Rp = R2, pa = 1, pb = 1, Eab = 0, Eba = 0;
        for (int k = 0; k <= P; k++) {
          Eab = mac_term(Eab, Pi[P - k], pb, a.comb[k]);
          Eba = mac_term(Eba, Pj[P - k], pa, a.comb[k]);
          pa *= Ri;
          pb *= Rj;
          Rp *= R;
        }
        const double fac = 8 * fmax(Ri, Rj) / s / Rp;
        Eab *= fac;
        f1 = Eab / (a.facc * c.min_acc[j]);
        Eba *= fac;
        f2 = Eba / (a.facc * c.min_acc[i]);
      }
      if (f2 < 1 && f1 < 1) {
        *nsplit = 0;
        const int ni = c.nbody[i], nj = c.nbody[j];
        if ((ni <= 2 && nj <= 8) || (nj <= 2 && ni <= 8)) return 2;  // P2P
        return a.steps ? 0 : 1;                                       // M2L, or nothing
      }
      accept = false;
    } else {
      const double R2 = mac_norm(dx, dy, dz);
      const double s = Ri + Rj;
      double f2 = 10;
      if (R2 * 0.4 > s * s) {
        const int P = a.p;
        const double R = sqrt(R2);
        const double Rp = pow(R, -(double)(P + 1));
        const double *Pn = a.Pn + (size_t)j * (P + 1);
        double pa = 1, E = 0;
        for (int k = 0; k <= P; k++) {
          E = mac_term(E, Pn[P - k], pa, a.comb[k]);
          pa *= Ri;
        }
        E *= 8 * fmax(Ri, Rj) * Rp;
        f2 = E / (a.facc * c.min_acc[i]);
      }
      accept = f2 < 1;
    }
  }
  *nsplit = 0;
  if (accept) return 1;                       // M2L
  if (nci == 0 && ncj == 0) return 2;         // P2P
  if (ncj == 0 || (Ri >= Rj && nci != 0)) {   // split the target
    *nsplit = nci;
    return 3;
  }
  *nsplit = ncj;  // split the source
  return 4;
}

// Exclusive scan of three ints over a block of NW waves; totals to *tot.
template <int NW>
__device__ __forceinline__ int3 block_excl_scan3_n(int3 v, int3 *tot) {
  __shared__ int3 wsum[NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int tx, ty, tz;
  int ex = wave_excl_scan(v.x, &tx), ey = wave_excl_scan(v.y, &ty), ez = wave_excl_scan(v.z, &tz);
  if (lane == 0) wsum[w] = make_int3(tx, ty, tz);
  __syncthreads();
  int3 base = make_int3(0, 0, 0), all = make_int3(0, 0, 0);
  for (int k = 0; k < NW; k++) {
    int3 q = wsum[k];
    if (k < w) base = make_int3(base.x + q.x, base.y + q.y, base.z + q.z);
    all = make_int3(all.x + q.x, all.y + q.y, all.z + q.z);
  }
  *tot = all;
  __syncthreads();
  return make_int3(base.x + ex, base.y + ey, base.z + ez);
}

// Steps of the walk queued without host reads: the frontier size, the list
// lengths and the step's offsets live in TravState.  Block b of the
// kTravBlocks blocks owns frontier pairs [b R, (b + 1) R), R = ceil(nin /
// kTravBlocks), walked 256 at a time, so the outputs keep frontier order.
__device__ __forceinline__ void trav_range(int nin, int &lo, int &hi) {
  const int R = (nin + kTravBlocks - 1) / kTravBlocks;
  const long long l = (long long)blockIdx.x * R;
  lo = l < nin ? (int)l : nin;
  hi = nin - lo < R ? nin : lo + R;
}

// pass 1: per-block totals of M2L pairs, P2P pairs and child pairs; each
// pair's class (kind << 4 | child pairs) is kept for pass 2
__global__ __launch_bounds__(256) void k_trav_count(CellsDev c, const int2 *__restrict__ fin,
                                                    const TravState *__restrict__ st, int par,
                                                    MacArgs mac, int *__restrict__ blk,
                                                    unsigned char *__restrict__ cls) {
  int lo, hi;
  trav_range(st->nin[par], lo, hi);
  int3 acc = make_int3(0, 0, 0);
  for (int base = lo; base < hi; base += 256) {
    const int t = base + (int)threadIdx.x;
    int3 v = make_int3(0, 0, 0);
    if (t < hi) {
      const int2 pr = fin[t];
      int ns;
      const int kind = trav_classify(c, mac, pr.x, pr.y, &ns);
      v = make_int3(kind == 1, kind == 2, ns);
      cls[t] = (unsigned char)(kind << 4 | ns);
    }
    int3 tot;
    block_excl_scan3_n<4>(v, &tot);
    acc = make_int3(acc.x + tot.x, acc.y + tot.y, acc.z + tot.z);
  }
  if (threadIdx.x == 0) {
    blk[blockIdx.x] = acc.x;
    blk[kTravBlocks + blockIdx.x] = acc.y;
    blk[2 * kTravBlocks + blockIdx.x] = acc.z;
  }
}

// the block offsets (exclusive scans of pass 1's totals, to blk[3 G ..]) and
// the state for pass 2 and the next step; a step whose output would not fit
// sets the overflow flag, records what it needed and ends the walk.  One
// wave: it must find room on a CU while the upward moments (second stream)
// fill the machine with small blocks, which a 1024-thread block waits out.
__global__ __launch_bounds__(64) void k_trav_scan(int *__restrict__ blk, TravState *st, int par,
                                                  long long cap_m2l, long long cap_p2p,
                                                  long long cap_front) {
  constexpr int G = kTravBlocks, PER = G / kWave;
  static_assert(G % kWave == 0, "whole entries per lane");
  const int t = threadIdx.x;
  int3 sum = make_int3(0, 0, 0);
  for (int k = 0; k < PER; k++) {
    const int i = t * PER + k;
    sum = make_int3(sum.x + blk[i], sum.y + blk[G + i], sum.z + blk[2 * G + i]);
  }
  int3 tot;
  int3 run;
  run.x = wave_excl_scan(sum.x, &tot.x);
  run.y = wave_excl_scan(sum.y, &tot.y);
  run.z = wave_excl_scan(sum.z, &tot.z);
  for (int k = 0; k < PER; k++) {
    const int i = t * PER + k;
    const int3 a = make_int3(blk[i], blk[G + i], blk[2 * G + i]);
    blk[3 * G + i] = run.x;
    blk[4 * G + i] = run.y;
    blk[5 * G + i] = run.z;
    run = make_int3(run.x + a.x, run.y + a.y, run.z + a.z);
  }
  if (t == 0) {
    const int nin = st->nin[par];
    const long long nm = st->nm2l, np = st->np2p;
    const bool fits = !st->overflow && nm + tot.x <= cap_m2l && np + tot.y <= cap_p2p &&
                      tot.z <= cap_front;
    if (fits) {
      st->base_m2l = (int)nm;
      st->base_p2p = (int)np;
      st->nm2l = (int)(nm + tot.x);
      st->np2p = (int)(np + tot.y);
      st->nin[par ^ 1] = tot.z;
      st->write = 1;
      if (nin > 0) st->steps++;
    } else {
      if (!st->overflow) {
        const long long big = 0x7fffffff;
        st->need_m2l = (int)(nm + tot.x < big ? nm + tot.x : big);
        st->need_p2p = (int)(np + tot.y < big ? np + tot.y : big);
        st->need_front = tot.z;
      }
      st->overflow = 1;
      st->write = 0;
      st->nin[par ^ 1] = 0;
    }
  }
}

// pass 2: every output at its place (block offset + running count within the
// block + rank within the 256-pair slice)
__global__ __launch_bounds__(256) void k_trav_write(CellsDev c, TravArgs a,
                                                    const TravState *__restrict__ st, int par,
                                                    const int *__restrict__ blk) {
  if (!st->write) return;
  int lo, hi;
  trav_range(st->nin[par], lo, hi);
  const int b = blockIdx.x;
  int3 run = make_int3(blk[3 * kTravBlocks + b], blk[4 * kTravBlocks + b],
                       blk[5 * kTravBlocks + b]);
  const int bm = st->base_m2l, bp = st->base_p2p;
  for (int base = lo; base < hi; base += 256) {
    const int t = base + (int)threadIdx.x;
    int kind = 0, ns = 0;
    int2 pr = make_int2(0, 0);
    if (t < hi) {
      pr = a.fin[t];
      const int k = a.cls[t];
      kind = k >> 4;
      ns = k & 15;
    }
    int3 tot;
    const int3 ex = block_excl_scan3_n<4>(make_int3(kind == 1, kind == 2, ns), &tot);
    if (kind == 1) a.m2l[bm + run.x + ex.x] = pr;
    if (kind == 2) a.p2p[bp + run.y + ex.y] = pr;
    if (kind >= 3) {
      const int o = run.z + ex.z;
      const int c0 = c.child[kind == 3 ? pr.x : pr.y];
      for (int k = 0; k < ns; k++)
        a.fout[o + k] = kind == 3 ? make_int2(c0 + k, pr.y) : make_int2(pr.x, c0 + k);
    }
    run = make_int3(run.x + tot.x, run.y + tot.y, run.z + tot.z);
  }
}

int launch_trav_step(CellsDev c, const TravArgs &a, TravState *st, int par, int *blk,
                     long long cap_m2l, long long cap_p2p, long long cap_front, hipStream_t s) {
  hipLaunchKernelGGL(k_trav_count, dim3(kTravBlocks), dim3(256), 0, s, c, a.fin,
                     (const TravState *)st, par, a.mac, blk, a.cls);
  hipLaunchKernelGGL(k_trav_scan, dim3(1), dim3(kWave), 0, s, blk, st, par, cap_m2l, cap_p2p,
                     cap_front);
  hipLaunchKernelGGL(k_trav_write, dim3(kTravBlocks), dim3(256), 0, s, c, a,
                     (const TravState *)st, par, (const int *)blk);
  return 0;
}

// Group (target, source) pairs by target: stable radix sort on the target,
// then the start and end of each target's run.
size_t pair_sort_temp_bytes(int n) {
  size_t b = 0, bk = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, b, (int *)nullptr, (int *)nullptr, (int *)nullptr,
                                     (int *)nullptr, n, 0, 32);
  hipcub::DeviceRadixSort::SortKeys(nullptr, bk, (unsigned *)nullptr, (unsigned *)nullptr, n, 0,
                                    32);
  return std::max(b, bk);
}

__global__ void k_split_pairs(const int2 *__restrict__ pairs, int n, int *__restrict__ tg,
                              int *__restrict__ sr) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int2 p = pairs[k];
  tg[k] = p.x;
  sr[k] = p.y;
}

// Marks runs of equal targets: start[t] = first index, endp[t] = one past the
// last (both zero-initialized, so targets without entries keep 0 / 0).
__global__ void k_runs(const int *__restrict__ tg, int n, int *__restrict__ start,
                       int *__restrict__ endp) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int t = tg[k];
  if (k == 0 || tg[k - 1] != t) start[t] = k;
  if (k == n - 1 || tg[k + 1] != t) endp[t] = k + 1;
}

__global__ void k_run_counts(const int *__restrict__ start, int *__restrict__ cnt, int ncells) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c < ncells) cnt[c] -= start[c];
}

// Cell ids below 2^16: one 32-bit word per pair, target above source, so the
// sort moves half the bytes of a key + value sort (the same stable order).
__global__ void k_pack_pairs(const int2 *__restrict__ pairs, int n, unsigned *__restrict__ w) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int2 p = pairs[k];
  w[k] = ((unsigned)p.x << 16) | (unsigned)p.y;
}

// k_runs on the sorted words, writing the sources as it goes
__global__ void k_runs_packed(const unsigned *__restrict__ w, int n, int *__restrict__ start,
                              int *__restrict__ endp, int *__restrict__ src) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const unsigned v = w[k];
  const int t = (int)(v >> 16);
  src[k] = (int)(v & 0xffffu);
  if (k == 0 || (int)(w[k - 1] >> 16) != t) start[t] = k;
  if (k == n - 1 || (int)(w[k + 1] >> 16) != t) endp[t] = k + 1;
}

// After the call: src[] holds sources grouped by target (stable), start[] /
// count[] (zeroed by the caller) the run of every target.
int launch_group_pairs(const int2 *pairs, int n, int key_bits, int ncells, int *tg_a, int *tg_b,
                       int *src_a, int *src, int *start, int *count, void *temp,
                       size_t temp_bytes, hipStream_t s) {
  if (n <= 0) return 0;
  if (ncells <= (1 << 16) && key_bits <= 16) {
    unsigned *wa = reinterpret_cast<unsigned *>(tg_a), *wb = reinterpret_cast<unsigned *>(tg_b);
    hipLaunchKernelGGL(k_pack_pairs, dim3((n + 255) / 256), dim3(256), 0, s, pairs, n, wa);
    size_t tb = temp_bytes;
    if (hipcub::DeviceRadixSort::SortKeys(temp, tb, wa, wb, n, 16, 16 + key_bits, s) !=
        hipSuccess)
      return -1;
    hipLaunchKernelGGL(k_runs_packed, dim3((n + 255) / 256), dim3(256), 0, s, wb, n, start,
                       count, src);
    hipLaunchKernelGGL(k_run_counts, dim3((ncells + 255) / 256), dim3(256), 0, s, start, count,
                       ncells);
    return 0;
  }
  hipLaunchKernelGGL(k_split_pairs, dim3((n + 255) / 256), dim3(256), 0, s, pairs, n, tg_a, src_a);
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, tg_a, tg_b, src_a, src, n, 0, key_bits, s) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_runs, dim3((n + 255) / 256), dim3(256), 0, s, tg_b, n, start, count);
  hipLaunchKernelGGL(k_run_counts, dim3((ncells + 255) / 256), dim3(256), 0, s, start, count,
                     ncells);
  return 0;
}

// M2L chunks in a stable order by their target's order (per-cell orders):
// each order's chunks keep the target order of the list, which keeps
// neighbouring targets (sharing sources in L2) on one XCD.
__global__ void k_chunk_keys(const int4 *__restrict__ ch, int n, const int *__restrict__ q,
                             int *__restrict__ key, int *__restrict__ idx) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  key[k] = q[ch[k].x];
  idx[k] = k;
}

__global__ void k_chunk_gather(const int4 *__restrict__ ch, const int *__restrict__ idx, int n,
                               int4 *__restrict__ out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) out[k] = ch[idx[k]];
}

int launch_sort_chunks_by_order(const int4 *chunks, int n, const int *q, int *key_a, int *key_b,
                                int *idx_a, int *idx_b, int4 *out, void *temp, size_t temp_bytes,
                                hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_chunk_keys, dim3((n + 255) / 256), dim3(256), 0, s, chunks, n, q, key_a,
                     idx_a);
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, key_a, key_b, idx_a, idx_b, n, 0, 5, s) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_chunk_gather, dim3((n + 255) / 256), dim3(256), 0, s, chunks, idx_b, n,
                     out);
  return 0;
}

// chunk_offs = exclusive scan of ceil(count/64) over cells
__global__ void k_make_chunks(const int *__restrict__ offs, const int *__restrict__ count,
                              const int *__restrict__ chunk_offs, int ncells,
                              int4 *__restrict__ chunks) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncells) return;
  int n = count[c], o = offs[c], q = chunk_offs[c];
  for (int k = 0; k * 64 < n; k++) {
    int cnt = n - 64 * k;
    if (cnt > 64) cnt = 64;
    chunks[q + k] = make_int4(c, o + 64 * k, cnt, 0);
  }
}

int launch_make_chunks(const int *offs, const int *count, const int *chunk_offs, int ncells,
                       int4 *chunks, hipStream_t s) {
  if (ncells <= 0) return 0;
  hipLaunchKernelGGL(k_make_chunks, dim3((ncells + 255) / 256), dim3(256), 0, s, offs, count,
                     chunk_offs, ncells, chunks);
  return 0;
}

}  // namespace tcfmm
