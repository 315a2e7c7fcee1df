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
//   Appends are wave-aggregated atomics; the lists are then grouped by target
//   cell (counting sort) and the M2L list is cut into chunks of <= 64 sources.
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

__global__ void k_bounds_final(const double *__restrict__ part, int nb,
                               const double *__restrict__ shift_frac, TreeParams *tp) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  double mn[3], mx[3];
  for (int d = 0; d < 3; d++) mn[d] = part[d], mx[d] = part[3 + d];
  for (int b = 1; b < nb; b++)
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(part[6 * b + d], mn[d]);
      mx[d] = fmax(part[6 * b + 3 + d], mx[d]);
    }
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
                                                const TreeParams *__restrict__ tp,
                                                uint64_t *__restrict__ key_hi,
                                                uint64_t *__restrict__ key_lo,
                                                int *__restrict__ idx) {
#pragma clang fp contract(off)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
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
                  uint64_t *key_lo, int *idx, hipStream_t s) {
  hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, s, x3, n, d_tp, key_hi,
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
__global__ void k_root(CellsDev c, int n, const TreeParams *__restrict__ tp) {
  if (threadIdx.x != 0) return;
  c.level[0] = 0;
  c.nchild[0] = 0;
  c.child[0] = -1;
  c.body[0] = 0;
  c.nbody[0] = n;
  for (int d = 0; d < 3; d++) c.Xg[d] = tp->X0[d];
  c.Rg[0] = tp->R0;
}

int launch_root(CellsDev c, int n, const TreeParams *d_tp, hipStream_t s) {
  hipLaunchKernelGGL(k_root, dim3(1), dim3(64), 0, s, c, n, d_tp);
  return 0;
}

__device__ __forceinline__ int octant_of(const BodiesDev &b, int i, double X0, double X1,
                                         double X2) {
  return (b.x[i] >= X0) + ((b.y[i] >= X1) << 1) + ((b.z[i] >= X2) << 2);
}

// offs: 8 ints per cell of the level (octant starts), nch: children per cell
__global__ __launch_bounds__(256) void k_split(CellsDev c, BodiesDev b, int ls, int le, int ncrit,
                                               int *__restrict__ offs, int *__restrict__ nch) {
  int k = blockIdx.x * 256 + threadIdx.x;
  int ci = ls + k;
  if (ci >= le) return;
  int nb = c.nbody[ci];
  if (nb <= ncrit) {
    nch[k] = 0;
    c.nchild[ci] = 0;
    c.child[ci] = -1;
    return;
  }
  int begin = c.body[ci], end = begin + nb;
  double X0 = c.Xg[3 * ci], X1 = c.Xg[3 * ci + 1], X2 = c.Xg[3 * ci + 2];
  int o[8], size[8];
  if (nb <= 32) {
    for (int i = 0; i < 8; i++) size[i] = 0;
    for (int i = begin; i < end; i++) size[octant_of(b, i, X0, X1, X2)]++;
    int acc = begin;
    for (int i = 0; i < 8; i++) {
      o[i] = acc;
      acc += size[i];
    }
  } else {
    o[0] = begin;
    for (int i = 1; i < 8; i++) {
      int l = begin, r = end;
      while (l < r - 1) {
        int mid = (l + r) / 2;
        if (octant_of(b, mid, X0, X1, X2) >= i)
          r = mid;
        else
          l = mid;
      }
      o[i] = l + 1;
    }
    for (int i = 0; i < 7; i++) size[i] = o[i + 1] - o[i];
    size[7] = end - o[7];
  }
  int cnt = 0;
  for (int i = 0; i < 8; i++) {
    offs[8 * k + i] = o[i];
    if (size[i]) cnt++;
  }
  nch[k] = cnt;
  c.nchild[ci] = cnt;
}

int launch_split(CellsDev c, BodiesDev b, int ls, int le, int ncrit, int *offs, int *nch,
                 hipStream_t s) {
  int n = le - ls;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_split, dim3((n + 255) / 256), dim3(256), 0, s, c, b, ls, le, ncrit, offs,
                     nch);
  return 0;
}

__global__ __launch_bounds__(256) void k_children(CellsDev c, int ls, int le,
                                                  const int *__restrict__ offs,
                                                  const int *__restrict__ nch_scan,
                                                  const TreeParams *__restrict__ tp) {
#pragma clang fp contract(off)
  int k = blockIdx.x * 256 + threadIdx.x;
  int ci = ls + k;
  if (ci >= le) return;
  int nc = c.nchild[ci];
  if (nc == 0) return;
  int first = le + nch_scan[k];
  c.child[ci] = first;
  int level = c.level[ci];
  int begin = c.body[ci], end = begin + c.nbody[ci];
  double R0 = tp->R0;
  double r = ldexp(R0, -(level + 1));
  int j = 0;
  for (int i = 0; i < 8; i++) {
    int o = offs[8 * k + i];
    int e = (i < 7) ? offs[8 * k + i + 1] : end;
    int size = e - o;
    if (!size) continue;
    int cc = first + j;
    for (int d = 0; d < 3; d++) c.Xg[3 * cc + d] = c.Xg[3 * ci + d] + r * (((i >> d) & 1) * 2 - 1);
    c.Rg[cc] = r;
    c.level[cc] = level + 1;
    c.body[cc] = o;
    c.nbody[cc] = size;
    c.nchild[cc] = 0;
    c.child[cc] = -1;
    j++;
  }
}

int launch_children(CellsDev c, int ls, int le, const int *offs, const int *nch_scan,
                    const TreeParams *d_tp, hipStream_t s) {
  int n = le - ls;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_children, dim3((n + 255) / 256), dim3(256), 0, s, c, ls, le, offs,
                     nch_scan, d_tp);
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
// Appends `k` items for this lane to a list with one atomic per wave; returns
// the lane's first slot (may exceed the capacity; callers check).
__device__ __forceinline__ int wave_reserve(int *counter, int k) {
  int tot;
  int pre = wave_excl_scan(k, &tot);
  int base = 0;
  int lane = threadIdx.x & (kWave - 1);
  if (lane == 0 && tot > 0) base = atomicAdd(counter, tot);
  base = __shfl(base, 0, kWave);
  return base + pre;
}

__global__ __launch_bounds__(256) void k_traverse(CellsDev c, TravArgs a) {
#pragma clang fp contract(off)
  int t = blockIdx.x * 256 + threadIdx.x;
  bool live = t < a.nin;
  int2 pr = live ? a.fin[t] : make_int2(0, 0);
  int i = pr.x, j = pr.y;
  int kind = 0;  // 1 m2l, 2 p2p, 3 split i, 4 split j
  int nsplit = 0;
  if (live) {
    double dx = c.X[3 * i] - c.X[3 * j];
    double dy = c.X[3 * i + 1] - c.X[3 * j + 1];
    double dz = c.X[3 * i + 2] - c.X[3 * j + 2];
    double R2 = (dx * dx + dy * dy + dz * dz) * a.theta * a.theta;
    double Ri = c.R[i], Rj = c.R[j];
    int nci = c.nchild[i], ncj = c.nchild[j];
    if (R2 > (Ri + Rj) * (Ri + Rj)) {
      kind = 1;
    } else if (nci == 0 && ncj == 0) {
      kind = 2;
    } else if (ncj == 0 || (Ri >= Rj && nci != 0)) {
      kind = 3;
      nsplit = nci;
    } else {
      kind = 4;
      nsplit = ncj;
    }
  }
  int s1 = wave_reserve(&a.cnt[1], kind == 1 ? 1 : 0);
  if (kind == 1 && s1 < a.cap_m2l) a.m2l[s1] = make_int2(i, j);
  int s2 = wave_reserve(&a.cnt[2], kind == 2 ? 1 : 0);
  if (kind == 2 && s2 < a.cap_p2p) a.p2p[s2] = make_int2(i, j);
  int s3 = wave_reserve(&a.cnt[0], nsplit);
  if (kind == 3) {
    int c0 = c.child[i];
    for (int k = 0; k < nsplit; k++)
      if (s3 + k < a.cap_front) a.fout[s3 + k] = make_int2(c0 + k, j);
  } else if (kind == 4) {
    int c0 = c.child[j];
    for (int k = 0; k < nsplit; k++)
      if (s3 + k < a.cap_front) a.fout[s3 + k] = make_int2(i, c0 + k);
  }
}

int launch_traverse(CellsDev c, const TravArgs &a, hipStream_t s) {
  if (a.nin <= 0) return 0;
  hipLaunchKernelGGL(k_traverse, dim3((a.nin + 255) / 256), dim3(256), 0, s, c, a);
  return 0;
}

__global__ void k_count_targets(const int2 *__restrict__ pairs, int n, int *__restrict__ count) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) atomicAdd(&count[pairs[k].x], 1);
}

int launch_count_targets(const int2 *pairs, int n, int *count, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_count_targets, dim3((n + 255) / 256), dim3(256), 0, s, pairs, n, count);
  return 0;
}

__global__ void k_scatter_csr(const int2 *__restrict__ pairs, int n, const int *__restrict__ offs,
                              int *__restrict__ cursor, int *__restrict__ src) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int2 p = pairs[k];
  int pos = offs[p.x] + atomicAdd(&cursor[p.x], 1);
  src[pos] = p.y;
}

int launch_scatter_csr(const int2 *pairs, int n, const int *offs, int *cursor, int *src,
                       hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_scatter_csr, dim3((n + 255) / 256), dim3(256), 0, s, pairs, n, offs,
                     cursor, src);
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
