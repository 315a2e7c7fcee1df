// M2L: rotation-based multipole-to-local translation, FP64, gfx950.
//
// Follows M2L_rotate (taichi_with_jerks/kernel_fast_lean.h:1130-1371): per
// (target Ci, source Cj) pair the source moments are scaled to homogeneous
// form, rotated about z by alpha_z, x<->z swapped, rotated by alpha_x, swapped
// back (so the pair axis lies along z), translated along z, and rotated back
// with the same four steps in reverse; the result is added to Ci's local
// expansion.  The same operator maps Mdot to Ldot.
//
// GPU form.  The swap S_n is the same matrix for every pair, so it is applied
// as a product of constant sparse blocks against each lane's coefficients:
// one lane = one (target, source) pair, the swap entries are folded into the
// instruction stream as constant operands of v_fma_f64, and only the per-pair
// z-rotations and translation factors live in registers.  The forward half's
// scaling by (n+|m|)!(n-|m|)! is folded into its two swaps' constants.  Work that the translation cannot see is skipped: after the
// second forward swap only components |m| <= min(n, p-n) are formed, and the
// first backward swap only reads those.  The quirk of the reference's
// factorial table (factorial_table[2p] == 0, kernel_fast_lean.h:129) removes
// M_p^{+-p} from the sum; with quirk=1 those inputs are zeroed.
//
// The order P of a launch is the truncation order of its pairs: the cells'
// coefficient rows may be longer (stride >= (P+1)^2 doubles, an expansion of
// a higher order), and only the terms up to order P are read and added, so
// one set of coefficient arrays serves pairs of different orders.
//
// Work decomposition: a block = one wave = one chunk of <= 64 sources of a
// single target cell (the chunk list is grouped by target) for one array
// (M -> L or Mdot -> Ldot); each XCD takes a contiguous range of the
// (array, chunk) items so neighbouring targets share its L2.  Each order's 2n+1 contributions are
// summed across the wave through an LDS transpose; the chunk adds its sum
// into the target's L with one atomic per coefficient.
#include "tc_common.h"
#include "tc_kernels.h"
#include "swap_tables.h"

#ifndef TC_SHARD
#error "compiled once per order shard, with -DTC_SHARD=k (see TC_NSHARDS in tc_kernels.h)"
#endif

namespace tcfmm {
namespace TC_SHARD_NS(TC_SHARD) {

// n! for the translation factors, loaded with scalar loads (uniform across the
// wave).  Filled by m2l_init_tables().
__constant__ double c_fact[2 * kMaxP + 2];

// z-rotation of one order by e^{i m alpha} (forward) or e^{-i m alpha}
// (backward), m = 1..N.  The powers come from the recurrence
// e^{i(m+1)a} = e^{ima} e^{ia} (kernel_fast_lean.h:1163-1169).
//
// rotate(): the chain is recomputed inside the call from an opaque copy of
// e^{ia}.  Without the copy the compiler merges the identical chains of all
// unrolled orders into one table of 4p powers that stays live across the
// kernel (p = 12: ~90 extra registers, one wave per SIMD instead of two).
// The forward half uses it: its orders run p..0, so a table would be largest
// exactly where the working set is.  The backward half runs 0..p and grows a
// table by one power per order instead (rotate_tab), while the rotated
// moments it consumes die off.
template <int N, bool kBack>
__device__ __forceinline__ void rotate(double *v, double c1, double s1) {
  asm volatile("" : "+v"(c1), "+v"(s1));
  // Chebyshev form of the power recurrence: cos((m+1)a) = 2 cos a cos(ma) -
  // cos((m-1)a), and the same for sin, one FMA per component per power
  const double t = c1 + c1;
  double cq = 1.0, sq = 0.0, c = c1, s = s1;
#pragma unroll
  for (int m = 1; m <= N; m++) {
    double im = v[N - m], re = v[N + m];
    if (kBack) {
      v[N - m] = c * im - s * re;
      v[N + m] = c * re + s * im;
    } else {
      v[N - m] = c * im + s * re;
      v[N + m] = c * re - s * im;
    }
    if (m < N) {
      const double cn = fma(t, c, -cq), sn = fma(t, s, -sq);
      cq = c;
      sq = s;
      c = cn;
      s = sn;
    }
  }
}

template <int N, bool kBack>
__device__ __forceinline__ void rotate_tab(double *v, const double *cp, const double *sp) {
#pragma unroll
  for (int m = 1; m <= N; m++) {
    const double c = cp[m], s = sp[m];
    double im = v[N - m], re = v[N + m];
    if (kBack) {
      v[N - m] = c * im - s * re;
      v[N + m] = c * re + s * im;
    } else {
      v[N - m] = c * im + s * re;
      v[N + m] = c * re - s * im;
    }
  }
}

// Powers e^{iNa} from the two below (the Chebyshev recurrence rotate() uses),
// stored at index N; index 0 holds e^0.
template <int N>
__device__ __forceinline__ void extend_powers(double *cp, double *sp, double c1, double s1) {
  if constexpr (N == 1) {
    cp[0] = 1.0;
    sp[0] = 0.0;
    cp[1] = c1;
    sp[1] = s1;
  } else if constexpr (N > 1) {
    const double t = c1 + c1;
    cp[N] = fma(t, cp[N - 1], -cp[N - 2]);
    sp[N] = fma(t, sp[N - 1], -sp[N - 2]);
  }
}

// out rows |r-N| <= QR from input columns |c-N| <= QC (everything else 0).
// S_n only couples components of one parity class (Re/Im part and parity of
// m); rows are produced class by class so each class's inputs die early.
// kScaled: D^-1 S_n D with D = diag((n+|m|)!(n-|m|)!) instead of S_n.
template <int N, int QR, int QC, bool kScaled = false>
__device__ __forceinline__ void swap_block(const double *in, double *out) {
#pragma unroll
  for (int cls = 0; cls < 4; cls++) {
#pragma unroll
    for (int r = 0; r < 2 * N + 1; r++) {
      const int m = r - N;
      const int rc = (m >= 0 ? 0 : 2) + ((m < 0 ? -m : m) & 1);
      if (rc != cls) continue;
      double acc = 0.0;
      if (m <= QR && -m <= QR) {
#pragma unroll
        for (int e = kSwapRowPtr[N * N + r]; e < kSwapRowPtr[N * N + r + 1]; e++) {
          const int c = kSwapCol[e];
          if (c - N <= QC && N - c <= QC) acc += (kScaled ? kSwapValScaled[e] : kSwapVal[e]) * in[c];
        }
      }
      out[r] = acc;
    }
  }
}

// Row stride (doubles) of the LDS reduction buffer: 64 lanes + 16 padding,
// so rows r and r+1 start 32 banks apart.
constexpr int kRedStride = kWave + 16;
// One reduction buffer holds the rows of the largest order.  Orders up to 12
// fit two waves per SIMD in registers, and a second buffer would then halve
// occupancy through LDS; from 14 on, one wave per SIMD runs anyway and two
// buffers alternate between consecutive orders so the reduction of one order
// overlaps the next order's arithmetic.
template <int P>
struct RedBuf {
  static constexpr int size = (2 * P + 1) * kRedStride;
  static constexpr bool pipelined = P >= 14;
  static constexpr int count = pipelined ? 2 : 1;
};

template <int P>
struct M2LRegs {
  static constexpr int q(int n) { return n < P - n ? n : P - n; }
  // offset of order n in the packed rotated-multipole store (|m| <= q(n))
  static constexpr int off(int n) {
    int o = 0;
    for (int k = 0; k < n; k++) o += 2 * q(k) + 1;
    return o;
  }
  static constexpr int total = off(P + 1);
};

template <int N>
__device__ __forceinline__ void load_order(const double *__restrict__ src, double *v) {
#pragma unroll
  for (int r = 0; r < 2 * N + 1; r++) v[r] = src[N * N + r];
}

// Prefetch distance of the forward half, in orders: the source gathers of
// orders N-1 .. N-kAhead are in flight while order N computes (one wave per
// SIMD leaves registers to spare, and a 64-lane gather takes thousands of
// cycles).
#ifndef TC_M2L_AHEAD
#define TC_M2L_AHEAD 1
#endif
constexpr int kAhead = TC_M2L_AHEAD;
static_assert(kAhead >= 1 && kAhead <= 4, "prefetch distance");

template <int N>
struct OrderBuf {
  static constexpr int size = N >= 0 ? 2 * N + 1 : 1;
};

// Forward half for order N of the source moments: scale, rotate by alpha_z,
// swap, rotate by alpha_x, swap (rows |m| <= min(N, p-N) only), unscale into
// the rotated-moment store.  b0 holds order N; b1..b3 hold orders N-1..N-3
// when loaded (kAhead of them); the gather of order N-kAhead is issued first,
// fenced with scheduling barriers so the compiler neither hoists every
// order's gathers to the top nor sinks them to their use.
template <int P, int N>
__device__ __forceinline__ void m2l_forward_order(const double *__restrict__ src, bool quirk,
                                                  double ezr, double ezi, double exr,
                                                  double exi, double *b0, double *b1,
                                                  double *b2, double *b3, double *mp) {
  constexpr int K = 2 * N + 1;
  constexpr int Q = M2LRegs<P>::q(N);
  constexpr int O = M2LRegs<P>::off(N);
  double nb[OrderBuf<N - kAhead>::size];
  if constexpr (N - kAhead >= 0) {
    __builtin_amdgcn_sched_barrier(0);
    load_order<N - kAhead>(src, nb);
    __builtin_amdgcn_sched_barrier(0);
  }
  // D^-1 S Rx S D Rz with D = diag((n+|m|)!(n-|m|)!) (the reference's scale
  // and unscale, kernel_fast_lean.h:1175-1181, 1247-1251): the z-rotations
  // only mix +-m, so they commute with D, and the product is
  // (D^-1 S D) Rx (D^-1 S D) Rz -- the scaling lives in the swap constants
  double v[K], w[K];
#pragma unroll
  for (int r = 0; r < K; r++) v[r] = b0[r];
  if (N == P && quirk) {
    v[0] = 0.0;
    v[2 * N] = 0.0;
  }
  rotate<N, false>(v, ezr, ezi);
  swap_block<N, N, N, true>(v, w);
  rotate<N, false>(w, exr, exi);
  swap_block<N, Q, N, true>(w, v);
#pragma unroll
  for (int m = -Q; m <= Q; m++) mp[O + Q + m] = v[N + m];
  if constexpr (N > 0) {
    if constexpr (kAhead == 1)
      m2l_forward_order<P, N - 1>(src, quirk, ezr, ezi, exr, exi, nb, nullptr, nullptr, nullptr,
                                  mp);
    else if constexpr (kAhead == 2)
      m2l_forward_order<P, N - 1>(src, quirk, ezr, ezi, exr, exi, b1, nb, nullptr, nullptr, mp);
    else if constexpr (kAhead == 3)
      m2l_forward_order<P, N - 1>(src, quirk, ezr, ezi, exr, exi, b1, b2, nb, nullptr, mp);
    else
      m2l_forward_order<P, N - 1>(src, quirk, ezr, ezi, exr, exi, b1, b2, b3, nb, mp);
  }
}

// Sum order N's K = 2N+1 contributions over the wave: the lanes' values sit
// in LDS rows (row stride kRedStride); groups of 8 lanes each sum one row
// (8 values per lane, a 3-step butterfly inside the group).  A group reads
// its row as four contiguous 128-byte runs and the padded stride puts
// neighbouring rows in opposite halves of the 64 banks, so the eight groups of
// one read fill the banks evenly.  Rows past K are read as row K-1 and
// written to a per-lane spare slot, which keeps the loop free of divergent
// branches.
template <int P, int N>
__device__ __forceinline__ void reduce_order(const double *red, double *acc_lds, int lane) {
  constexpr int K = 2 * N + 1;
  __syncthreads();
  const int seg = lane & 7, grp = lane >> 3;
#pragma unroll
  for (int g = 0; g < (K + 7) / 8; g++) {
    const int r = g * 8 + grp;
    const int rr = r < K ? r : K - 1;
    double sum = 0.0;
    const double2 *row = reinterpret_cast<const double2 *>(red + rr * kRedStride);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double2 t = row[seg + 8 * q];
      sum += t.x + t.y;
    }
    sum += __shfl_xor(sum, 1, kWave);
    sum += __shfl_xor(sum, 2, kWave);
    sum += __shfl_xor(sum, 4, kWave);
    acc_lds[(seg == 0 && r < K) ? N * N + r : nterm(P) + lane] = sum;
  }
}

template <int P, int N>
__device__ __forceinline__ void m2l_backward_order(const double *mp, double &rpn, double ri,
                                                   double ezr, double ezi, double exr,
                                                   double exi, double *czp, double *szp,
                                                   double *cxp, double *sxp, bool active,
                                                   double *acc_lds, double *red, int lane) {
  constexpr int K = 2 * N + 1;
  constexpr int Q = M2LRegs<P>::q(N);
  // one order at a time: keeps the scheduler from interleaving the
  // register-only work of several orders (which multiplies live state)
  __builtin_amdgcn_sched_barrier(0);
  // translation factors f[N+k] = (N+k)! r^-(N+k+1), k = 0..P-N, from the
  // power chain r^-(j+1) = r^-j * r^-1 (rpn = r^-(N+1) on entry)
  double f[P - N + 1];
  {
    double rp = rpn;
#pragma unroll
    for (int k = 0; k <= P - N; k++) {
      f[k] = c_fact[N + k] * rp;
      rp *= ri;
    }
    rpn *= ri;
  }
  double u[K], w[K];
#pragma unroll
  for (int r = 0; r < K; r++) u[r] = 0.0;
  // translation along z (kernel_fast_lean.h:1253-1295):
  // L_n^{+-m} += (-1)^m sum_{k=m}^{p-n} M_k^{+-m} (n+k)! r^-(n+k+1)
#pragma unroll
  for (int k = 0; k <= P - N; k++) {
    const int qk = M2LRegs<P>::q(k);
    const int ok = M2LRegs<P>::off(k);
    u[N] += mp[ok + qk] * f[k];
  }
#pragma unroll
  for (int m = 1; m <= Q; m++) {
#pragma unroll
    for (int k = m; k <= P - N; k++) {
      const int qk = M2LRegs<P>::q(k);
      const int ok = M2LRegs<P>::off(k);
      double fk = (m & 1) ? -f[k] : f[k];
      u[N - m] += mp[ok + qk - m] * fk;
      u[N + m] += mp[ok + qk + m] * fk;
    }
  }
  extend_powers<N>(cxp, sxp, exr, exi);
  extend_powers<N>(czp, szp, ezr, ezi);
  swap_block<N, N, Q>(u, w);
  rotate_tab<N, true>(w, cxp, sxp);
  swap_block<N, N, N>(w, u);
  rotate_tab<N, true>(u, czp, szp);
  if constexpr (RedBuf<P>::pipelined) {
    // The previous order's rows were written one order ago: reduce them now,
    // so the barrier finds those writes long complete; then store this
    // order's rows in the other buffer (reduced after the next order, or at
    // the end).
    if constexpr (N > 0)
      reduce_order<P, N - 1>(red + ((N - 1) & 1) * RedBuf<P>::size, acc_lds, lane);
#pragma unroll
    for (int r = 0; r < K; r++) red[(N & 1) * RedBuf<P>::size + r * kRedStride + lane] = u[r];
    if constexpr (N == P) reduce_order<P, N>(red + (N & 1) * RedBuf<P>::size, acc_lds, lane);
  } else {
#pragma unroll
    for (int r = 0; r < K; r++) red[r * kRedStride + lane] = u[r];
    reduce_order<P, N>(red, acc_lds, lane);
    __syncthreads();
  }
}

template <int P, int... Ns>
__device__ __forceinline__ void m2l_backward_all(const double *mp, double rpn, double ri,
                                                 double ezr, double ezi, double exr, double exi,
                                                 bool active, double *acc, double *red,
                                                 int lane, IntSeq<Ns...>) {
  double czp[P + 1], szp[P + 1], cxp[P + 1], sxp[P + 1];
  (m2l_backward_order<P, Ns>(mp, rpn, ri, ezr, ezi, exr, exi, czp, szp, cxp, sxp, active, acc,
                             red, lane),
   ...);
}

template <int P>
__global__ __launch_bounds__(64) void k_m2l(const int4 *__restrict__ chunks,
                                            const int *__restrict__ src_list,
                                            const double *__restrict__ cX,
                                            const unsigned char *__restrict__ has_sink,
                                            const double *__restrict__ M,
                                            const double *__restrict__ Md,
                                            double *__restrict__ L, double *__restrict__ Ld,
                                            int quirk, int stride) {
  constexpr int NT = nterm(P);
  __shared__ double acc[NT + kWave];  // + one spare slot per lane
  __shared__ __attribute__((aligned(16))) double red[RedBuf<P>::count * RedBuf<P>::size];
  const int lane = threadIdx.x;
  // (chunk, array) from the linear block id, contiguous per XCD
  const int nchunk = gridDim.x;
  const int item = xcd_contiguous(blockIdx.y * nchunk + blockIdx.x, 2 * nchunk);
  const int which = item >= nchunk;
  const int4 ch = chunks[item - which * nchunk];
  const int tgt = ch.x, start = ch.y, cnt = ch.z;
  if (!has_sink[tgt]) return;
  const double *Msrc = which ? Md : M;
  double *Lout = which ? Ld : L;
  const bool active = lane < cnt;
  const int j = src_list[start + (active ? lane : 0)];

  double dx = cX[3 * tgt + 0] - cX[3 * j + 0];
  double dy = cX[3 * tgt + 1] - cX[3 * j + 1];
  double dz = cX[3 * tgt + 2] - cX[3 * j + 2];
  double rxy2 = dx * dx + dy * dy;
  double rxy = sqrt(rxy2);
  double r = sqrt(rxy2 + dz * dz);
  double ri = 1.0 / r;
  // e^{i alpha_z} = (dY + i dX)/r_xy (1 on the z axis), e^{i alpha_x} = (dZ - i r_xy)/r
  double ezr = 1.0, ezi = 0.0;
  const double exr = dz * ri, exi = -rxy * ri;
  if (rxy > 0) {
    double q = 1.0 / rxy;
    ezr = dy * q;
    ezi = dx * q;
  }
  double mp[M2LRegs<P>::total];
  const double *src = Msrc + (size_t)j * stride;
  // Orders run from p down to 0: the large orders (biggest working set) run
  // while few rotated moments are live.
  // orders P .. P-kAhead+1 are gathered up front
  double v0[OrderBuf<P>::size], v1[OrderBuf<P - 1>::size], v2[OrderBuf<P - 2>::size],
      v3[OrderBuf<P - 3>::size];
  load_order<P>(src, v0);
  if constexpr (kAhead >= 2 && P - 1 >= 0) load_order<P - 1>(src, v1);
  if constexpr (kAhead >= 3 && P - 2 >= 0) load_order<P - 2>(src, v2);
  if constexpr (kAhead >= 4 && P - 3 >= 0) load_order<P - 3>(src, v3);
  m2l_forward_order<P, P>(src, quirk != 0, ezr, ezi, exr, exi, v0, v1, v2, v3, mp);
  // Fence the rotated moments: without it, IR-level code motion interleaves
  // the forward and backward halves and roughly doubles the live registers
  // (p = 12: 512 registers plus scratch spills instead of ~310, no spills).
#pragma unroll
  for (int t = 0; t < M2LRegs<P>::total; t++) asm volatile("" : "+v"(mp[t]));
  __builtin_amdgcn_sched_barrier(0);
  // translation factors (n+k)! r^-(n+k+1) are formed per order from the power
  // chain starting at r^-1; zero on idle lanes, which zeroes their whole
  // contribution
  const double rp0 = active ? ri : 0.0;
  m2l_backward_all<P>(mp, rp0, ri, ezr, ezi, exr, exi, active, acc, red, lane,
                      MakeIntSeq<P + 1>{});
  __syncthreads();
  double *dst = Lout + (size_t)tgt * stride;
  for (int t = lane; t < NT; t += kWave) atomicAdd(dst + t, acc[t]);
}

void m2l_init_tables(int p, int quirk) {
  static int done_p = -1, done_q = -1;
  if (done_p == p && done_q == quirk) return;
  double fact[2 * kMaxP + 2];
  fact[0] = 1.0;
  for (int i = 1; i < 2 * kMaxP + 2; i++) fact[i] = fact[i - 1] * i;
  hipMemcpyToSymbol(HIP_SYMBOL(c_fact), fact, sizeof(fact));
  done_p = p;
  done_q = quirk;
}

template <int P>
static void launch(const M2LArgs &a, hipStream_t s) {
  if (a.nchunks <= 0) return;
  dim3 grid(a.nchunks, 2);
  hipLaunchKernelGGL(k_m2l<P>, grid, dim3(64), 0, s, a.chunks, a.src, a.cX, a.has_sink, a.M,
                     a.Md, a.L, a.Ld, a.quirk, a.stride > 0 ? a.stride : nterm(P));
}

int launch_m2l(int p, const M2LArgs &a, hipStream_t s) {
  m2l_init_tables(p, a.quirk);
  switch (p) {
#define TC_CASE(P)\
  case P:\
    if constexpr (in_shard(P, TC_SHARD)) {\
      launch<P>(a, s);\
      return 0;\
    } else {\
      return -1;\
    }
    TC_FOR_EACH_P(TC_CASE)
#undef TC_CASE
    default:
      return -1;
  }
}

}  // namespace TC_SHARD_NS(TC_SHARD)
}  // namespace tcfmm
