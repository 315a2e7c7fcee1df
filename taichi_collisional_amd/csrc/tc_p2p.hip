// Near field on the device, FP64: P2P between leaf cells (taichi_with_jerks/
// kernel_fast_lean.h:499-699, the Vec8d double-precision branch) and an
// all-pairs direct sum used to measure the FMM error.
//
// Per pair (target i, source j), dx = x_j - x_i, dv = v_j - v_i:
//   pot += m/r,  a += m dx / r^3,  j += m dv / r^3 - 3 m (dx.dv) dx / r^5,
// pairs with r^2 == 0 contribute nothing (kernel_fast_lean.h:583-584).
//
// P2P kernel: one wave per target group (<= 64 bodies of a leaf).  With n_t
// targets the wave is split into G = floor(64/n_t) slices of n_t lanes; the
// source leaves of the leaf's list are copied tile by tile into LDS, slice s
// walks its own contiguous block of the tile's bodies, and the slices are
// summed through LDS at the end.  How a leaf is cut into groups (which sets
// n_t, so the share of the 64 lanes that work) is chosen on the host
// (p2p_split below).
#include "tc_common.h"
#include "tc_kernels.h"

#include <algorithm>
#include <vector>

namespace tcfmm {

struct Acc7 {
  double p, ax, ay, az, jx, jy, jz;
};

// Per pair (FP64, the all-pairs kernel): the jerk in one accumulator.
__device__ __forceinline__ void pair_fp64(double xi, double yi, double zi, double vxi, double vyi,
                                          double vzi, double xj, double yj, double zj, double vxj,
                                          double vyj, double vzj, double mj, Acc7 &a) {
  double dx = xj - xi, dy = yj - yi, dz = zj - zi;
  double dvx = vxj - vxi, dvy = vyj - vyi, dvz = vzj - vzi;
  double r2 = dx * dx + dy * dy + dz * dz;
  double vr = 3.0 * (dx * dvx + dy * dvy + dz * dvz);
  const double rq = rsqrt_fp64(r2);
  double ri = r2 > 0.0 ? rq : 0.0;
  double ri2 = ri * ri;
  double f = mj * ri;
  a.p += f;
  f *= ri2;
  a.ax += dx * f;
  a.ay += dy * f;
  a.az += dz * f;
  a.jx += dvx * f;
  a.jy += dvy * f;
  a.jz += dvz * f;
  f *= -ri2 * vr;
  a.jx += dx * f;
  a.jy += dy * f;
  a.jz += dz * f;
}

// Per pair (FP64, the P2P kernel's inner loop), 35 VALU instructions: the
// jerk's second term m (dx.dv) dx / r^5 goes to its own accumulator k without
// the factor -3 (applied once per target), and r^2 = 0 is handled by clearing
// the high word of the reciprocal square root: rsqrt_fp64(0) is +inf or a
// NaN, whose low word is zero for the hardware's default NaN, so the result
// is +0 (and at worst a denormal whose square underflows to 0: the potential
// then gains less than 1e-300 m, the force and jerk nothing) -- one 32-bit
// select where a double select takes two.
__device__ __forceinline__ void pair_fp64k(double xi, double yi, double zi, double vxi,
                                           double vyi, double vzi, const double4 &P,
                                           const double4 &V, Acc7 &a, double &kx, double &ky,
                                           double &kz) {
  const double dx = P.x - xi, dy = P.y - yi, dz = P.z - zi;
  const double dvx = V.x - vxi, dvy = V.y - vyi, dvz = V.z - vzi;
  const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const double vr = fma(dz, dvz, fma(dy, dvy, dx * dvx));
  const double rq = rsqrt_fp64(r2);
  const int hi = r2 > 0.0 ? __double2hiint(rq) : 0;
  const double ri = __hiloint2double(hi, __double2loint(rq));
  const double ri2 = ri * ri;
  double f = P.w * ri;
  a.p += f;
  f *= ri2;
  a.ax = fma(dx, f, a.ax);
  a.ay = fma(dy, f, a.ay);
  a.az = fma(dz, f, a.az);
  a.jx = fma(dvx, f, a.jx);
  a.jy = fma(dvy, f, a.jy);
  a.jz = fma(dvz, f, a.jz);
  const double g = f * (ri2 * vr);
  kx = fma(dx, g, kx);
  ky = fma(dy, g, ky);
  kz = fma(dz, g, kz);
}

// Tile sizes the P2P kernel is built for (source bodies per LDS tile).  LDS
// per wave: 64 (FP64) or 48 (mixed forms) bytes per slot for tile + 64 slots
// (64 slots of padding let every slice's block be whole, see the kernel),
// plus 512.
constexpr int kPadSlots = kWave;
#ifndef TC_P2P_UNROLL
#define TC_P2P_UNROLL 4
#endif
#ifndef TC_P2P_UNROLL2
#define TC_P2P_UNROLL2 2
#endif
#ifndef TC_P2P_ATTR
#define TC_P2P_ATTR
#endif

// Mixed-precision pair (FP32): source and target coordinates are offsets
// from a reference point of the target group (its first body), so they are
// small numbers and the FP32 differences keep ~7 digits of the separation.
// kHiLo: each offset is held as a float pair (hi + lo) and the difference is
// (hi_j - hi_i) + (lo_j - lo_i), accurate to FP32 rounding of the separation
// itself; otherwise the offsets are rounded to one float each.  Pairs with
// r^2 <= the close radius squared (so r = 0 always) contribute nothing here
// and set `close`; the caller evaluates them in FP64.  The jerk's second
// term is accumulated without the factor -3 (applied once at the end).
struct Acc10f {
  float p, ax, ay, az, jx, jy, jz, kx, ky, kz;
};

template <bool kHiLo>
__device__ __forceinline__ void pair_fp32(float xh, float yh, float zh, float xl, float yl,
                                          float zl, float vx, float vy, float vz,
                                          const float4 &S, const float4 &Sl, const float4 &V,
                                          float rc2, Acc10f &a, bool &close) {
  float dx = S.x - xh, dy = S.y - yh, dz = S.z - zh;
  if (kHiLo) {
    dx += Sl.x - xl;
    dy += Sl.y - yl;
    dz += Sl.z - zl;
  }
  const float dvx = V.x - vx, dvy = V.y - vy, dvz = V.z - vz;
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  const bool c = r2 <= rc2;  // r = 0 included, also when rc2 == 0
  close |= c;
  const float ri = c ? 0.0f : __builtin_amdgcn_rsqf(r2);
  const float ri2 = ri * ri;
  float f = S.w * ri;
  a.p += f;
  f *= ri2;
  a.ax = fmaf(dx, f, a.ax);
  a.ay = fmaf(dy, f, a.ay);
  a.az = fmaf(dz, f, a.az);
  a.jx = fmaf(dvx, f, a.jx);
  a.jy = fmaf(dvy, f, a.jy);
  a.jz = fmaf(dvz, f, a.jz);
  const float g = f * ri2 * fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
  a.kx = fmaf(dx, g, a.kx);
  a.ky = fmaf(dy, g, a.ky);
  a.kz = fmaf(dz, g, a.kz);
}

// The FP32 partial sums of one tile into the FP64 accumulators (blocked
// summation: FP32 runs are at most one tile long per lane).
__device__ __forceinline__ void flush_fp32(Acc10f &f, Acc7 &a, double &kx, double &ky,
                                           double &kz) {
  a.p += (double)f.p;
  a.ax += (double)f.ax;
  a.ay += (double)f.ay;
  a.az += (double)f.az;
  a.jx += (double)f.jx;
  a.jy += (double)f.jy;
  a.jz += (double)f.jz;
  kx += (double)f.kx;
  ky += (double)f.ky;
  kz += (double)f.kz;
  f = Acc10f{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
}

// One block (one wave) per work item: a target group (<= 64 kR consecutive
// bodies of one leaf) and a slice of at most a fixed number of entries of the
// leaf's P2P list (long lists -- a sparse halo leaf can touch most of the
// system -- are split so no single wave serializes them; partial sums are
// added atomically).  The leaf's P2P list is a list of source body ranges
// (one per source leaf); up to 64 ranges are fetched at once (one per lane),
// a wave prefix sum over their lengths places as many whole ranges as fit
// into an LDS tile of kTile bodies (position+mass, velocity as double4), and
// every lane then copies tile slots in parallel.  A single range longer than
// the tile (only when ncrit > kTile) is streamed through the tile in pieces.
//
// Lanes: each lane holds kR targets (u, u + nl, ..., nl = ceil(n_t / kR)
// lanes per slice), so one LDS read of a source body serves kR pairs; the
// wave runs G = floor(64 / nl) slices.  With fill bodies in the tile, slice
// s takes the B = ceil(fill / G) slots [s B, s B + B) (B rounded up to odd
// for 2 <= G <= 16, so the slices' first slots fall in different LDS banks);
// the slots past fill are padded with massless bodies.  The trip count B is
// the same on every lane (a scalar loop) and consecutive iterations read
// consecutive slots (immediate LDS offsets, no per-pair address arithmetic).
//
// kMode: kP2PFp64 (the reference's double-precision P2P), or the mixed forms
// kP2PMixed / kP2PMixedHiLo (kR = 1): sources are staged as FP32 offsets from
// the group's first body (three float4 per body), pairs are evaluated in FP32
// and summed in FP32 within a tile, tile sums in FP64; pairs closer than the
// close radius (r^2 < rc2) are redone in FP64 from the FP64 bodies.
enum { kP2PFp64 = 0, kP2PMixed = 1, kP2PMixedHiLo = 2 };

template <int kMode, int kTile, int kR>
__global__ __launch_bounds__(64) TC_P2P_ATTR void k_p2p(CellsDev c, BodiesDev b, ResultsDev r,
                                                        const int4 *__restrict__ groups,
                                                        const int *__restrict__ offs,
                                                        const int *__restrict__ count,
                                                        const int2 *__restrict__ ranges,
                                                        unsigned long long *__restrict__ npairs,
                                                        float rc2) {
  static_assert(kMode == kP2PFp64 || kR == 1, "mixed forms hold one target per lane");
  static_assert(kR == 1 || kR == 2, "targets per lane");
  constexpr int kCap = kTile + kPadSlots;  // slots, padding included
  // FP64: two double4 per slot; mixed forms: three float4 per slot
  constexpr int kBuf = kMode == kP2PFp64 ? 2 * kCap : (3 * kCap + 1) / 2;
  static_assert(kTile >= 64 && kBuf * 32 >= 7 * kR * kWave * 8, "tile holds the reduction");
  __shared__ double4 sbuf[kBuf];
  __shared__ int rstart[kWave], rexcl[kWave];
  double4 *sp = sbuf, *sv = sbuf + kCap;
  // mixed forms: three float4 per body in the same buffer
  float4 *fs = reinterpret_cast<float4 *>(sbuf), *fl = fs + kCap, *fv = fs + 2 * kCap;
  double(*red)[kWave] = reinterpret_cast<double(*)[kWave]>(sbuf);  // after the loop
  const int lane = threadIdx.x;
  // work item: target leaf, first target, list slice start, slice length | group size << 16
  // round-robin dispatch spreads the consecutive, heavy items of one large
  // list over all XCDs; a contiguous per-XCD range would pile them on one
  const int4 gw = groups[blockIdx.x];
  const int ci = gw.x;
  if (!c.has_sink[ci]) return;
  const int nt = gw.w >> 16;
  const int nl = (nt + kR - 1) / kR;
  const int G = kWave / nl;
  const int s = lane / nl, u = lane - s * nl;
  const bool act = s < G;
  // the lane's targets u + k nl; a slot past the group repeats target u (its
  // sums are dropped)
  int ib[kR];
  double xi[kR], yi[kR], zi[kR], vxi[kR], vyi[kR], vzi[kR];
  Acc7 a[kR];
  double kx[kR], ky[kR], kz[kR];  // jerk terms m (dx.dv) dx / r^5, factor -3 applied at the end
#pragma unroll
  for (int k = 0; k < kR; k++) {
    const int t = u + k * nl;
    ib[k] = c.body[ci] + gw.y + (t < nt ? t : u);
    xi[k] = b.x[ib[k]];
    yi[k] = b.y[ib[k]];
    zi[k] = b.z[ib[k]];
    vxi[k] = b.vx[ib[k]];
    vyi[k] = b.vy[ib[k]];
    vzi[k] = b.vz[ib[k]];
    a[k] = Acc7{0, 0, 0, 0, 0, 0, 0};
    kx[k] = ky[k] = kz[k] = 0.0;
  }
  // mixed forms: the group's reference point and the target's offsets
  const int i0 = c.body[ci] + gw.y;
  const double x0 = b.x[i0], y0 = b.y[i0], z0 = b.z[i0];
  const double u0 = b.vx[i0], v0 = b.vy[i0], w0 = b.vz[i0];
  float txh = 0, tyh = 0, tzh = 0, txl = 0, tyl = 0, tzl = 0, tvx = 0, tvy = 0, tvz = 0;
  Acc10f af = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (kMode != kP2PFp64) {
    const double ox = xi[0] - x0, oy = yi[0] - y0, oz = zi[0] - z0;
    txh = (float)ox;
    tyh = (float)oy;
    tzh = (float)oz;
    txl = (float)(ox - (double)txh);
    tyl = (float)(oy - (double)tyh);
    tzl = (float)(oz - (double)tzh);
    tvx = (float)(vxi[0] - u0);
    tvy = (float)(vyi[0] - v0);
    tvz = (float)(vzi[0] - w0);
  }
  // stage source body bq in tile slot q; bq < 0 stages a massless padding
  // body (FP64: at the group's first body, where the r^2 = 0 branch drops it;
  // mixed forms: 1e18 away, far outside any close radius, so it never sends
  // a wave into the close-pair pass, and its index -1 is skipped there)
  auto stage = [&](int q, int bq) {
    if (kMode == kP2PFp64) {
      if (bq < 0) {
        sp[q] = make_double4(x0, y0, z0, 0.0);
        sv[q] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
      }
      const double mq = (b.flags[bq] & kSource) ? b.q[bq] : 0.0;
      sp[q] = make_double4(b.x[bq], b.y[bq], b.z[bq], mq);
      sv[q] = make_double4(b.vx[bq], b.vy[bq], b.vz[bq], 0.0);
    } else {
      if (bq < 0) {
        fs[q] = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
        fl[q] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
        fv[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
      }
      const double mq = (b.flags[bq] & kSource) ? b.q[bq] : 0.0;
      const double ox = b.x[bq] - x0, oy = b.y[bq] - y0, oz = b.z[bq] - z0;
      const float hx = (float)ox, hy = (float)oy, hz = (float)oz;
      fs[q] = make_float4(hx, hy, hz, (float)mq);
      fl[q] = make_float4((float)(ox - (double)hx), (float)(oy - (double)hy),
                          (float)(oz - (double)hz), __int_as_float(bq));
      fv[q] = make_float4((float)(b.vx[bq] - u0), (float)(b.vy[bq] - v0),
                          (float)(b.vz[bq] - w0), 0.0f);
    }
  };
  const int l0 = offs[ci] + gw.z, ln = gw.w & 0xffff;
  long long nsrc = 0;
  int e = 0, off = 0;
  // list entries e + lane, loaded one tile ahead (the next tile's entries are
  // in flight while this tile is staged and evaluated)
  int2 rgn = (lane < ln) ? ranges[l0 + lane] : make_int2(0, 0);
  while (e < ln) {
    const int2 rg = rgn;
    int tot;
    const int excl = wave_excl_scan(rg.y, &tot);
    const bool fits = (e + lane < ln) && (excl + rg.y <= kTile);
    const unsigned long long fm = __ballot(fits);
    const int nfit = __popcll(fm);  // ranges 0..nfit-1 fit (prefix sums grow)
    int fill;
    if (nfit > 0) {
      rstart[lane] = rg.x;
      rexcl[lane] = excl;
      fill = __shfl(excl + rg.y, nfit - 1, kWave);
      __syncthreads();
      e += nfit;
      rgn = (e + lane < ln) ? ranges[l0 + e + lane] : make_int2(0, 0);
      for (int q = lane; q < fill; q += kWave) {
        int lo = 0, hi = nfit - 1;  // last range with rexcl <= q
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (rexcl[mid] <= q)
            lo = mid;
          else
            hi = mid - 1;
        }
        stage(q, rstart[lo] + q - rexcl[lo]);
      }
    } else {
      // first range longer than the tile: stream it piece by piece
      const int base = __shfl(rg.x, 0, kWave), len = __shfl(rg.y, 0, kWave);
      fill = min(kTile, len - off);
      for (int q = lane; q < fill; q += kWave) stage(q, base + off + q);
      off += fill;
      if (off >= len) {
        off = 0;
        e += 1;
        rgn = (e + lane < ln) ? ranges[l0 + e + lane] : make_int2(0, 0);
      }
    }
    nsrc += fill;
    // slice blocks: G * B - fill < 2 G <= kPadSlots + G slots of padding for
    // G <= 16 (B rounded to odd), < G otherwise; both within kCap
    int B = __builtin_amdgcn_readfirstlane((fill + G - 1) / G);
    if (G > 1 && G <= 16) B |= 1;
    if (lane < G * B - fill) stage(fill + lane, -1);
    __syncthreads();
    if (act) {
      const int q0 = s * B;
      if (kMode == kP2PFp64) {
        const double4 *P = sp + q0, *V = sv + q0;
#pragma unroll(kR == 1 ? TC_P2P_UNROLL : TC_P2P_UNROLL2)
        for (int i = 0; i < B; i++) {
          const double4 Pi = P[i], Vi = V[i];
#pragma unroll
          for (int k = 0; k < kR; k++)
            pair_fp64k(xi[k], yi[k], zi[k], vxi[k], vyi[k], vzi[k], Pi, Vi, a[k], kx[k], ky[k],
                       kz[k]);
        }
      } else {
        constexpr bool hl = kMode == kP2PMixedHiLo;
        bool close = false;
        const float4 *S = fs + q0, *Sl = fl + q0, *Vv = fv + q0;
#pragma unroll 2
        for (int i = 0; i < B; i++)
          pair_fp32<hl>(txh, tyh, tzh, txl, tyl, tzl, tvx, tvy, tvz, S[i], Sl[i], Vv[i], rc2,
                        af, close);
        flush_fp32(af, a[0], kx[0], ky[0], kz[0]);
        if (close) {
          // the few pairs inside the close radius (or at r = 0), in FP64 from
          // the bodies themselves; the same FP32 test picks them
          for (int i = 0; i < B; i++) {
            Acc10f dummy = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            bool cq = false;
            pair_fp32<hl>(txh, tyh, tzh, txl, tyl, tzl, tvx, tvy, tvz, S[i], Sl[i], Vv[i], rc2,
                          dummy, cq);
            const int bq = __float_as_int(Sl[i].w);
            if (cq && bq >= 0) {
              const double mq = (b.flags[bq] & kSource) ? b.q[bq] : 0.0;
              pair_fp64(xi[0], yi[0], zi[0], vxi[0], vyi[0], vzi[0], b.x[bq], b.y[bq], b.z[bq],
                        b.vx[bq], b.vy[bq], b.vz[bq], mq, a[0]);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (lane == 0) atomicAdd(npairs, (unsigned long long)(nsrc * nt));
#pragma unroll
  for (int k = 0; k < kR; k++) {
    a[k].jx -= 3.0 * kx[k];
    a[k].jy -= 3.0 * ky[k];
    a[k].jz -= 3.0 * kz[k];
    red[7 * k + 0][lane] = a[k].p;
    red[7 * k + 1][lane] = a[k].ax;
    red[7 * k + 2][lane] = a[k].ay;
    red[7 * k + 3][lane] = a[k].az;
    red[7 * k + 4][lane] = a[k].jx;
    red[7 * k + 5][lane] = a[k].jy;
    red[7 * k + 6][lane] = a[k].jz;
  }
  __syncthreads();
  if (s == 0) {
#pragma unroll
    for (int k = 0; k < kR; k++) {
      if (u + k * nl >= nt || !(b.flags[ib[k]] & kSink)) continue;
      double o[7];
      for (int q = 0; q < 7; q++) {
        double v = 0;
        for (int g = 0; g < G; g++) v += red[7 * k + q][g * nl + u];
        o[q] = v;
      }
      // several work items may cover one target group (long lists are split)
      atomicAdd(&r.pot[ib[k]], o[0]);
      atomicAdd(&r.ax[ib[k]], o[1]);
      atomicAdd(&r.ay[ib[k]], o[2]);
      atomicAdd(&r.az[ib[k]], o[3]);
      atomicAdd(&r.jx[ib[k]], o[4]);
      atomicAdd(&r.jy[ib[k]], o[5]);
      atomicAdd(&r.jz[ib[k]], o[6]);
    }
  }
}

// Source body range (first body, count) of every P2P list entry.
__global__ void k_p2p_ranges(CellsDev c, const int *__restrict__ src, int n,
                             int2 *__restrict__ ranges) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int j = src[k];
  ranges[k] = make_int2(c.body[j], c.nbody[j]);
}

int launch_p2p_ranges(CellsDev c, const int *src, int n, int2 *ranges, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_p2p_ranges, dim3((n + 255) / 256), dim3(256), 0, s, c, src, n, ranges);
  return 0;
}

template <int kMode, int kTile, int kR>
static void launch_p2p_t(CellsDev c, BodiesDev b, ResultsDev r, const int4 *groups, int ngroups,
                         const int *offs, const int *count, const int2 *ranges,
                         unsigned long long *npairs, float rc2, hipStream_t s) {
  hipLaunchKernelGGL((k_p2p<kMode, kTile, kR>), dim3(ngroups), dim3(64), 0, s, c, b, r, groups,
                     offs, count, ranges, npairs, rc2);
}

template <int kMode, int kR>
static int launch_p2p_m(int tile, CellsDev c, BodiesDev b, ResultsDev r, const int4 *groups,
                        int ngroups, const int *offs, const int *count, const int2 *ranges,
                        unsigned long long *npairs, float rc2, hipStream_t s) {
  switch (tile) {
    case 64:
      launch_p2p_t<kMode, 64, kR>(c, b, r, groups, ngroups, offs, count, ranges, npairs, rc2, s);
      return 0;
    case 128:
      launch_p2p_t<kMode, 128, kR>(c, b, r, groups, ngroups, offs, count, ranges, npairs, rc2,
                                   s);
      return 0;
    case 192:
      launch_p2p_t<kMode, 192, kR>(c, b, r, groups, ngroups, offs, count, ranges, npairs, rc2,
                                   s);
      return 0;
    default:
      return -1;
  }
}

bool p2p_tile_supported(int tile) { return tile == 64 || tile == 128 || tile == 192; }

int p2p_targets_per_lane(int mode, int requested) {
  return mode == kP2PFp64 && requested == 2 ? 2 : 1;
}

int launch_p2p(CellsDev c, BodiesDev b, ResultsDev r, const int4 *groups, int ngroups,
               const int *offs, const int *count, const int2 *ranges,
               unsigned long long *npairs, hipStream_t s, int mode, double close_radius,
               int tile, int per_lane) {
  if (ngroups <= 0) return 0;
  const float rc2 = (float)(close_radius * close_radius);
  if (mode == kP2PMixed)
    return launch_p2p_m<kP2PMixed, 1>(tile, c, b, r, groups, ngroups, offs, count, ranges, npairs,
                                      rc2, s);
  if (mode == kP2PMixedHiLo)
    return launch_p2p_m<kP2PMixedHiLo, 1>(tile, c, b, r, groups, ngroups, offs, count, ranges,
                                          npairs, rc2, s);
  if (per_lane == 2)
    return launch_p2p_m<kP2PFp64, 2>(tile, c, b, r, groups, ngroups, offs, count, ranges, npairs,
                                     0.0f, s);
  return launch_p2p_m<kP2PFp64, 1>(tile, c, b, r, groups, ngroups, offs, count, ranges, npairs,
                                   0.0f, s);
}

// ---- target groups of a leaf (host)
//
// With R targets per lane a g-body group takes nl = ceil(g / R) lanes per
// slice and the wave runs G = floor(64 / nl) slices.  Per source tile a
// group costs R ceil(fill / G) pair rounds plus the tile's staging
// (stage_rounds, in pair rounds); the leaf of nb bodies is cut into the
// groups of least total cost, found by dynamic programming over nb with
// group sizes 1..64 R.  E.g. 48 bodies at R = 1: 32 + 16 (two groups, every
// lane busy) rather than 24 + 24 (48 of 64 lanes).
void P2PSplit::configure(double stage_rounds, int fill, int per_lane) {
  if (stage_rounds == stage_ && fill == fill_ && per_lane == per_lane_) return;
  stage_ = stage_rounds;
  fill_ = fill;
  per_lane_ = per_lane;
  first_.assign(1, 0);
  sizes_.clear();
  cost_.assign(1, 0.0);
  choice_.assign(1, 0);
}

void P2PSplit::grow(int nmax) {
  const int have = (int)cost_.size() - 1;
  if (nmax <= have) return;
  cost_.resize(nmax + 1);
  choice_.resize(nmax + 1);
  for (int n = have + 1; n <= nmax; n++) {
    double best = 1e300;
    int arg = 1;
    for (int g = 1; g <= std::min(n, kWave * per_lane_); g++) {
      const int G = kWave / ((g + per_lane_ - 1) / per_lane_);
      const double c = cost_[n - g] + (double)(per_lane_ * ((fill_ + G - 1) / G)) + stage_;
      if (c < best - 1e-9) {
        best = c;
        arg = g;
      }
    }
    cost_[n] = best;
    choice_[n] = arg;
  }
  // flattened group sizes per nb, largest group first
  first_.resize(nmax + 2);
  for (int n = have + 1; n <= nmax; n++) {
    first_[n] = (int)sizes_.size();
    std::vector<int> g;
    for (int m = n; m > 0; m -= choice_[m]) g.push_back(choice_[m]);
    std::sort(g.rbegin(), g.rend());
    sizes_.insert(sizes_.end(), g.begin(), g.end());
    first_[n + 1] = (int)sizes_.size();
  }
}

// All-pairs sum for a list of targets (sorted-order indices), FP64.  The
// sources are split into gridDim.y slabs so enough blocks run; each slab adds
// its partial sums into the (zeroed) outputs.
__global__ __launch_bounds__(256) void k_direct(BodiesDev b, int n, const int *__restrict__ tg,
                                                int nt, int slab, double *__restrict__ pot,
                                                double *__restrict__ acc,
                                                double *__restrict__ jerk) {
  __shared__ double4 sp[256], sv[256];
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int i = k < nt ? tg[k] : 0;
  const double xi = b.x[i], yi = b.y[i], zi = b.z[i], vxi = b.vx[i], vyi = b.vy[i],
               vzi = b.vz[i];
  Acc7 a = {0, 0, 0, 0, 0, 0, 0};
  const int j0 = blockIdx.y * slab, j1 = min(n, j0 + slab);
  for (int base = j0; base < j1; base += 256) {
    __syncthreads();
    int j = base + threadIdx.x;
    if (j < j1) {
      double mq = (b.flags[j] & kSource) ? b.q[j] : 0.0;
      sp[threadIdx.x] = make_double4(b.x[j], b.y[j], b.z[j], mq);
      sv[threadIdx.x] = make_double4(b.vx[j], b.vy[j], b.vz[j], 0.0);
    }
    __syncthreads();
    int lim = min(256, j1 - base);
    for (int q = 0; q < lim; q++) {
      const double4 P = sp[q], V = sv[q];
      pair_fp64(xi, yi, zi, vxi, vyi, vzi, P.x, P.y, P.z, V.x, V.y, V.z, P.w, a);
    }
  }
  if (k < nt) {
    atomicAdd(&pot[k], a.p);
    atomicAdd(&acc[3 * k], a.ax);
    atomicAdd(&acc[3 * k + 1], a.ay);
    atomicAdd(&acc[3 * k + 2], a.az);
    atomicAdd(&jerk[3 * k], a.jx);
    atomicAdd(&jerk[3 * k + 1], a.jy);
    atomicAdd(&jerk[3 * k + 2], a.jz);
  }
}

__global__ void k_zero_f64(double *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

void launch_zero_f64(double *p, size_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

int launch_direct(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                  double *jerk, hipStream_t s) {
  if (nt <= 0) return 0;
  hipMemsetAsync(pot, 0, sizeof(double) * nt, s);
  hipMemsetAsync(acc, 0, sizeof(double) * 3 * nt, s);
  hipMemsetAsync(jerk, 0, sizeof(double) * 3 * nt, s);
  return launch_direct_sum(b, n, targets, nt, pot, acc, jerk, s);
}

// The all-pairs sum alone: the outputs must already be zero.
int launch_direct_sum(BodiesDev b, int n, const int *targets, int nt, double *pot, double *acc,
                      double *jerk, hipStream_t s) {
  if (nt <= 0) return 0;
  int bx = (nt + 255) / 256;
  int slabs = (2048 + bx - 1) / bx;  // aim for >= 2048 blocks
  int slab = (n + slabs - 1) / slabs;
  slab = ((slab + 255) / 256) * 256;
  slabs = (n + slab - 1) / slab;
  hipLaunchKernelGGL(k_direct, dim3(bx, slabs), dim3(256), 0, s, b, n, targets, nt, slab, pot,
                     acc, jerk);
  return 0;
}

// ---- results back to the caller's body order (after the downward pass)
__global__ void k_scatter_results(BodiesDev b, ResultsDev r, int n, double *__restrict__ pot,
                                  double *__restrict__ acc, double *__restrict__ jerk) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int i = b.order[k];
  if (pot) pot[i] = r.pot[k];
  if (acc) {
    acc[3 * (size_t)i] = r.ax[k];
    acc[3 * (size_t)i + 1] = r.ay[k];
    acc[3 * (size_t)i + 2] = r.az[k];
  }
  if (jerk) {
    jerk[3 * (size_t)i] = r.jx[k];
    jerk[3 * (size_t)i + 1] = r.jy[k];
    jerk[3 * (size_t)i + 2] = r.jz[k];
  }
}

// near + far result sets, summed per body in tree order, to the caller's order
__global__ void k_scatter_results_sum(BodiesDev b, ResultsDev r, ResultsDev f, int n,
                                      double *__restrict__ pot, double *__restrict__ acc,
                                      double *__restrict__ jerk) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int i = b.order[k];
  if (pot) pot[i] = r.pot[k] + f.pot[k];
  if (acc) {
    acc[3 * (size_t)i] = r.ax[k] + f.ax[k];
    acc[3 * (size_t)i + 1] = r.ay[k] + f.ay[k];
    acc[3 * (size_t)i + 2] = r.az[k] + f.az[k];
  }
  if (jerk) {
    jerk[3 * (size_t)i] = r.jx[k] + f.jx[k];
    jerk[3 * (size_t)i + 1] = r.jy[k] + f.jy[k];
    jerk[3 * (size_t)i + 2] = r.jz[k] + f.jz[k];
  }
}

int launch_scatter_results_sum(BodiesDev b, ResultsDev r, ResultsDev f, int n, double *pot,
                               double *acc3, double *jerk3, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_scatter_results_sum, dim3((n + 255) / 256), dim3(256), 0, s, b, r, f, n,
                     pot, acc3, jerk3);
  return 0;
}

int launch_scatter_results(BodiesDev b, ResultsDev r, int n, double *pot, double *acc3,
                           double *jerk3, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_scatter_results, dim3((n + 255) / 256), dim3(256), 0, s, b, r, n, pot,
                     acc3, jerk3);
  return 0;
}

}  // namespace tcfmm
