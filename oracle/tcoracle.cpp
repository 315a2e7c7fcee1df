// tcoracle -- CPU restatement of taichi_collisional's FMM force + potential +
// jerk evaluation.  TEST INFRASTRUCTURE ONLY: it is the checker the GPU path is
// compared with (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).
// The product never links or calls it.
//
// Every routine restates the reference algorithm in its own code; the
// reference files and lines each one follows are cited (paths relative to the
// reference's taichi_with_jerks/ directory unless stated).  FP64 throughout,
// OpenMP over target cells (the reference uses OpenMP tasks + atomics; the
// results differ only in summation order).
//
// Pinning (what fixes each part beyond this restatement), exercised by
// tests/test_oracle_*.py and tests/test_master_oracle.py:
//   harmonics      closed forms (Legendre) in tests
//   swap matrices  exact rational derivation + reference file hashes
//   M2L_rotate     == O(p^4) M2L to round-off; == reference M2L_rotate (KAT)
//   M2M/L2L        exactness invariants + reference KATs
//   P2P/direct     long-double brute force + reference KATs
//   tree/miniball  reference tree (same shift) + minimum-ball optimality checks
//   whole FMM      reference end-to-end goldens + FP64 direct sum
//   master direct + time step, -DFMM passes, step loop (orc_master_*)
//                  goldens from the compiled master reference (direct sums,
//                  walk/M2L/P2P passes, do_evolve at N = 100, 400 and 5000)
//                  + numpy brute force for the forces
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

#include "../taichi_collisional_amd/csrc/swap_tables.h"
#include "../taichi_collisional_amd/csrc/tdesign.h"

namespace orc {

typedef double real;
typedef std::complex<double> cplx;

static inline int ix(int n, int m) { return n * n + n + m; }

// The walks' squared distance norm(dX) (kernel_fast_lean.h:113-116, called by
// every MAC in traverse_eager.h) as the reference's x86-64 build evaluates it:
// GCC at -O3 contracts x*x + y*y + z*z into fma(z, z, fma(x, x, y*y)).  The
// p = 2 pre-pass compares it with 3(Ri+Rj)^2 on grid-aligned cells, where
// diagonal neighbours tie in exact arithmetic and this rounding decides them.
static inline real mac_norm(const real *d) {
  return std::fma(d[2], d[2], std::fma(d[0], d[0], d[1] * d[1]));
}
// One term of the error-MAC sums (traverse_eager.h:80-88; master :82-90), as
// the same build contracts it: E = fma(Pn * power, comb, E).
static inline real mac_term(real E, real pn, real pw, real comb) {
  return std::fma(pn * pw, comb, E);
}

// ---------------------------------------------------------------- tables ----
// initKernel (kernel_fast_lean.h:124-162).  With quirk=1 the factorial table
// is filled only below index 2p, as the reference does (:129), so fact[2p]=0.
struct Tables {
  int P, NTERM;
  std::vector<real> fact, scale, scale_inv;  // scale = (n+|m|)!(n-|m|)! per term
  std::vector<real> comb;                     // C(P,k), k=0..P
  explicit Tables(int p, int quirk) : P(p), NTERM((p + 1) * (p + 1)) {
    fact.assign(2 * p + 1, 0.0);
    fact[0] = 1.0;
    int top = quirk ? 2 * p : 2 * p + 1;
    for (int i = 1; i < top; i++) fact[i] = (real)i * fact[i - 1];
    scale.assign(NTERM, 0.0);
    scale_inv.assign(NTERM, 0.0);
    for (int n = 0; n <= p; n++)
      for (int m = -n; m <= n; m++) {
        int a = std::abs(m);
        real f = fact[n + a] * fact[n - a];
        scale[ix(n, m)] = f;
        scale_inv[ix(n, m)] = 1.0 / f;
      }
    comb.assign(p + 1, 0.0);
    for (int k = 0; k <= p; k++) comb[k] = fact[p] / fact[k] / fact[p - k];
  }
};

// ------------------------------------------------------------ harmonics ----
// make_Gnm (kernel_fast_lean.h:237-268): regular solid harmonics.
static void make_G(const real *d, cplx *G, int P) {
  real r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  G[0] = cplx(1, 0);
  for (int n = 1; n <= P; n++) {
    G[ix(n, n)] = G[ix(n - 1, n - 1)] * cplx(d[0] / (2 * n), d[1] / (2 * n));
    for (int m = 0; m < n - 1; m++) {
      real q = (real)(n * n - m * m);
      G[ix(n, m)] = (2 * n - 1.) / q * d[2] * G[ix(n - 1, m)] - r2 / q * G[ix(n - 2, m)];
    }
    G[ix(n, n - 1)] = d[2] * G[ix(n - 1, n - 1)];
    real s = -1;
    for (int m = 1; m <= n; m++) {
      G[ix(n, -m)] = std::conj(G[ix(n, m)] * s);
      s = -s;
    }
  }
}

// make_Gnmdot (kernel_fast_lean.h:272-304): time derivative of the harmonics
// for a point moving with velocity v.  The reference reads G[-1] for n = 0, 1
// (a slot before the array); the harmonic Y_{n-1}^{-1} it stands for does not
// exist, so it is taken as 0 here.
static void make_Gdot(const real *v, const cplx *G, cplx *Gd, int P) {
  cplx etadot(-v[0] * 0.5, v[1] * 0.5), zdot(v[2], 0), ksidot(v[0] * 0.5, v[1] * 0.5);
  int NT = (P + 1) * (P + 1);
  for (int i = 0; i < NT; i++) Gd[i] = 0;
  for (int n = 0; n <= P; n++) {
    int s = ix(n, 0), sl = (n - 1) * (n - 1) + (n - 1);
    for (int m = 0; m <= n - 2; m++) Gd[s + m] += G[sl + m + 1] * etadot;
    for (int m = 0; m <= n - 1; m++) Gd[s + m] += G[sl + m] * zdot;
    for (int m = 0; m <= n; m++) {
      int k = sl + m - 1;
      if (k >= 0) Gd[s + m] += G[k] * ksidot;
    }
    real sg = -1;
    for (int m = 1; m <= n; m++) {
      Gd[s - m] = std::conj(Gd[s + m] * sg);
      sg = -sg;
    }
  }
}

// make_Tnm (kernel_fast_lean.h:195-235): singular solid harmonics.
static void make_T(const real *d, cplx *T, int P) {
  real r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  real ir2 = 1 / r2;
  T[0] = cplx(std::sqrt(ir2), 0);
  for (int n = 1; n <= P; n++) {
    T[ix(n, n)] = T[ix(n - 1, n - 1)] * cplx(d[0] * (2 * n - 1) * ir2, d[1] * (2 * n - 1) * ir2);
    for (int m = 0; m < n - 1; m++)
      T[ix(n, m)] = ir2 * (2 * n - 1.) * d[2] * T[ix(n - 1, m)] -
                    ir2 * (real)((n - 1) * (n - 1) - m * m) * T[ix(n - 2, m)];
    T[ix(n, n - 1)] = ir2 * (2 * n - 1.) * d[2] * T[ix(n - 1, n - 1)];
    real s = -1;
    for (int m = 1; m <= n; m++) {
      T[ix(n, -m)] = std::conj(T[ix(n, m)] * s);
      s = -s;
    }
  }
}

// real_2_complex / complex_2_real (kernel_fast_lean.h:164-193).
static void r2c(const real *a, cplx *c, int P) {
  for (int n = 0; n <= P; n++) {
    int s = ix(n, 0);
    c[s] = cplx(a[s], 0);
    real f = -1;
    for (int m = 1; m <= n; m++) {
      c[s - m] = cplx(a[s + m] * f, -a[s - m] * f);
      c[s + m] = cplx(a[s + m], a[s - m]);
      f = -f;
    }
  }
}
static void c2r(const cplx *c, real *a, int P) {
  for (int n = 0; n <= P; n++) {
    int s = ix(n, 0);
    a[s] = c[s].real();
    for (int m = 1; m <= n; m++) {
      a[s - m] = c[s + m].imag();
      a[s + m] = c[s + m].real();
    }
  }
}

// swap_x_z (kernel_rotate.h:6-53): apply S_n of every order n <= P in place.
static void swap_xz(real *a, int P) {
  real tmp[64];
  for (int n = 0; n <= P; n++) {
    int base = n * n, k = 2 * n + 1;
    for (int r = 0; r < k; r++) {
      real acc = 0;
      for (int e = tcfmm::kSwapRowPtr[base + r]; e < tcfmm::kSwapRowPtr[base + r + 1]; e++)
        acc += tcfmm::kSwapVal[e] * a[base + tcfmm::kSwapCol[e]];
      tmp[r] = acc;
    }
    for (int r = 0; r < k; r++) a[base + r] = tmp[r];
  }
}

// ------------------------------------------------------------- miniball ----
// Smallest enclosing ball (the semantics of the Miniball the reference uses,
// miniball.hpp, at kernel_fast_lean.h:733-763 and :875-935).  Own algorithm:
// iterative support-set refinement.  Start from one point; repeatedly take
// the point of largest excess |q-c|^2 - r^2; if it is outside, replace the
// ball by the smallest ball of (support + q), found by trying every subset of
// the support together with q (at most 15 circumballs of <= 4 points) and
// keeping the smallest one that encloses support + q.  The radius grows
// strictly, so the loop ends; at the end every point is inside and the ball is
// the minimum ball of its own support, hence of all points.
struct Ball {
  real c[3];
  real r2;
};

// Circumball of k <= 4 points with centre in their affine hull.
static bool circumball(const real (*p)[3], int k, Ball &b) {
  if (k == 1) {
    for (int d = 0; d < 3; d++) b.c[d] = p[0][d];
    b.r2 = 0;
    return true;
  }
  real v[3][3];
  for (int i = 1; i < k; i++)
    for (int d = 0; d < 3; d++) v[i - 1][d] = p[i][d] - p[0][d];
  int q = k - 1;
  real A[3][3], rhs[3];
  for (int i = 0; i < q; i++) {
    for (int j = 0; j < q; j++) A[i][j] = v[i][0] * v[j][0] + v[i][1] * v[j][1] + v[i][2] * v[j][2];
    rhs[i] = 0.5 * A[i][i];
  }
  real lam[3] = {0, 0, 0};
  if (q == 1) {
    if (!(A[0][0] > 0)) return false;
    lam[0] = 0.5;
  } else if (q == 2) {
    real det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    if (!(std::fabs(det) > 1e-13 * A[0][0] * A[1][1])) return false;
    lam[0] = (rhs[0] * A[1][1] - A[0][1] * rhs[1]) / det;
    lam[1] = (A[0][0] * rhs[1] - rhs[0] * A[1][0]) / det;
  } else {
    real c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    real c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    real c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    real det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    if (!(std::fabs(det) > 1e-13 * A[0][0] * A[1][1] * A[2][2])) return false;
    real inv[3][3];
    inv[0][0] = c00;
    inv[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    inv[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    inv[1][0] = c01;
    inv[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    inv[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    inv[2][0] = c02;
    inv[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    inv[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    for (int i = 0; i < 3; i++)
      lam[i] = (inv[i][0] * rhs[0] + inv[i][1] * rhs[1] + inv[i][2] * rhs[2]) / det;
  }
  for (int d = 0; d < 3; d++) {
    real c = p[0][d];
    for (int i = 0; i < q; i++) c += lam[i] * v[i][d];
    b.c[d] = c;
  }
  real r2 = 0;
  for (int i = 0; i < k; i++) {
    real dd = 0;
    for (int d = 0; d < 3; d++) dd += (p[i][d] - b.c[d]) * (p[i][d] - b.c[d]);
    r2 = std::max(r2, dd);
  }
  b.r2 = r2;
  return true;
}

static const real kBallTol = 1e-13;

template <class Get>
static Ball minball(int n, Get get) {
  Ball b;
  real sup[4][3];
  int ns = 1;
  get(0, sup[0]);
  for (int d = 0; d < 3; d++) b.c[d] = sup[0][d];
  b.r2 = 0;
  for (int iter = 0; iter < 10000; iter++) {
    real best = -1;
    real q[3], bq[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
      get(i, q);
      real e = 0;
      for (int d = 0; d < 3; d++) e += (q[d] - b.c[d]) * (q[d] - b.c[d]);
      e -= b.r2;
      if (e > best) {
        best = e;
        bq[0] = q[0], bq[1] = q[1], bq[2] = q[2];
      }
    }
    if (best <= kBallTol * b.r2 || best <= 0) break;
    // smallest ball of support + bq with bq on the boundary
    Ball bestb;
    bestb.r2 = -1;
    int bestmask = 0;
    for (int mask = 0; mask < (1 << ns); mask++) {
      int cnt = __builtin_popcount(mask);
      if (cnt > 3) continue;
      real pts[4][3];
      int k = 0;
      for (int d = 0; d < 3; d++) pts[k][d] = bq[d];
      k++;
      for (int s = 0; s < ns; s++)
        if (mask & (1 << s)) {
          for (int d = 0; d < 3; d++) pts[k][d] = sup[s][d];
          k++;
        }
      Ball cb;
      if (!circumball(pts, k, cb)) continue;
      if (bestb.r2 >= 0 && cb.r2 >= bestb.r2) continue;
      bool ok = true;
      for (int s = 0; s < ns && ok; s++) {
        if (mask & (1 << s)) continue;
        real dd = 0;
        for (int d = 0; d < 3; d++) dd += (sup[s][d] - cb.c[d]) * (sup[s][d] - cb.c[d]);
        if (dd > cb.r2 * (1 + 1e-12)) ok = false;
      }
      if (!ok) continue;
      bestb = cb;
      bestmask = mask;
    }
    if (bestb.r2 < 0) break;  // numerically degenerate; keep current ball
    real nsup[4][3];
    int k = 0;
    for (int d = 0; d < 3; d++) nsup[k][d] = bq[d];
    k++;
    for (int s = 0; s < ns; s++)
      if (bestmask & (1 << s)) {
        for (int d = 0; d < 3; d++) nsup[k][d] = sup[s][d];
        k++;
      }
    std::memcpy(sup, nsup, sizeof(sup));
    ns = k;
    if (bestb.r2 <= b.r2) {  // no progress possible in floating point
      b = bestb;
      break;
    }
    b = bestb;
  }
  return b;
}

// 64-point spherical design used to sample grandchild spheres in M2M.
static const real *tdesign() { return tcfmm::kTDesign; }

// ------------------------------------------------------------------ tree ----
struct Body {
  real X[3], V[3], q;
  real F[3], J[3], p, acc_old;
  real ts;  // master variant: time-step sum
  unsigned Morton[3];
  int index;
  bool issink, issource;
};

struct Cell {
  int NCHILD, NBODY, CHILD, BODY, level;
  real X[3], R;
  real Xg[3], Rg;  // geometric centre/half-width from the tree build
  real min_acc;
  bool has_sink;
  real *M, *L, *Md, *Ld, *Pn;
};

struct Params {
  int p, ncrit;
  double theta;
  int mac;
  double force_accuracy;
  int quirk;
  double shift_frac[3];
  int direct_below;
  int nthreads;
};

struct Stats {
  long long n_m2l, n_p2p_cells, n_p2p_pairs;
  int ncells, nleaves, depth;
};

// Body::operator< (exafmm.h:31-40): compare the axis holding the most
// significant differing bit; ties go to z, then y, then x.
static inline bool morton_less(const Body &a, const Body &b) {
  unsigned x = a.Morton[0] ^ b.Morton[0];
  unsigned y = a.Morton[1] ^ b.Morton[1];
  unsigned z = a.Morton[2] ^ b.Morton[2];
  auto hi = [](unsigned u) { return u ? 31 - __builtin_clz(u) : -1; };
  int hx = hi(x), hy = hi(y), hz = hi(z);
  if (hz >= 0 && hz >= hx && hz >= hy) return a.Morton[2] < b.Morton[2];
  if (hy >= 0 && hy >= hx) return a.Morton[1] < b.Morton[1];
  return a.Morton[0] < b.Morton[0];
}

struct Fmm {
  Params prm;
  Tables tab;
  std::vector<Body> B;
  std::vector<Cell> C;
  std::vector<real> coef;  // per cell: M, L, Md, Ld (NTERM each) + Pn (P+1)
  Stats st;
  real X0[3], R0;
  // theta/error-MAC interaction lists, grouped later by target
  std::vector<std::pair<int, int>> m2l, p2p;

  explicit Fmm(const Params &p) : prm(p), tab(p.p, p.quirk) { std::memset(&st, 0, sizeof(st)); }

  // getBounds (build_tree.h:239-276) with the three (drand48()-0.5) values given.
  void bounds() {
    real mn[3], mx[3];
    for (int d = 0; d < 3; d++) mn[d] = mx[d] = B[0].X[d];
    for (auto &b : B)
      for (int d = 0; d < 3; d++) {
        mn[d] = std::fmin(b.X[d], mn[d]);
        mx[d] = std::fmax(b.X[d], mx[d]);
      }
    for (int d = 0; d < 3; d++) X0[d] = (mx[d] + mn[d]) / 2;
    real box = -1;
    for (int d = 0; d < 3; d++)
      if (mx[d] - mn[d] > box) box = mx[d] - mn[d];
    for (int d = 0; d < 3; d++) X0[d] += prm.shift_frac[d] * box / 25;
    R0 = 0;
    for (int d = 0; d < 3; d++) {
      R0 = std::fmax(X0[d] - mn[d], R0);
      R0 = std::fmax(mx[d] - X0[d], R0);
    }
    R0 *= 1.00001;
  }

  // buildCellsSorted (build_tree.h:360-451): top-down octree over the
  // Morton-sorted bodies; octant boundaries by binary search on x >= centre.
  // The search always leaves the first body in the first non-empty slot
  // range (offsets[i] >= begin+1), exactly as the reference does.
  void build_cells(int ci, int begin, int end, const real *X, int level) {
    Cell &c = C[ci];
    c.BODY = begin;
    c.NBODY = end - begin;
    c.NCHILD = 0;
    c.CHILD = -1;
    c.level = level;
    for (int d = 0; d < 3; d++) c.Xg[d] = X[d];
    c.Rg = R0 / (1 << level);
    st.depth = std::max(st.depth, level);
    if (end - begin <= prm.ncrit) return;
    int size[8] = {0}, offs[8];
    auto oct = [&](int i) {
      return (B[i].X[0] >= X[0]) + ((B[i].X[1] >= X[1]) << 1) + ((B[i].X[2] >= X[2]) << 2);
    };
    if (end - begin <= 32) {
      for (int i = begin; i < end; i++) size[oct(i)]++;
      int o = begin;
      for (int i = 0; i < 8; i++) {
        offs[i] = o;
        o += size[i];
      }
    } else {
      offs[0] = begin;
      for (int i = 1; i < 8; i++) {
        int l = begin, r = end;
        while (l < r - 1) {
          int mid = (l + r) / 2;
          if (oct(mid) >= i)
            r = mid;
          else
            l = mid;
        }
        offs[i] = l + 1;
      }
      for (int i = 0; i < 7; i++) size[i] = offs[i + 1] - offs[i];
      size[7] = end - offs[7];
    }
    int nch = 0;
    for (int i = 0; i < 8; i++)
      if (size[i]) nch++;
    int first = (int)C.size();
    C.resize(C.size() + nch);
    C[ci].NCHILD = nch;
    C[ci].CHILD = first;
    int k = 0;
    for (int i = 0; i < 8; i++) {
      real Xc[3];
      real r = R0 / (1 << (level + 1));
      for (int d = 0; d < 3; d++) Xc[d] = X[d] + r * (((i & 1 << d) >> d) * 2 - 1);
      if (size[i]) {
        build_cells(first + k, offs[i], offs[i] + size[i], Xc, level + 1);
        k++;
      }
    }
  }

  // buildTree (build_tree.h:454-487)
  void build_tree() {
    bounds();
    unsigned maxint = 1u << 28;
    for (auto &b : B)
      for (int d = 0; d < 3; d++) {
        real f = (b.X[d] - (X0[d] - R0)) / (2.0 * R0);
        b.Morton[d] = (unsigned)std::floor(f * (real)maxint);
      }
    std::stable_sort(B.begin(), B.end(), morton_less);
    C.clear();
    C.reserve(B.size() / 2 + 16);
    C.resize(1);
    build_cells(0, 0, (int)B.size(), X0, 0);
    int NT = tab.NTERM, P = prm.p;
    size_t per = 4 * (size_t)NT + P + 1;
    coef.assign(per * C.size(), 0.0);
    for (size_t i = 0; i < C.size(); i++) {
      real *base = &coef[per * i];
      C[i].M = base;
      C[i].L = base + NT;
      C[i].Md = base + 2 * NT;
      C[i].Ld = base + 3 * NT;
      C[i].Pn = base + 4 * NT;
      C[i].min_acc = HUGE_VAL;
      C[i].has_sink = false;
      C[i].X[0] = C[i].Xg[0], C[i].X[1] = C[i].Xg[1], C[i].X[2] = C[i].Xg[2];
      C[i].R = C[i].Rg;
    }
    st.ncells = (int)C.size();
    st.nleaves = 0;
    for (auto &c : C)
      if (c.NCHILD == 0) st.nleaves++;
  }

  // P2M (kernel_fast_lean.h:701-824), Miniball branch.
  void P2M(Cell &c) {
    int P = prm.p, NT = tab.NTERM;
    real min_acc = 1e30;
    if (c.NBODY > 0) {
      Ball b = minball(c.NBODY, [&](int i, real *q) {
        for (int d = 0; d < 3; d++) q[d] = B[c.BODY + i].X[d];
      });
      for (int d = 0; d < 3; d++) c.X[d] = b.c[d];
      c.R = std::sqrt(b.r2) + 1e-6;
    } else {
      c.R = 1e-6;
    }
    std::vector<cplx> cm(NT, 0), cmd(NT, 0), G(NT), Gd(NT);
    for (int i = c.BODY; i < c.BODY + c.NBODY; i++) {
      Body &b = B[i];
      real dx[3] = {b.X[0] - c.X[0], b.X[1] - c.X[1], b.X[2] - c.X[2]};
      min_acc = std::fmin(b.acc_old, min_acc);
      if (b.issink) c.has_sink = true;
      make_G(dx, G.data(), P);
      make_Gdot(b.V, G.data(), Gd.data(), P);
      real w = b.q * (b.issource ? 1 : 0);
      for (int t = 0; t < NT; t++) {
        cm[t] += w * G[t];
        cmd[t] += w * Gd[t];
      }
    }
    std::vector<real> r(NT);
    c2r(cm.data(), r.data(), P);
    for (int t = 0; t < NT; t++) c.M[t] += r[t];
    c2r(cmd.data(), r.data(), P);
    for (int t = 0; t < NT; t++) c.Md[t] += r[t];
    c.min_acc = min_acc;
  }

  // Geometry of M2M (kernel_fast_lean.h:875-935): smallest ball of the bodies
  // of leaf children and of 64 design points on each grandchild sphere.
  void m2m_geometry(Cell &c) {
    std::vector<real> pts;
    for (int k = c.CHILD; k < c.CHILD + c.NCHILD; k++) {
      Cell &ch = C[k];
      if (ch.NCHILD == 0) {
        for (int i = ch.BODY; i < ch.BODY + ch.NBODY; i++)
          for (int d = 0; d < 3; d++) pts.push_back(B[i].X[d]);
      } else {
        const real *td = tdesign();
        for (int g = ch.CHILD; g < ch.CHILD + ch.NCHILD; g++) {
          Cell &gc = C[g];
          for (int t = 0; t < 64; t++)
            for (int d = 0; d < 3; d++) pts.push_back(gc.X[d] + gc.R * td[3 * t + d]);
        }
      }
    }
    int n = (int)pts.size() / 3;
    Ball b = minball(n, [&](int i, real *q) {
      for (int d = 0; d < 3; d++) q[d] = pts[3 * i + d];
    });
    for (int d = 0; d < 3; d++) c.X[d] = b.c[d];
    c.R = std::sqrt(b.r2) + 1e-6;
  }

  // M2M (kernel_fast_lean.h:826-999)
  void M2M(Cell &c) {
    int P = prm.p, NT = tab.NTERM;
    real min_acc = 1e30;
    for (int k = c.CHILD; k < c.CHILD + c.NCHILD; k++) {
      min_acc = std::fmin(C[k].min_acc, min_acc);
      if (C[k].has_sink) c.has_sink = true;
    }
    if (c.NCHILD > 0)
      m2m_geometry(c);
    else
      c.R = 1e-6;
    c.min_acc = min_acc;
    std::vector<cplx> cm(NT, 0), cmd(NT, 0), cj(NT), cjd(NT), G(NT);
    for (int k = c.CHILD; k < c.CHILD + c.NCHILD; k++) {
      Cell &ch = C[k];
      real dx[3] = {ch.X[0] - c.X[0], ch.X[1] - c.X[1], ch.X[2] - c.X[2]};
      r2c(ch.M, cj.data(), P);
      r2c(ch.Md, cjd.data(), P);
      make_G(dx, G.data(), P);
      for (int n = 0; n <= P; n++)
        for (int m = 0; m <= n; m++) {
          int t = ix(n, m);
          for (int kk = 0; kk <= n; kk++)
            for (int l = std::max(-kk, m - n + kk); l <= std::min(kk, m + n - kk); l++) {
              cm[t] += cj[ix(n - kk, m - l)] * G[ix(kk, l)];
              cmd[t] += cjd[ix(n - kk, m - l)] * G[ix(kk, l)];
            }
        }
    }
    std::vector<real> r(NT);
    c2r(cm.data(), r.data(), P);
    for (int t = 0; t < NT; t++) c.M[t] += r[t];
    c2r(cmd.data(), r.data(), P);
    for (int t = 0; t < NT; t++) c.Md[t] += r[t];
  }

  // upwardPass (traverse_eager.h:9-32): post-order, children before parents.
  void upward() {
    for (auto &c : C) c.has_sink = false;
    // cells are created parent-before-child, so reverse index order is post-order safe
    std::vector<int> order(C.size());
    for (size_t i = 0; i < C.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return C[a].level > C[b].level; });
    size_t i = 0;
    while (i < order.size()) {
      size_t j = i;
      while (j < order.size() && C[order[j]].level == C[order[i]].level) j++;
#pragma omp parallel for schedule(dynamic, 4)
      for (size_t k = i; k < j; k++) {
        Cell &c = C[order[k]];
        if (c.NCHILD == 0)
          P2M(c);
        else
          M2M(c);
      }
      i = j;
    }
  }

  // Pn (integrator.h:387-399): per-order multipole norms for the error MAC.
  void pn() {
    int P = prm.p;
    for (auto &c : C)
      for (int n = 0; n <= P; n++) {
        real s = 0;
        for (int m = -n; m <= n; m++) {
          int t = ix(n, m);
          s += c.M[t] * c.M[t] * tab.fact[n + m] * tab.fact[n - m];
        }
        c.Pn[n] = std::sqrt(s);
      }
  }

  // horizontalPass, theta-MAC (traverse_eager.h:125-156)
  void trav_theta(int i, int j) {
    Cell &ci = C[i], &cj = C[j];
    real dx[3] = {ci.X[0] - cj.X[0], ci.X[1] - cj.X[1], ci.X[2] - cj.X[2]};
    real R2 = mac_norm(dx) * prm.theta * prm.theta;
    if (R2 > (ci.R + cj.R) * (ci.R + cj.R)) {
      m2l.push_back({i, j});
    } else if (ci.NCHILD == 0 && cj.NCHILD == 0) {
      p2p.push_back({i, j});
    } else if (cj.NCHILD == 0 || (ci.R >= cj.R && ci.NCHILD != 0)) {
      for (int k = ci.CHILD; k < ci.CHILD + ci.NCHILD; k++) trav_theta(k, j);
    } else {
      for (int k = cj.CHILD; k < cj.CHILD + cj.NCHILD; k++) trav_theta(i, k);
    }
  }

  // horizontalPassHigh, error-controlled MAC (traverse_eager.h:59-122)
  void trav_error(int i, int j) {
    Cell &ci = C[i], &cj = C[j];
    int P = prm.p;
    real dx[3] = {ci.X[0] - cj.X[0], ci.X[1] - cj.X[1], ci.X[2] - cj.X[2]};
    real R2 = mac_norm(dx);
    real f2 = 10;
    real Ra = ci.R, Rb = cj.R, s = Ra + Rb;
    if (R2 * 0.4 > s * s) {
      real R = std::sqrt(R2);
      real Rp = std::pow(R, -(P + 1));
      real pa = 1, E = 0;
      for (int k = 0; k <= P; k++) {
        E = mac_term(E, cj.Pn[P - k], pa, tab.comb[k]);
        pa *= Ra;
      }
      E *= 8 * std::fmax(Ra, Rb) * Rp;
      f2 = E / (prm.force_accuracy * ci.min_acc);
    }
    if (f2 < 1) {
      m2l.push_back({i, j});
    } else if (ci.NCHILD == 0 && cj.NCHILD == 0) {
      p2p.push_back({i, j});
    } else if (cj.NCHILD == 0 || (ci.R >= cj.R && ci.NCHILD != 0)) {
      for (int k = ci.CHILD; k < ci.CHILD + ci.NCHILD; k++) trav_error(k, j);
    } else {
      for (int k = cj.CHILD; k < cj.CHILD + cj.NCHILD; k++) trav_error(i, k);
    }
  }

  // M2L_rotate (kernel_fast_lean.h:1130-1371) for one pair, into L/Ld.
  void m2l_rotate(const Cell &ci, const Cell &cj, real *L, real *Ld) const {
    int P = prm.p, NT = tab.NTERM;
    real dx[3] = {ci.X[0] - cj.X[0], ci.X[1] - cj.X[1], ci.X[2] - cj.X[2]};
    m2l_rotate_raw(P, tab, dx, cj.M, cj.Md, L, Ld);
    (void)NT;
  }

  static void m2l_rotate_raw(int P, const Tables &tab, const real *dx, const real *Mj,
                             const real *Mdj, real *Lout, real *Ldout) {
    int NT = (P + 1) * (P + 1);
    real rxy2 = dx[0] * dx[0] + dx[1] * dx[1];
    real rxy = std::sqrt(rxy2);
    real r = std::sqrt(rxy2 + dx[2] * dx[2]);
    real ri = 1.0 / r;
    std::vector<real> ezr(P + 1), ezi(P + 1), exr(P + 1), exi(P + 1);
    ezr[0] = 1, ezi[0] = 0, exr[0] = 1, exi[0] = 0;
    ezr[1] = 1, ezi[1] = 0;
    exr[1] = dx[2] * ri;
    exi[1] = -rxy * ri;
    if (rxy > 0) {
      real q = 1.0 / rxy;
      ezr[1] = dx[1] * q;
      ezi[1] = dx[0] * q;
    }
    for (int m = 1; m < P; m++) {
      ezr[m + 1] = ezr[m] * ezr[1] - ezi[m] * ezi[1];
      ezi[m + 1] = ezr[m] * ezi[1] + ezi[m] * ezr[1];
      exr[m + 1] = exr[m] * exr[1] - exi[m] * exi[1];
      exi[m + 1] = exr[m] * exi[1] + exi[m] * exr[1];
    }
    std::vector<real> a(NT), ad(NT), l(NT, 0), ld(NT, 0);
    for (int t = 0; t < NT; t++) {
      a[t] = Mj[t] * tab.scale[t];
      ad[t] = Mdj[t] * tab.scale[t];
    }
    auto rot = [&](real *v, const std::vector<real> &er, const std::vector<real> &ei, bool back) {
      for (int n = 1; n <= P; n++) {
        int s = ix(n, 0);
        for (int m = 1; m <= n; m++) {
          real c = er[m], sn = back ? -ei[m] : ei[m];
          real im = v[s - m], re = v[s + m];
          v[s - m] = c * im + sn * re;
          v[s + m] = c * re - sn * im;
        }
      }
    };
    rot(a.data(), ezr, ezi, false);
    rot(ad.data(), ezr, ezi, false);
    swap_xz(a.data(), P);
    swap_xz(ad.data(), P);
    rot(a.data(), exr, exi, false);
    rot(ad.data(), exr, exi, false);
    swap_xz(a.data(), P);
    swap_xz(ad.data(), P);
    for (int t = 0; t < NT; t++) {
      a[t] *= tab.scale_inv[t];
      ad[t] *= tab.scale_inv[t];
    }
    real prn = ri;
    for (int n = 0; n <= P; n++) {
      int s = n * (n + 1);
      real pr = prn;
      for (int k = 0; k <= P - n; k++) {
        real f = tab.fact[n + k] * pr;
        l[s] += a[k * (k + 1)] * f;
        ld[s] += ad[k * (k + 1)] * f;
        pr *= ri;
      }
      real prm = ri, sg = -1;
      for (int m = 1; m <= std::min(n, P - n); m++) {
        pr = prn * prm;
        for (int k = m; k <= P - n; k++) {
          int sk = k * (k + 1);
          real f = tab.fact[n + k] * pr * sg;
          l[s - m] += a[sk - m] * f;
          l[s + m] += a[sk + m] * f;
          ld[s - m] += ad[sk - m] * f;
          ld[s + m] += ad[sk + m] * f;
          pr *= ri;
        }
        sg = -sg;
        prm *= ri;
      }
      prn *= ri;
    }
    swap_xz(l.data(), P);
    swap_xz(ld.data(), P);
    rot(l.data(), exr, exi, true);
    rot(ld.data(), exr, exi, true);
    swap_xz(l.data(), P);
    swap_xz(ld.data(), P);
    rot(l.data(), ezr, ezi, true);
    rot(ld.data(), ezr, ezi, true);
    for (int t = 0; t < NT; t++) {
      Lout[t] += l[t];
      Ldout[t] += ld[t];
    }
  }

  // M2L, O(p^4) complex form (kernel_fast_lean.h:1373-1438).
  static void m2l_direct_raw(int P, const real *dx, const real *Mj, const real *Mdj, real *Lout,
                             real *Ldout) {
    int NT = (P + 1) * (P + 1);
    std::vector<real> a(Mj, Mj + NT), ad(Mdj, Mdj + NT);
    for (int n = 1; n <= P; n++)
      for (int m = 1; m <= n; m++) {
        a[ix(n, -m)] *= -1;
        ad[ix(n, -m)] *= -1;
      }
    std::vector<cplx> c(NT), cd(NT), T(NT), l(NT, 0), ld(NT, 0);
    r2c(a.data(), c.data(), P);
    r2c(ad.data(), cd.data(), P);
    make_T(dx, T.data(), P);
    for (int n = 0; n <= P; n++)
      for (int m = -n; m <= n; m++) {
        int t = ix(n, m);
        for (int k = 0; k <= P - n; k++)
          for (int q = -k; q <= k; q++) {
            l[t] += c[ix(k, q)] * T[ix(n + k, m + q)];
            ld[t] += cd[ix(k, q)] * T[ix(n + k, m + q)];
          }
      }
    std::vector<real> r(NT);
    c2r(l.data(), r.data(), P);
    for (int t = 0; t < NT; t++) Lout[t] += r[t];
    c2r(ld.data(), r.data(), P);
    for (int t = 0; t < NT; t++) Ldout[t] += r[t];
  }

  // P2P (kernel_fast_lean.h:499-699; the Vec8d FP64 branch) of target cell i
  // with source cell j; accumulates into the bodies of i.
  void P2P(const Cell &ci, const Cell &cj) {
    for (int a = ci.BODY; a < ci.BODY + ci.NBODY; a++) {
      Body &bi = B[a];
      real ax = 0, ay = 0, az = 0, pot = 0, jx = 0, jy = 0, jz = 0;
      for (int b = cj.BODY; b < cj.BODY + cj.NBODY; b++) {
        const Body &bj = B[b];
        if (!bj.issource) continue;
        p2p_pair(bi.X, bi.V, bj.X, bj.V, bj.q, pot, ax, ay, az, jx, jy, jz);
      }
      if (bi.issink) {
        bi.p += pot;
        bi.F[0] += ax, bi.F[1] += ay, bi.F[2] += az;
        bi.J[0] += jx, bi.J[1] += jy, bi.J[2] += jz;
      }
    }
  }

  static inline void p2p_pair(const real *xi, const real *vi, const real *xj, const real *vj,
                              real mj, real &pot, real &ax, real &ay, real &az, real &jx,
                              real &jy, real &jz) {
    real dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
    real dvx = vj[0] - vi[0], dvy = vj[1] - vi[1], dvz = vj[2] - vi[2];
    real r2 = dx * dx + dy * dy + dz * dz;
    real vr = 3 * (dx * dvx + dy * dvy + dz * dvz);
    if (!(r2 > 0)) return;
    real r = std::sqrt(r2);
    real ir = 1 / r;
    real ir2 = ir * ir;
    real f = mj * ir;
    pot += f;
    f *= ir2;
    ax += dx * f, ay += dy * f, az += dz * f;
    jx += dvx * f, jy += dvy * f, jz += dvz * f;
    f *= -ir2 * vr;
    jx += dx * f, jy += dy * f, jz += dz * f;
  }

  // L2L (kernel_fast_lean.h:1001-1058) from parent c into every child.
  void L2L(const Cell &c) {
    int P = prm.p, NT = tab.NTERM;
    std::vector<cplx> cl(NT), cld(NT), G(NT), o(NT), od(NT);
    r2c(c.L, cl.data(), P);
    r2c(c.Ld, cld.data(), P);
    std::vector<real> r(NT);
    for (int k = c.CHILD; k < c.CHILD + c.NCHILD; k++) {
      Cell &ch = C[k];
      real dx[3] = {c.X[0] - ch.X[0], c.X[1] - ch.X[1], c.X[2] - ch.X[2]};
      make_G(dx, G.data(), P);
      for (auto &g : G) g = std::conj(g);
      std::fill(o.begin(), o.end(), cplx(0));
      std::fill(od.begin(), od.end(), cplx(0));
      for (int n = 0; n <= P; n++)
        for (int m = 0; m <= n; m++) {
          int t = ix(n, m);
          for (int kk = 0; kk <= P - n; kk++)
            for (int l = -kk; l <= kk; l++) {
              o[t] += cl[ix(n + kk, m + l)] * G[ix(kk, l)];
              od[t] += cld[ix(n + kk, m + l)] * G[ix(kk, l)];
            }
        }
      c2r(o.data(), r.data(), P);
      for (int t = 0; t < NT; t++) ch.L[t] += r[t];
      c2r(od.data(), r.data(), P);
      for (int t = 0; t < NT; t++) ch.Ld[t] += r[t];
    }
  }

  // L2P (kernel_fast_lean.h:1060-1128)
  static void l2p_raw(int P, const real *Xc, const real *L, const real *Ld, const real *x,
                      const real *v, real &pot, real *acc, real *jerk) {
    int NT = (P + 1) * (P + 1);
    std::vector<cplx> cl(NT), cld(NT), G(NT), Gd(NT);
    r2c(L, cl.data(), P);
    r2c(Ld, cld.data(), P);
    real dx[3] = {Xc[0] - x[0], Xc[1] - x[1], Xc[2] - x[2]};
    real dv[3] = {-v[0], -v[1], -v[2]};
    make_G(dx, G.data(), P);
    make_Gdot(dv, G.data(), Gd.data(), P);
    for (int t = 0; t < NT; t++) {
      G[t] = std::conj(G[t]);
      Gd[t] = std::conj(Gd[t]);
    }
    cplx phi[4], phid[4];
    for (int n = 0; n <= 1; n++)
      for (int m = 0; m <= n; m++) {
        int t = ix(n, m);
        for (int k = 0; k <= P - n; k++)
          for (int l = -k; l <= k; l++) {
            int a = ix(n + k, m + l), b = ix(k, l);
            phi[t] += cl[a] * G[b];
            phid[t] += cl[a] * Gd[b];
            phid[t] += cld[a] * G[b];
          }
      }
    pot += phi[0].real();
    acc[0] -= phi[3].real();
    acc[1] -= phi[3].imag();
    acc[2] -= phi[2].real();
    jerk[0] -= phid[3].real();
    jerk[1] -= phid[3].imag();
    jerk[2] -= phid[2].real();
  }

  void L2P(const Cell &c) {
    for (int i = c.BODY; i < c.BODY + c.NBODY; i++) {
      Body &b = B[i];
      if (!b.issink) continue;
      l2p_raw(prm.p, c.X, c.L, c.Ld, b.X, b.V, b.p, b.F, b.J);
    }
  }

  // Evaluate the interaction lists, grouped by target (no atomics needed).
  void horizontal_eval() {
    int NT = tab.NTERM;
    std::vector<std::vector<int>> lm(C.size()), lp(C.size());
    for (auto &e : m2l) lm[e.first].push_back(e.second);
    for (auto &e : p2p) lp[e.first].push_back(e.second);
    st.n_m2l = (long long)m2l.size();
    st.n_p2p_cells = (long long)p2p.size();
    long long pairs = 0;
    for (auto &e : p2p) pairs += (long long)C[e.first].NBODY * C[e.second].NBODY;
    st.n_p2p_pairs = pairs;
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)C.size(); i++) {
      Cell &ci = C[i];
      if (!ci.has_sink) continue;
      for (int j : lm[i]) m2l_rotate(ci, C[j], ci.L, ci.Ld);
      for (int j : lp[i]) P2P(ci, C[j]);
    }
    (void)NT;
  }

  // downwardPass (traverse_eager.h:272-292): pre-order L2L, L2P at leaves.
  void downward() {
    std::vector<int> order(C.size());
    for (size_t i = 0; i < C.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return C[a].level < C[b].level; });
    size_t i = 0;
    while (i < order.size()) {
      size_t j = i;
      while (j < order.size() && C[order[j]].level == C[order[i]].level) j++;
#pragma omp parallel for schedule(dynamic, 4)
      for (size_t k = i; k < j; k++) {
        Cell &c = C[order[k]];
        if (c.NCHILD == 0)
          L2P(c);
        else
          L2L(c);
      }
      i = j;
    }
  }

  // ----------------------------------------- p=2 pre-pass (error MAC) ----
  // P2M_low / M2M_low / M2L_low / P2P_low / L2L_low / L2P_low
  // (kernel_fast_lean.h:1446-1780) with the passes of traverse_eager.h:34-56,
  // :172-212, :249-269.  Geometry is the tree's (centre, half-width).  The
  // reference runs P2P_low in FP32 (Vec16f); so does this restatement.
  // master = true: the master sources' P2P_low and L2P_low (master
  // kernel_fast_lean.h:1479-1556, 1609-1646), which add to acc_old instead of
  // forming |a|: per listed source cell q_last / r^2 to the source cell's last
  // body (the vector loop overwrites its register every source, FP64 under
  // DOUBLE_P2P; r = 0 gives 0), and the magnitude of the far-field gradient.
  void low_pass(bool master = false) {
    const int P = 2, NT = 9;
    std::vector<real> lowM(9 * C.size(), 0), lowL(9 * C.size(), 0);
    std::vector<int> order(C.size());
    for (size_t i = 0; i < C.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return C[a].level > C[b].level; });
    for (int ci : order) {
      Cell &c = C[ci];
      cplx cm[9];
      for (auto &v : cm) v = 0;
      cplx G[9];
      if (c.NCHILD == 0) {
        for (int i = c.BODY; i < c.BODY + c.NBODY; i++) {
          real dx[3] = {B[i].X[0] - c.Xg[0], B[i].X[1] - c.Xg[1], B[i].X[2] - c.Xg[2]};
          make_G(dx, G, P);
          for (int t = 0; t < NT; t++) cm[t] += B[i].q * G[t];
        }
      } else {
        for (int k = c.CHILD; k < c.CHILD + c.NCHILD; k++) {
          real dx[3] = {C[k].Xg[0] - c.Xg[0], C[k].Xg[1] - c.Xg[1], C[k].Xg[2] - c.Xg[2]};
          cplx cj[9];
          r2c(&lowM[9 * k], cj, P);
          make_G(dx, G, P);
          for (int n = 0; n <= P; n++)
            for (int m = 0; m <= n; m++) {
              int t = ix(n, m);
              for (int kk = 0; kk <= n; kk++)
                for (int l = std::max(-kk, m - n + kk); l <= std::min(kk, m + n - kk); l++)
                  cm[t] += cj[ix(n - kk, m - l)] * G[ix(kk, l)];
            }
        }
      }
      real r[9];
      c2r(cm, r, P);
      for (int t = 0; t < NT; t++) lowM[9 * ci + t] += r[t];
    }
    std::vector<std::pair<int, int>> lm, lp;
    std::vector<std::pair<int, int>> stack = {{0, 0}};
    while (!stack.empty()) {
      auto pr = stack.back();
      stack.pop_back();
      Cell &a = C[pr.first], &b = C[pr.second];
      real dx[3] = {a.Xg[0] - b.Xg[0], a.Xg[1] - b.Xg[1], a.Xg[2] - b.Xg[2]};
      // Cells of the tree sit on a grid, so diagonal neighbours tie with this
      // threshold in exact arithmetic and the rounding of R2 decides them:
      // mac_norm reproduces the reference build's FMA contraction.
      real R2 = mac_norm(dx);
      if (R2 > (a.Rg + b.Rg) * (a.Rg + b.Rg) * 3) {
        lm.push_back(pr);
      } else if (a.NCHILD == 0 && b.NCHILD == 0) {
        lp.push_back(pr);
      } else if (b.NCHILD == 0 || (a.Rg >= b.Rg && a.NCHILD != 0)) {
        for (int k = a.CHILD; k < a.CHILD + a.NCHILD; k++) stack.push_back({k, pr.second});
      } else {
        for (int k = b.CHILD; k < b.CHILD + b.NCHILD; k++) stack.push_back({pr.first, k});
      }
    }
    std::vector<std::vector<int>> gm(C.size()), gp(C.size());
    for (auto &e : lm) gm[e.first].push_back(e.second);
    for (auto &e : lp) gp[e.first].push_back(e.second);
    std::vector<real> F(3 * B.size(), 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)C.size(); i++) {
      Cell &a = C[i];
      for (int j : gm[i]) {
        Cell &b = C[j];
        real dx[3] = {a.Xg[0] - b.Xg[0], a.Xg[1] - b.Xg[1], a.Xg[2] - b.Xg[2]};
        real mm[9];
        for (int t = 0; t < 9; t++) mm[t] = lowM[9 * j + t];
        for (int n = 1; n <= 2; n++)
          for (int m = 1; m <= n; m++) mm[ix(n, -m)] *= -1;
        cplx cj[9], T[9], l[9];
        for (auto &v : l) v = 0;
        r2c(mm, cj, P);
        make_T(dx, T, P);
        for (int n = 0; n <= P; n++)
          for (int m = 0; m <= n; m++)
            for (int k = 0; k <= P - n; k++)
              for (int q = -k; q <= k; q++) l[ix(n, m)] += cj[ix(k, q)] * T[ix(n + k, m + q)];
        real r[9];
        c2r(l, r, P);
        for (int t = 0; t < 9; t++) lowL[9 * i + t] += r[t];
      }
      for (int j : gp[i]) {
        Cell &b = C[j];
        if (master) {
          const Body &bl = B[b.BODY + b.NBODY - 1];
          for (int x = a.BODY; x < a.BODY + a.NBODY; x++) {
            real dx = B[x].X[0] - bl.X[0], dy = B[x].X[1] - bl.X[1], dz = B[x].X[2] - bl.X[2];
            real r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0) F[3 * x] += bl.q / r2;
          }
          continue;
        }
        for (int x = a.BODY; x < a.BODY + a.NBODY; x++) {
          float xi = -(float)B[x].X[0], yi = -(float)B[x].X[1], zi = -(float)B[x].X[2];
          float ax = 0, ay = 0, az = 0;
          for (int y = b.BODY; y < b.BODY + b.NBODY; y++) {
            float mj = (float)B[y].q;
            float dx = xi + (float)B[y].X[0], dy = yi + (float)B[y].X[1],
                  dz = zi + (float)B[y].X[2];
            float r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > 0)) continue;
            float rr = std::sqrt(r2);
            float ir = 1 / rr;
            float f = mj * ir * ir * ir;
            ax += dx * f, ay += dy * f, az += dz * f;
          }
          F[3 * x] += ax, F[3 * x + 1] += ay, F[3 * x + 2] += az;
        }
      }
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) { return C[a].level < C[b].level; });
    for (int ci : order) {
      Cell &c = C[ci];
      cplx cl[9], G[9];
      r2c(&lowL[9 * ci], cl, P);
      if (c.NCHILD == 0) {
        for (int i = c.BODY; i < c.BODY + c.NBODY; i++) {
          real dx[3] = {c.Xg[0] - B[i].X[0], c.Xg[1] - B[i].X[1], c.Xg[2] - B[i].X[2]};
          make_G(dx, G, P);
          for (auto &g : G) g = std::conj(g);
          cplx phi[4];
          for (int n = 0; n <= 1; n++)
            for (int m = 0; m <= n; m++)
              for (int k = 0; k <= P - n; k++)
                for (int l = -k; l <= k; l++) phi[ix(n, m)] += cl[ix(n + k, m + l)] * G[ix(k, l)];
          if (master) {
            F[3 * i + 1] += std::sqrt(phi[3].real() * phi[3].real() +
                                      phi[3].imag() * phi[3].imag() +
                                      phi[2].real() * phi[2].real());
            continue;
          }
          F[3 * i] -= phi[3].real();
          F[3 * i + 1] -= phi[3].imag();
          F[3 * i + 2] -= phi[2].real();
        }
      } else {
        for (int k = c.CHILD; k < c.CHILD + c.NCHILD; k++) {
          real dx[3] = {c.Xg[0] - C[k].Xg[0], c.Xg[1] - C[k].Xg[1], c.Xg[2] - C[k].Xg[2]};
          make_G(dx, G, P);
          for (auto &g : G) g = std::conj(g);
          cplx o[9];
          for (auto &v : o) v = 0;
          for (int n = 0; n <= P; n++)
            for (int m = 0; m <= n; m++)
              for (int kk = 0; kk <= P - n; kk++)
                for (int l = -kk; l <= kk; l++) o[ix(n, m)] += cl[ix(n + kk, m + l)] * G[ix(kk, l)];
          real r[9];
          c2r(o, r, P);
          for (int t = 0; t < 9; t++) lowL[9 * k + t] += r[t];
        }
      }
    }
    if (master) {  // F[3i]: near terms, F[3i+1]: far-field magnitude
      for (size_t i = 0; i < B.size(); i++) B[i].acc_old += F[3 * i] + F[3 * i + 1];
      return;
    }
    for (size_t i = 0; i < B.size(); i++)
      B[i].acc_old = std::sqrt(F[3 * i] * F[3 * i] + F[3 * i + 1] * F[3 * i + 1] +
                               F[3 * i + 2] * F[3 * i + 2]);
  }

  // All-pairs sum (direct(), traverse_eager.h:295-304: every leaf pair by P2P).
  void direct_all() {
    int n = (int)B.size();
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      real ax = 0, ay = 0, az = 0, pot = 0, jx = 0, jy = 0, jz = 0;
      for (int j = 0; j < n; j++) {
        if (!B[j].issource) continue;
        p2p_pair(B[i].X, B[i].V, B[j].X, B[j].V, B[j].q, pot, ax, ay, az, jx, jy, jz);
      }
      if (B[i].issink) {
        B[i].p += pot;
        B[i].F[0] += ax, B[i].F[1] += ay, B[i].F[2] += az;
        B[i].J[0] += jx, B[i].J[1] += jy, B[i].J[2] += jz;
      }
    }
  }

  // ---- master (non-jerk) variant ------------------------------------------
  // Timestep estimate of one (target, source) pair, folded into the running
  // minimum ts.  Two forms (master kernel_fast_lean.h, the file at the top of
  // the reference tree):
  //  * vector form (:474-592, SIMD_P2P, the master Makefile's build): the
  //    target mass register accumulates every source met with r2 > 0 within
  //    the current source leaf (mi += mj), r2 == 0 is replaced by 1e38, v2
  //    carries +1e-20 and both estimates are always taken;
  //  * scalar form (:593-672): pairs with r2 == 0 skipped, pair mass
  //    q_i + q_j, the fly-by estimate only for v2 > 0.
  // tau_ff = dt/sqrt2 * sqrt(r^3 / M) and tau_fb = dt * r / |dv|, each divided
  // by (1 - dtau/2) with dtau = its relative rate of change capped at 1.
  static inline void master_tau(const real *xi, const real *vi, real qi, const real *xj,
                                const real *vj, real qj, bool pair_mass, real dt, real &mi,
                                real &ts) {
    real dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
    real dvx = vj[0] - vi[0], dvy = vj[1] - vi[1], dvz = vj[2] - vi[2];
    real r2 = dx * dx + dy * dy + dz * dz;
    if (!pair_mass) {
      real v2 = dvx * dvx + dvy * dvy + dvz * dvz + 1e-20;
      if (r2 > 0) mi += qj;
      else r2 = 1e38;
      real vdr = (dx * dvx + dy * dvy + dz * dvz) / r2;
      real r = std::sqrt(r2);
      real tau = dt / M_SQRT2 * r * std::sqrt(r / mi);
      real dtau = 3 * tau * vdr / 2;
      dtau = dtau < 1 ? dtau : 1;
      tau /= (1 - dtau / 2);
      ts = std::fmin(tau, ts);
      tau = dt * r / std::sqrt(v2);
      dtau = tau * vdr * (1 + mi / (v2 * r));
      dtau = dtau < 1 ? dtau : 1;
      tau /= (1 - dtau / 2);
      ts = std::fmin(tau, ts);
    } else {
      if (!(r2 > 0)) return;
      real r = std::sqrt(r2), v2 = dvx * dvx + dvy * dvy + dvz * dvz;
      real vdr = (dx * dvx + dy * dvy + dz * dvz) / r2, mm = qi + qj;
      real tau = dt / M_SQRT2 * std::sqrt(r * r2 / mm);
      real dtau = 3 * tau * vdr / 2;
      if (dtau > 1) dtau = 1;
      tau /= (1 - dtau / 2);
      if (tau < ts) ts = tau;
      if (v2 > 0) {
        tau = dt * r / std::sqrt(v2);
        dtau = tau * vdr * (1 + mm / (v2 * r));
        if (dtau > 1) dtau = 1;
        tau /= (1 - dtau / 2);
        if (tau < ts) ts = tau;
      }
    }
  }

  // Master get_force_and_potential in the master Makefile's default build
  // (FMM undefined, integrator.h:557-559): direct() (traverse_eager.h:349-370)
  // builds the tree (getBounds without the root shift) and runs P2P_simple
  // (kernel_fast_lean.h:268-460) or, with get_steps, P2P (:462-674) on every
  // pair of leaves (directPass, :262-297).  Per sink body: p += sum m/r,
  // F += sum m dx/r^3 over sources with r > 0; with get_steps,
  // timestep += (min over the source leaf of tau)^-4 for every leaf.
  void master_direct(bool steps, bool pair_mass, real dt) {
    std::memset(&st, 0, sizeof(st));
    for (auto &b : B) {
      b.F[0] = b.F[1] = b.F[2] = 0;
      b.J[0] = b.J[1] = b.J[2] = 0;
      b.p = 0;
      b.acc_old = 0;  // holds the timestep sum
    }
    build_tree();
    std::vector<int> leaves;
    for (int c = 0; c < (int)C.size(); c++)
      if (C[c].NCHILD == 0) leaves.push_back(c);
    int n = (int)B.size();
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
      Body &bi = B[i];
      if (!bi.issink) continue;
      real ax = 0, ay = 0, az = 0, pot = 0, tsum = 0;
      for (int c : leaves) {
        real mi = bi.q, ts = 1e38;
        for (int j = C[c].BODY; j < C[c].BODY + C[c].NBODY; j++) {
          const Body &bj = B[j];
          if (steps) master_tau(bi.X, bi.V, bi.q, bj.X, bj.V, bj.q, pair_mass, dt, mi, ts);
          if (!bj.issource) continue;
          real dx = bj.X[0] - bi.X[0], dy = bj.X[1] - bi.X[1], dz = bj.X[2] - bi.X[2];
          real r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 > 0)) continue;
          real ir = 1 / std::sqrt(r2), f = bj.q * ir;
          pot += f;
          f *= ir * ir;
          ax += dx * f, ay += dy * f, az += dz * f;
        }
        if (steps) {
          ts *= ts;
          ts *= ts;
          tsum += 1 / ts;
        }
      }
      bi.p = pot;
      bi.F[0] = ax, bi.F[1] = ay, bi.F[2] = az;
      bi.acc_old = tsum;
    }
  }

  // master horizontalPassHigh (traverse_eager.h:61-149): the symmetric
  // error-controlled walk.  f1 bounds the error of Ci's expansion at Cj, f2
  // that of Cj's at Ci, both with R^-(P+3) and a 1/(Ra+Rb) factor; the pair
  // is accepted when both are below 1.  Accepted pairs of tiny cells (2 and 8
  // bodies) go to P2P; with steps, accepted pairs add nothing else.
  void trav_master(int i, int j, bool steps) {
    Cell &ci = C[i], &cj = C[j];
    int P = prm.p;
    real dx[3] = {ci.X[0] - cj.X[0], ci.X[1] - cj.X[1], ci.X[2] - cj.X[2]};
    real R2 = mac_norm(dx);
    real f1 = 10, f2 = 10;
    real Ra = ci.R, Rb = cj.R, s = Ra + Rb;
    if (R2 > s * s) {
      real R = std::sqrt(R2), Rp = R2, pa = 1, pb = 1, Eab = 0, Eba = 0;
      for (int k = 0; k <= P; k++) {
        Eab = mac_term(Eab, ci.Pn[P - k], pb, tab.comb[k]);
        Eba = mac_term(Eba, cj.Pn[P - k], pa, tab.comb[k]);
        pa *= Ra;
        pb *= Rb;
        Rp *= R;
      }
      real fac = 8 * std::fmax(Ra, Rb) / s / Rp;
      Eab *= fac;
      f1 = Eab / (prm.force_accuracy * cj.min_acc);
      Eba *= fac;
      f2 = Eba / (prm.force_accuracy * ci.min_acc);
    }
    if (f2 < 1 && f1 < 1) {
      if ((ci.NBODY <= 2 && cj.NBODY <= 8) || (cj.NBODY <= 2 && ci.NBODY <= 8))
        p2p.push_back({i, j});
      else if (!steps)
        m2l.push_back({i, j});
    } else if (ci.NCHILD == 0 && cj.NCHILD == 0) {
      p2p.push_back({i, j});
    } else if (cj.NCHILD == 0 || (ci.R >= cj.R && ci.NCHILD != 0)) {
      for (int k = ci.CHILD; k < ci.CHILD + ci.NCHILD; k++) trav_master(k, j, steps);
    } else {
      for (int k = cj.CHILD; k < cj.CHILD + cj.NCHILD; k++) trav_master(i, k, steps);
    }
  }

  // master P2P_simple / P2P of target cell i with source cell j
  // (kernel_fast_lean.h:268-674): force and potential on the sink bodies and,
  // with steps, one (min tau)^-4 term per body.
  void master_p2p(const Cell &ci, const Cell &cj, bool steps, bool pair_mass, real dt) {
    for (int a = ci.BODY; a < ci.BODY + ci.NBODY; a++) {
      Body &bi = B[a];
      if (!bi.issink) continue;
      real ax = 0, ay = 0, az = 0, pot = 0, mi = bi.q, ts = 1e38;
      for (int b = cj.BODY; b < cj.BODY + cj.NBODY; b++) {
        const Body &bj = B[b];
        if (steps) master_tau(bi.X, bi.V, bi.q, bj.X, bj.V, bj.q, pair_mass, dt, mi, ts);
        if (!bj.issource) continue;
        real dx = bj.X[0] - bi.X[0], dy = bj.X[1] - bi.X[1], dz = bj.X[2] - bi.X[2];
        real r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 > 0)) continue;
        real ir = 1 / std::sqrt(r2), f = bj.q * ir;
        pot += f;
        f *= ir * ir;
        ax += dx * f, ay += dy * f, az += dz * f;
      }
      bi.p += pot;
      bi.F[0] += ax, bi.F[1] += ay, bi.F[2] += az;
      if (steps) {
        ts *= ts;
        ts *= ts;
        bi.ts += 1 / ts;
      }
    }
  }

  // Master get_force_and_potential built with -DFMM (integrator.h:461-556):
  // below ncrit bodies direct(); else the master p = 2 pre-pass (acc_old
  // accumulates onto the caller's values), the upward pass, Pn, the symmetric
  // walk, symmetric M2L_rotate (kernel_fast_lean.h:1034-1356: each accepted
  // pair with Ci < Cj, in cell order, updates the locals of both cells; the
  // pair in the other order does nothing), P2P and the downward pass.
  // prepass = false: skip the pre-pass and use acc_old as given (the
  // reference harness's mode 2 does the same).
  void master_fmm(bool steps, bool pair_mass, real dt, bool prepass = true) {
    if ((int)B.size() < prm.ncrit) {
      for (auto &b : B) b.ts = 0;
      master_direct(steps, pair_mass, dt);
      for (auto &b : B) b.ts = b.acc_old, b.acc_old = 0;
      return;
    }
    std::memset(&st, 0, sizeof(st));
    for (auto &b : B) {
      b.F[0] = b.F[1] = b.F[2] = 0;
      b.J[0] = b.J[1] = b.J[2] = 0;
      b.p = 0;
      b.ts = 0;
    }
    build_tree();
    if (prepass) low_pass(true);
    upward();
    pn();
    m2l.clear();
    p2p.clear();
    trav_master(0, 0, steps);
    std::vector<std::vector<int>> lm(C.size()), lp(C.size());
    for (auto &e : m2l) {
      int i = e.first, j = e.second;
      if (i > j || (!C[i].has_sink && !C[j].has_sink)) continue;
      // each local is written only when its cell holds a sink (:1276-1294)
      if (C[i].has_sink) lm[i].push_back(j);
      if (C[j].has_sink) lm[j].push_back(i);
    }
    for (auto &e : p2p) lp[e.first].push_back(e.second);
    st.n_m2l = (long long)m2l.size();
    st.n_p2p_cells = (long long)p2p.size();
    long long pairs = 0;
    for (auto &e : p2p) pairs += (long long)C[e.first].NBODY * C[e.second].NBODY;
    st.n_p2p_pairs = pairs;
    // P2P lists may hold a tiny non-leaf target whose bodies also belong to
    // its descendants: run targets level by level so no body is written by
    // two threads at once
    std::vector<int> order(C.size());
    for (size_t i = 0; i < C.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return C[a].level < C[b].level; });
    size_t i0 = 0;
    while (i0 < order.size()) {
      size_t i1 = i0;
      while (i1 < order.size() && C[order[i1]].level == C[order[i0]].level) i1++;
#pragma omp parallel for schedule(dynamic, 1)
      for (size_t k = i0; k < i1; k++) {
        Cell &ci = C[order[k]];
        for (int j : lm[order[k]]) m2l_rotate(ci, C[j], ci.L, ci.Ld);
        if (!ci.has_sink) continue;
        for (int j : lp[order[k]]) master_p2p(ci, C[j], steps, pair_mass, dt);
      }
      i0 = i1;
    }
    downward();
  }

  // get_force_and_potential (integrator.h:308-432) with the MAC chosen by prm.mac.
  void eval() {
    std::memset(&st, 0, sizeof(st));
    for (auto &b : B) {
      b.F[0] = b.F[1] = b.F[2] = 0;
      b.J[0] = b.J[1] = b.J[2] = 0;
      b.p = 0;
    }
    if (prm.direct_below > 0 && (int)B.size() < prm.direct_below) {
      direct_all();
      return;
    }
    build_tree();
    if (prm.mac == 1) low_pass();
    upward();
    m2l.clear();
    p2p.clear();
    if (prm.mac == 1) {
      pn();
      trav_error(0, 0);
    } else {
      trav_theta(0, 0);
    }
    horizontal_eval();
    downward();
  }
};

}  // namespace orc

// ------------------------------------------------ master step loop ----
// do_evolve with GRADIENT (master integrator.h:864-884): kick_naive, then
// evolve_split_omelyan (:720-812) with kick (:378-457), kick_cpu (:65-98),
// kick_gradient (:313-376), split (:592-640) and drift with COMPSUM (:642-655,
// :42).  Subsystems are index lists (split is a stable partition; the
// reference swaps records, which changes no force or step).  Forces: the
// master default build (direct over leaf pairs) or, with fmm, the -DFMM
// build, each on a fresh body set of sinks ++ sources as the reference's
// kicks build them.
namespace orc {
struct MasterLoop {
  int n;
  real *pos, *pos_e, *vel, *mass, *ts, *acc, *pot, *acc_old;
  Params prm;
  bool fmm, pair_mass;
  real dtp, simtime;

  void force(const std::vector<int> &si, const std::vector<int> &sj, bool sinks_are_sources,
             real shift, bool steps, std::vector<real> &F, std::vector<real> &P,
             std::vector<real> &T, std::vector<real> &AO) {
    Fmm f(prm);
    int nu = (int)(si.size() + sj.size());
    f.B.resize(nu);
    for (int k = 0; k < nu; k++) {
      bool sink = k < (int)si.size();
      int i = sink ? si[k] : sj[k - si.size()];
      Body &b = f.B[k];
      for (int d = 0; d < 3; d++) {
        b.X[d] = shift != 0 ? pos[3 * i + d] + acc[3 * i + d] * shift : pos[3 * i + d];
        b.V[d] = vel[3 * i + d];
      }
      b.q = mass[i];
      b.index = k;
      b.issink = sink;
      b.issource = sink ? sinks_are_sources : true;
      b.acc_old = 0;
    }
    if (fmm)
      f.master_fmm(steps, pair_mass, dtp);
    else
      f.master_direct(steps, pair_mass, dtp);
    F.assign(3 * nu, 0);
    P.assign(nu, 0);
    T.assign(nu, 0);
    AO.assign(nu, 0);
    for (auto &b : f.B) {
      int k = b.index;
      for (int d = 0; d < 3; d++) F[3 * k + d] = b.F[d];
      P[k] = b.p;
      T[k] = fmm ? b.ts : b.acc_old;  // master_direct keeps the sum in acc_old
      AO[k] = fmm ? b.acc_old : 0;
    }
  }
  void kick(const std::vector<int> &si, const std::vector<int> &sj, real dt, bool sinks_fast) {
    if (si.empty()) return;
    std::vector<real> F, P, T, AO;
    force(si, sj, !sinks_fast, 0, false, F, P, T, AO);
    for (size_t k = 0; k < si.size(); k++) {
      for (int d = 0; d < 3; d++) vel[3 * si[k] + d] += F[3 * k + d] * dt;
      pot[si[k]] = P[k];
    }
  }
  void kick_cpu(const std::vector<int> &s1, const std::vector<int> &s2, real dt) {
    for (int i : s1) {
      real a[3] = {0, 0, 0};
      for (int j : s2) {
        real dx[3];
        for (int d = 0; d < 3; d++) dx[d] = pos[3 * i + d] - pos[3 * j + d];
        real dr2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        if (dr2 > 0) {
          real dr = std::sqrt(dr2), dr3 = dr * dr2, acci = mass[j] / dr3;
          for (int d = 0; d < 3; d++) a[d] -= dx[d] * acci;
        }
      }
      for (int d = 0; d < 3; d++) vel[3 * i + d] += dt * a[d];
    }
  }
  void kick_gradient(const std::vector<int> &si, const std::vector<int> &sj, real dt,
                     bool sinks_fast) {
    if (si.empty()) return;
    std::vector<real> F, P, T, AO;
    force(si, sj, !sinks_fast, dt * dt / 24, false, F, P, T, AO);
    for (size_t k = 0; k < si.size(); k++)
      for (int d = 0; d < 3; d++) vel[3 * si[k] + d] += F[3 * k + d] * dt * 2 / 3;
  }
  void kick_naive(const std::vector<int> &s) {
    if (s.empty()) return;
    std::vector<real> F, P, T, AO;
    force(s, {}, true, 0, true, F, P, T, AO);
    real mn = HUGE_VAL;
    for (size_t k = 0; k < s.size(); k++) {
      int i = s[k];
      pot[i] = P[k];
      for (int d = 0; d < 3; d++) acc[3 * i + d] = F[3 * k + d];
      acc_old[i] = AO[k];
      real t = 1 / std::sqrt(std::sqrt(T[k]));
      if (mn > t) mn = t;
    }
    for (int i : s) ts[i] = mn;
  }
  void drift(const std::vector<int> &s, real dt) {
    for (int i : s)
      for (int d = 0; d < 3; d++) {
        real a = pos[3 * i + d];
        pos_e[3 * i + d] += dt * vel[3 * i + d];
        pos[3 * i + d] = a + pos_e[3 * i + d];
        pos_e[3 * i + d] += a - pos[3 * i + d];
      }
  }
  void omelyan(const std::vector<int> &s, real stime, real etime, real dt, bool calc) {
    if (calc) kick_naive(s);
    std::vector<int> fast, slow;
    for (int i : s)
      if (ts[i] < std::fabs(dt)) fast.push_back(i);
    for (int i : s)
      if (!(ts[i] < std::fabs(dt))) slow.push_back(i);
    if (fast.size() == 1) {
      slow.insert(slow.begin(), fast[0]);
      fast.clear();
    }
    if (fast.empty()) simtime += dt;
    auto outer = [&]() {
      if (s.size() > 128) {
        kick(slow, fast, dt / 6, false);
        if (!fast.empty()) kick(fast, slow, dt / 6, true);
      } else if (!slow.empty()) {
        std::vector<int> joined = fast;
        joined.insert(joined.end(), slow.begin(), slow.end());
        kick_cpu(slow, joined, dt / 6);
        if (!fast.empty()) kick_cpu(fast, slow, dt / 6);
      }
    };
    outer();
    if (!fast.empty()) omelyan(fast, stime, stime + dt / 2, dt / 2, false);
    drift(slow, dt / 2);
    if (!s.empty()) {
      kick(slow, {}, 0, false);
      kick_gradient(slow, {}, dt, false);
      kick(slow, fast, 0, true);
      kick(fast, slow, 0, true);
      kick_gradient(slow, fast, dt, true);
      kick_gradient(fast, slow, dt, true);
    }
    drift(slow, dt / 2);
    if (!fast.empty()) omelyan(fast, stime + dt / 2, etime, dt / 2, true);
    outer();
  }
  void do_evolve(real dt) {
    if (dt == 0) return;
    std::vector<int> all(n);
    for (int i = 0; i < n; i++) all[i] = i;
    kick_naive(all);
    omelyan(all, simtime, simtime + dt, dt, true);
  }
};
}  // namespace orc

// ------------------------------------------------------------------ C API ----
extern "C" {

typedef struct {
  int p, ncrit;
  double theta;
  int mac;
  double force_accuracy;
  int quirk;
  double shift_frac[3];
  int direct_below;
  int nthreads;
} orc_params;

typedef struct {
  long long n_m2l, n_p2p_cells, n_p2p_pairs;
  int ncells, nleaves, depth;
} orc_stats;

void *orc_create(const orc_params *p) {
  orc::Params q;
  std::memcpy(&q, p, sizeof(q));
  return new orc::Fmm(q);
}

void orc_destroy(void *h) { delete (orc::Fmm *)h; }

int orc_eval(void *h, int n, const double *x, const double *v, const double *m,
             const unsigned char *issink, const unsigned char *issource, double *pot,
             double *acc, double *jerk, double *acc_old, orc_stats *st) {
  orc::Fmm &F = *(orc::Fmm *)h;
  if (n <= 0) return 1;
  if (F.prm.nthreads > 0) omp_set_num_threads(F.prm.nthreads);
  F.B.resize(n);
  for (int i = 0; i < n; i++) {
    orc::Body &b = F.B[i];
    for (int d = 0; d < 3; d++) {
      b.X[d] = x[3 * i + d];
      b.V[d] = v[3 * i + d];
    }
    b.q = m[i];
    b.index = i;
    b.issink = issink ? issink[i] != 0 : true;
    b.issource = issource ? issource[i] != 0 : true;
    b.acc_old = 0;
  }
  F.eval();
  for (auto &b : F.B) {
    int i = b.index;
    if (pot) pot[i] = b.p;
    for (int d = 0; d < 3; d++) {
      if (acc) acc[3 * i + d] = b.F[d];
      if (jerk) jerk[3 * i + d] = b.J[d];
    }
    if (acc_old) acc_old[i] = b.acc_old;
  }
  if (st) std::memcpy(st, &F.st, sizeof(*st));
  return 0;
}

// Master variant, default build: all-pairs force and potential over the leaf
// pairs of the master tree (no root shift; the caller passes shift_frac = 0)
// and, with steps, the per-body timestep sum.  pair_mass selects the scalar
// form of the timestep estimate.  Outputs in the caller's order.
int orc_master_direct(void *h, int n, const double *x, const double *v, const double *m,
                      const unsigned char *issink, const unsigned char *issource, int steps,
                      int pair_mass, double dt_param, double *pot, double *acc,
                      double *timestep) {
  orc::Fmm &F = *(orc::Fmm *)h;
  if (n <= 0) return 1;
  if (F.prm.nthreads > 0) omp_set_num_threads(F.prm.nthreads);
  F.B.resize(n);
  for (int i = 0; i < n; i++) {
    orc::Body &b = F.B[i];
    for (int d = 0; d < 3; d++) {
      b.X[d] = x[3 * i + d];
      b.V[d] = v[3 * i + d];
    }
    b.q = m[i];
    b.index = i;
    b.issink = issink ? issink[i] != 0 : true;
    b.issource = issource ? issource[i] != 0 : true;
  }
  F.master_direct(steps != 0, pair_mass != 0, dt_param);
  for (auto &b : F.B) {
    int i = b.index;
    if (pot) pot[i] = b.p;
    for (int d = 0; d < 3; d++)
      if (acc) acc[3 * i + d] = b.F[d];
    if (timestep) timestep[i] = b.acc_old;
  }
  return 0;
}

// Master variant built with -DFMM: acc_old [n] is read (the caller's
// values, onto which the pre-pass adds) and written; outputs in the caller's
// order.  Uses the handle's p, ncrit and force_accuracy; shift_frac must be 0.
int orc_master_fmm(void *h, int n, const double *x, const double *v, const double *m,
                   const unsigned char *issink, const unsigned char *issource, int steps,
                   int pair_mass, int prepass, double dt_param, double *acc_old, double *pot,
                   double *acc, double *timestep, orc_stats *st) {
  orc::Fmm &F = *(orc::Fmm *)h;
  if (n <= 0) return 1;
  if (F.prm.nthreads > 0) omp_set_num_threads(F.prm.nthreads);
  F.B.resize(n);
  for (int i = 0; i < n; i++) {
    orc::Body &b = F.B[i];
    for (int d = 0; d < 3; d++) {
      b.X[d] = x[3 * i + d];
      b.V[d] = v[3 * i + d];
    }
    b.q = m[i];
    b.index = i;
    b.issink = issink ? issink[i] != 0 : true;
    b.issource = issource ? issource[i] != 0 : true;
    b.acc_old = acc_old ? acc_old[i] : 0;
  }
  const bool direct = n < F.prm.ncrit;
  F.master_fmm(steps != 0, pair_mass != 0, dt_param, prepass != 0);
  for (auto &b : F.B) {
    int i = b.index;
    if (pot) pot[i] = b.p;
    for (int d = 0; d < 3; d++)
      if (acc) acc[3 * i + d] = b.F[d];
    if (timestep) timestep[i] = b.ts;
    if (acc_old && !direct) acc_old[i] = b.acc_old;
  }
  if (st) std::memcpy(st, &F.st, sizeof(*st));
  return 0;
}

// nsteps calls of the master do_evolve on device-free host arrays (in/out).
int orc_master_evolve(int n, const double *mass, double *pos, double *pos_e, double *vel,
                      double *ts, double *acc, double *pot, double *acc_old, int p, int ncrit,
                      double force_accuracy, int fmm, int pair_mass, double dt_param, int nsteps,
                      double dt, double *simtime) {
  orc::MasterLoop L;
  L.n = n;
  L.mass = const_cast<double *>(mass);
  L.pos = pos, L.pos_e = pos_e, L.vel = vel, L.ts = ts, L.acc = acc, L.pot = pot;
  L.acc_old = acc_old;
  std::memset(&L.prm, 0, sizeof(L.prm));
  L.prm.p = p;
  L.prm.ncrit = ncrit;
  L.prm.theta = 0.5;
  L.prm.mac = 1;
  L.prm.force_accuracy = force_accuracy;
  L.prm.quirk = 1;
  L.fmm = fmm != 0;
  L.pair_mass = pair_mass != 0;
  L.dtp = dt_param;
  L.simtime = *simtime;
  for (int k = 0; k < nsteps; k++) L.do_evolve(dt);
  *simtime = L.simtime;
  return 0;
}

// Tree of the last evaluation: per cell level, child/body ranges (sorted
// order), geometric and Miniball geometry.  order[k] = original index of the
// k-th sorted body.  Returns the cell count, or -count if maxc is too small.
int orc_cells(void *h, int maxc, int *level, int *nchild, int *child, int *body, int *nbody,
              double *Xg, double *Rg, double *X, double *R, int *order) {
  orc::Fmm &F = *(orc::Fmm *)h;
  int nc = (int)F.C.size();
  if (maxc < nc) return -nc;
  for (int c = 0; c < nc; c++) {
    orc::Cell &C = F.C[c];
    level[c] = C.level;
    nchild[c] = C.NCHILD;
    child[c] = C.CHILD;
    body[c] = C.BODY;
    nbody[c] = C.NBODY;
    for (int d = 0; d < 3; d++) {
      Xg[3 * c + d] = C.Xg[d];
      X[3 * c + d] = C.X[d];
    }
    Rg[c] = C.Rg;
    R[c] = C.R;
  }
  if (order)
    for (size_t b = 0; b < F.B.size(); b++) order[b] = F.B[b].index;
  return nc;
}

// Multipole / local coefficients of the last evaluation (cell-major, NTERM each).
int orc_coefs(void *h, double *M, double *Md, double *L, double *Ld) {
  orc::Fmm &F = *(orc::Fmm *)h;
  int NT = F.tab.NTERM;
  for (size_t c = 0; c < F.C.size(); c++)
    for (int t = 0; t < NT; t++) {
      if (M) M[c * NT + t] = F.C[c].M[t];
      if (Md) Md[c * NT + t] = F.C[c].Md[t];
      if (L) L[c * NT + t] = F.C[c].L[t];
      if (Ld) Ld[c * NT + t] = F.C[c].Ld[t];
    }
  return (int)F.C.size();
}

// Interaction lists of the last evaluation as (target, source) cell pairs.
long long orc_lists(void *h, int which, long long maxn, int *tgt, int *src) {
  orc::Fmm &F = *(orc::Fmm *)h;
  auto &l = which == 0 ? F.m2l : F.p2p;
  if ((long long)l.size() > maxn) return -(long long)l.size();
  for (size_t k = 0; k < l.size(); k++) {
    tgt[k] = l[k].first;
    src[k] = l[k].second;
  }
  return (long long)l.size();
}

// FP64 (or long double) all-pairs sum for a subset of targets.
int orc_direct(int n, const double *x, const double *v, const double *m, int nt, const int *tgt,
               int use_long_double, int nthreads, double *pot, double *acc, double *jerk) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < nt; k++) {
    int i = tgt[k];
    if (!use_long_double) {
      double ax = 0, ay = 0, az = 0, p = 0, jx = 0, jy = 0, jz = 0;
      for (int j = 0; j < n; j++)
        orc::Fmm::p2p_pair(x + 3 * i, v + 3 * i, x + 3 * j, v + 3 * j, m[j], p, ax, ay, az, jx,
                           jy, jz);
      pot[k] = p;
      acc[3 * k] = ax, acc[3 * k + 1] = ay, acc[3 * k + 2] = az;
      jerk[3 * k] = jx, jerk[3 * k + 1] = jy, jerk[3 * k + 2] = jz;
    } else {
      long double ax = 0, ay = 0, az = 0, p = 0, jx = 0, jy = 0, jz = 0;
      for (int j = 0; j < n; j++) {
        long double dx = (long double)x[3 * j] - x[3 * i], dy = (long double)x[3 * j + 1] - x[3 * i + 1],
                    dz = (long double)x[3 * j + 2] - x[3 * i + 2];
        long double dvx = (long double)v[3 * j] - v[3 * i], dvy = (long double)v[3 * j + 1] - v[3 * i + 1],
                    dvz = (long double)v[3 * j + 2] - v[3 * i + 2];
        long double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 > 0)) continue;
        long double ir = 1 / sqrtl(r2), ir2 = ir * ir, f = m[j] * ir;
        p += f;
        f *= ir2;
        ax += dx * f, ay += dy * f, az += dz * f;
        long double vr = 3 * (dx * dvx + dy * dvy + dz * dvz);
        jx += dvx * f - dx * f * ir2 * vr;
        jy += dvy * f - dy * f * ir2 * vr;
        jz += dvz * f - dz * f * ir2 * vr;
      }
      pot[k] = (double)p;
      acc[3 * k] = (double)ax, acc[3 * k + 1] = (double)ay, acc[3 * k + 2] = (double)az;
      jerk[3 * k] = (double)jx, jerk[3 * k + 1] = (double)jy, jerk[3 * k + 2] = (double)jz;
    }
  }
  return 0;
}

// ---- single kernels (same conventions as the reference harness) ----------
int orc_k_p2m(int p, int nb, const double *x, const double *v, const double *m, double *Xc,
              double *Rc, double *M, double *Md) {
  orc::Params q{};
  q.p = p;
  q.quirk = 1;
  orc::Fmm F(q);
  F.B.resize(nb);
  for (int i = 0; i < nb; i++) {
    for (int d = 0; d < 3; d++) {
      F.B[i].X[d] = x[3 * i + d];
      F.B[i].V[d] = v[3 * i + d];
    }
    F.B[i].q = m[i];
    F.B[i].issink = F.B[i].issource = true;
    F.B[i].acc_old = 0;
  }
  int NT = (p + 1) * (p + 1);
  std::vector<double> buf(5 * NT + p + 1, 0);
  orc::Cell c{};
  c.BODY = 0;
  c.NBODY = nb;
  c.NCHILD = 0;
  c.M = buf.data();
  c.L = c.M + NT;
  c.Md = c.M + 2 * NT;
  c.Ld = c.M + 3 * NT;
  c.Pn = c.M + 4 * NT;
  F.P2M(c);
  for (int d = 0; d < 3; d++) Xc[d] = c.X[d];
  *Rc = c.R;
  for (int t = 0; t < NT; t++) {
    M[t] = c.M[t];
    Md[t] = c.Md[t];
  }
  return 0;
}

// Multipole (and its time derivative) of bodies about a fixed centre Xc.
int orc_k_p2m_at(int p, int nb, const double *x, const double *v, const double *m,
                 const double *Xc, double *M, double *Md) {
  int NT = (p + 1) * (p + 1);
  std::vector<orc::cplx> cm(NT, 0), cmd(NT, 0), G(NT), Gd(NT);
  for (int i = 0; i < nb; i++) {
    double dx[3] = {x[3 * i] - Xc[0], x[3 * i + 1] - Xc[1], x[3 * i + 2] - Xc[2]};
    orc::make_G(dx, G.data(), p);
    orc::make_Gdot(v + 3 * i, G.data(), Gd.data(), p);
    for (int t = 0; t < NT; t++) {
      cm[t] += m[i] * G[t];
      cmd[t] += m[i] * Gd[t];
    }
  }
  orc::c2r(cm.data(), M, p);
  orc::c2r(cmd.data(), Md, p);
  return 0;
}

int orc_k_m2l(int p, int quirk, int rotate, const double *Xi, const double *Xj, const double *Mj,
              const double *Mdj, double *L, double *Ld) {
  orc::Tables tab(p, quirk);
  double dx[3] = {Xi[0] - Xj[0], Xi[1] - Xj[1], Xi[2] - Xj[2]};
  if (rotate)
    orc::Fmm::m2l_rotate_raw(p, tab, dx, Mj, Mdj, L, Ld);
  else
    orc::Fmm::m2l_direct_raw(p, dx, Mj, Mdj, L, Ld);
  return 0;
}

// Translation of child moments (centres Xc) to a parent centre Xp (no geometry).
int orc_k_m2m(int p, int nchild, const double *Xc, const double *Mc, const double *Mdc,
              const double *Xp, double *M, double *Md) {
  int NT = (p + 1) * (p + 1);
  std::vector<orc::cplx> cm(NT, 0), cmd(NT, 0), cj(NT), cjd(NT), G(NT);
  for (int k = 0; k < nchild; k++) {
    double dx[3] = {Xc[3 * k] - Xp[0], Xc[3 * k + 1] - Xp[1], Xc[3 * k + 2] - Xp[2]};
    orc::r2c(Mc + (size_t)k * NT, cj.data(), p);
    orc::r2c(Mdc + (size_t)k * NT, cjd.data(), p);
    orc::make_G(dx, G.data(), p);
    for (int n = 0; n <= p; n++)
      for (int m = 0; m <= n; m++) {
        int t = orc::ix(n, m);
        for (int kk = 0; kk <= n; kk++)
          for (int l = std::max(-kk, m - n + kk); l <= std::min(kk, m + n - kk); l++) {
            cm[t] += cj[orc::ix(n - kk, m - l)] * G[orc::ix(kk, l)];
            cmd[t] += cjd[orc::ix(n - kk, m - l)] * G[orc::ix(kk, l)];
          }
      }
  }
  std::vector<double> r(NT);
  orc::c2r(cm.data(), r.data(), p);
  for (int t = 0; t < NT; t++) M[t] += r[t];
  orc::c2r(cmd.data(), r.data(), p);
  for (int t = 0; t < NT; t++) Md[t] += r[t];
  return 0;
}

int orc_k_l2l(int p, const double *Xp, const double *L, const double *Ld, const double *Xc,
              double *Lc, double *Ldc) {
  int NT = (p + 1) * (p + 1);
  std::vector<orc::cplx> cl(NT), cld(NT), G(NT), o(NT, 0), od(NT, 0);
  orc::r2c(L, cl.data(), p);
  orc::r2c(Ld, cld.data(), p);
  double dx[3] = {Xp[0] - Xc[0], Xp[1] - Xc[1], Xp[2] - Xc[2]};
  orc::make_G(dx, G.data(), p);
  for (auto &g : G) g = std::conj(g);
  for (int n = 0; n <= p; n++)
    for (int m = 0; m <= n; m++)
      for (int k = 0; k <= p - n; k++)
        for (int l = -k; l <= k; l++) {
          o[orc::ix(n, m)] += cl[orc::ix(n + k, m + l)] * G[orc::ix(k, l)];
          od[orc::ix(n, m)] += cld[orc::ix(n + k, m + l)] * G[orc::ix(k, l)];
        }
  std::vector<double> r(NT);
  orc::c2r(o.data(), r.data(), p);
  for (int t = 0; t < NT; t++) Lc[t] += r[t];
  orc::c2r(od.data(), r.data(), p);
  for (int t = 0; t < NT; t++) Ldc[t] += r[t];
  return 0;
}

int orc_k_l2p(int p, const double *Xc, const double *L, const double *Ld, int nb, const double *x,
              const double *v, double *pot, double *acc, double *jerk) {
  for (int i = 0; i < nb; i++) {
    pot[i] = 0;
    for (int d = 0; d < 3; d++) acc[3 * i + d] = jerk[3 * i + d] = 0;
    orc::Fmm::l2p_raw(p, Xc, L, Ld, x + 3 * i, v + 3 * i, pot[i], acc + 3 * i, jerk + 3 * i);
  }
  return 0;
}

int orc_k_p2p(int ni, const double *xi, const double *vi, int nj, const double *xj,
              const double *vj, const double *mj, double *pot, double *acc, double *jerk) {
  for (int i = 0; i < ni; i++) {
    double ax = 0, ay = 0, az = 0, p = 0, jx = 0, jy = 0, jz = 0;
    for (int j = 0; j < nj; j++)
      orc::Fmm::p2p_pair(xi + 3 * i, vi + 3 * i, xj + 3 * j, vj + 3 * j, mj[j], p, ax, ay, az,
                         jx, jy, jz);
    pot[i] = p;
    acc[3 * i] = ax, acc[3 * i + 1] = ay, acc[3 * i + 2] = az;
    jerk[3 * i] = jx, jerk[3 * i + 1] = jy, jerk[3 * i + 2] = jz;
  }
  return 0;
}

int orc_minball(int n, const double *pts, double *c, double *r2) {
  orc::Ball b = orc::minball(n, [&](int i, double *q) {
    for (int d = 0; d < 3; d++) q[d] = pts[3 * i + d];
  });
  for (int d = 0; d < 3; d++) c[d] = b.c[d];
  *r2 = b.r2;
  return 0;
}

int orc_harmonics(int p, const double *dx, const double *v, double *G, double *Gd) {
  int NT = (p + 1) * (p + 1);
  std::vector<orc::cplx> g(NT), gd(NT);
  orc::make_G(dx, g.data(), p);
  orc::make_Gdot(v, g.data(), gd.data(), p);
  for (int t = 0; t < NT; t++) {
    G[2 * t] = g[t].real(), G[2 * t + 1] = g[t].imag();
    Gd[2 * t] = gd[t].real(), Gd[2 * t + 1] = gd[t].imag();
  }
  return 0;
}

int orc_swap(int p, double *a) {
  orc::swap_xz(a, p);
  return 0;
}

// srand48(seed) followed by three drand48() calls, as getBounds consumes them
// (build_tree.h:262-264), restated from the documented 48-bit LCG
// X' = (0x5DEECE66D X + 0xB) mod 2^48, seeded X = seed<<16 | 0x330E,
// value X / 2^48.  Returns the three (drand48() - 0.5) values.
void orc_shift_from_seed(long seed, int nskip, double *u) {
  uint64_t X = (((uint64_t)(uint32_t)seed) << 16) | 0x330E;
  const uint64_t a = 0x5DEECE66DULL, c = 0xB, mask = (1ULL << 48) - 1;
  for (int k = 0; k < nskip + 3; k++) {
    X = (a * X + c) & mask;
    if (k >= nskip) u[k - nskip] = std::ldexp((double)X, -48) - 0.5;
  }
}

}  // extern "C"
