// Reference harness -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the taichi_collisional reference (taichi_with_jerks/) from its own
// headers where they lie under /root/reference and exposes a C interface so
// the test suite and bench.py's CPU-baseline leg can drive it.  Nothing of the
// reference is copied into this repository; this file only calls it.  The
// built library lands in oracle/_ref/ (git-ignored).  Built by
// oracle/build_ref.sh with -DEXPANSION=<p>, one library per expansion order.
//
// Entry points:
//   ref_eval      one force+jerk evaluation, the passes driven directly:
//                 mode 0: buildTree, upwardPass, horizontalPass(theta-MAC),
//                         downwardPass      (taichi_with_jerks/traverse_eager.h:125-156)
//                 mode 1: get_force_and_potential (taichi_with_jerks/integrator.h:308-432):
//                         direct sum below 8000 bodies, else the p=2 pre-pass
//                         plus the error-controlled MAC
//                 mode 2: direct(bodies, bodies)   (taichi_with_jerks/traverse_eager.h:295-304)
//                 The root shift comes from drand48() after srand48(seed)
//                 (taichi_with_jerks/build_tree.h:256-267).
//   ref_cells     geometry/structure of the tree of the last mode-0 ref_eval
//   ref_k_*       single-kernel calls of P2M, M2M, M2L_rotate, M2L, L2L, L2P, P2P
#include <cstring>
#include <cmath>
#include <vector>
#include <omp.h>

#include "exafmm.h"
#include "integrator.h"
#include "kernel_fast_lean.h"
#include "traverse_eager.h"

using namespace exafmm;

struct diagnostics global_diag;
struct diagnostics *diag;
struct sys zerosys = {0, NULL, NULL};

namespace {
Cells g_cells;
Bodies g_bodies;
bool g_init = false;

void ensure_init() {
  if (!g_init) {
    P = EXPANSION;
    initKernel();
    g_init = true;
  }
}

void load_bodies(Bodies &bodies, int n, const double *x, const double *v, const double *m,
                 const unsigned char *issink, const unsigned char *issource) {
  bodies.clear();
  bodies.resize(n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      bodies[i].X[d] = x[3 * i + d];
      bodies[i].V[d] = v[3 * i + d];
    }
    bodies[i].q = m[i];
    bodies[i].index = i;
    bodies[i].issink = issink ? (issink[i] != 0) : true;
    bodies[i].issource = issource ? (issource[i] != 0) : true;
    bodies[i].timestep = HUGE;
    bodies[i].acc_old = 0;
  }
}

// Count the pairs the theta-MAC dual traversal visits, mirroring the
// decisions of taichi_with_jerks/traverse_eager.h:125-156 on the reference's cells.
void count_theta(Cell *Ci, Cell *Cj, long long *cnt) {
  real_t d[3];
  for (int k = 0; k < 3; k++) d[k] = Ci->X[k] - Cj->X[k];
  real_t R2 = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * theta * theta;
  if (R2 > (Ci->R + Cj->R) * (Ci->R + Cj->R)) {
    cnt[0]++;
  } else if (Ci->NCHILD == 0 && Cj->NCHILD == 0) {
    cnt[1]++;
    cnt[2] += (long long)Ci->NBODY * Cj->NBODY;
  } else if (Cj->NCHILD == 0 || (Ci->R >= Cj->R && Ci->NCHILD != 0)) {
    for (Cell *ci = Ci->CHILD; ci != Ci->CHILD + Ci->NCHILD; ci++) count_theta(ci, Cj, cnt);
  } else {
    for (Cell *cj = Cj->CHILD; cj != Cj->CHILD + Cj->NCHILD; cj++) count_theta(Ci, cj, cnt);
  }
}
}  // namespace

extern "C" {

int ref_expansion() { return EXPANSION; }

// Returns 0 on success.  counts[0..2] = M2L pairs, P2P cell pairs, P2P body
// pairs (mode 0 only); shift_out = the three (drand48()-0.5) values consumed
// by the first getBounds call.
int ref_eval(int n, const double *x, const double *v, const double *m,
             const unsigned char *issink, const unsigned char *issource, long seed,
             int ncrit_, double theta_, int mode, int nthreads, double *pot, double *acc,
             double *jerk, double *acc_old, long long *counts, double *shift_out) {
  ensure_init();
  if (nthreads > 0) omp_set_num_threads(nthreads);
  ncrit = ncrit_;
  theta = theta_;
  load_bodies(g_bodies, n, x, v, m, issink, issource);
  srand48(seed);
  if (shift_out) {
    for (int d = 0; d < 3; d++) shift_out[d] = drand48() - 0.5;
    srand48(seed);
  }
  for (int k = 0; k < 3 && counts; k++) counts[k] = 0;
  for (auto &b : g_bodies) {
    b.F[0] = b.F[1] = b.F[2] = 0;
    b.J[0] = b.J[1] = b.J[2] = 0;
    b.p = 0;
  }
  if (mode == 0) {
    g_cells = buildTree(g_bodies);
    for (size_t i = 0; i < g_cells.size(); i++) {
      for (int t = 0; t < NTERM; t++) {
        g_cells[i].M[t] = g_cells[i].L[t] = g_cells[i].Mdot[t] = g_cells[i].Ldot[t] = 0;
      }
      g_cells[i].min_acc = HUGE;
      g_cells[i].has_sink = false;
    }
    upwardPass(g_cells);
    if (counts) count_theta(&g_cells[0], &g_cells[0], counts);
    horizontalPass(g_cells, g_cells, 0, false);
    downwardPass(g_cells);
  } else if (mode == 1) {
    get_force_and_potential(g_bodies, true);
  } else if (mode == 2) {
    direct(g_bodies, g_bodies, false);
  } else {
    return 1;
  }
  for (size_t b = 0; b < g_bodies.size(); b++) {
    int i = g_bodies[b].index;
    if (pot) pot[i] = g_bodies[b].p;
    for (int d = 0; d < 3; d++) {
      if (acc) acc[3 * i + d] = g_bodies[b].F[d];
      if (jerk) jerk[3 * i + d] = g_bodies[b].J[d];
    }
    if (acc_old) acc_old[i] = g_bodies[b].acc_old;
  }
  return 0;
}

// Tree of the last mode-0 evaluation.  Cells are reported in the reference's
// own creation order; child = index of the first child (-1 for leaves); body
// ranges are offsets into the sorted body array; order[k] = original index of
// sorted body k.  Returns the number of cells (or -needed if maxc too small).
int ref_cells(int maxc, int *nchild, int *child, int *body, int *nbody, double *X, double *R,
              int *order) {
  int nc = (int)g_cells.size();
  if (maxc < nc) return -nc;
  for (int c = 0; c < nc; c++) {
    Cell &C = g_cells[c];
    nchild[c] = C.NCHILD;
    child[c] = C.NCHILD ? (int)(C.CHILD - &g_cells[0]) : -1;
    body[c] = (int)(C.BODY - &g_bodies[0]);
    nbody[c] = C.NBODY;
    for (int d = 0; d < 3; d++) X[3 * c + d] = C.X[d];
    R[c] = C.R;
  }
  if (order)
    for (size_t b = 0; b < g_bodies.size(); b++) order[b] = g_bodies[b].index;
  return nc;
}

// ---- single kernels -------------------------------------------------------

// P2M on one leaf of nb bodies (Miniball geometry): returns centre, radius, M, Mdot.
int ref_k_p2m(int nb, const double *x, const double *v, const double *m, double *Xc,
              double *Rc, double *M, double *Md) {
  ensure_init();
  Bodies b;
  load_bodies(b, nb, x, v, m, nullptr, nullptr);
  std::vector<Cell> c(1);
  std::memset((void *)&c[0], 0, sizeof(Cell));
  c[0].BODY = &b[0];
  c[0].NBODY = nb;
  c[0].NCHILD = 0;
  P2M(&c[0]);
  for (int d = 0; d < 3; d++) Xc[d] = c[0].X[d];
  *Rc = c[0].R;
  for (int t = 0; t < NTERM; t++) {
    M[t] = c[0].M[t];
    Md[t] = c[0].Mdot[t];
  }
  return 0;
}

static void fill_cell(Cell &c, const double *X, const double *M, const double *Md) {
  std::memset((void *)&c, 0, sizeof(Cell));
  for (int d = 0; d < 3; d++) c.X[d] = X[d];
  if (M)
    for (int t = 0; t < NTERM; t++) c.M[t] = M[t];
  if (Md)
    for (int t = 0; t < NTERM; t++) c.Mdot[t] = Md[t];
  c.has_sink = true;
}

// M2L_rotate(Ci, Cj) and the O(p^4) M2L(Ci, Cj): L, Ld += contribution of Cj at Xi.
int ref_k_m2l(int rotate, const double *Xi, const double *Xj, const double *Mj,
              const double *Mdj, double *L, double *Ld) {
  ensure_init();
  std::vector<Cell> c(2);
  fill_cell(c[0], Xi, nullptr, nullptr);
  fill_cell(c[1], Xj, Mj, Mdj);
  if (rotate)
    M2L_rotate(&c[0], &c[1]);
  else
    M2L(&c[0], &c[1]);
  for (int t = 0; t < NTERM; t++) {
    L[t] = c[0].L[t];
    Ld[t] = c[0].Ldot[t];
  }
  return 0;
}

// Translation part of M2M: children (leaves, centres Xc, moments Mc/Mdc) into a
// parent whose centre is fixed to Xp.  The reference recomputes the parent's
// Miniball from the children's bodies; here each child holds one body at its
// centre with zero mass so the parent ball is known: caller passes the bodies.
int ref_k_m2m(int nchild, const double *Xc, const double *Rc, const double *Mc,
              const double *Mdc, int nb, const double *bx, const int *bchild, double *Xp,
              double *Rp, double *M, double *Md) {
  ensure_init();
  Bodies b(nb);
  for (int i = 0; i < nb; i++) {
    std::memset((void *)&b[i], 0, sizeof(Body));
    for (int d = 0; d < 3; d++) b[i].X[d] = bx[3 * i + d];
  }
  std::vector<Cell> c(1 + nchild);
  std::memset((void *)&c[0], 0, sizeof(Cell));
  c[0].CHILD = &c[1];
  c[0].NCHILD = nchild;
  int off = 0;
  for (int k = 0; k < nchild; k++) {
    fill_cell(c[1 + k], Xc + 3 * k, Mc + (size_t)k * NTERM, Mdc + (size_t)k * NTERM);
    c[1 + k].R = Rc[k];
    c[1 + k].NCHILD = 0;
    int cnt = 0;
    while (off + cnt < nb && bchild[off + cnt] == k) cnt++;
    c[1 + k].BODY = &b[off];
    c[1 + k].NBODY = cnt;
    off += cnt;
  }
  M2M(&c[0]);
  for (int d = 0; d < 3; d++) Xp[d] = c[0].X[d];
  *Rp = c[0].R;
  for (int t = 0; t < NTERM; t++) {
    M[t] = c[0].M[t];
    Md[t] = c[0].Mdot[t];
  }
  return 0;
}

// L2L from a parent at Xp to one child at Xc (adds into Lc/Ldc).
int ref_k_l2l(const double *Xp, const double *L, const double *Ld, const double *Xc, double *Lc,
              double *Ldc) {
  ensure_init();
  std::vector<Cell> c(2);
  fill_cell(c[0], Xp, nullptr, nullptr);
  for (int t = 0; t < NTERM; t++) {
    c[0].L[t] = L[t];
    c[0].Ldot[t] = Ld[t];
  }
  fill_cell(c[1], Xc, nullptr, nullptr);
  c[0].CHILD = &c[1];
  c[0].NCHILD = 1;
  L2L(&c[0]);
  for (int t = 0; t < NTERM; t++) {
    Lc[t] = c[1].L[t];
    Ldc[t] = c[1].Ldot[t];
  }
  return 0;
}

// L2P of a leaf at Xc with local L/Ld onto nb bodies.
int ref_k_l2p(const double *Xc, const double *L, const double *Ld, int nb, const double *x,
              const double *v, double *pot, double *acc, double *jerk) {
  ensure_init();
  Bodies b;
  std::vector<double> m(nb, 1.0);
  load_bodies(b, nb, x, v, m.data(), nullptr, nullptr);
  for (auto &q : b) {
    q.p = 0;
    q.F[0] = q.F[1] = q.F[2] = 0;
    q.J[0] = q.J[1] = q.J[2] = 0;
  }
  std::vector<Cell> c(1);
  fill_cell(c[0], Xc, nullptr, nullptr);
  for (int t = 0; t < NTERM; t++) {
    c[0].L[t] = L[t];
    c[0].Ldot[t] = Ld[t];
  }
  c[0].BODY = &b[0];
  c[0].NBODY = nb;
  L2P(&c[0]);
  for (int i = 0; i < nb; i++) {
    pot[i] = b[i].p;
    for (int d = 0; d < 3; d++) {
      acc[3 * i + d] = b[i].F[d];
      jerk[3 * i + d] = b[i].J[d];
    }
  }
  return 0;
}

// P2P(Ci, Cj) between two body sets.
int ref_k_p2p(int ni, const double *xi, const double *vi, int nj, const double *xj,
              const double *vj, const double *mj, double *pot, double *acc, double *jerk) {
  ensure_init();
  Bodies bi, bj;
  std::vector<double> mi(ni, 1.0);
  load_bodies(bi, ni, xi, vi, mi.data(), nullptr, nullptr);
  load_bodies(bj, nj, xj, vj, mj, nullptr, nullptr);
  for (auto &q : bi) {
    q.p = 0;
    q.F[0] = q.F[1] = q.F[2] = 0;
    q.J[0] = q.J[1] = q.J[2] = 0;
  }
  std::vector<Cell> c(2);
  std::memset((void *)&c[0], 0, 2 * sizeof(Cell));
  c[0].BODY = &bi[0];
  c[0].NBODY = ni;
  c[0].has_sink = true;
  c[1].BODY = &bj[0];
  c[1].NBODY = nj;
  P2P(&c[0], &c[1]);
  for (int i = 0; i < ni; i++) {
    pot[i] = bi[i].p;
    for (int d = 0; d < 3; d++) {
      acc[3 * i + d] = bi[i].F[d];
      jerk[3 * i + d] = bi[i].J[d];
    }
  }
  return 0;
}

// Layout probe: sizeof/offsetof of the reference's Body and Cell at this EXPANSION.
int ref_layout(long long *out) {
  out[0] = sizeof(Body);
  out[1] = offsetof(Body, X);
  out[2] = offsetof(Body, q);
  out[3] = offsetof(Body, F);
  out[4] = offsetof(Body, p);
  out[5] = offsetof(Body, V);
  out[6] = offsetof(Body, acc_old);
  out[7] = offsetof(Body, Morton);
  out[8] = offsetof(Body, timestep);
  out[9] = offsetof(Body, J);
  out[10] = offsetof(Body, index);
  out[11] = offsetof(Body, issink);
  out[12] = offsetof(Body, issource);
  out[13] = sizeof(Cell);
  out[14] = offsetof(Cell, NCHILD);
  out[15] = offsetof(Cell, NBODY);
  out[16] = offsetof(Cell, CHILD);
  out[17] = offsetof(Cell, BODY);
  out[18] = offsetof(Cell, X);
  out[19] = offsetof(Cell, R);
  out[20] = offsetof(Cell, min_acc);
  out[21] = offsetof(Cell, cell_mass);
  out[22] = offsetof(Cell, has_sink);
  out[23] = offsetof(Cell, M);
  out[24] = offsetof(Cell, L);
  out[25] = offsetof(Cell, Mdot);
  out[26] = offsetof(Cell, Ldot);
  out[27] = offsetof(Cell, Pn);
  return 28;
}

}  // extern "C"
