// Reference harness for the master (non-jerk) variant -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the taichi_collisional reference's top-level ("master") sources
// from their own headers where they lie under /root/reference and exposes a C
// interface for the test suite's golden-vector script.  Nothing of the
// reference is copied into this repository; this file only calls it.  The
// built library lands in oracle/_ref/ (git-ignored).  Built by
// oracle/build_ref.sh with the master Makefile's flags (Makefile:1-12) plus
// -DFMM, so both of get_force_and_potential's branches are reachable:
//
//   refm_eval  mode 0: direct(bodies, bodies, get_steps)    (traverse_eager.h:349-370),
//                      what get_force_and_potential runs in the Makefile's
//                      default build (FMM undefined, integrator.h:557-559)
//              mode 1: get_force_and_potential(bodies, get_steps) with FMM
//                      (integrator.h:459-556): direct below ncrit bodies, else
//                      the p=2 pre-pass, the symmetric error-controlled walk
//                      and symmetric M2L_rotate
//              mode 2: the FMM branch's passes after the pre-pass, driven
//                      directly with the caller's acc_old (integrator.h:
//                      491-552: reset, upwardPass, Pn, horizontalPass(.., 1,
//                      get_steps), downwardPass), so the symmetric walk, M2L
//                      and P2P can be compared without the pre-pass
//   refm_evolve        nsteps calls of do_evolve(mainsys, dt) (integrator.h:864-884,
//                      the GRADIENT scheme evolve_split_omelyan) on a particle
//                      system, outputs by particle id (split reorders records)
//   refm_kick_steps    kick_cpu_with_steps(0, s, s, 0, true) (integrator.h:101-180):
//                      the integrator's own all-pairs time step, min over every
//                      source of the pair-mass estimate
//
// Bodies are set up as kick_naive does (integrator.h:190-205): value-initialised
// records, then X, V, q, index, issink, issource, timestep = 0.
#include <cstddef>
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
bool g_init = false;

void ensure_init() {
  if (!g_init) {
    P = EXPANSION;
    initKernel();
    g_init = true;
  }
}
}  // namespace

extern "C" {

int refm_expansion() { return EXPANSION; }
int refm_body_size() { return (int)sizeof(Body); }

// Master Body and Cell layouts as compiled (sizes and byte offsets), for the
// test that pins tc_master_body and tc_get_master_cell_layout.
int refm_layout(long long *out) {
  const long long v[] = {(long long)sizeof(Body), (long long)offsetof(Body, X),
                         (long long)offsetof(Body, q), (long long)offsetof(Body, F),
                         (long long)offsetof(Body, p), (long long)offsetof(Body, V),
                         (long long)offsetof(Body, acc_old), (long long)offsetof(Body, Morton),
                         (long long)offsetof(Body, timestep), (long long)offsetof(Body, index),
                         (long long)offsetof(Body, issink), (long long)offsetof(Body, issource),
                         (long long)sizeof(Cell), (long long)offsetof(Cell, M),
                         (long long)offsetof(Cell, L), (long long)offsetof(Cell, Pn),
                         (long long)offsetof(Cell, NCHILD), (long long)offsetof(Cell, NBODY),
                         (long long)offsetof(Cell, CHILD), (long long)offsetof(Cell, BODY),
                         (long long)offsetof(Cell, X), (long long)offsetof(Cell, R),
                         (long long)offsetof(Cell, min_acc), (long long)offsetof(Cell, cell_mass),
                         (long long)offsetof(Cell, has_sink)};
  const int n = (int)(sizeof(v) / sizeof(v[0]));
  for (int k = 0; k < n; k++) out[k] = v[k];
  return n;
}

int refm_eval(int n, const double *x, const double *v, const double *m,
              const unsigned char *issink, const unsigned char *issource, int ncrit_,
              double dt_param_, double force_accuracy_, int mode, int get_steps, int nthreads,
              double *pot, double *acc, double *timestep, double *acc_old) {
  ensure_init();
  if (nthreads > 0) omp_set_num_threads(nthreads);
  ncrit = ncrit_;
  dt_param = dt_param_;
  force_accuracy = force_accuracy_;
  Bodies bodies;
  bodies.resize(n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      bodies[i].X[d] = x[3 * i + d];
      bodies[i].V[d] = v[3 * i + d];
    }
    bodies[i].index = i;
    bodies[i].q = m[i];
    bodies[i].issink = issink ? (issink[i] != 0) : true;
    bodies[i].issource = issource ? (issource[i] != 0) : true;
    bodies[i].timestep = 0;
  }
  if (mode == 0) {
    direct(bodies, bodies, get_steps != 0);
  } else if (mode == 1) {
    get_force_and_potential(bodies, get_steps != 0);
  } else if (mode == 2) {
    for (int i = 0; i < n; i++) bodies[i].acc_old = acc_old ? acc_old[i] : 0;
    Cells cells = buildTree(bodies);
    for (size_t i = 0; i < cells.size(); i++) {
      for (int t = 0; t < NTERM; t++) cells[i].L[t] = cells[i].M[t] = 0;
      for (int k = 0; k <= P; k++) cells[i].Pn[k] = 0;
      cells[i].min_acc = HUGE;
      cells[i].has_sink = false;
    }
    upwardPass(cells);
    for (size_t i = 0; i < cells.size(); i++) {
      for (int k = 0; k <= P; k++) {
        for (int m = -k; m <= k; m++) {
          int index = k * k + k + m;
          cells[i].Pn[k] += cells[i].M[index] * cells[i].M[index] * factorial_table[k + m] *
                            factorial_table[k - m];
        }
        cells[i].Pn[k] = sqrt(cells[i].Pn[k]);
      }
    }
    horizontalPass(cells, cells, 1, get_steps != 0);
    downwardPass(cells);
  } else {
    return 1;
  }
  for (size_t b = 0; b < bodies.size(); b++) {
    int i = bodies[b].index;
    if (pot) pot[i] = bodies[b].p;
    for (int d = 0; d < 3; d++)
      if (acc) acc[3 * i + d] = bodies[b].F[d];
    if (timestep) timestep[i] = bodies[b].timestep;
    if (acc_old) acc_old[i] = bodies[b].acc_old;
  }
  return 0;
}

int refm_kick_steps(int n, const double *x, const double *v, const double *m, double dt_param_,
                    double *timestep) {
  dt_param = dt_param_;
  std::vector<particle> part(n);
  std::memset((void *)part.data(), 0, sizeof(particle) * n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      part[i].pos[d] = x[3 * i + d];
      part[i].vel[d] = v[3 * i + d];
    }
    part[i].mass = m[i];
    part[i].id = i;
  }
  struct sys s = {(size_t)n, part.data(), part.data() + n - 1};
  // kick of zero length: velocities stay put (the loop reads the sources'
  // velocities while it updates the sinks', and here sinks are the sources)
  kick_cpu_with_steps(0, s, s, 0.0, true);
  for (int i = 0; i < n; i++) timestep[i] = part[i].timestep;
  return 0;
}

int refm_evolve(int n, const double *m, double *pos, double *vel, double *pos_e, double *pot,
                double *timestep, double *acc, int ncrit_, double dt_param_,
                double force_accuracy_, int nsteps, double dt, double *simtime) {
  ensure_init();
  ncrit = ncrit_;
  dt_param = dt_param_;
  force_accuracy = force_accuracy_;
  diag = &global_diag;
  diag->simtime = *simtime;
  std::vector<particle> part(n);
  std::memset((void *)part.data(), 0, sizeof(particle) * n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      part[i].pos[d] = pos[3 * i + d];
      part[i].vel[d] = vel[3 * i + d];
    }
    part[i].mass = m[i];
    part[i].id = i;
  }
  struct sys s = {(size_t)n, part.data(), part.data() + n - 1};
  for (int k = 0; k < nsteps; k++) do_evolve(s, dt);
  for (int k = 0; k < n; k++) {
    const particle &q = part[k];
    const int i = (int)q.id;
    for (int d = 0; d < 3; d++) {
      pos[3 * i + d] = q.pos[d];
      vel[3 * i + d] = q.vel[d];
      pos_e[3 * i + d] = q.pos_e[d];
      acc[3 * i + d] = q.acc[d];
    }
    pot[i] = q.pot;
    timestep[i] = q.timestep;
  }
  *simtime = diag->simtime;
  return 0;
}

}  // extern "C"
