/*
 * tcfmm -- MI355X (gfx950) fast-multipole force, potential and jerk evaluation
 * for the taichi_collisional N-body code.  C ABI: plain pointers and sizes,
 * int status codes (0 = success), no exceptions across the boundary.
 *
 * Reference interfaces replaced (paths relative to the reference's
 * taichi_with_jerks/ directory):
 *   tc_eval_bodies  <- void get_force_and_potential(Bodies&, bool)   integrator.h:308-432
 *   tc_eval_soa     <- the same, on device-resident structure-of-arrays bodies
 *   tc_config       <- globals P/EXPANSION, ncrit, theta, force_accuracy  exafmm.h:99,113-116
 *   tc_set_seed     <- srand48(seed) before buildTree's drand48() root shift
 *                      (taichi.cxx:32, build_tree.h:256-267)
 *   tc_body         <- struct Body (exafmm.h:22-46), byte-for-byte
 *   tc_direct_soa   <- direct(bodies, bodies) / directPass  traverse_eager.h:215-304
 *   tc_master_*     <- the master sources' get_force_and_potential (top-level
 *                      integrator.h:459-562) on their 128-byte Body (see below)
 *
 * Semantics kept from the reference: outputs are accelerations (no target
 * mass factor, G = 1), potential is +sum m/r, jerk is da/dt; bodies with
 * issink == false receive nothing, bodies with issource == false exert
 * nothing; coincident bodies (r == 0) do not interact; below
 * config.direct_below bodies the evaluation is an all-pairs sum
 * (integrator.h:311).
 */
#ifndef TCFMM_H
#define TCFMM_H

#include <stddef.h>

#ifdef __cplusplus
#define TC_BOOL bool
extern "C" {
#else
#define TC_BOOL _Bool
#endif

/* Body record of the reference (tj exafmm.h:22-46), 152 bytes on x86-64. */
typedef struct tc_body {
  double X[3];             /* position            @0   */
  double q;                /* mass                @24  */
  double F[3];             /* acceleration (out)  @32  */
  double p;                /* potential (out)     @56  */
  double V[3];             /* velocity            @64  */
  double acc_old;          /* |a| estimate        @88  */
  unsigned int Morton[3];  /* tree keys (scratch) @96  */
  double timestep;         /*                     @112 */
  double J[3];             /* jerk (out)          @120 */
  int index;               /* caller's index      @144 */
  TC_BOOL issink;            /*                     @148 */
  TC_BOOL issource;          /*                     @149 */
} tc_body;

enum { TC_MAC_THETA = 0, TC_MAC_ERROR = 1 };

typedef struct tc_config {
  int p;                 /* expansion order (EXPANSION); compiled: every p in 2..24 */
  int ncrit;             /* max bodies per leaf (ncrit, exafmm.h:115)               */
  double theta;          /* opening angle of the theta-MAC (exafmm.h:116)          */
  int mac;               /* TC_MAC_THETA (horizontalPass) or TC_MAC_ERROR           */
  double force_accuracy; /* error-MAC target (exafmm.h:99, 2e-8)                    */
  int direct_below;      /* N below this -> all-pairs sum (integrator.h:311: 8000); 0 = never */
  int quirk_factorial;   /* 1: drop M_p^{+-p} in M2L as factorial_table[2p]==0 does
                            (kernel_fast_lean.h:129,1179); 0: exact truncated sum  */
} tc_config;

typedef struct tc_stats {
  long long n_m2l;       /* accepted (target, source) cell pairs                   */
  long long n_p2p_cells; /* leaf pairs evaluated directly                          */
  long long n_p2p_pairs; /* body pairs in those leaf pairs (incl. self pairs)      */
  int ncells, nleaves, depth, nlevels;
  double ms_tree, ms_upward, ms_traverse, ms_lists, ms_m2l, ms_p2p, ms_downward, ms_total;
  long long n_m2l_low;   /* TC_MAC_ERROR: pre-pass (p = 2) M2L cell pairs         */
  long long n_p2p_low;   /* TC_MAC_ERROR: pre-pass leaf pairs                      */
  double ms_prepass;     /* TC_MAC_ERROR: pre-pass time (not part of ms_tree)      */
} tc_stats;

typedef struct tc_context tc_context;

/* Create a context on the current HIP device.  stream: a hipStream_t (NULL =
 * the context creates its own).  Returns 0, or a negative code. */
int tc_create(const tc_config *cfg, void *stream, tc_context **out);
int tc_destroy(tc_context *ctx);

/* Root-shift control.  The reference shifts the root centre by
 * (drand48()-0.5)*boxsize/25 per axis on every tree build.  tc_set_seed(s)
 * reproduces srand48(s) followed by the reference's drand48() draws (three
 * per evaluation); tc_set_shift fixes the three (drand48()-0.5) values for
 * every following evaluation until the next tc_set_seed. */
int tc_set_seed(tc_context *ctx, long seed);
int tc_set_shift(tc_context *ctx, const double frac[3]);
/* The three values used by the most recent evaluation. */
int tc_last_shift(const tc_context *ctx, double frac[3]);

/* Precision of the near field (P2P) of the following evaluations.
 * TC_P2P_FP64 (default): the reference's double-precision P2P
 * (-DDOUBLE_P2P, kernel_fast_lean.h:541-548).  TC_P2P_MIXED: body pairs in
 * FP32 on coordinates taken relative to the target group's first body, FP32
 * sums within an LDS tile and FP64 sums across tiles, and pairs closer than
 * close_radius in FP64 (the reference's own FP32 P2P is its Vec16f branch,
 * kernel_fast_lean.h:549-607, without a close-pair rule).
 * TC_P2P_MIXED_HILO: as MIXED, with the relative coordinates held as float
 * hi + lo pairs so the FP32 separations are exact to FP32 rounding.
 * Returns -1 for an unknown mode or a negative radius. */
enum { TC_P2P_FP64 = 0, TC_P2P_MIXED = 1, TC_P2P_MIXED_HILO = 2 };
int tc_set_precision(tc_context *ctx, int p2p_mode, double close_radius);

/* ---- Launch shape of the near field (no reference counterpart: the
 * reference's P2P, kernel_fast_lean.h:499-699, is a CPU loop over leaf
 * pairs) ----------------------------------------------------------------
 * tile: source bodies per LDS tile (64, 128 or 192); list_slice: P2P-list
 * entries per work item (a leaf's list is split into slices that run as
 * separate waves; 1..65535); stage_rounds: the cost of staging one tile, in
 * rounds of pair evaluations, that the cut of each leaf into target groups
 * trades against idle lanes (0 favours groups that fill the wave, larger
 * values fewer groups); targets_per_lane: 1 or 2 (one LDS read of a source
 * body serves that many pairs; the FP64 form only, the mixed forms always
 * run 1).  Defaults 128, 128, 8, 1.  Results do not depend on the shape
 * beyond summation order (FP64; the mixed forms' FP32 tile sums also follow
 * the tile's contents).
 * tc_tune_p2p: after an evaluation, time that evaluation's near field under
 * every shape of the grid {1,2} x {64,128,192} x {64,128,256} x {0,2,4,8}
 * (targets per lane x tile x list slice x staging cost), reps launches each,
 * and keep the fastest for the following evaluations; best_ms / default_ms
 * (may be NULL) receive its time and that of the shape set before the call;
 * grid_ms (NULL or 72 doubles) every shape's time in that nesting order.
 * Returns -1 without a completed tree evaluation. */
int tc_set_p2p_shape(tc_context *ctx, int tile, int list_slice, double stage_rounds,
                     int targets_per_lane);
int tc_get_p2p_shape(tc_context *ctx, int *tile, int *list_slice, double *stage_rounds,
                     int *targets_per_lane);
int tc_tune_p2p(tc_context *ctx, int reps, double *best_ms, double *default_ms,
                double *grid_ms);

/* ---- Per-cell expansion order and opening angle (no reference counterpart;
 * the reference's analogue is its error-controlled walk,
 * traverse_eager.h:59-122, which sets the MAC from a formula) ------------
 * With per-cell orders, every leaf carries its own (q, theta): a pair
 * (Ci, Cj) is accepted when |dX|^2 theta_i^2 > (Ri + Rj)^2 and its M2L is
 * truncated at order q_i; internal cells take the largest q and smallest
 * theta of the leaves below them.  Moments and locals are formed at the
 * context's order p (q <= p).  A probe walk at probe_theta supplies each
 * leaf's list lengths (features).  theta MAC only.
 *   mode 0: off (every pair at cfg.p / cfg.theta);
 *   mode 1: every leaf at (q, theta) through the per-cell path;
 *   mode 2: the learned choice (needs tc_set_order_model). */
int tc_set_adaptive(tc_context *ctx, int mode, int q, double theta, double probe_theta);
/* The MLP (tc_mlp_param_count() floats: W1[64][11] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[64] b4, tanh hidden units) maps a leaf's 9
 * features and a candidate (q / 20, theta) to log10 of the leaf's rms
 * relative force error; each leaf takes the cheapest of the ncand candidates
 * (relative costs cost[c]) predicted at or below target_rms, else the most
 * accurate.  Each feature is clamped to [feat_lo, feat_hi] (9 values each,
 * the training range) before the network sees it.  Host memory. */
int tc_set_order_model(tc_context *ctx, const float *params, int nparams, const int *q,
                       const float *theta, const float *cost, int ncand, const float *feat_lo,
                       const float *feat_hi, double target_rms);
/* After an evaluation with per-cell orders: the number of leaves; per leaf
 * the 9 features (level/16, log10 R/R_root, log2(bodies)/6, log10 mass/mean
 * body mass, log10 of the order-2 and order-4 multipole norms over M_0 R^n,
 * log10(1 + M2L / P2P list length in the probe walk), log10 of the distance
 * from the root centre over R_root), the leaf index of
 * every body (caller order) and, in mode 2, each leaf's predicted log10
 * error.  Device pointers; any may be NULL. */
int tc_leaf_count(const tc_context *ctx);
int tc_leaf_data(tc_context *ctx, float *features, int *leaf_of_body, float *pred);
/* log10 of each leaf's rms relative force error of acc against acc_ref
 * ([n][3] device arrays, caller order) on the last evaluation's tree: the
 * measurement the MLP is trained on. */
int tc_leaf_errors(tc_context *ctx, const double *acc, const double *acc_ref, float *out);
/* M2L cell pairs of the last evaluation per truncation order, counts[0..maxp]. */
int tc_m2l_order_counts(const tc_context *ctx, long long *counts, int maxp);
/* Training kernels, FP32, device pointers.  X: [B][11] inputs, T: [B]
 * targets; tc_mlp_grad writes the gradient of the expectile loss
 * mean(w (Y - T)^2), w = tau where Y < T and 1 - tau elsewhere (0 < tau < 1),
 * into G (nparams floats) and the loss into *loss; tc_mlp_adam applies one
 * Adam step (beta 0.9 / 0.999, eps 1e-8, step >= 1) to params with moments
 * m, v. */
int tc_mlp_param_count(void);
int tc_mlp_input_count(void);
int tc_mlp_forward(tc_context *ctx, const float *params, const float *X, int B, float *Y);
int tc_mlp_grad(tc_context *ctx, const float *params, const float *X, const float *T, int B,
                float tau, float *G, float *loss);
int tc_mlp_adam(tc_context *ctx, float *params, const float *G, float *m, float *v, float lr,
                int step);

/* Drop-in for get_force_and_potential on the reference's host Body array:
 * F, J, p are overwritten (the reference driver zeroes them before the
 * call, integrator.h:350-366), acc_old is written under TC_MAC_ERROR, the
 * array is not reordered.  Copies host<->device inside. */
int tc_eval_bodies(tc_context *ctx, tc_body *bodies, int n);

/* Device-resident evaluation.  x, v: [n][3] doubles; m: [n]; flags: [n]
 * bytes, bit0 = sink, bit1 = source (NULL = all both); pot: [n]; acc, jerk:
 * [n][3] (any output may be NULL).  All pointers are device pointers; work
 * is queued on the context's stream and the call returns after it finishes. */
int tc_eval_soa(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                const unsigned char *flags, double *pot, double *acc, double *jerk);

/* Many independent systems in one evaluation (a forest): cluster k is the
 * bodies [starts[k], starts[k+1]) of the device arrays (starts: host, ncl + 1
 * increasing ints from 0, at least one body each).  Every cluster gets its
 * own root (own bounds and root shift: three drand48() values per cluster in
 * cluster order, or the fixed tc_set_shift values), no pair of cells of two
 * clusters interacts, and each phase (tree levels, walk steps, M2L, P2P,
 * L2L, L2P) runs once for all clusters.  The result equals evaluating every
 * cluster on its own context (per cluster, the reference's
 * get_force_and_potential, integrator.h:308-432), to summation round-off.
 * No all-pairs branch; not with per-cell orders. */
int tc_eval_clusters(tc_context *ctx, int ncl, const int *starts, const double *x,
                     const double *v, const double *m, const unsigned char *flags, double *pot,
                     double *acc, double *jerk);

/* |a| estimates of the last TC_MAC_ERROR evaluation's pre-pass (the values
 * the reference stores in Body.acc_old, integrator.h:350-356), in the
 * caller's body order, into acc_old[n] (device pointer).  Returns -2 when the
 * last evaluation did not run the pre-pass (theta MAC or the all-pairs
 * branch). */
int tc_get_acc_old(tc_context *ctx, double *acc_old);

/* FP64 all-pairs sum for nt targets (indices into the n bodies), device
 * pointers; used to measure the FMM error. */
int tc_direct_soa(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                  const unsigned char *flags, int nt, const int *targets, double *pot,
                  double *acc, double *jerk);

/* tc_direct_soa without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  Same arguments and results as
 * tc_direct_soa, for n <= 262144 bodies (larger n returns -1 with an error
 * message).  The bodies are staged in a fixed device array of the library
 * (~19 MiB, held on each device used until the process ends) instead of the
 * context's buffers, and the outputs are cleared by a kernel.  The work is
 * queued on the context's stream and the call returns without waiting.  All
 * calls on one device share the fixed array: calls that overlap on one
 * device (other streams, other threads) race. */
int tc_direct_soa_noalloc(tc_context *ctx, int n, const double *x, const double *v,
                          const double *m, const unsigned char *flags, int nt,
                          const int *targets, double *pot, double *acc, double *jerk);

/* ---- Per-kernel entry points over the reference's Cell records -----------
 * The reference's kernels take Cell pointers (kernel_fast_lean.h): P2M(C)
 * :701, M2M(C) :826, M2L_rotate(Ci, Cj) :1130, L2L(C) :1001, L2P(C) :1060,
 * P2P(Ci, Cj) :499.  Each entry point below applies that kernel to every
 * listed cell (ids) or (targets[k], sources[k]) pair on the GPU, with the
 * result of calling the reference kernel once per entry in list order:
 *   tc_P2M  sets X, R (smallest ball + 1e-6), min_acc; ORs has_sink;
 *           M, Mdot += (bodies of the cell);
 *   tc_M2M  the same from the children (and grandchildren's balls);
 *   tc_M2L  L, Ldot of the target += (skipped when !has_sink);
 *   tc_L2L  L, Ldot of the children += ;
 *   tc_L2P  p, F, J of the sink bodies += / -= ;
 *   tc_P2P  p, F, J of the target's sink bodies += (skipped when !has_sink).
 * `cells` holds ncells records laid out as the reference's struct Cell for
 * the context's order p (tc_get_cell_layout); CHILD and BODY hold pointers
 * into `cells` and `bodies` (nbodies Body records), as buildTree leaves them.
 * Host memory; the call returns when the records are updated.  Returns 0, -1
 * for bad arguments (pointers outside the arrays, ids out of range). */
typedef struct tc_cell_layout {
  int size;               /* sizeof(Cell)                                    */
  int nchild, nbody, child, body, X, R, min_acc, cell_mass, has_sink;
  int M, L, Mdot, Ldot, Pn; /* byte offsets                                  */
} tc_cell_layout;
int tc_get_cell_layout(int p, tc_cell_layout *out);

int tc_P2M(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids);
int tc_M2M(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids);
int tc_M2L(tc_context *ctx, void *cells, int ncells, const int *targets, const int *sources,
           int npairs);
int tc_L2L(tc_context *ctx, void *cells, int ncells, const int *ids, int nids);
int tc_L2P(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids);
int tc_P2P(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *targets, const int *sources, int npairs);

/* ---- Master (non-jerk) variant ----------------------------------------------
 * The reference's top-level sources (the "master" code, its own exafmm.h /
 * integrator.h / traverse_eager.h / kernel_fast_lean.h) evaluate force and
 * potential without jerk, plus an optional per-body time-step estimate:
 *   tc_master_eval_bodies <- void get_force_and_potential(Bodies&, bool get_steps)
 *                            (master integrator.h:459-562)
 *   tc_master_body        <- master struct Body (exafmm.h:21-48), byte-for-byte
 * mode bits:
 *   TC_MASTER_STEPS      get_steps = true: timestep += (min over each source leaf
 *                        of tau)^-4, tau = min(free-fall, fly-by) estimate
 *                        (master P2P, kernel_fast_lean.h:462-674).  The value
 *                        written is that sum; the caller converts it with
 *                        1/sqrt(sqrt(.)) as kick_naive does (integrator.h:265).
 *   TC_MASTER_FMM        the build with -DFMM (integrator.h:461-556): below
 *                        config.ncrit bodies direct(), else the p = 2
 *                        pre-pass, the symmetric error-controlled walk
 *                        (traverse_eager.h:61-149, both f1 and f2 < 1 at :103) with
 *                        symmetric M2L_rotate (kernel_fast_lean.h:1034-1356);
 *                        with STEPS the walk's accepted far pairs add nothing,
 *                        as in the reference (traverse_eager.h:116-120).
 *                        Without this bit: the master Makefile's default build
 *                        (FMM undefined, integrator.h:557-559): direct() over
 *                        every pair of leaves.
 *   TC_MASTER_PAIR_MASS  the scalar form of the time-step estimate (pair mass
 *                        q_i + q_j, kernel_fast_lean.h:593-672) instead of the
 *                        vector form the master build compiles (SIMD_P2P,
 *                        :474-592), whose mass term accumulates along the
 *                        source leaf.
 * The tree is the master one: ncrit = config.ncrit, no root shift.  dt_param
 * is the global of the same name (master exafmm.h:109, set from argv by
 * main.cxx:75); it must be > 0 when TC_MASTER_STEPS is set. */
enum { TC_MASTER_STEPS = 1, TC_MASTER_FMM = 2, TC_MASTER_PAIR_MASS = 4 };

typedef struct tc_master_body {
  double X[3];             /* position            @0   */
  double q;                /* mass                @24  */
  double F[3];             /* acceleration (out)  @32  */
  double p;                /* potential (out)     @56  */
  double V[3];             /* velocity            @64  */
  double acc_old;          /* |a| estimate        @88  */
  unsigned int Morton[3];  /* tree keys (scratch) @96  */
  double timestep;         /* step sum (out)      @112 */
  int index;               /* caller's index      @120 */
  TC_BOOL issink;          /*                     @124 */
  TC_BOOL issource;        /*                     @125 */
} tc_master_body;

/* Host Body records: F, p and (with TC_MASTER_STEPS) timestep are
 * overwritten, acc_old is written by the TC_MASTER_FMM pre-pass; the array is
 * not reordered.  Copies host<->device inside. */
int tc_master_eval_bodies(tc_context *ctx, tc_master_body *bodies, int n, int mode,
                          double dt_param);

/* Device-resident form: x, v [n][3], m [n], flags [n] (bit0 sink, bit1
 * source; NULL = all both), pot [n], acc [n][3], timestep [n] (needed with
 * TC_MASTER_STEPS, else ignored), acc_old [n] (TC_MASTER_FMM above ncrit
 * bodies: read as the values the pre-pass adds to, then overwritten with the
 * result; NULL = start from 0 and return nothing).  Returns after the work is
 * done. */
int tc_master_eval_soa(tc_context *ctx, int n, const double *x, const double *v, const double *m,
                       const unsigned char *flags, int mode, double dt_param, double *pot,
                       double *acc, double *timestep, double *acc_old);

/* With TC_MASTER_STEPS and without TC_MASTER_FMM, systems of at most
 * max_bodies bodies (capped at 4096; the default) build their tree in a single
 * launch instead of the level-by-level build: the same tree, the same body
 * order and leaves; the time-step sum then runs in slabs that cut leaves, so
 * the results agree with the level-by-level build's to round-off (the order
 * of the sums).  The context keeps the last such tree's sorted order and
 * starts the next one of the same size from it.  0 selects the level-by-level
 * build at every size.  No reference counterpart (a launch-count choice). */
int tc_set_small_tree(tc_context *ctx, int max_bodies);

/* Per-kernel entry points over the master Cell (exafmm.h:50-71: M[NTERM],
 * L[NTERM], Pn[P+1] first, then NCHILD ... has_sink; no Mdot / Ldot) and
 * tc_master_body records, with the list-order semantics of the tc_P2M ...
 * tc_P2P family above.  Master kernels (kernel_fast_lean.h):
 *   tc_master_P2M  <- P2M(C) :676      X, R, min_acc, has_sink; M +=
 *   tc_master_M2M  <- M2M(C) :784      the same from the children
 *   tc_master_M2L  <- M2L_rotate(Ci, Cj) :1034, symmetric: entry (ci[k],
 *                     cj[k]) does nothing unless ci[k] < cj[k] (the
 *                     reference compares the Cell pointers) and one of the
 *                     two has a sink; then L of each sink cell of the two +=
 *                     the other's field.  ci[k] == cj[k] is rejected (-1).
 *   tc_master_L2L  <- L2L(C) :941      L of the children +=
 *   tc_master_L2P  <- L2P(C) :989      p, F of the sink bodies += / -=
 *   tc_master_P2P  <- P2P_simple(Ci, Cj) :268, or with TC_MASTER_STEPS in
 *                     mode P2P(Ci, Cj) :462, which also adds (min over Cj's
 *                     bodies of tau)^-4 to each sink body's timestep
 *                     (TC_MASTER_PAIR_MASS: the scalar form, as for
 *                     tc_master_eval_bodies).
 * tc_get_master_cell_layout gives the offsets (Mdot = Ldot = -1): at p = 12,
 * 2888 bytes, L@1352, Pn@2704, NCHILD@2808. */
int tc_get_master_cell_layout(int p, tc_cell_layout *out);
int tc_master_P2M(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *ids, int nids);
int tc_master_M2M(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *ids, int nids);
int tc_master_M2L(tc_context *ctx, void *cells, int ncells, const int *ci, const int *cj,
                  int npairs);
int tc_master_L2L(tc_context *ctx, void *cells, int ncells, const int *ids, int nids);
int tc_master_L2P(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *ids, int nids);
int tc_master_P2P(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *targets, const int *sources, int npairs, int mode,
                  double dt_param);

/* ---- The master variant's step loop ------------------------------------------
 * tc_master_evolve <- do_evolve(sys, dt) (master integrator.h:864-884) as the
 * master Makefile builds it (GRADIENT): kick_naive for the time steps, then
 * evolve_split_omelyan (:720-812), the fourth-order force-gradient scheme over
 * a hierarchy of time-step levels (split :592-640, kick :378-457, kick_cpu
 * :65-98 for subsystems of <= 128 bodies, kick_gradient :313-376, drift with
 * compensated summation :642-655).  Every force is the master variant's
 * get_force_and_potential (default build, or the FMM with TC_MASTER_FMM; mode
 * as for tc_master_eval_*).  One call advances the particles by dt; *simtime
 * is advanced as the reference advances diag->simtime.
 * Particle arrays are device pointers, the fields of the reference's struct
 * particle (exafmm.h:73-86) that the loop reads or writes; the arrays keep
 * their order (the reference's split swaps records; identity is the index).
 * timestep, acc, pot, acc_old are outputs (as kick_naive leaves them); pos_e
 * holds the compensated-summation error terms and must start at 0.
 * On an error return the particle arrays may be partly updated: the time-step
 * trees of subsystems of up to 4096 bodies report "tree deeper than 64
 * levels" at the loop's next split, not before the kick that follows. */
typedef struct tc_particles {
  int n;
  double *pos;       /* [n][3] */
  double *pos_e;     /* [n][3] */
  double *vel;       /* [n][3] */
  double *mass;      /* [n]    */
  double *timestep;  /* [n]    */
  double *acc;       /* [n][3] */
  double *pot;       /* [n]    */
  double *acc_old;   /* [n]    */
} tc_particles;

/* TC_EVOLVE_JERK_FMM (mode bit): every force of the loop is instead the jerk
 * build's get_force_and_potential (tc_eval_soa: the context's theta-MAC or
 * error-MAC FMM, or its per-cell orders), which has no time-step estimate:
 * every body keeps time step +inf, so nothing is split off and the loop is
 * the single-level fourth-order scheme (five force evaluations per call).
 * Not with TC_MASTER_FMM / TC_MASTER_PAIR_MASS. */
enum { TC_EVOLVE_JERK_FMM = 8 };
int tc_master_evolve(tc_context *ctx, const tc_particles *p, double *simtime, double dt,
                     int mode, double dt_param);

/* The reference's struct particle (master exafmm.h:73-86), 168 bytes. */
typedef struct tc_particle {
  double pos[3];
  double pos_e[3];
  double mass;
  double vel[3];
  double timestep;
  double acc[3];
  double pot;
  double jerk[3];   /* not used by the master loop */
  double postime;   /* end time of the last drift: the call's end time */
  double acc_old;
  unsigned int id;
} tc_particle;

/* Host-record form of tc_master_evolve on the reference's particle array
 * (copies host<->device inside; the array is not reordered). */
int tc_master_evolve_particles(tc_context *ctx, tc_particle *part, int n, double *simtime,
                               double dt, int mode, double dt_param);

int tc_get_stats(const tc_context *ctx, tc_stats *out);
const char *tc_last_error(const tc_context *ctx);

/* Expansion orders this build was compiled for: writes up to maxn values,
 * returns the count. */
int tc_supported_orders(int *orders, int maxn);

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus)
static_assert(sizeof(tc_body) == 152, "tc_body must match the reference Body (152 bytes)");
static_assert(__builtin_offsetof(tc_body, q) == 24 && __builtin_offsetof(tc_body, F) == 32 && __builtin_offsetof(tc_body, p) == 56,
              "tc_body layout");
static_assert(__builtin_offsetof(tc_body, V) == 64 && __builtin_offsetof(tc_body, acc_old) == 88 &&
                  __builtin_offsetof(tc_body, Morton) == 96 && __builtin_offsetof(tc_body, timestep) == 112,
              "tc_body layout");
static_assert(__builtin_offsetof(tc_body, J) == 120 && __builtin_offsetof(tc_body, index) == 144 &&
                  __builtin_offsetof(tc_body, issink) == 148 && __builtin_offsetof(tc_body, issource) == 149,
              "tc_body layout");
static_assert(sizeof(tc_master_body) == 128, "tc_master_body must match the master Body (128 bytes)");
static_assert(__builtin_offsetof(tc_master_body, acc_old) == 88 &&
                  __builtin_offsetof(tc_master_body, timestep) == 112 &&
                  __builtin_offsetof(tc_master_body, index) == 120 &&
                  __builtin_offsetof(tc_master_body, issink) == 124 &&
                  __builtin_offsetof(tc_master_body, issource) == 125,
              "tc_master_body layout");
static_assert(sizeof(tc_particle) == 168 && __builtin_offsetof(tc_particle, acc_old) == 152 &&
                  __builtin_offsetof(tc_particle, id) == 160,
              "tc_particle must match the reference's struct particle (168 bytes)");
#else
_Static_assert(sizeof(tc_particle) == 168, "tc_particle layout");
_Static_assert(sizeof(tc_master_body) == 128, "tc_master_body layout");
_Static_assert(sizeof(tc_body) == 152, "tc_body must match the reference Body (152 bytes)");
_Static_assert(__builtin_offsetof(tc_body, J) == 120 && __builtin_offsetof(tc_body, issink) == 148, "tc_body layout");
#endif

#endif /* TCFMM_H */
