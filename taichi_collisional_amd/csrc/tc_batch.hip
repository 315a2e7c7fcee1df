// Per-kernel entry points over the reference's own Cell and Body records
// (tc_P2M, tc_M2M, tc_M2L, tc_L2L, tc_L2P, tc_P2P and their tc_master_*
// counterparts in include/tcfmm.h).
//
// The reference's kernels take Cell pointers (taichi_with_jerks/
// kernel_fast_lean.h): P2M(C) :701, M2M(C) :826, M2L_rotate(Ci, Cj) :1130,
// L2L(C) :1001, L2P(C) :1060, P2P(Ci, Cj) :499.  These entry points apply the
// same kernel to every listed cell or pair, on the GPU, with the semantics of
// calling the reference kernel once per entry in list order: accumulating
// fields accumulate (M, Mdot, L, Ldot, p, F, J), P2M / M2M set X, R and
// min_acc and OR into has_sink, M2L and P2P skip targets without sinks.
//
// The master sources (top-level kernel_fast_lean.h) have the same kernels
// without jerk over their own Cell (M, L, Pn first, no Mdot / Ldot) and Body
// (128 bytes, no J): P2M :676, M2M :784, L2L :941, L2P :989, the symmetric
// M2L_rotate(Ci, Cj) :1034 and P2P_simple :268 / P2P :462 (time-step sum).
//
// The records are read as they lie in host memory (layout of Cell for the
// context's order p, see tc_get_cell_layout / tc_get_master_cell_layout;
// CHILD and BODY are pointers into `cells` and `bodies`), copied to the
// device as structure-of-arrays, run through the same HIP kernels as the full
// evaluation, and the written fields are copied back.  Entries that depend on
// each other (a cell and its parent in M2M / L2L, the same cell twice) go to
// separate launches in list order.
#include <cstring>

#include "tc_engine.h"

namespace {

constexpr size_t kNone = ~(size_t)0;

struct Layout {
  size_t size, nchild, nbody, child, body, X, R, min_acc, cell_mass, has_sink, M, L, Md, Ld, Pn;
};

// struct Cell of taichi_with_jerks/exafmm.h:50-71 on x86-64 SysV.
Layout layout_for(int p) {
  const size_t nt = (size_t)(p + 1) * (p + 1);
  Layout l;
  l.nchild = 0;
  l.nbody = 4;
  l.child = 8;
  l.body = 16;
  l.X = 24;
  l.R = 48;
  l.min_acc = 56;
  l.cell_mass = 64;
  l.has_sink = 72;
  l.M = 80;
  l.L = l.M + 8 * nt;
  l.Md = l.L + 8 * nt;
  l.Ld = l.Md + 8 * nt;
  l.Pn = l.Ld + 8 * nt;
  l.size = l.Pn + 8 * (size_t)(p + 1);
  return l;
}

// struct Cell of the master exafmm.h:50-71: M[NTERM], L[NTERM], Pn[P+1],
// then the same scalar fields as above; no Mdot / Ldot.
Layout master_layout_for(int p) {
  const size_t nt = (size_t)(p + 1) * (p + 1);
  Layout l;
  l.M = 0;
  l.L = 8 * nt;
  l.Pn = 16 * nt;
  l.Md = l.Ld = kNone;
  const size_t b = l.Pn + 8 * (size_t)(p + 1);
  l.nchild = b;
  l.nbody = b + 4;
  l.child = b + 8;
  l.body = b + 16;
  l.X = b + 24;
  l.R = b + 48;
  l.min_acc = b + 56;
  l.cell_mass = b + 64;
  l.has_sink = b + 72;
  l.size = b + 80;  // bool at +72, padded to the 8-byte alignment of the doubles
  return l;
}

template <class T>
T rd(const char *rec, size_t off) {
  T v;
  std::memcpy(&v, rec + off, sizeof(T));
  return v;
}
template <class T>
void wr(char *rec, size_t off, const T &v) {
  std::memcpy(rec + off, &v, sizeof(T));
}

// Jerk of a body record: the master Body has none (reads as zero, writes drop).
inline void get_j(const tc_body &b, double *j) {
  for (int d = 0; d < 3; d++) j[d] = b.J[d];
}
inline void get_j(const tc_master_body &, double *j) { j[0] = j[1] = j[2] = 0; }
inline void put_j(tc_body &b, const double *j) {
  for (int d = 0; d < 3; d++) b.J[d] = j[d];
}
inline void put_j(tc_master_body &, const double *) {}

// Device copy of a batch: cells and (optionally) bodies with their outputs.
struct Batch {
  int p = 0, nc = 0, nb = 0;
  size_t NT = 0;
  Layout lay;
  std::vector<int> nchild, child, body, nbody, parent;
  std::vector<unsigned char> sink;
  DBuf level, dnchild, dchild, dbody, dnbody, dX, dR, dmin, dsink, dM, dMd, dL, dLd;
  DBuf bx, by, bz, bvx, bvy, bvz, bq, bfl, bord, bacc, bts;
  DBuf rp, rax, ray, raz, rjx, rjy, rjz;
  DBuf ids, aux1, aux2, aux3, npairs;
  ~Batch() {
    for (DBuf *b : {&level, &dnchild, &dchild, &dbody, &dnbody, &dX, &dR, &dmin, &dsink, &dM,
                    &dMd, &dL, &dLd, &bx, &by, &bz, &bvx, &bvy, &bvz, &bq, &bfl, &bord, &bacc,
                    &bts, &rp, &rax, &ray, &raz, &rjx, &rjy, &rjz, &ids, &aux1, &aux2, &aux3,
                    &npairs})
      b->release();
  }
  CellsDev cells() const {
    // the tree geometry (Xg, Rg) is not part of a Cell record; the ball stands in
    return {level.as<int>(), dnchild.as<int>(), dchild.as<int>(), dbody.as<int>(),
            dnbody.as<int>(), dX.as<double>(), dR.as<double>(), dX.as<double>(),
            dR.as<double>(), dmin.as<double>(), dsink.as<unsigned char>(), dM.as<double>(),
            dMd.as<double>(), dL.as<double>(), dLd.as<double>()};
  }
  BodiesDev bodies() const {
    return {bx.as<double>(),  by.as<double>(), bz.as<double>(), bvx.as<double>(),
            bvy.as<double>(), bvz.as<double>(), bq.as<double>(), bfl.as<unsigned char>(),
            bord.as<int>(),   bacc.as<double>()};
  }
  ResultsDev results() const {
    return {rp.as<double>(),  rax.as<double>(), ray.as<double>(), raz.as<double>(),
            rjx.as<double>(), rjy.as<double>(), rjz.as<double>()};
  }
};

int fail(tc_context *ctx, const char *msg) {
  ctx->err = msg;
  return -1;
}

// Read and check the cell records (BODY pointers against `bodies`, records of
// bsize bytes, when given), then upload them.
int load_cells(tc_context *ctx, Batch &B, const Layout &L, const void *cells, int ncells,
               const void *bodies, size_t bsize, int nbodies) {
  hipStream_t s = ctx->stream;
  B.p = ctx->cfg.p;
  B.NT = (size_t)(B.p + 1) * (B.p + 1);
  B.lay = L;
  B.nc = ncells;
  B.nb = bodies ? nbodies : 0;
  if (!cells || ncells <= 0) return fail(ctx, "cells must be non-empty");
  if (bodies && nbodies <= 0) return fail(ctx, "nbodies must be positive");
  const char *base = (const char *)cells;
  const size_t nt = B.NT;
  const bool dots = L.Md != kNone;
  B.nchild.assign(ncells, 0);
  B.child.assign(ncells, 0);
  B.body.assign(ncells, 0);
  B.nbody.assign(ncells, 0);
  B.parent.assign(ncells, -1);
  B.sink.assign(ncells, 0);
  std::vector<double> X(3 * (size_t)ncells), R(ncells), mn(ncells), M(nt * ncells),
      Lc(nt * ncells), Md(dots ? nt * ncells : 0), Ldc(dots ? nt * ncells : 0);
  for (int i = 0; i < ncells; i++) {
    const char *rec = base + L.size * (size_t)i;
    const int nch = rd<int>(rec, L.nchild), nbd = rd<int>(rec, L.nbody);
    if (nch < 0 || nch > 8 || nbd < 0) return fail(ctx, "cell with NCHILD outside 0..8 or NBODY < 0");
    B.nchild[i] = nch;
    B.nbody[i] = nbd;
    if (nch > 0) {
      const char *cp = rd<const char *>(rec, L.child);
      const ptrdiff_t off = cp - base;
      if (off < 0 || off % (ptrdiff_t)L.size != 0 ||
          off / (ptrdiff_t)L.size + nch > (ptrdiff_t)ncells)
        return fail(ctx, "CHILD does not point into the cell array");
      B.child[i] = (int)(off / (ptrdiff_t)L.size);
    }
    if (bodies && nbd > 0) {
      const char *bp = rd<const char *>(rec, L.body);
      const ptrdiff_t off = bp - (const char *)bodies;
      if (off < 0 || off % (ptrdiff_t)bsize != 0 || off / (ptrdiff_t)bsize + nbd > (ptrdiff_t)nbodies)
        return fail(ctx, "BODY does not point into the body array");
      B.body[i] = (int)(off / (ptrdiff_t)bsize);
    }
    for (int d = 0; d < 3; d++) X[3 * i + d] = rd<double>(rec, L.X + 8 * d);
    R[i] = rd<double>(rec, L.R);
    mn[i] = rd<double>(rec, L.min_acc);
    B.sink[i] = rd<bool>(rec, L.has_sink) ? 1 : 0;
    std::memcpy(&M[nt * i], rec + L.M, 8 * nt);
    std::memcpy(&Lc[nt * i], rec + L.L, 8 * nt);
    if (dots) {
      std::memcpy(&Md[nt * i], rec + L.Md, 8 * nt);
      std::memcpy(&Ldc[nt * i], rec + L.Ld, 8 * nt);
    }
  }
  for (int i = 0; i < ncells; i++)
    for (int k = 0; k < B.nchild[i]; k++) B.parent[B.child[i] + k] = i;
  const size_t nc = ncells;
  for (DBuf *b : {&B.level, &B.dnchild, &B.dchild, &B.dbody, &B.dnbody})
    TC_CHECK(b->ensure(sizeof(int) * nc));
  TC_CHECK(B.dX.ensure(24 * nc));
  TC_CHECK(B.dR.ensure(8 * nc));
  TC_CHECK(B.dmin.ensure(8 * nc));
  TC_CHECK(B.dsink.ensure(nc));
  for (DBuf *b : {&B.dM, &B.dMd, &B.dL, &B.dLd}) TC_CHECK(b->ensure(8 * nt * nc));
  TC_CHECK(hipMemsetAsync(B.level.p, 0, sizeof(int) * nc, s));
  TC_CHECK(hipMemcpyAsync(B.dnchild.p, B.nchild.data(), 4 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dchild.p, B.child.data(), 4 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dbody.p, B.body.data(), 4 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dnbody.p, B.nbody.data(), 4 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dX.p, X.data(), 24 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dR.p, R.data(), 8 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dmin.p, mn.data(), 8 * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dsink.p, B.sink.data(), nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dM.p, M.data(), 8 * nt * nc, hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.dL.p, Lc.data(), 8 * nt * nc, hipMemcpyHostToDevice, s));
  if (dots) {
    TC_CHECK(hipMemcpyAsync(B.dMd.p, Md.data(), 8 * nt * nc, hipMemcpyHostToDevice, s));
    TC_CHECK(hipMemcpyAsync(B.dLd.p, Ldc.data(), 8 * nt * nc, hipMemcpyHostToDevice, s));
  } else {
    // the kernels also carry the time derivatives; without them in the record
    // they run on zeros and are dropped
    TC_CHECK(hipMemsetAsync(B.dMd.p, 0, 8 * nt * nc, s));
    TC_CHECK(hipMemsetAsync(B.dLd.p, 0, 8 * nt * nc, s));
  }
  return 0;
}

// Upload the body records: inputs, and the accumulating outputs p, F, J and
// timestep as the caller holds them.
template <class Rec>
int load_bodies(tc_context *ctx, Batch &B, const Rec *bodies, int nbodies) {
  hipStream_t s = ctx->stream;
  const size_t n = nbodies;
  std::vector<double> a[8];
  for (auto &v : a) v.resize(n);
  std::vector<unsigned char> fl(n);
  std::vector<int> ord(n);
  std::vector<double> out[8];
  for (auto &v : out) v.resize(n);
  for (size_t i = 0; i < n; i++) {
    const Rec &q = bodies[i];
    a[0][i] = q.X[0], a[1][i] = q.X[1], a[2][i] = q.X[2];
    a[3][i] = q.V[0], a[4][i] = q.V[1], a[5][i] = q.V[2];
    a[6][i] = q.q;
    a[7][i] = q.acc_old;
    fl[i] = (q.issink ? kSink : 0) | (q.issource ? kSource : 0);
    ord[i] = (int)i;
    double j[3];
    get_j(q, j);
    out[0][i] = q.p;
    for (int d = 0; d < 3; d++) {
      out[1 + d][i] = q.F[d];
      out[4 + d][i] = j[d];
    }
    out[7][i] = q.timestep;
  }
  DBuf *in[8] = {&B.bx, &B.by, &B.bz, &B.bvx, &B.bvy, &B.bvz, &B.bq, &B.bacc};
  for (int k = 0; k < 8; k++) {
    TC_CHECK(in[k]->ensure(8 * n));
    TC_CHECK(hipMemcpyAsync(in[k]->p, a[k].data(), 8 * n, hipMemcpyHostToDevice, s));
  }
  TC_CHECK(B.bfl.ensure(n));
  TC_CHECK(hipMemcpyAsync(B.bfl.p, fl.data(), n, hipMemcpyHostToDevice, s));
  TC_CHECK(B.bord.ensure(4 * n));
  TC_CHECK(hipMemcpyAsync(B.bord.p, ord.data(), 4 * n, hipMemcpyHostToDevice, s));
  DBuf *rs[8] = {&B.rp, &B.rax, &B.ray, &B.raz, &B.rjx, &B.rjy, &B.rjz, &B.bts};
  for (int k = 0; k < 8; k++) {
    TC_CHECK(rs[k]->ensure(8 * n));
    TC_CHECK(hipMemcpyAsync(rs[k]->p, out[k].data(), 8 * n, hipMemcpyHostToDevice, s));
  }
  return 0;
}

template <class Rec>
int load(tc_context *ctx, Batch &B, const Layout &L, const Rec *bodies, int nbodies,
         const void *cells, int ncells) {
  if (int e = load_cells(ctx, B, L, cells, ncells, bodies, sizeof(Rec), nbodies)) return e;
  if (bodies)
    if (int e = load_bodies(ctx, B, bodies, nbodies)) return e;
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// Copy back the cell fields a kernel writes, for the given cells.
enum { kGeomAndM = 1, kLocal = 2 };
int store_cells(tc_context *ctx, Batch &B, void *cells, const std::vector<int> &which, int what) {
  hipStream_t s = ctx->stream;
  const size_t nc = B.nc, nt = B.NT;
  const Layout &L = B.lay;
  const bool dots = L.Md != kNone;
  std::vector<double> X, R, mn, M, Md, Lc, Ldc;
  std::vector<unsigned char> hs;
  if (what & kGeomAndM) {
    X.resize(3 * nc), R.resize(nc), mn.resize(nc), M.resize(nt * nc);
    hs.resize(nc);
    TC_CHECK(hipMemcpyAsync(X.data(), B.dX.p, 24 * nc, hipMemcpyDeviceToHost, s));
    TC_CHECK(hipMemcpyAsync(R.data(), B.dR.p, 8 * nc, hipMemcpyDeviceToHost, s));
    TC_CHECK(hipMemcpyAsync(mn.data(), B.dmin.p, 8 * nc, hipMemcpyDeviceToHost, s));
    TC_CHECK(hipMemcpyAsync(hs.data(), B.dsink.p, nc, hipMemcpyDeviceToHost, s));
    TC_CHECK(hipMemcpyAsync(M.data(), B.dM.p, 8 * nt * nc, hipMemcpyDeviceToHost, s));
    if (dots) {
      Md.resize(nt * nc);
      TC_CHECK(hipMemcpyAsync(Md.data(), B.dMd.p, 8 * nt * nc, hipMemcpyDeviceToHost, s));
    }
  }
  if (what & kLocal) {
    Lc.resize(nt * nc);
    TC_CHECK(hipMemcpyAsync(Lc.data(), B.dL.p, 8 * nt * nc, hipMemcpyDeviceToHost, s));
    if (dots) {
      Ldc.resize(nt * nc);
      TC_CHECK(hipMemcpyAsync(Ldc.data(), B.dLd.p, 8 * nt * nc, hipMemcpyDeviceToHost, s));
    }
  }
  TC_CHECK(hipStreamSynchronize(s));
  for (int i : which) {
    char *rec = (char *)cells + L.size * (size_t)i;
    if (what & kGeomAndM) {
      for (int d = 0; d < 3; d++) wr<double>(rec, L.X + 8 * d, X[3 * i + d]);
      wr<double>(rec, L.R, R[i]);
      wr<double>(rec, L.min_acc, mn[i]);
      wr<bool>(rec, L.has_sink, hs[i] != 0);
      std::memcpy(rec + L.M, &M[nt * i], 8 * nt);
      if (dots) std::memcpy(rec + L.Md, &Md[nt * i], 8 * nt);
    }
    if (what & kLocal) {
      std::memcpy(rec + L.L, &Lc[nt * i], 8 * nt);
      if (dots) std::memcpy(rec + L.Ld, &Ldc[nt * i], 8 * nt);
    }
  }
  return 0;
}

// Copy back p, F, J (where the record has it) and, with `ts`, timestep of
// every body.
template <class Rec>
int store_bodies(tc_context *ctx, Batch &B, Rec *bodies, bool ts = false) {
  hipStream_t s = ctx->stream;
  const size_t n = B.nb;
  std::vector<double> o[8];
  const DBuf *rs[8] = {&B.rp, &B.rax, &B.ray, &B.raz, &B.rjx, &B.rjy, &B.rjz, &B.bts};
  for (int k = 0; k < (ts ? 8 : 7); k++) {
    o[k].resize(n);
    TC_CHECK(hipMemcpyAsync(o[k].data(), rs[k]->p, 8 * n, hipMemcpyDeviceToHost, s));
  }
  TC_CHECK(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; i++) {
    bodies[i].p = o[0][i];
    const double j[3] = {o[4][i], o[5][i], o[6][i]};
    for (int d = 0; d < 3; d++) bodies[i].F[d] = o[1 + d][i];
    put_j(bodies[i], j);
    if (ts) bodies[i].timestep = o[7][i];
  }
  return 0;
}

int check_ids(tc_context *ctx, const int *ids, int nids, int ncells) {
  if (nids < 0 || (nids > 0 && !ids)) return fail(ctx, "bad id list");
  for (int k = 0; k < nids; k++)
    if (ids[k] < 0 || ids[k] >= ncells) return fail(ctx, "cell id out of range");
  return 0;
}

// Split a list into rounds of mutually independent entries, in list order:
// an entry starts a new round when it writes a cell the round reads or
// writes, or reads a cell the round writes.  writes(c, f) / reads(c, f) call
// f on each cell entry c writes / reads.
template <class W, class R>
std::vector<std::vector<int>> rounds(const int *ids, int nids, int ncells, W writes, R reads) {
  std::vector<std::vector<int>> out(1);
  std::vector<int> wmark(ncells, -1), rmark(ncells, -1);
  int r = 0;
  for (int k = 0; k < nids; k++) {
    const int c = ids[k];
    bool clash = false;
    writes(c, [&](int t) {
      if (wmark[t] == r || rmark[t] == r) clash = true;
    });
    reads(c, [&](int t) {
      if (wmark[t] == r) clash = true;
    });
    if (clash) {
      out.emplace_back();
      r++;
    }
    writes(c, [&](int t) { wmark[t] = r; });
    reads(c, [&](int t) { rmark[t] = r; });
    out[r].push_back(c);
  }
  return out;
}

auto no_reads = [](int, auto) {};
auto self_only = [](int c, auto f) { f(c); };

int launch_ids(tc_context *ctx, Batch &B, const std::vector<int> &ids,
               int (*go)(tc_context *, Batch &, const int *, int)) {
  if (ids.empty()) return 0;
  TC_CHECK(B.ids.ensure(4 * ids.size()));
  TC_CHECK(hipMemcpyAsync(B.ids.p, ids.data(), 4 * ids.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  if (int e = go(ctx, B, B.ids.as<int>(), (int)ids.size())) return e;
  TC_CHECK(hipStreamSynchronize(ctx->stream));
  TC_CHECK(hipGetLastError());
  return 0;
}

// Group (target, source) pairs by target, in list order within a target.
void group_pairs(int ncells, const int *tg, const int *sc, int npairs, std::vector<int> &off,
                 std::vector<int> &cnt, std::vector<int> &src) {
  off.assign(ncells, 0);
  cnt.assign(ncells, 0);
  for (int k = 0; k < npairs; k++) cnt[tg[k]]++;
  int o = 0;
  for (int i = 0; i < ncells; i++) {
    off[i] = o;
    o += cnt[i];
  }
  std::vector<int> cur(off);
  src.assign(npairs, 0);
  for (int k = 0; k < npairs; k++) src[cur[tg[k]]++] = sc[k];
}

int check_pairs(tc_context *ctx, const int *tg, const int *sc, int npairs, int ncells) {
  if (npairs < 0 || (npairs > 0 && (!tg || !sc))) return fail(ctx, "bad pair lists");
  for (int k = 0; k < npairs; k++)
    if (tg[k] < 0 || tg[k] >= ncells || sc[k] < 0 || sc[k] >= ncells)
      return fail(ctx, "cell id out of range");
  return 0;
}

// ---- the kernels, shared by both record layouts ------------------------------

template <class Rec>
int do_P2M(tc_context *ctx, const Layout &lay, Rec *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids) {
  if (!bodies) return fail(ctx, "bodies are required");
  if (int e = check_ids(ctx, ids, nids, ncells)) return e;
  Batch B;
  if (int e = load(ctx, B, lay, bodies, nbodies, cells, ncells)) return e;
  auto rs = rounds(ids, nids, ncells, self_only, no_reads);
  for (auto &r : rs)
    if (int e = launch_ids(ctx, B, r, [](tc_context *cx, Batch &b, const int *d, int n) {
          return launch_p2m(cx->cfg.p, b.cells(), b.bodies(), d, n, cx->stream) ? -2 : 0;
        }))
      return e;
  return store_cells(ctx, B, cells, std::vector<int>(ids, ids + nids), kGeomAndM);
}

template <class Rec>
int do_M2M(tc_context *ctx, const Layout &lay, Rec *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids) {
  if (!bodies) return fail(ctx, "bodies are required (leaf children's bodies fix the ball)");
  if (int e = check_ids(ctx, ids, nids, ncells)) return e;
  Batch B;
  if (int e = load(ctx, B, lay, bodies, nbodies, cells, ncells)) return e;
  // M2M(C) reads its children and grandchildren and writes C
  auto rs = rounds(ids, nids, ncells, self_only, [&](int c, auto f) {
    for (int k = 0; k < B.nchild[c]; k++) {
      const int ch = B.child[c] + k;
      f(ch);
      for (int g = 0; g < B.nchild[ch]; g++) f(B.child[ch] + g);
    }
  });
  for (auto &r : rs)
    if (int e = launch_ids(ctx, B, r, [](tc_context *cx, Batch &b, const int *d, int n) {
          return launch_m2m_ids(cx->cfg.p, b.cells(), b.bodies(), d, n, cx->stream) ? -2 : 0;
        }))
      return e;
  return store_cells(ctx, B, cells, std::vector<int>(ids, ids + nids), kGeomAndM);
}

// One-sided translations target <- source, in list order, on a loaded batch;
// `tgts` receives the cells whose locals were written.
int run_m2l(tc_context *ctx, Batch &B, const int *targets, const int *sources, int npairs,
            std::vector<int> &tgts) {
  const int ncells = B.nc;
  std::vector<int> off, cnt, src;
  group_pairs(ncells, targets, sources, npairs, off, cnt, src);
  std::vector<int4> chunks;
  for (int i = 0; i < ncells; i++) {
    if (cnt[i]) tgts.push_back(i);
    for (int k = 0; k < cnt[i]; k += 64)
      chunks.push_back(make_int4(i, off[i] + k, std::min(64, cnt[i] - k), 0));
  }
  if (chunks.empty()) return 0;
  hipStream_t s = ctx->stream;
  TC_CHECK(B.aux1.ensure(sizeof(int4) * chunks.size()));
  TC_CHECK(B.aux2.ensure(4 * src.size()));
  TC_CHECK(hipMemcpyAsync(B.aux1.p, chunks.data(), sizeof(int4) * chunks.size(),
                          hipMemcpyHostToDevice, s));
  TC_CHECK(hipMemcpyAsync(B.aux2.p, src.data(), 4 * src.size(), hipMemcpyHostToDevice, s));
  M2LArgs a;
  a.chunks = B.aux1.as<int4>();
  a.nchunks = (int)chunks.size();
  a.src = B.aux2.as<int>();
  a.cX = B.dX.as<double>();
  a.has_sink = B.dsink.as<unsigned char>();
  a.M = B.dM.as<double>();
  a.Md = B.dMd.as<double>();
  a.L = B.dL.as<double>();
  a.Ld = B.dLd.as<double>();
  a.quirk = ctx->cfg.quirk_factorial;
  if (launch_m2l(ctx->cfg.p, a, s)) return fail(ctx, "expansion order not compiled");
  TC_CHECK(hipStreamSynchronize(s));
  TC_CHECK(hipGetLastError());
  return 0;
}

int do_L2L(tc_context *ctx, const Layout &lay, void *cells, int ncells, const int *ids, int nids) {
  if (int e = check_ids(ctx, ids, nids, ncells)) return e;
  Batch B;
  if (int e = load<tc_body>(ctx, B, lay, nullptr, 0, cells, ncells)) return e;
  // L2L(C) reads C and adds into its children
  auto rs = rounds(
      ids, nids, ncells,
      [&](int c, auto f) {
        for (int k = 0; k < B.nchild[c]; k++) f(B.child[c] + k);
      },
      self_only);
  for (auto &r : rs)
    if (int e = launch_ids(ctx, B, r, [](tc_context *cx, Batch &b, const int *d, int n) {
          return launch_l2l_ids(cx->cfg.p, b.cells(), d, n, cx->stream) ? -2 : 0;
        }))
      return e;
  std::vector<int> out;
  for (int k = 0; k < nids; k++)
    for (int c = 0; c < B.nchild[ids[k]]; c++) out.push_back(B.child[ids[k]] + c);
  return store_cells(ctx, B, cells, out, kLocal);
}

template <class Rec>
int do_L2P(tc_context *ctx, const Layout &lay, Rec *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids) {
  if (!bodies) return fail(ctx, "bodies are required");
  if (int e = check_ids(ctx, ids, nids, ncells)) return e;
  Batch B;
  if (int e = load(ctx, B, lay, bodies, nbodies, cells, ncells)) return e;
  auto rs = rounds(ids, nids, ncells, self_only, no_reads);
  for (auto &r : rs)
    if (int e = launch_ids(ctx, B, r, [](tc_context *cx, Batch &b, const int *d, int n) {
          return launch_l2p(cx->cfg.p, b.cells(), b.bodies(), b.results(), d, n, cx->stream)
                     ? -2
                     : 0;
        }))
      return e;
  return store_bodies(ctx, B, bodies);
}

void fill_layout(const Layout &l, tc_cell_layout *out) {
  auto o = [](size_t v) { return v == kNone ? -1 : (int)v; };
  out->size = o(l.size);
  out->nchild = o(l.nchild);
  out->nbody = o(l.nbody);
  out->child = o(l.child);
  out->body = o(l.body);
  out->X = o(l.X);
  out->R = o(l.R);
  out->min_acc = o(l.min_acc);
  out->cell_mass = o(l.cell_mass);
  out->has_sink = o(l.has_sink);
  out->M = o(l.M);
  out->L = o(l.L);
  out->Mdot = o(l.Md);
  out->Ldot = o(l.Ld);
  out->Pn = o(l.Pn);
}

}  // namespace

extern "C" {

int tc_get_cell_layout(int p, tc_cell_layout *out) {
  if (!out || p < 1 || p > kMaxP) return -1;
  fill_layout(layout_for(p), out);
  return 0;
}

int tc_P2M(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids) {
  if (!ctx) return -1;
  return do_P2M(ctx, layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells, ids, nids);
}

int tc_M2M(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids) {
  if (!ctx) return -1;
  return do_M2M(ctx, layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells, ids, nids);
}

int tc_M2L(tc_context *ctx, void *cells, int ncells, const int *targets, const int *sources,
           int npairs) {
  if (!ctx) return -1;
  if (int e = check_pairs(ctx, targets, sources, npairs, ncells)) return e;
  Batch B;
  if (int e = load<tc_body>(ctx, B, layout_for(ctx->cfg.p), nullptr, 0, cells, ncells)) return e;
  std::vector<int> tgts;
  if (int e = run_m2l(ctx, B, targets, sources, npairs, tgts)) return e;
  return store_cells(ctx, B, cells, tgts, kLocal);
}

int tc_L2L(tc_context *ctx, void *cells, int ncells, const int *ids, int nids) {
  if (!ctx) return -1;
  return do_L2L(ctx, layout_for(ctx->cfg.p), cells, ncells, ids, nids);
}

int tc_L2P(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *ids, int nids) {
  if (!ctx) return -1;
  return do_L2P(ctx, layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells, ids, nids);
}

int tc_P2P(tc_context *ctx, tc_body *bodies, int nbodies, void *cells, int ncells,
           const int *targets, const int *sources, int npairs) {
  if (!ctx) return -1;
  if (!bodies) return fail(ctx, "bodies are required");
  if (int e = check_pairs(ctx, targets, sources, npairs, ncells)) return e;
  Batch B;
  if (int e = load(ctx, B, layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells)) return e;
  std::vector<int> off, cnt, src;
  group_pairs(ncells, targets, sources, npairs, off, cnt, src);
  std::vector<int2> ranges(src.size());
  for (size_t k = 0; k < src.size(); k++) ranges[k] = make_int2(B.body[src[k]], B.nbody[src[k]]);
  // work items: target groups (the context's split of each cell's bodies) x
  // list slices of <= p2p_list entries
  std::vector<int4> items;
  P2PSplit &sp = ctx->p2p_split;
  const int per_lane = p2p_targets_per_lane(TC_P2P_FP64, ctx->p2p_per_lane);
  sp.configure(ctx->p2p_stage, (ctx->p2p_tile * 7) / 8, per_lane);
  const int kList = ctx->p2p_list;
  for (int i = 0; i < ncells; i++) {
    if (!cnt[i] || B.nbody[i] == 0) continue;
    const int nb = B.nbody[i];
    sp.grow(nb);
    const int *gs = sp.sizes(nb);
    int o = 0;
    for (int g = 0; g < sp.count(nb); g++) {
      for (int e0 = 0; e0 < cnt[i]; e0 += kList)
        items.push_back(make_int4(i, o, e0, std::min(kList, cnt[i] - e0) | (gs[g] << 16)));
      o += gs[g];
    }
  }
  if (!items.empty()) {
    hipStream_t s = ctx->stream;
    TC_CHECK(B.aux1.ensure(sizeof(int4) * items.size()));
    TC_CHECK(B.aux2.ensure(4 * (size_t)ncells * 2));
    TC_CHECK(B.aux3.ensure(sizeof(int2) * ranges.size()));
    TC_CHECK(B.npairs.ensure(8));
    TC_CHECK(hipMemcpyAsync(B.aux1.p, items.data(), sizeof(int4) * items.size(),
                            hipMemcpyHostToDevice, s));
    TC_CHECK(hipMemcpyAsync(B.aux2.p, off.data(), 4 * (size_t)ncells, hipMemcpyHostToDevice, s));
    TC_CHECK(hipMemcpyAsync(B.aux2.as<int>() + ncells, cnt.data(), 4 * (size_t)ncells,
                            hipMemcpyHostToDevice, s));
    TC_CHECK(hipMemcpyAsync(B.aux3.p, ranges.data(), sizeof(int2) * ranges.size(),
                            hipMemcpyHostToDevice, s));
    TC_CHECK(hipMemsetAsync(B.npairs.p, 0, 8, s));
    launch_p2p(B.cells(), B.bodies(), B.results(), B.aux1.as<int4>(), (int)items.size(),
               B.aux2.as<int>(), B.aux2.as<int>() + ncells, B.aux3.as<int2>(),
               B.npairs.as<unsigned long long>(), s, TC_P2P_FP64, 0.0, ctx->p2p_tile,
               per_lane);
    TC_CHECK(hipStreamSynchronize(s));
    TC_CHECK(hipGetLastError());
  }
  return store_bodies(ctx, B, bodies);
}

// ---- master Cell / Body --------------------------------------------------------

int tc_get_master_cell_layout(int p, tc_cell_layout *out) {
  if (!out || p < 1 || p > kMaxP) return -1;
  fill_layout(master_layout_for(p), out);
  return 0;
}

int tc_master_P2M(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *ids, int nids) {
  if (!ctx) return -1;
  return do_P2M(ctx, master_layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells, ids, nids);
}

int tc_master_M2M(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *ids, int nids) {
  if (!ctx) return -1;
  return do_M2M(ctx, master_layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells, ids, nids);
}

// Symmetric M2L_rotate (master kernel_fast_lean.h:1034-1296): an entry (i, j)
// does nothing unless i < j (the reference compares the Cell pointers) and
// one of the two holds a sink; it then adds j's field to i's local when i
// holds a sink and i's field to j's local when j does.  As one-sided
// translations: (i <- j) and (j <- i), each when its target holds a sink.
int tc_master_M2L(tc_context *ctx, void *cells, int ncells, const int *ci, const int *cj,
                  int npairs) {
  if (!ctx) return -1;
  if (int e = check_pairs(ctx, ci, cj, npairs, ncells)) return e;
  for (int k = 0; k < npairs; k++)
    if (ci[k] == cj[k]) return fail(ctx, "M2L of a cell with itself is undefined");
  Batch B;
  if (int e = load<tc_master_body>(ctx, B, master_layout_for(ctx->cfg.p), nullptr, 0, cells,
                                   ncells))
    return e;
  std::vector<int> tg, sc;
  for (int k = 0; k < npairs; k++) {
    const int i = ci[k], j = cj[k];
    if (i > j || (!B.sink[i] && !B.sink[j])) continue;
    if (B.sink[i]) tg.push_back(i), sc.push_back(j);
    if (B.sink[j]) tg.push_back(j), sc.push_back(i);
  }
  std::vector<int> tgts;
  if (int e = run_m2l(ctx, B, tg.data(), sc.data(), (int)tg.size(), tgts)) return e;
  return store_cells(ctx, B, cells, tgts, kLocal);
}

int tc_master_L2L(tc_context *ctx, void *cells, int ncells, const int *ids, int nids) {
  if (!ctx) return -1;
  return do_L2L(ctx, master_layout_for(ctx->cfg.p), cells, ncells, ids, nids);
}

int tc_master_L2P(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *ids, int nids) {
  if (!ctx) return -1;
  return do_L2P(ctx, master_layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells, ids, nids);
}

// P2P_simple (master kernel_fast_lean.h:268-460) or, with TC_MASTER_STEPS,
// P2P (:462-674), whose every call adds (min over the source cell of tau)^-4
// to each sink body's timestep.
int tc_master_P2P(tc_context *ctx, tc_master_body *bodies, int nbodies, void *cells, int ncells,
                  const int *targets, const int *sources, int npairs, int mode,
                  double dt_param) {
  if (!ctx) return -1;
  if (!bodies) return fail(ctx, "bodies are required");
  if (mode & ~(TC_MASTER_STEPS | TC_MASTER_PAIR_MASS)) return fail(ctx, "unknown mode bits");
  const bool steps = mode & TC_MASTER_STEPS;
  if (steps && !(dt_param > 0)) return fail(ctx, "TC_MASTER_STEPS needs dt_param > 0");
  if (int e = check_pairs(ctx, targets, sources, npairs, ncells)) return e;
  Batch B;
  if (int e = load(ctx, B, master_layout_for(ctx->cfg.p), bodies, nbodies, cells, ncells))
    return e;
  if (npairs > 0) {
    hipStream_t s = ctx->stream;
    std::vector<int2> pairs(npairs);
    for (int k = 0; k < npairs; k++) pairs[k] = make_int2(targets[k], sources[k]);
    TC_CHECK(B.aux1.ensure(sizeof(int2) * (size_t)npairs));
    TC_CHECK(hipMemcpyAsync(B.aux1.p, pairs.data(), sizeof(int2) * (size_t)npairs,
                            hipMemcpyHostToDevice, s));
    launch_master_p2p(B.cells(), B.bodies(), B.aux1.as<int2>(), npairs, steps,
                      (mode & TC_MASTER_PAIR_MASS) != 0, dt_param, B.results(),
                      B.bts.as<double>(), s);
    TC_CHECK(hipStreamSynchronize(s));
    TC_CHECK(hipGetLastError());
  }
  return store_bodies(ctx, B, bodies, steps);
}

}  // extern "C"
