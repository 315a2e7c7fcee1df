// The taichi_collisional HDF5 snapshot format: write_snapshot / read_snapshot
// (taichi_with_jerks/io.h:11-334, master io.h:11-303) over the reference's
// struct particle.  Host code only, built with g++ against libhdf5 into
// libtcio.so, separate from the HIP library so that libtcfmm.so does not
// depend on HDF5.
#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/tcio.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &what) {
  g_err = what;
  return code;
}

// Turns libhdf5's own error printing off for the duration of one call and
// restores the caller's handler after it.
struct QuietH5 {
  H5E_auto2_t fn = nullptr;
  void *data = nullptr;
  QuietH5() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// Closes an HDF5 handle when it leaves scope.
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
  bool ok() const { return id >= 0; }
};

// One scalar attribute of /Header (io.h:124-136).
int write_attr(hid_t grp, const char *name, hid_t type, const void *v) {
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  if (!sp.ok()) return fail(-3, std::string("H5Screate for ") + name);
  Hid a(H5Acreate2(grp, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return fail(-3, std::string("H5Acreate ") + name);
  if (H5Awrite(a, type, v) < 0) return fail(-3, std::string("H5Awrite ") + name);
  return 0;
}

// One 1-D dataset of n elements (io.h:162-170 and its repeats).
int write_dset(hid_t file, const char *name, hid_t type, hsize_t n, const void *data) {
  hsize_t dims[1] = {n};
  Hid sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
  if (!sp.ok()) return fail(-3, std::string("H5Screate_simple for ") + name);
  Hid d(H5Dcreate2(file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(-3, std::string("H5Dcreate ") + name);
  if (H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(-3, std::string("H5Dwrite ") + name);
  return 0;
}

// Reads the 1-D dataset `name` (n elements) converted to `memtype`.
int read_dset(hid_t file, const char *name, hid_t memtype, unsigned n, void *out) {
  Hid d(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(-4, std::string("no dataset ") + name);
  Hid sp(H5Dget_space(d), H5Sclose);
  hssize_t len = sp.ok() ? H5Sget_simple_extent_npoints(sp) : -1;
  if (len != (hssize_t)n)
    return fail(-4, std::string("dataset ") + name + " holds " + std::to_string((long long)len) +
                        " elements, expected " + std::to_string(n));
  if (n && H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    return fail(-4, std::string("H5Dread ") + name);
  return 0;
}

int read_header(hid_t file, unsigned *n, double *t) {
  Hid g(H5Gopen2(file, "/Header", H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail(-4, "no /Header group");
  Hid a(H5Aopen(g, "NumPartTotal", H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Aread(a, H5T_NATIVE_UINT, n) < 0) return fail(-4, "no Header/NumPartTotal");
  Hid b(H5Aopen(g, "Time", H5P_DEFAULT), H5Aclose);
  if (!b.ok() || H5Aread(b, H5T_NATIVE_DOUBLE, t) < 0) return fail(-4, "no Header/Time");
  return 0;
}

}  // namespace

extern "C" {

const char *tc_io_last_error(void) { return g_err.c_str(); }

int tc_write_snapshot(const char *path, const tc_particle *part, unsigned n, double t_now,
                      int fields) {
  if (!path || (n && !part)) return fail(-1, "null path or particle array");
  if (fields & ~TC_SNAP_JERK_BUILD) return fail(-1, "unknown field bits");
  QuietH5 quiet;
  // particles in increasing id (io.h:154-160); equal ids keep their order
  std::vector<unsigned> ids(n);
  std::iota(ids.begin(), ids.end(), 0u);
  std::stable_sort(ids.begin(), ids.end(),
                   [part](unsigned a, unsigned b) { return part[a].id < part[b].id; });
  Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot create ") + path);
  {
    Hid hdr(H5Gcreate2(file, "/Header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!hdr.ok()) return fail(-3, "H5Gcreate /Header");
    int rc = write_attr(hdr, "NumPartTotal", H5T_NATIVE_UINT, &n);
    if (!rc) rc = write_attr(hdr, "Time", H5T_NATIVE_DOUBLE, &t_now);
    if (rc) return rc;
  }
  std::vector<double> col(n);
  // field of a record, as a byte offset into tc_particle
  auto put = [&](const char *name, size_t off) {
    const char *base = (const char *)part;
    for (unsigned b = 0; b < n; ++b)
      std::memcpy(&col[b], base + (size_t)ids[b] * sizeof(tc_particle) + off, sizeof(double));
    return write_dset(file, name, H5T_NATIVE_DOUBLE, n, col.data());
  };
  struct Col {
    const char *name;
    size_t off;
    int need;
  };
  // dataset order of io.h:162-320
  const Col cols[] = {
      {"Posx", offsetof(tc_particle, pos[0]), 0},
      {"Posy", offsetof(tc_particle, pos[1]), 0},
      {"Posz", offsetof(tc_particle, pos[2]), 0},
      {"Velx", offsetof(tc_particle, vel[0]), 0},
      {"Vely", offsetof(tc_particle, vel[1]), 0},
      {"Velz", offsetof(tc_particle, vel[2]), 0},
      {"Mass", offsetof(tc_particle, mass), 0},
      {"Potential", offsetof(tc_particle, pot), TC_SNAP_POT},
      {"Stepsize", offsetof(tc_particle, timestep), TC_SNAP_DEBUG},
      {"Scalarforce", offsetof(tc_particle, acc_old), TC_SNAP_DEBUG},
      {"Accx", offsetof(tc_particle, acc[0]), TC_SNAP_ACC},
      {"Accy", offsetof(tc_particle, acc[1]), TC_SNAP_ACC},
      {"Accz", offsetof(tc_particle, acc[2]), TC_SNAP_ACC},
      {"Jerkx", offsetof(tc_particle, jerk[0]), TC_SNAP_JERK},
      {"Jerky", offsetof(tc_particle, jerk[1]), TC_SNAP_JERK},
      {"Jerkz", offsetof(tc_particle, jerk[2]), TC_SNAP_JERK},
  };
  for (const Col &c : cols) {
    if (c.need && !(fields & c.need)) continue;
    if (int rc = put(c.name, c.off)) return rc;
  }
  std::vector<unsigned> idv(n);
  for (unsigned b = 0; b < n; ++b) idv[b] = part[ids[b]].id;
  if (int rc = write_dset(file, "ID", H5T_NATIVE_UINT, n, idv.data())) return rc;
  return 0;
}

int tc_read_snapshot_header(const char *path, unsigned *n, double *t_now) {
  if (!path || !n || !t_now) return fail(-1, "null argument");
  QuietH5 quiet;
  Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot open ") + path);
  return read_header(file, n, t_now);
}

int tc_read_snapshot(const char *path, tc_particle *part, unsigned n) {
  if (!path || (n && !part)) return fail(-1, "null argument");
  QuietH5 quiet;
  Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot open ") + path);
  unsigned nf = 0;
  double t = 0;
  if (int rc = read_header(file, &nf, &t)) return rc;
  if (nf != n)
    return fail(-1, "NumPartTotal is " + std::to_string(nf) + ", n is " + std::to_string(n));
  std::memset(part, 0, sizeof(tc_particle) * (size_t)n);
  std::vector<double> col(n);
  struct Col {
    const char *name;
    size_t off;
  };
  // the datasets io.h:35-93 reads, in its order
  const Col cols[] = {{"/Posx/", offsetof(tc_particle, pos[0])},
                      {"/Posy/", offsetof(tc_particle, pos[1])},
                      {"/Posz/", offsetof(tc_particle, pos[2])},
                      {"/Velx/", offsetof(tc_particle, vel[0])},
                      {"/Vely/", offsetof(tc_particle, vel[1])},
                      {"/Velz/", offsetof(tc_particle, vel[2])},
                      {"/Mass/", offsetof(tc_particle, mass)}};
  for (const Col &c : cols) {
    if (int rc = read_dset(file, c.name, H5T_NATIVE_DOUBLE, n, col.data())) return rc;
    char *base = (char *)part;
    for (unsigned b = 0; b < n; ++b)
      std::memcpy(base + (size_t)b * sizeof(tc_particle) + c.off, &col[b], sizeof(double));
  }
  std::vector<unsigned> idv(n);
  if (int rc = read_dset(file, "/ID/", H5T_NATIVE_UINT, n, idv.data())) return rc;
  for (unsigned b = 0; b < n; ++b) part[b].id = idv[b];
  return 0;
}

int tc_snapshot_number(const char *path) {
  if (!path) return -1;
  const char *base = std::strrchr(path, '/');
  base = base ? base + 1 : path;
  const size_t lp = 9, ls = 5;  // "snapshot_", ".hdf5"
  size_t len = std::strlen(base);
  if (len <= lp + ls || std::strncmp(base, "snapshot_", lp) || std::strcmp(base + len - ls, ".hdf5"))
    return -1;
  int v = 0;
  for (size_t i = lp; i < len - ls; ++i) {
    if (base[i] < '0' || base[i] > '9') return -1;
    v = v * 10 + (base[i] - '0');
  }
  return v;
}

int tc_read_snapshot_dataset(const char *path, const char *name, void *out, unsigned n,
                             int as_uint) {
  if (!path || !name || (n && !out)) return fail(-1, "null argument");
  QuietH5 quiet;
  Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot open ") + path);
  return read_dset(file, name, as_uint ? H5T_NATIVE_UINT : H5T_NATIVE_DOUBLE, n, out);
}

int tc_snapshot_describe(const char *path, const char *name, int *cls, int *size,
                         int *is_signed, long long *length) {
  if (!path || !name || !cls || !size || !is_signed || !length) return fail(-1, "null argument");
  QuietH5 quiet;
  Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot open ") + path);
  hid_t type = -1, space = -1;
  const char *slash = std::strrchr(name, '/');
  Hid obj(-1, nullptr);
  if (slash && slash != name) {  // "<group>/<attribute>"
    std::string grp(name, slash - name);
    Hid g(H5Gopen2(file, grp.c_str(), H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(-4, "no group " + grp);
    Hid a(H5Aopen(g, slash + 1, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(-4, std::string("no attribute ") + name);
    type = H5Aget_type(a);
    space = H5Aget_space(a);
  } else {
    Hid d(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(-4, std::string("no dataset ") + name);
    type = H5Dget_type(d);
    space = H5Dget_space(d);
  }
  Hid t(type, H5Tclose), s(space, H5Sclose);
  if (!t.ok() || !s.ok()) return fail(-4, std::string("cannot inspect ") + name);
  H5T_class_t c = H5Tget_class(t);
  *cls = c == H5T_FLOAT ? 0 : c == H5T_INTEGER ? 1 : 2;
  *size = (int)H5Tget_size(t);
  *is_signed = c == H5T_INTEGER ? (H5Tget_sign(t) == H5T_SGN_2) : 1;
  *length = (long long)H5Sget_simple_extent_npoints(s);
  return 0;
}

int tc_snapshot_list(const char *path, char *buf, int cap) {
  if (!path || !buf || cap <= 0) return fail(-1, "null argument");
  QuietH5 quiet;
  Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(-2, std::string("cannot open ") + path);
  std::string names;
  auto cb = [](hid_t, const char *nm, const H5L_info_t *, void *op) -> herr_t {
    std::string *s = (std::string *)op;
    if (!s->empty()) s->push_back('\n');
    s->append(nm);
    return 0;
  };
  if (H5Literate(file, H5_INDEX_NAME, H5_ITER_INC, nullptr, cb, &names) < 0)
    return fail(-4, "H5Literate");
  if ((int)names.size() >= cap) return fail(-1, "buffer too small");
  std::memcpy(buf, names.c_str(), names.size() + 1);
  return 0;
}

}  // extern "C"
