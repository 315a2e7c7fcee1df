// Test infrastructure only: the reference's own snapshot I/O
// (taichi_with_jerks/io.h:11-334, or the master io.h:11-303) compiled from
// its headers where they lie (oracle/build_ref.sh), to pin libtcio against.
// write_snapshot writes "snapshot_<num>.hdf5" into the working directory and
// read_snapshot reads the global input_fname from it, so both calls run with
// the working directory set to `dir`.
#include <unistd.h>

#include <climits>
#include <cstdlib>
#include <cstring>

#include "exafmm.h"
#include "io.h"

using namespace exafmm;

namespace {
struct InDir {
  char old[PATH_MAX];
  bool ok = false;
  explicit InDir(const char *dir) {
    ok = getcwd(old, sizeof old) != nullptr && chdir(dir) == 0;
  }
  ~InDir() {
    if (ok && chdir(old) != 0) std::abort();
  }
};
}  // namespace

extern "C" {

int refio_particle_size(void) { return (int)sizeof(particle); }

// write_snapshot(num, s) over n records (io.h:106): the header takes the
// globals numBodies and t_now.
int refio_write(const char *dir, int num, const particle *p, unsigned n, double t) {
  InDir d(dir);
  if (!d.ok) return -1;
  sys s;
  s.n = n;
  s.part = const_cast<particle *>(p);
  s.last = n ? s.part + n - 1 : s.part;
  numBodies = n;
  t_now = t;
  write_snapshot(num, s);
  return 0;
}

// read_snapshot (io.h:11): input_fname is the file name relative to dir; the
// records land in a fresh mainsys.part, copied out here (at most cap).
int refio_read(const char *dir, const char *fname, particle *out, unsigned cap, unsigned *n,
               double *t, int *num) {
  InDir d(dir);
  if (!d.ok) return -1;
  std::strncpy(input_fname, fname, sizeof input_fname - 1);
  read_snapshot(0);
  *n = numBodies;
  *t = t_now;
  *num = snapnum;
  unsigned m = numBodies < cap ? numBodies : cap;
  std::memcpy(out, mainsys.part, sizeof(particle) * m);
  std::free(mainsys.part);
  mainsys.part = nullptr;
  return 0;
}
}
