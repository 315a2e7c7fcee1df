/* tcio: the taichi_collisional HDF5 snapshot format (host code, libtcio.so).
 *
 * Replaces the reference's write_snapshot / read_snapshot
 * (taichi_with_jerks/io.h:11-103 and :106-334; master io.h:11-303), which
 * taichi.cxx:148,163 and main.cxx:146,172 call.  A snapshot file holds
 *
 *   /Header  group, attributes NumPartTotal (native unsigned int, scalar) and
 *            Time (native double, scalar);
 *   Posx Posy Posz Velx Vely Velz Mass          1-D native double [n];
 *   Potential                                   with TC_SNAP_POT   (OUTPUTPOT);
 *   Stepsize Scalarforce                        with TC_SNAP_DEBUG (DEBUG; timestep, acc_old);
 *   Accx Accy Accz                              with TC_SNAP_ACC   (OUTPUTACC);
 *   Jerkx Jerky Jerkz                           with TC_SNAP_JERK  (the jerk build's OUTPUTACC);
 *   ID                                          1-D native unsigned int [n];
 *
 * in that order, every dataset in increasing particle id (io.h:154-160).  Both
 * reference Makefiles define OUTPUTPOT, OUTPUTACC and DEBUG, so their files
 * hold TC_SNAP_JERK_BUILD (jerk) or TC_SNAP_MASTER_BUILD (master) fields.
 *
 * Records are the reference's struct particle (168 bytes, tc_particle in
 * tcfmm.h).  Every function returns 0 on success and a negative code on error
 * (tc_io_last_error names it); nothing is printed.  Not thread-safe: libhdf5
 * as built in this image is not.
 */
#ifndef TCIO_H
#define TCIO_H

#include "tcfmm.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TC_SNAP_POT = 1,   /* Potential */
  TC_SNAP_DEBUG = 2, /* Stepsize (timestep), Scalarforce (acc_old) */
  TC_SNAP_ACC = 4,   /* Accx, Accy, Accz */
  TC_SNAP_JERK = 8,  /* Jerkx, Jerky, Jerkz */
  TC_SNAP_MASTER_BUILD = TC_SNAP_POT | TC_SNAP_DEBUG | TC_SNAP_ACC,
  TC_SNAP_JERK_BUILD = TC_SNAP_MASTER_BUILD | TC_SNAP_JERK
};

/* write_snapshot (taichi_with_jerks/io.h:106-334): writes the n records of
 * part, sorted by id, and the header (n, t_now) to `path` (the reference
 * names it "snapshot_<num>.hdf5" in the working directory), replacing any
 * file there.  `fields` selects the optional datasets above.  Records with
 * equal ids keep their input order. */
int tc_write_snapshot(const char *path, const tc_particle *part, unsigned n, double t_now,
                      int fields);

/* The header of a snapshot (io.h:19-31): NumPartTotal and Time. */
int tc_read_snapshot_header(const char *path, unsigned *n, double *t_now);

/* read_snapshot (io.h:11-103): fills pos, vel, mass and id of the first n
 * records of part from the file's datasets (n must equal NumPartTotal); the
 * other fields are set to zero.  The reference reads only these. */
int tc_read_snapshot(const char *path, tc_particle *part, unsigned n);

/* The snapshot number the reference parses from its input file name
 * (io.h:15-17): the digits between "snapshot_" and ".hdf5" of the file's
 * base name, or -1 when the base name is not of that form.  (The reference
 * copies the digits into a static buffer without a terminator, so a name with
 * fewer digits than an earlier one keeps that one's trailing digits; this
 * function parses each name on its own.) */
int tc_snapshot_number(const char *path);

/* Any 1-D dataset of a snapshot, converted to double (as_uint = 0) or to
 * unsigned int (as_uint = 1); n must equal its length.  Reads the optional
 * datasets back (a restart that keeps acc_old, the error checks of
 * plot_jerk_error.py). */
int tc_read_snapshot_dataset(const char *path, const char *name, void *out, unsigned n,
                             int as_uint);

/* What is stored under `name` ("Header/Time" etc. for attributes):
 * *cls = 0 float, 1 integer; *size in bytes; *is_signed; *length = number of
 * elements (1 for a scalar).  Returns a negative code if there is none. */
int tc_snapshot_describe(const char *path, const char *name, int *cls, int *size,
                         int *is_signed, long long *length);

/* Names of the file's root-level objects in name order (the reference's
 * files do not record creation order), separated by '\n', written into buf
 * (cap bytes, NUL-terminated). */
int tc_snapshot_list(const char *path, char *buf, int cap);

const char *tc_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TCIO_H */
