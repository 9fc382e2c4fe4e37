/*
 * lagrbf_io.h -- file formats around the LagRBF hot path (liblagrbf.so, host code).
 *
 * In the reference these are C functions its drivers call directly (no Python
 * binding): PETSc binary Vec/IS files through PETSc's binary viewer, PVD
 * collections (src/io_utils.c), the swarm output/restart files
 * (src/swarm/swarm_output.c, src/swarm/swarm_load.c) and the data readers
 * (src/readers/).  Here they are C-ABI functions without PETSc; a Python or C
 * caller binds them as it binds lagrbf.h.
 *
 * PETSc binary layout: big-endian header words [class id, length] of PetscInt
 * width (int64_indices = 0: 4 bytes, the reference's 32-bit-index build; 1: 8
 * bytes), then the entries (double for Vec, PetscInt for IS).  A Vec of block
 * size bs > 1 has "-vecload_block_size bs" in "<file>.info".
 *
 * Every int function returns 0 or a PETSc error code (65 file open, 66 file
 * read, 67 file write, 60 size mismatch, 62 wrong argument, 63 out of range,
 * 83 user error, 85 null argument, 86 unknown type); the message is kept for
 * lagrbf_last_error().  Functions returning a handle return NULL on error.
 */
#ifndef LAGRBF_IO_H
#define LAGRBF_IO_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- PETSc binary Vec / IS ---------------- */

/* VecView on a binary viewer (used by SwarmOutputFieldPetscVec_FromName,
 * src/swarm/swarm_output.c:141-167): values [n], n divisible by bs. */
int lagrbf_io_vec_write(const char *path, int64_t n, const double *values, int32_t bs, int32_t int64_indices);

/* VecLoad (VecLoadFromFile, src/readers/petsc_loader/petsc_loader.c): with
 * values == NULL only the length n and the block size bs (from "<path>.info",
 * 1 when absent) are returned; else the n values are written to values
 * (capacity entries available). */
int lagrbf_io_vec_read(const char *path, int32_t int64_indices, int64_t *n, int32_t *bs, double *values,
                       int64_t capacity);

/* ISView on a binary viewer (SwarmOutputFieldPetscIS_FromName,
 * src/swarm/swarm_output.c:59-138); indices must fit the index width. */
int lagrbf_io_is_write(const char *path, int64_t n, const int64_t *idx, int32_t int64_indices);

/* ISLoad (SwarmLoadPidxFromFile, src/swarm/swarm_load.c:89-118); idx == NULL
 * returns the length only. */
int lagrbf_io_is_read(const char *path, int32_t int64_indices, int64_t *n, int64_t *idx, int64_t capacity);

/* ---------------- PVD collections ---------------- */

/* ParaviewPVDOpen, src/io_utils.c:91-115 */
int lagrbf_io_pvd_open(const char *path);
/* ParaviewPVDAppend, src/io_utils.c:117-168 (directory NULL or "" -> "./file") */
int lagrbf_io_pvd_append(const char *path, double time, const char *datafile, const char *directory);
/* ParaviewPVDOpenAppend, src/io_utils.c:170-209 */
int lagrbf_io_pvd_open_append(bool not_first, int64_t step, const char *path, double time, const char *datafile,
                              const char *directory);

/* ---------------- file names ---------------- */

/* BuildFilenameFromTemplate, src/io_utils.c:322-413: tokens {directory} (a '/'
 * is added when missing), {prefix}, {field}, {separator}, {step}, {step:N}. */
int lagrbf_io_filename(const char *template_name, const char *directory, const char *prefix, const char *field,
                       const char *separator, int64_t step, char *out, int32_t len);

/* GetFilelistFromOptions, src/io_utils.c:415-483, from the options channel of
 * lagrbf.h (-<prefix>fname_template, fname_prefix, fname_separator,
 * dirname_template, dirname_prefix, dirname_base, fname_nfiles, fname_fields,
 * fname_coordinates).  out receives the coordinate file name (empty when not
 * given) and then each field file, every name followed by '\n'. */
int lagrbf_io_filelist(const char *prefix, int64_t step, char *out, int32_t len, int32_t *nfiles);

/* ---------------- data readers ---------------- */

/* StagBinSwarmSetFromFiles, src/readers/stagYY/binary_loader.c:483-531:
 * StagYY binary files, is_vp[f] != 0 for the velocity-pressure file (4 fields,
 * moved first), else one scalar field each; Cartesian Yin-Yang coordinates,
 * overlap points removed.  Returns a dataset handle. */
void *lagrbf_io_stag_load(const char **files, const int32_t *is_vp, int32_t nfiles, int32_t nfields);

/* SwarmSetFromPetscFiles, src/readers/petsc_loader/petsc_loader.c:85-101:
 * coordinates (block size dim) and field files appended block after block. */
void *lagrbf_io_petsc_load(const char *coords_file, const char **files, int32_t nfiles, int32_t dim,
                           int32_t nfields);

/* SwarmDataLoad, src/swarm/swarm_data.c:161-180: -data_format 0 (StagYY,
 * options -data_fname_*) or 1 (PETSc binary, options -swarm_data_fname_*). */
void *lagrbf_io_load_data(int64_t step, int32_t dim, int32_t nfields);

int lagrbf_io_dataset_info(void *dataset, int64_t *n, int32_t *dim, int32_t *nfields);
/* coords [n*dim], fields [n*nfields]; either may be NULL */
int lagrbf_io_dataset_copy(void *dataset, double *coords, double *fields);
void lagrbf_io_dataset_free(void *dataset);

/* ---------------- context-level ---------------- */

/* SwarmOutput, src/swarm/swarm_output.c:169-234, for the particle swarm:
 * <root>/step<step>/step%06d[_forward|_backward].swarm_interp_<field>.pbvec for the
 * coordinates (DMSwarmPIC_coor) and each listed field ("fields", the velocity
 * name, "weights"; "DMSwarm_pid" as .pbis), and an entry of
 * <root>/timeseries_X.pvd.  direction: 0 none, 1 forward, 2 backward.  Rows are
 * the live particles in fill order (the pid file keys them). */
int lagrbf_swarm_output(void *ctx, const char *root, int64_t step, double time, int32_t nfields,
                        const char **fields, int32_t direction, bool permute);

/* SwarmLoadFromFiles, src/swarm/swarm_load.c:133-201: refills the particle swarm
 * from the files lagrbf_swarm_output wrote in dirname (coordinates, plus
 * "fields" and "DMSwarm_pid" when listed). */
int lagrbf_swarm_load(void *ctx, const char *dirname, int64_t step, int32_t direction, int32_t nfields,
                      const char **fields);

/* SwarmDataLoad + fill of the data swarm (time_loop.c:146-147 order): reads the
 * step's files named by the options and refills the node swarm. */
int lagrbf_load_data(void *ctx, int64_t step);

/* Fill the data (is_data) or particle swarm from a dataset handle. */
int lagrbf_fill_swarm_from_dataset(void *ctx, bool is_data, void *dataset);

#ifdef __cplusplus
}
#endif

#endif /* LAGRBF_IO_H */
