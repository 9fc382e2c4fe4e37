// Host-side file formats around the hot path: PETSc binary Vec/IS files, PVD
// collections, option-driven file lists and the StagYY Yin-Yang binary ingest.
// Every function returns 0 or a PETSc error code (the message goes through
// set_err, as for the rest of the library).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lrb {

struct OptView;

namespace io {

// PETSc error codes of the file layer
enum { E_FILE_OPEN = 65, E_FILE_READ = 66, E_FILE_WRITE = 67, E_FILE_UNEXPECTED = 79, E_USER = 83,
       E_UNKNOWN_TYPE = 86 };

// Object class ids in the PETSc binary format (first header word of a record)
constexpr int64_t VEC_FILE_CLASSID = 1211214;
constexpr int64_t MAT_FILE_CLASSID = 1211216;
constexpr int64_t IS_FILE_CLASSID = 1211218;

// PETSC_MAX_PATH_LEN: expanded file names are limited to this length minus one
constexpr int MAX_PATH_LEN = 4096;

// Vec: [classid, n] header words of `iw` bytes (4 = 32-bit PetscInt, 8 = 64-bit),
// then n big-endian doubles.  `path.info` receives "-vecload_block_size bs" when
// bs > 1 (created empty otherwise), as PETSc's binary viewer writes it.
int vec_write(const char* path, int64_t n, const double* v, int bs, int iw);
// Header only (n) and block size from `path.info` (1 when absent).
int vec_info(const char* path, int iw, int64_t& n, int& bs);
int vec_read(const char* path, int iw, std::vector<double>& out, int& bs);

// IS: [classid, n] then n indices, all words `iw` bytes, big-endian.  Writing
// checks that every index fits the word width.
int is_write(const char* path, int64_t n, const int64_t* idx, int iw);
int is_read(const char* path, int iw, std::vector<int64_t>& out);

// PVD collection files (ParaView), byte-for-byte as io_utils.c:91-209 writes them.
int pvd_open(const char* path);
int pvd_append(const char* path, double time, const char* datafile, const char* directory);
int pvd_open_append(bool not_first, int64_t step, const char* path, double time, const char* datafile,
                    const char* directory);

// "{directory}{prefix}{separator}{field}{step:05}" -> file name (io_utils.c:322-413)
int filename_from_template(const std::string& tmpl, const std::string& directory, const std::string& prefix,
                           const std::string& field, const std::string& separator, int64_t step, std::string& out);

// File list from options -<prefix>fname_* / -<prefix>dirname_* (io_utils.c:415-483).
// `coords` is empty when -<prefix>fname_coordinates is not given.
int filelist_from_options(const OptView& o, const std::string& prefix, int64_t step, std::vector<std::string>& files,
                          std::string& coords);

// Points with coordinates [n*dim] and field values [n*nfields] (row-major).
struct Dataset {
  int dim = 0;
  int nfields = 0;
  std::vector<double> coords, fields;
  int64_t n() const { return dim > 0 ? (int64_t)(coords.size() / dim) : 0; }
};

// StagYY binary files (binary_loader.c:48-531): is_vp[f] marks the
// velocity-pressure file (4 components, first in the field blocks), the others
// are scalar fields.  Yin-Yang coordinates are rebuilt in Cartesian form and the
// points of the overlap are removed.
int stag_load(std::vector<std::string> files, std::vector<int> is_vp, int nfields, Dataset& ds);
// PETSc binary files (petsc_loader.c:38-101): coordinates from `coords`, fields
// appended block after block in file order.
int petsc_load(const std::string& coords, const std::vector<std::string>& files, int dim, int nfields, Dataset& ds);

// -data_format 0 (StagYY, options -data_fname_*) or 1 (PETSc binary, options
// -swarm_data_fname_*), as SwarmDataLoad dispatches (swarm_data.c:161-180).
int load_data_from_options(const OptView& o, int64_t step, int dim, int nfields, Dataset& ds);

}  // namespace io
}  // namespace lrb
