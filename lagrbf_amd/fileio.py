"""File formats of the C side (include/lagrbf_io.h in liblagrbf.so), with the
reference's C function names: PETSc binary Vec/IS files as PETSc's binary
viewer writes them (swarm_output.c:59-167, swarm_load.c:36-118), PVD
collections (io_utils.c:91-209), option-driven file lists (io_utils.c:322-483)
and the data readers (src/readers/: StagYY Yin-Yang binary, PETSc binary).

These are host-side functions: they need the built library but no GPU.
"""
import ctypes

import numpy as np

from ._lib import cfuncs, last_error

__all__ = ["VecWrite", "VecLoad", "ISWrite", "ISLoad", "ParaviewPVDOpen", "ParaviewPVDAppend",
           "ParaviewPVDOpenAppend", "BuildFilenameFromTemplate", "GetFilelistFromOptions",
           "ReadersGetRankDistribution", "Dataset", "StagBinSwarmSetFromFiles", "SwarmSetFromPetscFiles",
           "SwarmDataLoad"]


def _b(s):
    return None if s is None else str(s).encode()


def _check(r, name):
    if r != 0:
        raise RuntimeError(f"{name} failed with error code {r}: {last_error()}")


def _strings(items):
    arr = (ctypes.c_char_p * max(len(items), 1))()
    for i, s in enumerate(items):
        arr[i] = _b(s)
    return arr


def VecWrite(path, values, bs=None, int64_indices=False):
    """PETSc binary Vec of ``values`` (row-major; bs = number of columns of a
    2-D array unless given) and its ".info" block-size file."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    if bs is None:
        bs = v.shape[1] if v.ndim > 1 else 1
    _check(cfuncs()["lagrbf_io_vec_write"](_b(path), v.size, v.reshape(-1), int(bs), int(int64_indices)), "VecView")


def VecLoad(path, int64_indices=False):
    """Entries of a PETSc binary Vec, shaped (n / bs, bs) when ".info" gives bs > 1."""
    f = cfuncs()["lagrbf_io_vec_read"]
    n, bs = ctypes.c_int64(0), ctypes.c_int32(1)
    _check(f(_b(path), int(int64_indices), ctypes.byref(n), ctypes.byref(bs), None, 0), "VecLoad")
    out = np.empty(n.value, dtype=np.float64)
    _check(f(_b(path), int(int64_indices), ctypes.byref(n), ctypes.byref(bs), out, out.size), "VecLoad")
    return out.reshape(-1, bs.value) if bs.value > 1 else out


def ISWrite(path, indices, int64_indices=False):
    idx = np.ascontiguousarray(indices, dtype=np.int64).reshape(-1)
    _check(cfuncs()["lagrbf_io_is_write"](_b(path), idx.size, idx, int(int64_indices)), "ISView")


def ISLoad(path, int64_indices=False):
    f = cfuncs()["lagrbf_io_is_read"]
    n = ctypes.c_int64(0)
    _check(f(_b(path), int(int64_indices), ctypes.byref(n), None, 0), "ISLoad")
    out = np.empty(n.value, dtype=np.int64)
    _check(f(_b(path), int(int64_indices), ctypes.byref(n), out, out.size), "ISLoad")
    return out


def ParaviewPVDOpen(path):
    _check(cfuncs()["lagrbf_io_pvd_open"](_b(path)), "ParaviewPVDOpen")


def ParaviewPVDAppend(path, time, datafile, directory=None):
    _check(cfuncs()["lagrbf_io_pvd_append"](_b(path), float(time), _b(datafile), _b(directory)), "ParaviewPVDAppend")


def ParaviewPVDOpenAppend(not_first, step, path, time, datafile, directory=None):
    _check(cfuncs()["lagrbf_io_pvd_open_append"](bool(not_first), int(step), _b(path), float(time), _b(datafile),
                                                 _b(directory)), "ParaviewPVDOpenAppend")


def BuildFilenameFromTemplate(template, directory="", prefix="", field="", separator="", step=0):
    buf = ctypes.create_string_buffer(4096)
    _check(cfuncs()["lagrbf_io_filename"](_b(template), _b(directory), _b(prefix), _b(field), _b(separator), int(step),
                                          buf, len(buf)), "BuildFilenameFromTemplate")
    return buf.value.decode()


def GetFilelistFromOptions(prefix, step):
    """(field file names, coordinate file name or None) from the options
    -<prefix>fname_* / -<prefix>dirname_* set through lagrbf_set_option."""
    buf = ctypes.create_string_buffer(1 << 20)
    n = ctypes.c_int32(0)
    _check(cfuncs()["lagrbf_io_filelist"](_b(prefix), int(step), buf, len(buf), ctypes.byref(n)),
           "GetFilelistFromOptions")
    lines = buf.value.decode().split("\n")
    return lines[1:1 + n.value], (lines[0] or None)


def ReadersGetRankDistribution(nfiles, rank, nranks):
    """(start, end, count) of the contiguous block of files rank reads
    (reader_utils.c:30-61): the first nfiles % nranks ranks take one more."""
    per, rem = divmod(int(nfiles), int(nranks))
    if rank < rem:
        start = rank * (per + 1)
        end = start + per + 1
    else:
        start = rank * per + rem
        end = start + per
    return start, end, end - start


class Dataset:
    """Points loaded by a reader, held by the library until freed: coordinates
    [n, dim] and field values [n, nfields]; ``LagRBFCtx.fill_from_dataset``
    moves them into a swarm without a copy through Python."""

    def __init__(self, handle, name):
        if not handle:
            raise RuntimeError(f"{name} failed: {last_error()}")
        self.handle = handle
        n, d, nf = ctypes.c_int64(0), ctypes.c_int32(0), ctypes.c_int32(0)
        _check(cfuncs()["lagrbf_io_dataset_info"](handle, ctypes.byref(n), ctypes.byref(d), ctypes.byref(nf)), name)
        self.n, self.dim, self.nfields = n.value, d.value, nf.value

    def arrays(self):
        coords = np.empty((self.n, self.dim), dtype=np.float64)
        fields = np.empty((self.n, self.nfields), dtype=np.float64)
        _check(cfuncs()["lagrbf_io_dataset_copy"](self.handle, coords.reshape(-1), fields.reshape(-1)),
               "lagrbf_io_dataset_copy")
        return coords, fields

    def free(self):
        if getattr(self, "handle", None):
            cfuncs()["lagrbf_io_dataset_free"](self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def StagBinSwarmSetFromFiles(files, is_vp, nfields):
    """StagYY binary files -> Dataset (3-D Yin-Yang points, overlap removed)."""
    files = list(files)
    vp = (ctypes.c_int32 * max(len(files), 1))(*[int(bool(x)) for x in is_vp])
    if len(is_vp) != len(files):
        raise ValueError("one velocity-pressure flag per file is required")
    return Dataset(cfuncs()["lagrbf_io_stag_load"](_strings(files), vp, len(files), int(nfields)),
                   "StagBinSwarmSetFromFiles")


def SwarmSetFromPetscFiles(coords_file, files, dim, nfields):
    """PETSc binary coordinates + field files -> Dataset."""
    files = list(files)
    return Dataset(cfuncs()["lagrbf_io_petsc_load"](_b(coords_file), _strings(files), len(files), int(dim),
                                                    int(nfields)), "SwarmSetFromPetscFiles")


def SwarmDataLoad(step, dim, nfields):
    """Data files of ``step`` named by the options (-data_format 0: StagYY with
    -data_fname_*, 1: PETSc binary with -swarm_data_fname_*) -> Dataset."""
    return Dataset(cfuncs()["lagrbf_io_load_data"](int(step), int(dim), int(nfields)), "SwarmDataLoad")
