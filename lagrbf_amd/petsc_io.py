"""PETSc binary Vec / IS / Mat files with numpy alone (no PETSc installation).

Same functions, arguments and file contents as python_api/lagrbfpy/petsc_io/
(petscvec_loader.py:33-70, petscvec_writer.py:30-60), which go through
PETSc's PetscBinaryIO.  The layout is PETSc's binary format: big-endian header
words of PetscInt width -- [class id, length] for Vec and IS, [class id, M, N,
nz] for Mat -- then the entries.  Keyword arguments as PetscBinaryIO's:
precision ("double" | "single"), indices ("int32" | "int64"), complexscalars.

Loaded arrays are returned in native byte order.
"""
import numpy as np

VEC_FILE_CLASSID = 1211214
MAT_FILE_CLASSID = 1211216
IS_FILE_CLASSID = 1211218
MATRIX_BINARY_FORMAT_DENSE = -1

__all__ = ["PetscVecLoadFromFile", "PetscVecLoadBlockSizeFromFile", "PetscVecLoad", "PetscISLoadFromFile",
           "PetscMatDenseLoadFromFile", "PetscVecWriteToFile", "PetscVecWriteBlockSizeToFile", "PetscVecWrite",
           "PetscISWriteToFile", "VEC_FILE_CLASSID", "MAT_FILE_CLASSID", "IS_FILE_CLASSID"]


def _types(precision="double", indices="int32", complexscalars=False):
    if precision not in ("double", "single"):
        raise ValueError(f"precision must be 'double' or 'single', got {precision!r}")
    if indices not in ("int32", "int64"):
        raise ValueError(f"indices must be 'int32' or 'int64', got {indices!r}")
    it = np.dtype(">i4" if indices == "int32" else ">i8")
    if complexscalars:
        st = np.dtype(">c16" if precision == "double" else ">c8")
    else:
        st = np.dtype(">f8" if precision == "double" else ">f4")
    return it, st


def _read(fp, dtype, count, fname, what):
    a = np.fromfile(fp, dtype=dtype, count=count)
    if a.size != count:
        raise IOError(f"{fname}: file ended inside the {what} ({a.size} of {count} entries)")
    return a.astype(dtype.newbyteorder("="))


def _header(fp, it, classid, nwords, fname, what):
    h = _read(fp, it, nwords, fname, what + " header")
    if int(h[0]) != classid:
        raise ValueError(f"{fname}: expected a {what} (class id {classid}), found class id {int(h[0])}; "
                         f"check the indices width")
    return [int(x) for x in h[1:]]


def PetscVecLoadFromFile(fname, **kwargs):
    """Entries of the Vec stored in fname (1-D)."""
    it, st = _types(**kwargs)
    with open(fname, "rb") as fp:
        (n,) = _header(fp, it, VEC_FILE_CLASSID, 2, fname, "Vec")
        if n < 0:
            raise IOError(f"{fname}: negative Vec length {n}")
        return _read(fp, st, n, fname, "Vec entries")


def PetscVecLoadBlockSizeFromFile(fname):
    """Block size from a ".info" file whose first line is "-vecload_block_size <bs>"."""
    with open(fname) as fp:
        line = fp.readline()
        if not line.startswith("-vecload_block_size"):
            raise ValueError(f'Expected first line to be "-vecload_block_size <size>", got: {line}')
        return int(line.split()[1])


def PetscVecLoad(fname, order="C", **kwargs):
    """Vec reshaped to (-1, bs) with the block size of fname + ".info" when that
    file exists; a ".info" without the block-size line raises ValueError."""
    v = PetscVecLoadFromFile(fname, **kwargs)
    try:
        bs = PetscVecLoadBlockSizeFromFile(fname + ".info")
    except FileNotFoundError:
        return v
    return np.reshape(v, (-1, bs), order=order)


def PetscISLoadFromFile(fname, **kwargs):
    """Indices of the IS stored in fname."""
    it, _ = _types(**kwargs)
    with open(fname, "rb") as fp:
        (n,) = _header(fp, it, IS_FILE_CLASSID, 2, fname, "IS")
        if n < 0:
            raise IOError(f"{fname}: negative IS length {n}")
        return _read(fp, it, n, fname, "IS indices")


def PetscMatDenseLoadFromFile(fname, **kwargs):
    """Dense (M, N) array of the Mat stored in fname, from either the dense
    layout (nz = -1, row-major values) or the AIJ layout (row lengths, column
    indices, values)."""
    it, st = _types(**kwargs)
    with open(fname, "rb") as fp:
        M, N, nz = _header(fp, it, MAT_FILE_CLASSID, 4, fname, "Mat")
        if M < 0 or N < 0:
            raise IOError(f"{fname}: bad Mat size {M} x {N}")
        if nz == MATRIX_BINARY_FORMAT_DENSE:
            return _read(fp, st, M * N, fname, "dense Mat values").reshape(M, N)
        rowlen = _read(fp, it, M, fname, "Mat row lengths").astype(np.int64)
        if int(rowlen.sum()) != nz:
            raise IOError(f"{fname}: row lengths sum to {int(rowlen.sum())}, header says {nz} nonzeros")
        cols = _read(fp, it, nz, fname, "Mat column indices").astype(np.int64)
        vals = _read(fp, st, nz, fname, "Mat values")
    if nz and (cols.min() < 0 or cols.max() >= N):
        raise IOError(f"{fname}: column index out of range [0, {N})")
    mat = np.zeros((M, N), dtype=vals.dtype)
    rows = np.repeat(np.arange(M), rowlen)
    mat[rows, cols] = vals
    return mat


def PetscVecWriteToFile(fname, v, **kwargs):
    """Vec record [class id, n] + the n entries of v (flattened)."""
    it, st = _types(**kwargs)
    v = np.asarray(v).reshape(-1)
    with open(fname, "wb") as fp:
        np.array([VEC_FILE_CLASSID, v.size], dtype=it).tofile(fp)
        v.astype(st).tofile(fp)


def PetscVecWriteBlockSizeToFile(fname, v):
    """Write "-vecload_block_size <bs>" (bs = number of columns of v, 1 for a
    1-D v) to fname and return bs."""
    v = np.asarray(v)
    bs = v.shape[1] if v.ndim > 1 else 1
    with open(fname, "w") as fp:
        fp.write(f"-vecload_block_size {bs}\n")
    return bs


def PetscVecWrite(fname, v, **kwargs):
    """fname + ".info" with the block size, then the Vec of v in row-major order."""
    PetscVecWriteBlockSizeToFile(fname + ".info", v)
    PetscVecWriteToFile(fname, np.asarray(v).reshape(-1), **kwargs)


def PetscISWriteToFile(fname, iset, **kwargs):
    """IS record [class id, n] + the n indices."""
    it, _ = _types(**kwargs)
    iset = np.asarray(iset).reshape(-1)
    if iset.size and (iset.min() < np.iinfo(it).min or iset.max() > np.iinfo(it).max):
        raise OverflowError(f"indices do not fit {it.name}")
    with open(fname, "wb") as fp:
        np.array([IS_FILE_CLASSID, iset.size], dtype=it).tofile(fp)
        iset.astype(it).tofile(fp)
