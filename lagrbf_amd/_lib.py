"""ctypes loader of liblagrbf.so with the reference's argument types.

Mirrors python_api/lagrbfpy/interface/register_cfunc.py:44-195: the same
symbols, the same argtypes (PetscInt = int32, PetscReal = float64, PetscInt64 =
int64, PetscBool = c_bool) plus the PETSc-free option channel.
"""
import ctypes
import os

import numpy as np
from numpy.ctypeslib import ndpointer

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liblagrbf.so")

PetscInt = ctypes.c_int32
PetscReal = ctypes.c_double
REAL = np.dtype(np.float64)
INT64 = np.dtype(np.int64)

VELOCITY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(PetscReal), ctypes.POINTER(PetscReal), ctypes.c_void_p)
VELOCITY_FN_BATCH = ctypes.CFUNCTYPE(ctypes.c_int, PetscInt, PetscInt, ctypes.POINTER(PetscReal),
                                     ctypes.POINTER(PetscReal), ctypes.c_void_p)


def wrap_ndpointer(dtype=None, ndim=None, shape=None, flags=None):
    """ndpointer that also accepts None (NULL), as register_cfunc.py:33-41."""
    base = ndpointer(dtype=dtype, ndim=ndim, shape=shape, flags=flags)

    class WrappedPointer(base):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return obj
            return super().from_param(obj)
    return WrappedPointer


_clib = None
_cfuncs = None


def load_shared_lib(path=None):
    path = path or os.environ.get("LAGRBF_LIB", LIBPATH)
    if not os.path.exists(path):
        raise ImportError(f"{path} not found: build the HIP library first (python -m lagrbf_amd.build)")
    return ctypes.CDLL(path)


def register_functions(clib):
    f = {}
    c = clib.lagrbf_interpolate_driver
    c.restype = None
    c.argtypes = [PetscInt, ndpointer(REAL, flags="C_CONTIGUOUS"), ndpointer(REAL, flags="C_CONTIGUOUS"), PetscInt,
                  ndpointer(REAL, flags="C_CONTIGUOUS"), ndpointer(REAL, flags="C_CONTIGUOUS"), PetscInt, ctypes.c_char_p]
    f["lagrbf_interpolate_driver"] = c
    c = clib.lagrbf_create
    c.restype = ctypes.c_void_p
    c.argtypes = []
    f["lagrbf_create"] = c
    c = clib.lagrbf_destroy
    c.restype = None
    c.argtypes = [ctypes.c_void_p]
    f["lagrbf_destroy"] = c
    c = clib.lagrbf_fill_swarm
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, ctypes.c_bool, PetscInt, ndpointer(REAL, flags="C_CONTIGUOUS"),
                  wrap_ndpointer(REAL, flags="C_CONTIGUOUS"), wrap_ndpointer(INT64, flags="C_CONTIGUOUS")]
    f["lagrbf_fill_swarm"] = c
    c = clib.lagrbf_interpolation
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p]
    f["lagrbf_interpolation"] = c
    c = clib.lagrbf_output
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, ctypes.c_bool, PetscInt, wrap_ndpointer(REAL, flags="C_CONTIGUOUS"),
                  wrap_ndpointer(REAL, flags="C_CONTIGUOUS"), wrap_ndpointer(INT64, flags="C_CONTIGUOUS"), ctypes.c_char_p]
    f["lagrbf_output"] = c
    c = clib.lagrbf_avect_analytical
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, PetscReal, PetscReal]
    f["lagrbf_avect_analytical"] = c
    c = clib.lagrbf_advection_function
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, VELOCITY_FN, ctypes.c_void_p]
    f["lagrbf_advection_function"] = c
    c = clib.lagrbf_advection_function_batch
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, VELOCITY_FN_BATCH, ctypes.c_void_p]
    f["lagrbf_advection_function_batch"] = c
    c = clib.lagrbf_advect_interpolate
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, PetscReal, PetscReal]
    f["lagrbf_avect_interpolate"] = c          # key as in register_cfunc.py:178-185
    f["lagrbf_advect_interpolate"] = c
    c = clib.lagrbf_reset_coefficients
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p]
    f["lagrbf_reset_coefficients"] = c
    # extensions
    c = clib.lagrbf_set_option
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    f["lagrbf_set_option"] = c
    c = clib.lagrbf_clear_options
    c.restype = ctypes.c_int
    c.argtypes = []
    f["lagrbf_clear_options"] = c
    c = clib.lagrbf_options_insert_file
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_char_p]
    f["lagrbf_options_insert_file"] = c
    c = clib.lagrbf_advect_steps
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, PetscInt, PetscReal]
    f["lagrbf_advect_steps"] = c
    c = clib.lagrbf_get_info
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
    f["lagrbf_get_info"] = c
    c = clib.lagrbf_debug_get
    c.restype = ctypes.c_int64
    c.argtypes = [ctypes.c_void_p, ctypes.c_char_p, PetscInt, ctypes.c_void_p, ctypes.c_int64]
    f["lagrbf_debug_get"] = c
    c = clib.lagrbf_last_error
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_char_p, PetscInt]
    f["lagrbf_last_error"] = c
    # serving
    D = ndpointer(REAL, flags="C_CONTIGUOUS")
    for name, res, args in (
            ("lagrbf_serve_create", ctypes.c_void_p, [ctypes.c_double]),
            ("lagrbf_serve_destroy", None, [ctypes.c_void_p]),
            ("lagrbf_serve_add_dataset", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, PetscInt, D, D]),
            ("lagrbf_serve_submit", ctypes.c_int64, [ctypes.c_void_p, PetscInt, PetscInt, PetscInt, PetscReal, PetscInt, D]),
            ("lagrbf_serve_step", ctypes.c_int32, [ctypes.c_void_p]),
            ("lagrbf_serve_poll", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, wrap_ndpointer(REAL, flags="C_CONTIGUOUS"),
                                                   wrap_ndpointer(REAL, flags="C_CONTIGUOUS")]),
            ("lagrbf_serve_release", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
            ("lagrbf_serve_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)])):
        c = getattr(clib, name)
        c.restype = res
        c.argtypes = args
        f[name] = c
    # file formats (include/lagrbf_io.h)
    CP = ctypes.c_char_p
    SP = ctypes.POINTER(ctypes.c_char_p)
    I64 = ctypes.c_int64
    I32 = ctypes.c_int32
    VP = ctypes.c_void_p
    D0 = wrap_ndpointer(REAL, flags="C_CONTIGUOUS")
    L0 = wrap_ndpointer(INT64, flags="C_CONTIGUOUS")
    for name, res, args in (
            ("lagrbf_io_vec_write", I32, [CP, I64, D0, I32, I32]),
            ("lagrbf_io_vec_read", I32, [CP, I32, ctypes.POINTER(I64), ctypes.POINTER(I32), D0, I64]),
            ("lagrbf_io_is_write", I32, [CP, I64, L0, I32]),
            ("lagrbf_io_is_read", I32, [CP, I32, ctypes.POINTER(I64), L0, I64]),
            ("lagrbf_io_pvd_open", I32, [CP]),
            ("lagrbf_io_pvd_append", I32, [CP, PetscReal, CP, CP]),
            ("lagrbf_io_pvd_open_append", I32, [ctypes.c_bool, I64, CP, PetscReal, CP, CP]),
            ("lagrbf_io_filename", I32, [CP, CP, CP, CP, CP, I64, CP, I32]),
            ("lagrbf_io_filelist", I32, [CP, I64, CP, I32, ctypes.POINTER(I32)]),
            ("lagrbf_io_stag_load", VP, [SP, ctypes.POINTER(I32), I32, I32]),
            ("lagrbf_io_petsc_load", VP, [CP, SP, I32, I32, I32]),
            ("lagrbf_io_load_data", VP, [I64, I32, I32]),
            ("lagrbf_io_dataset_info", I32, [VP, ctypes.POINTER(I64), ctypes.POINTER(I32), ctypes.POINTER(I32)]),
            ("lagrbf_io_dataset_copy", I32, [VP, D0, D0]),
            ("lagrbf_io_dataset_free", None, [VP]),
            ("lagrbf_swarm_output", I32, [VP, CP, I64, PetscReal, I32, SP, I32, ctypes.c_bool]),
            ("lagrbf_swarm_load", I32, [VP, CP, I64, I32, I32, SP]),
            ("lagrbf_load_data", I32, [VP, I64]),
            ("lagrbf_advect_adaptive", I32, [VP, PetscReal, PetscReal, I32]),
            ("lagrbf_fill_swarm_from_dataset", I32, [VP, ctypes.c_bool, VP])):
        c = getattr(clib, name)
        c.restype = res
        c.argtypes = args
        f[name] = c
    return f


def noalloc_fn():
    """lagrbf_interpolate_device_noalloc (device pointers as integers)."""
    cfuncs()
    c = _clib.lagrbf_interpolate_device_noalloc
    c.restype = ctypes.c_int
    c.argtypes = [ctypes.c_void_p, PetscInt, ctypes.c_void_p, ctypes.c_void_p]
    return c


def cfuncs():
    global _clib, _cfuncs
    if _cfuncs is None:
        _clib = load_shared_lib()
        _cfuncs = register_functions(_clib)
    return _cfuncs


def last_error():
    buf = ctypes.create_string_buffer(1024)
    cfuncs()["lagrbf_last_error"](buf, 1024)
    return buf.value.decode(errors="replace")


EXPORTED_SYMBOLS = [
    "lagrbf_create", "lagrbf_destroy", "lagrbf_fill_swarm", "lagrbf_interpolation", "lagrbf_output",
    "lagrbf_avect_analytical", "lagrbf_advect_interpolate", "lagrbf_advection_function",
    "lagrbf_advection_function_batch", "lagrbf_reset_coefficients", "lagrbf_interpolate_driver",
    "lagrbf_set_option", "lagrbf_clear_options", "lagrbf_options_insert_file", "lagrbf_advect_steps",
    "lagrbf_get_info", "lagrbf_debug_get", "lagrbf_last_error",
    "lagrbf_serve_create", "lagrbf_serve_destroy", "lagrbf_serve_add_dataset", "lagrbf_serve_submit",
    "lagrbf_serve_step", "lagrbf_serve_poll", "lagrbf_serve_release", "lagrbf_serve_info",
    "lagrbf_io_vec_write", "lagrbf_io_vec_read", "lagrbf_io_is_write", "lagrbf_io_is_read", "lagrbf_io_pvd_open",
    "lagrbf_io_pvd_append", "lagrbf_io_pvd_open_append", "lagrbf_io_filename", "lagrbf_io_filelist",
    "lagrbf_io_stag_load", "lagrbf_io_petsc_load", "lagrbf_io_load_data", "lagrbf_io_dataset_info",
    "lagrbf_io_dataset_copy", "lagrbf_io_dataset_free", "lagrbf_swarm_output", "lagrbf_swarm_load",
    "lagrbf_load_data", "lagrbf_fill_swarm_from_dataset", "lagrbf_advect_adaptive",
    "lagrbf_interpolate_device_noalloc",
]
