"""lagrbf_amd: MI355X-native PU-RBF interpolation and Lagrangian advection.

Python mirror of lagrbfpy's public classes over liblagrbf.so (HIP, gfx950).
The library runs on the GPU only; importing works without it, but creating a
context requires the built liblagrbf.so and a HIP device.
"""
from .utils import PointContainer, FunctionData, FunctionPointer, halton_sequence, structured_points, halton_points
from .context import LagRBFCtx
from .interpolation import LagRBFInterpolator
from .advection import Advection, AdvectionAnalytical, AdvectionInterpolation
from .serve import Server
from .petsc_io import (PetscVecLoadFromFile, PetscVecLoadBlockSizeFromFile, PetscVecLoad, PetscISLoadFromFile,
                       PetscMatDenseLoadFromFile, PetscVecWriteToFile, PetscVecWriteBlockSizeToFile, PetscVecWrite,
                       PetscISWriteToFile)
from .timeseries import (timeseries_process, timeseries_compare, timeseries_write_step, timeseries_write_new,
                         timeseries_append, timeseries_write)
from .vtp import write_vtp, read_vtp
from .restart import restart_from_pvd

__all__ = ["PointContainer", "FunctionData", "FunctionPointer", "halton_sequence", "structured_points", "halton_points",
           "LagRBFCtx", "LagRBFInterpolator", "Advection", "AdvectionAnalytical", "AdvectionInterpolation", "Server",
           "PetscVecLoadFromFile", "PetscVecLoadBlockSizeFromFile", "PetscVecLoad", "PetscISLoadFromFile",
           "PetscMatDenseLoadFromFile", "PetscVecWriteToFile", "PetscVecWriteBlockSizeToFile", "PetscVecWrite",
           "PetscISWriteToFile", "timeseries_process", "timeseries_compare", "timeseries_write_step",
           "timeseries_write_new", "timeseries_append", "timeseries_write", "write_vtp", "read_vtp",
           "restart_from_pvd"]
