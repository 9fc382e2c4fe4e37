"""Advection drivers mirroring python_api/lagrbfpy/advection.py:37-496.

The call sequences, file names and PVD entries are the reference's:
``advect_step`` fills ``output`` and writes ``step{n:05d}{prefix}.vtp`` into
``output_path`` plus an entry of the ``pvdfilename`` collection (pyvista is
replaced by lagrbf_amd.vtp).  ``comm`` is accepted where the reference takes a
PETSc communicator; there is one rank.  Extension: ``write_output=False`` keeps
everything in memory.
"""
import os

import numpy as np

from .context import LagRBFCtx
from .timeseries import timeseries_write


class Advection(LagRBFCtx):
    def __init__(self, petsc_options=None, **kwargs):
        kwargs["advection"] = True
        kwargs["advection_scheme"] = kwargs.get("advection_scheme", 2)
        kwargs["advection_cfl"] = kwargs.get("advection_cfl", 0.3)
        kwargs["dt"] = kwargs.get("dt", 1e-7)
        super().__init__(petsc_options=petsc_options, **kwargs)
        time_start = kwargs.pop("time_start", None)
        time_end = kwargs.pop("time_end", None)
        nsteps = kwargs.pop("nsteps", None)
        time_steps = kwargs.pop("time_steps", None)
        steps = kwargs.pop("steps", None)
        self.start_idx = 0
        if time_steps is None:
            if time_start is None or time_end is None or nsteps is None:
                raise ValueError("time_start, time_end and nsteps must be provided if time_steps is not provided.")
            self.time_steps = np.linspace(time_start, time_end, nsteps, dtype=np.float64)
            self.nsteps = nsteps
        else:
            self.time_steps = np.asarray(time_steps, dtype=np.float64)
            self.nsteps = self.time_steps.shape[0]
        self.steps = np.arange(self.nsteps, dtype=np.int32) if steps is None else np.asarray(steps, dtype=np.int32)
        if self.steps.shape[0] != self.nsteps:
            raise ValueError("steps must have the same length as time_steps.")
        self.pvdfname = kwargs.get("pvdfilename", "timeseries-advection.pvd")
        self.write_output = bool(kwargs.get("write_output", True))

    def _write_step(self, comm, output, fields, fieldname, file_step, time, prefix):
        """step{file_step:05d}{prefix}.vtp and its PVD entry (advection.py:247-264, 399-414)."""
        if not self.write_output:
            self.output(coords=output.coords, fvals=output.values, pidx=output.pidx, fieldname=fieldname,
                        is_data=False)
            return
        os.makedirs(self.output_path, exist_ok=True)
        fname = f"step{int(file_step):05d}{prefix}.vtp"
        self.save(comm, fname=os.path.join(self.output_path, fname), fields_size=fields, output_points=output,
                  fieldname=fieldname, is_data=False)
        ts = {"time": [f"{time:.6e}"], "step": [f"{int(file_step):05d}"]}
        timeseries_write(fname=os.path.join(self.output_path, self.pvdfname), timeseries=ts, prefix=prefix,
                         extension="vtp", erase=False)

    def advect(self, *args, **kwargs):
        raise NotImplementedError("The advect method must be implemented in a subclass.")


class AdvectionAnalytical(Advection):
    """Advection with an analytical velocity callback (advection.py:171-320)."""

    def advect_step(self, step, output, comm=None, prefix=""):
        time, time_next = self.time_steps[step], self.time_steps[step + 1]
        func_data = self.get_ref("advection_data")
        if hasattr(func_data, "time"):
            func_data.time = float(time)
        if hasattr(func_data, "step"):
            func_data.step = int(self.steps[step])
        self.advect_analytical_c(time, time_next)
        self._write_step(comm, output, {"velocity": self.dim}, "velocity", self.steps[step + 1],
                         self.time_steps[step + 1], prefix)

    def advect(self, points, output, comm=None, prefix=""):
        try:
            self.get_ref("advection_fn")
        except KeyError as e:
            raise RuntimeError("No advection function has been registered.") from e
        self.fill_points(coords=points.coords, values=points.values, pidx=points.pidx)
        for step in range(self.nsteps - 1):
            self.advect_step(step, output, comm, prefix)


class AdvectionInterpolation(Advection):
    """Advection with the velocity interpolated from data points (advection.py:322-496)."""

    def advect_step(self, step, get_data, data_args, output, comm=None, fields=None, prefix=""):
        if "step" in data_args:
            data_args["step"] = np.int32(self.steps[step])
        if "time" in data_args:
            data_args["time"] = np.float64(self.time_steps[step])
        if "rank" in data_args:
            data_args["rank"] = np.int32(0)
        data = get_data(**data_args)
        self.fill_data(coords=data.coords, values=data.values, pidx=data.pidx)
        self.interpolate_c()
        fields = {"fields": self.nfields} if fields is None else fields
        self._write_step(comm, output, fields, "fields", self.steps[step], self.time_steps[step], prefix)
        if step == self.nsteps - 1:
            return
        self.advect_interpolate_c(self.time_steps[step], self.time_steps[step + 1])

    def advect(self, get_data, data_args, points, output, comm=None, fields=None, prefix="", callback=None):
        if not callable(get_data):
            raise ValueError("get_data must be a callable function that returns a PointContainer.")
        self.fill_points(coords=points.coords, values=points.values, pidx=points.pidx)
        for step in range(self.start_idx, self.nsteps):
            self.advect_step(step, get_data, data_args, output, comm, fields, prefix=prefix)
            if callback is not None:
                callback(step, output)
