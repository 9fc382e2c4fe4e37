"""LagRBFCtx: Python interface to liblagrbf.so.

Same constructor keywords, option names, defaults and methods as lagrbfpy's
LagRBFCtx (python_api/lagrbfpy/interface/context.py:39-682).  Options go to the
library's own options database instead of PETSc.Options: ``petsc_options`` may
be a dict of option name -> value; keyword arguments override it.
"""
import ctypes
import os

import numpy as np

from ._lib import cfuncs, last_error, VELOCITY_FN_BATCH, VELOCITY_FN
from .utils import FunctionData


def _opt_str(v):
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (list, tuple, np.ndarray)):
        return ",".join(_opt_str(x) for x in np.asarray(v).ravel().tolist())
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    return str(v)


class LagRBFCtx:
    """Context for PU-RBF interpolation and Lagrangian advection on the GPU.

    Keywords (context.py:113-206): dim, minCoords, maxCoords, npartitions, nfields,
    overlap, shape_factor, kernel, solve_method, advection, advection_scheme,
    advection_cfl, dt, rmin, rmax, output_path.  Extensions: precision
    ("fp64" | "mixed"), device.
    """

    def __init__(self, petsc_options=None, **kwargs):
        self.data_filled = False
        self.interp_filled = False
        self.PetscReal = np.dtype(np.float64)
        self.PetscRealCtype = ctypes.c_double
        o = {} if petsc_options is None else {k.lstrip("-"): v for k, v in dict(petsc_options).items()}
        self.options = o
        self.output_path = kwargs.get("output_path", o.get("output_path", "output"))
        self.dim = int(kwargs.get("dim", o.get("pu_mesh_dim", 3)))
        o["pu_mesh_dim"] = self.dim
        o["pu_mesh_O"] = kwargs.get("minCoords", o.get("pu_mesh_O", np.zeros(self.dim)))
        o["pu_mesh_L"] = kwargs.get("maxCoords", o.get("pu_mesh_L", np.ones(self.dim)))
        o["pu_mesh_n"] = kwargs.get("npartitions", o.get("pu_mesh_n", np.full(self.dim, 4, dtype=np.int32)))
        o["n_interpolation_fields"] = kwargs.get("nfields", o.get("n_interpolation_fields", 0))
        o["pu_mesh_overlap"] = kwargs.get("overlap", o.get("pu_mesh_overlap", 1.0))
        # the reference reads only "shape_factor" (context.py:161); "shape_fac" is ignored there too
        o["pu_mesh_shape_factor"] = kwargs.get("shape_factor", o.get("pu_mesh_shape_factor", 1.0))
        o["kernel_type"] = kwargs.get("kernel", o.get("kernel_type", 2))
        o["solve_method"] = kwargs.get("solve_method", o.get("solve_method", 2))
        o["advection"] = kwargs.get("advection", o.get("advection", False))
        o["advection_scheme"] = kwargs.get("advection_scheme", o.get("advection_scheme", 2))
        o["cfl_fac"] = kwargs.get("advection_cfl", o.get("cfl_fac", 0.3))
        o["dt"] = kwargs.get("dt", o.get("dt", 1e-4))
        o["rmin"] = kwargs.get("rmin", o.get("rmin", 0.0))
        o["rmax"] = kwargs.get("rmax", o.get("rmax", 1e132))
        if "precision" in kwargs:
            o["lagrbf_precision"] = kwargs["precision"]
        if "device" in kwargs:
            o["lagrbf_device"] = kwargs["device"]
        self.nfields = int(o["n_interpolation_fields"])
        f = cfuncs()
        f["lagrbf_clear_options"]()
        for k, v in o.items():
            f["lagrbf_set_option"](k.encode(), _opt_str(v).encode())
        self.ctx = f["lagrbf_create"]()
        if not self.ctx:
            raise RuntimeError(f"Failed to create LagRBF context: {last_error()}")
        self.refs = {}

    # -- reference helpers (context.py:208-287) --
    def get_ctx(self):
        if self.ctx is None:
            raise RuntimeError("LagRBF context has not been created or has been destroyed.")
        return self.ctx

    def get_stored_refs(self):
        return self.refs

    def get_ref(self, key):
        ref = self.refs.get(key, None)
        if ref is None:
            raise KeyError(f"No reference found for key '{key}'.")
        return ref

    def store_ref(self, **objs):
        self.refs.update(objs)

    def __bool__(self):
        return self.ctx is not None

    def __del__(self):
        if getattr(self, "ctx", None) is not None:
            try:
                cfuncs()["lagrbf_destroy"](self.ctx)
            except Exception:
                pass
            self.ctx = None

    def destroy(self):
        self.__del__()

    def _begin_c(self, array, dtype):
        if array is None:
            raise ValueError("Input array cannot be None.\nConsider using an empty array instead: np.empty((0, dim))\n")
        a = np.ascontiguousarray(array, dtype=dtype)
        return a.reshape((a.size,))

    def _check(self, ierr, name):
        if ierr != 0:
            raise RuntimeError(f"{name} failed with error code {ierr}: {last_error()}")

    def fill_swarm(self, coords, values=None, pidx=None, is_data=False):
        if coords is None:
            coords = np.empty((0, self.dim), dtype=np.float64)
        npoints = coords.shape[0]
        pc = self._begin_c(coords, np.float64)
        pv = None if values is None else self._begin_c(values, np.float64)
        pp = None if pidx is None else self._begin_c(pidx, np.int64)
        self._check(cfuncs()["lagrbf_fill_swarm"](self.ctx, bool(is_data), npoints, pc, pv, pp), "lagrbf_fill_swarm")

    def fill_data(self, coords, values, pidx=None):
        self.fill_swarm(coords=coords, values=values, pidx=pidx, is_data=True)
        self.data_filled = True

    def fill_points(self, coords, values=None, pidx=None):
        self.fill_swarm(coords=coords, values=values, pidx=pidx, is_data=False)
        self.interp_filled = True

    def output(self, coords=None, fvals=None, pidx=None, fieldname="fields", is_data=False):
        """Write swarm coordinates / field values / point ids into caller arrays (rows by pid)."""
        if coords is None and fvals is None and pidx is None:
            raise ValueError("At least one of coords, fvals, or pidx must be provided.")
        npoints = 0
        for a in (coords, fvals, pidx):
            if a is not None:
                npoints = a.shape[0]
        if npoints == 0:
            raise ValueError("Number of points must be greater than zero.")
        for a, dt in ((coords, np.float64), (fvals, np.float64), (pidx, np.int64)):
            if a is not None and (a.dtype != dt or not a.flags.c_contiguous):
                raise ValueError("output arrays must be C-contiguous float64 (int64 for pidx)")
        self._check(cfuncs()["lagrbf_output"](self.ctx, bool(is_data), npoints,
                                              None if coords is None else coords.reshape(-1),
                                              None if fvals is None else fvals.reshape(-1),
                                              None if pidx is None else pidx.reshape(-1),
                                              fieldname.encode("utf-8")), "lagrbf_output")

    def interpolate_c(self):
        if not self.data_filled:
            raise RuntimeError("Data have not been provided, use fill_data() to provide data points")
        if not self.interp_filled:
            raise RuntimeError("Interpolation swarm has not been provided, use fill_points() to provide the points on which to interpolate")
        self._check(cfuncs()["lagrbf_interpolation"](self.ctx), "lagrbf_interpolation")

    def reset_coefficients(self):
        self._check(cfuncs()["lagrbf_reset_coefficients"](self.ctx), "lagrbf_reset_coefficients")

    def set_advection_function(self, func, func_data=None):
        """Register a batch velocity callback (context.py:616-633)."""
        if not callable(func):
            raise ValueError("Provided advection function is not callable.")
        if func_data is None:
            func_data = FunctionData(time=0.0, step=0)
        elif not isinstance(func_data, ctypes.Structure):
            raise ValueError("Provided advection function data is not a ctypes.Structure.")
        if not isinstance(func, VELOCITY_FN_BATCH):
            func = VELOCITY_FN_BATCH(func)
        self.store_ref(advection_fn=func)
        self.store_ref(advection_data=func_data)
        self._check(cfuncs()["lagrbf_advection_function_batch"](self.ctx, func, ctypes.cast(ctypes.pointer(func_data), ctypes.c_void_p)),
                    "lagrbf_advection_function_batch")

    def set_advection_function_pointwise(self, func, func_data=None):
        if not isinstance(func, VELOCITY_FN):
            func = VELOCITY_FN(func)
        if func_data is None:
            func_data = FunctionData(time=0.0, step=0)
        self.store_ref(advection_fn=func, advection_data=func_data)
        self._check(cfuncs()["lagrbf_advection_function"](self.ctx, func, ctypes.cast(ctypes.pointer(func_data), ctypes.c_void_p)),
                    "lagrbf_advection_function")

    def advect_analytical_c(self, time_start, time_end):
        if not self.interp_filled:
            raise RuntimeError("Swarm to advect has not been provided, use fill_swarm() to provide the points to advect")
        self._check(cfuncs()["lagrbf_avect_analytical"](self.ctx, float(time_start), float(time_end)), "lagrbf_advect")

    def advect_interpolate_c(self, time_start, time_end):
        if not self.data_filled:
            raise RuntimeError("Data have not been provided, use fill_swarm() to provide data points")
        if not self.interp_filled:
            raise RuntimeError("Swarm to advect has not been provided, use fill_swarm() to provide the points to advect")
        self._check(cfuncs()["lagrbf_avect_interpolate"](self.ctx, float(time_start), float(time_end)), "lagrbf_advect")
        self.data_filled = False   # the data swarm is cleared at exit (time_loop.c:98)

    def save(self, comm, fname, fields_size, output_points, fieldname="fields", is_data=False):
        """Output the swarm into ``output_points`` and write it to the .vtp file
        fname: points, the columns of ``values`` split as ``fields_size``
        (name -> components) and "pidx" (context.py:459-511).  ``comm`` is
        accepted for signature compatibility; there is one rank."""
        from .vtp import write_vtp
        if not fields_size:
            raise ValueError("fields_size must name at least one field")
        if not str(fname).endswith(".vtp"):
            raise ValueError(f"output file name must end with .vtp, got {fname}")
        self.output(coords=output_points.coords, fvals=output_points.values, pidx=output_points.pidx,
                    fieldname=fieldname, is_data=is_data)
        data, idx = {}, 0
        for name, bs in fields_size.items():
            data[name] = output_points.values[:, idx:idx + bs]
            idx += bs
        data["pidx"] = output_points.pidx
        write_vtp(output_points.coords, data, str(fname))

    # -- file formats of the C side (include/lagrbf_io.h) --
    def swarm_output(self, root, step, time, fields=(), direction=0, permute=False):
        """SwarmOutput (swarm_output.c:169-234): PETSc binary files of the particle
        swarm under root/step<step>/ and an entry of root/timeseries_X.pvd.
        fields: "fields", the velocity name, "weights", "DMSwarm_pid";
        direction 0 / 1 (forward) / 2 (backward)."""
        from .fileio import _strings
        names = list(fields)
        self._check(cfuncs()["lagrbf_swarm_output"](self.ctx, str(root).encode(), int(step), float(time), len(names),
                                                    _strings(names), int(direction), bool(permute)), "SwarmOutput")

    def swarm_load(self, dirname, step, fields=(), direction=0):
        """SwarmLoadFromFiles (swarm_load.c:133-201): refill the particle swarm from
        the files ``swarm_output`` wrote in dirname (coordinates, plus "fields"
        and "DMSwarm_pid" when listed)."""
        from .fileio import _strings
        names = list(fields)
        self._check(cfuncs()["lagrbf_swarm_load"](self.ctx, str(dirname).encode(), int(step), int(direction),
                                                  len(names), _strings(names)), "SwarmLoadFromFiles")
        self.interp_filled = True

    def load_data(self, step):
        """SwarmDataLoad (swarm_data.c:161-180): read the data files of ``step``
        named by the options (data_format, data_fname_* / swarm_data_fname_*)
        into the data swarm."""
        self._check(cfuncs()["lagrbf_load_data"](self.ctx, int(step)), "SwarmDataLoad")
        self.data_filled = True

    def fill_from_dataset(self, dataset, is_data=True):
        """Fill the data (is_data) or particle swarm from a fileio.Dataset."""
        self._check(cfuncs()["lagrbf_fill_swarm_from_dataset"](self.ctx, bool(is_data), dataset.handle),
                    "lagrbf_fill_swarm_from_dataset")
        if is_data:
            self.data_filled = True
        else:
            self.interp_filled = True

    # -- extensions --
    def advect_adaptive_c(self, t_start, t_end, max_attempts=100000):
        """Per-particle adaptive DP45 from t_start to t_end (lagrbf_advect_adaptive);
        returns {attempts, accepted, rejected, active} of this call."""
        before = {k: self.info("dp_" + k) for k in ("attempts", "accepted", "rejected")}
        self._check(cfuncs()["lagrbf_advect_adaptive"](self.ctx, float(t_start), float(t_end), int(max_attempts)),
                    "lagrbf_advect_adaptive")
        out = {k: self.info("dp_" + k) - v for k, v in before.items()}
        out["active"] = self.info("dp_active")
        return out

    def advect_steps_c(self, nsteps, dt):
        self._check(cfuncs()["lagrbf_advect_steps"](self.ctx, int(nsteps), float(dt)), "lagrbf_advect_steps")

    def info(self, key):
        v = ctypes.c_double()
        self._check(cfuncs()["lagrbf_get_info"](self.ctx, key.encode(), ctypes.byref(v)), "lagrbf_get_info")
        return v.value

    def debug_get(self, what, index=0, dtype=np.int32):
        """Copy internal device state (tests): see lagrbf_debug_get in include/lagrbf.h."""
        f = cfuncs()["lagrbf_debug_get"]
        n = f(self.ctx, what.encode(), int(index), None, 0)
        if n < 0:
            raise RuntimeError(f"lagrbf_debug_get({what}) failed: {last_error()}")
        out = np.zeros(max(int(n), 1), dtype=dtype)
        r = f(self.ctx, what.encode(), int(index), out.ctypes.data_as(ctypes.c_void_p), out.size)
        if r < 0:
            raise RuntimeError(f"lagrbf_debug_get({what}) failed: {last_error()}")
        return out[: int(n)]
