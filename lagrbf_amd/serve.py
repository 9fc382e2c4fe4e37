"""Serving: continuous batching of advection requests against device-resident
datasets (liblagrbf.so lagrbf_serve_*).  No reference counterpart.

    srv = Server(hbm_budget=200e9)
    ds = srv.add_dataset(options, coords, velocities)    # all patches solved once
    rid = srv.submit(ds, scheme=2, nsteps=50, dt=1e-3, coords=P)
    while srv.step() > 0: ...                            # one RK step of every active request
    done, X, times = srv.poll(rid)

Requests on one dataset and scheme share one particle batch on the GPU; a
request's result does not depend on which others run beside it.
"""
import ctypes

import numpy as np

from ._lib import cfuncs, last_error
from .context import _opt_str

SCHEMES = {"rk1": 0, "rk2": 1, "rk4": 2, "dopri": 3}


def options_string(options):
    """Options dict -> "-name value ..." as lagrbf_set_option names them."""
    parts = []
    for k, v in options.items():
        parts.append("-" + k.lstrip("-"))
        parts.append(_opt_str(v))
    return " ".join(parts)


class Server:
    def __init__(self, hbm_budget=200e9):
        self._f = cfuncs()
        self.srv = self._f["lagrbf_serve_create"](float(hbm_budget))
        if not self.srv:
            raise RuntimeError(f"lagrbf_serve_create failed: {last_error()}")
        self.dims = []
        self.sizes = {}

    def close(self):
        if getattr(self, "srv", None):
            self._f["lagrbf_serve_destroy"](self.srv)
            self.srv = None

    def __del__(self):
        self.close()

    def add_dataset(self, options, coords, values):
        """Dataset of nodes coords [N, d] carrying the velocity values [N, d]
        (n_interpolation_fields = d); options as for LagRBFCtx (pu_mesh_*, kernel_type,
        lagrbf_precision, ...), and lagrbf_serve_tune_min_particles: batches of fewer
        particles (default 4194304) skip the autotuner's trials."""
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        values = np.ascontiguousarray(values, dtype=np.float64)
        o = dict(options)
        o["n_interpolation_fields"] = coords.shape[1]
        o.setdefault("pu_mesh_dim", coords.shape[1])
        r = self._f["lagrbf_serve_add_dataset"](self.srv, options_string(o).encode(), coords.shape[0],
                                                coords.reshape(-1), values.reshape(-1))
        if r < 0:
            raise RuntimeError(f"lagrbf_serve_add_dataset failed ({-r}): {last_error()}")
        self.dims.append(coords.shape[1])
        return r

    def submit(self, dataset, scheme, nsteps, dt, coords):
        s = SCHEMES.get(scheme, scheme) if isinstance(scheme, str) else int(scheme)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        rid = self._f["lagrbf_serve_submit"](self.srv, int(dataset), s, int(nsteps), float(dt), coords.shape[0],
                                             coords.reshape(-1))
        if rid < 0:
            raise RuntimeError(f"lagrbf_serve_submit failed ({-rid}): {last_error()}")
        self.sizes[rid] = (coords.shape[0], coords.shape[1])
        return rid

    def step(self):
        """One tick; returns the number of queued + active requests."""
        r = self._f["lagrbf_serve_step"](self.srv)
        if r < 0:
            raise RuntimeError(f"lagrbf_serve_step failed ({-r}): {last_error()}")
        return r

    def poll(self, rid, want_result=True):
        """(done, positions [n, d] or None, times [submit, admit, done] in seconds)."""
        n, d = self.sizes[rid]
        out = np.zeros((n, d)) if want_result else None
        t = np.zeros(3)
        r = self._f["lagrbf_serve_poll"](self.srv, int(rid), None if out is None else out.reshape(-1), t)
        if r < 0:
            raise RuntimeError(f"lagrbf_serve_poll failed ({-r}): {last_error()}")
        return bool(r), (out if r else None), t

    def release(self, rid):
        self._f["lagrbf_serve_release"](self.srv, int(rid))
        self.sizes.pop(rid, None)

    def info(self, key):
        v = ctypes.c_double(0.0)
        if self._f["lagrbf_serve_info"](self.srv, key.encode(), ctypes.byref(v)) != 0:
            raise RuntimeError(f"lagrbf_serve_info({key}) failed: {last_error()}")
        return v.value

    def run_until_idle(self, max_ticks=1000000):
        for _ in range(max_ticks):
            if self.step() == 0:
                return
        raise RuntimeError("server did not drain")
