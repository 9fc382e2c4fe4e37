"""The entry point that allocates nothing (lagrbf_interpolate_device_noalloc)
against the entry points it varies (lagrbf_fill_swarm + lagrbf_interpolation +
lagrbf_output): the same values bit for bit, on two calls in a row, the second
with fewer points, FP64 and mixed, Wendland C2 and r^3 + linear polynomial."""
import ctypes

import numpy as np
import pytest

import lagrbf_amd as L
from lagrbf_amd._lib import noalloc_fn, last_error
from tests import cases

pytestmark = pytest.mark.gpu


class DeviceArray:
    """Device memory from the HIP runtime liblagrbf.so itself links (libamdhip64.so.7),
    for the test's inputs and outputs."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            cfn = noalloc_fn()          # loads liblagrbf.so and with it its HIP runtime
            del cfn
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        h = self.hip()
        assert h.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert h.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0      # host to device

    def to_host(self):
        out = np.empty_like(self.host)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0         # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def reference_values(opts, X, V, P):
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
    out = np.zeros((P.shape[0], V.shape[1]))
    g.output(fvals=out)
    g.destroy()
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [2, 7])
def test_noalloc_matches_existing_entry_point(precision, kernel):
    opts, X, V, P = cases.fourier_3d(nlat=24, Np=50000, n=9, kernel=kernel)
    opts.update(lagrbf_setup_all=True, lagrbf_precision=precision, advection=False)
    if kernel == 7:
        opts["lagrbf_poly_degree"] = 1
    print(f"case: no-alloc kernel {kernel} {precision}", flush=True)
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V); g.fill_points(P[:10]); g.interpolate_c()     # every patch set up, table built
    fn = noalloc_fn()
    for n in (50000, 7001):
        Pn = np.ascontiguousarray(P[:n])
        ref = reference_values(opts, X, V, Pn)
        dx = DeviceArray(Pn)
        dv = DeviceArray(np.full((n, 3), -1.0))
        rc = fn(g.ctx, n, dx.ptr, dv.ptr)
        assert rc == 0, last_error()
        got = dv.to_host()
        dx.free(); dv.free()
        print(f"n={n}: max |diff| {np.abs(got - ref).max():.3e}")
        assert np.array_equal(got, ref)
    g.destroy()


def test_noalloc_rejects_beyond_limits():
    opts, X, V, P = cases.fourier_3d(nlat=12, Np=100, n=5)
    opts.update(lagrbf_setup_all=True, advection=False)
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
    fn = noalloc_fn()
    d = DeviceArray(np.zeros(3))
    assert fn(g.ctx, (1 << 18) + 1, d.ptr, d.ptr) == 60
    assert "exceed" in last_error()
    g2 = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_setup_all=False))
    g2.fill_data(X, V)
    assert fn(g2.ctx, 1, d.ptr, d.ptr) == 73
    d.free()
    g.destroy(); g2.destroy()
