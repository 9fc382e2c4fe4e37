"""ctypes wrapper of the FP64 CPU oracle (oracle/liblagrbf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package never
imports this module.

``OracleCtx`` exposes the same methods as the product's ``LagRBFCtx`` (itself a
mirror of lagrbfpy's python_api/lagrbfpy/interface/context.py:39-682) so the
parity tests can drive both through an identical call sequence.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblagrbf_oracle.so")


def build(force=False):
    """Compile the oracle with its Makefile (gcc, strict IEEE)."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("n", ctypes.c_int * 3),
        ("O", ctypes.c_double * 3), ("L", ctypes.c_double * 3),
        ("overlap", ctypes.c_double), ("shape_factor", ctypes.c_double),
        ("kernel_type", ctypes.c_int), ("solve_method", ctypes.c_int),
        ("n_ifields", ctypes.c_int), ("rmin", ctypes.c_double), ("rmax", ctypes.c_double),
        ("cfl_fac", ctypes.c_double), ("advection", ctypes.c_int), ("scheme", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("dopri_dt_fac", ctypes.c_double), ("dopri_dt_fac_min", ctypes.c_double),
        ("dopri_dt_fac_max", ctypes.c_double), ("dopri_atol", ctypes.c_double),
        ("dopri_rtol", ctypes.c_double),
        ("weights_brute", ctypes.c_int), ("nthreads", ctypes.c_int), ("setup_all", ctypes.c_int),
        ("poly_degree", ctypes.c_int), ("local_eps", ctypes.c_double), ("weights_corners", ctypes.c_int),
    ]


VELOCITY_BATCH = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.c_void_p)

_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB)
        D = ctypes.POINTER(ctypes.c_double)
        I32 = ctypes.POINTER(ctypes.c_int32)
        I64 = ctypes.POINTER(ctypes.c_int64)
        V = ctypes.c_void_p
        L.orc_config_default.argtypes = [ctypes.POINTER(OrcConfig)]
        L.orc_kernel.restype = ctypes.c_double
        L.orc_kernel.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_sytrf_lower.argtypes = [ctypes.c_int, D, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_sytrs_lower.argtypes = [ctypes.c_int, ctypes.c_int, D, ctypes.c_int, ctypes.POINTER(ctypes.c_int), D, ctypes.c_int]
        L.orc_potrf_lower.argtypes = [ctypes.c_int, D, ctypes.c_int]
        L.orc_potrs_lower.argtypes = [ctypes.c_int, ctypes.c_int, D, ctypes.c_int, D, ctypes.c_int]
        L.orc_locate.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), D, D, ctypes.c_int, D, I32]
        L.orc_cell_sort.argtypes = [ctypes.c_int, I32, ctypes.c_int, I32, I32]
        L.orc_create.restype = V
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig)]
        L.orc_destroy.argtypes = [V]
        L.orc_fill_swarm.argtypes = [V, ctypes.c_int, ctypes.c_int32, D, D, I64]
        L.orc_interpolation.argtypes = [V]
        L.orc_output.argtypes = [V, ctypes.c_int, ctypes.c_int32, D, D, I64, ctypes.c_char_p]
        L.orc_reset_coefficients.argtypes = [V]
        L.orc_set_velocity_batch.argtypes = [V, VELOCITY_BATCH, V]
        L.orc_advect_interpolate.argtypes = [V, ctypes.c_double, ctypes.c_double]
        L.orc_advect_analytical.argtypes = [V, ctypes.c_double, ctypes.c_double]
        L.orc_advect_steps.argtypes = [V, ctypes.c_int32, ctypes.c_double]
        L.orc_advect_adaptive.argtypes = [V, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, D]
        L.orc_get_pu.argtypes = [V, D, D, D]
        L.orc_get_cellids.argtypes = [V, ctypes.c_int, I32]
        L.orc_get_cell_lists.argtypes = [V, I32, I32]
        L.orc_get_weights.argtypes = [V, D]
        L.orc_patch_size.argtypes = [V, ctypes.c_int]
        L.orc_patch_members.argtypes = [V, ctypes.c_int, I32]
        L.orc_patch_coeffs.argtypes = [V, ctypes.c_int, D]
        L.orc_patch_flags.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_removed_count.restype = ctypes.c_int64
        L.orc_removed_count.argtypes = [V]
        L.orc_get_dt.restype = ctypes.c_double
        L.orc_get_dt.argtypes = [V]
        L.orc_get_alive.argtypes = [V, I32]
        L.orc_patch_poly.argtypes = [V, ctypes.c_int, D, D]
        L.orc_poly_terms.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_poly_eval.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
        L.orc_patch_degree.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _lp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _arr(v, n, dtype):
    if isinstance(v, str):
        v = [float(x) if dtype is float else int(x) for x in v.split(",")]
    v = list(np.atleast_1d(np.asarray(v)).tolist())
    return v + [0] * (n - len(v))


def _bool(v):
    if isinstance(v, str):
        return v.strip().lower() in ("1", "true", "yes", "on")
    return bool(v)


def config_from_options(options):
    """Translate an options dict (names as in the reference's PETSc options DB,
    lagrangian_rbf.c:135-195, partition_of_unity.c:68-111, dormand_prince.c:340-353)
    into the oracle's configuration struct."""
    c = OrcConfig()
    lib().orc_config_default(ctypes.byref(c))
    o = {k.lstrip("-"): v for k, v in options.items()}
    c.dim = int(o.get("pu_mesh_dim", c.dim))
    nn = _arr(o.get("pu_mesh_n", [4] * c.dim), 3, int)
    OO = _arr(o.get("pu_mesh_O", [0.0] * c.dim), 3, float)
    LL = _arr(o.get("pu_mesh_L", [1.0] * c.dim), 3, float)
    for d in range(3):
        c.n[d] = int(nn[d]); c.O[d] = float(OO[d]); c.L[d] = float(LL[d])
    c.overlap = float(o.get("pu_mesh_overlap", c.overlap))
    c.shape_factor = float(o.get("pu_mesh_shape_factor", c.shape_factor))
    c.kernel_type = int(o.get("kernel_type", c.kernel_type))
    c.solve_method = int(o.get("solve_method", c.solve_method))
    c.n_ifields = int(o.get("n_interpolation_fields", c.n_ifields))
    c.rmin = float(o.get("rmin", c.rmin))
    c.rmax = float(o.get("rmax", c.rmax))
    c.cfl_fac = float(o.get("cfl_fac", c.cfl_fac))
    c.advection = int(_bool(o.get("advection", False)))
    c.scheme = int(o.get("advection_scheme", 3))
    c.dt = float(o.get("dt", c.dt))
    c.dopri_dt_fac = float(o.get("dopri_dt_fac", c.dopri_dt_fac))
    c.dopri_dt_fac_min = float(o.get("dopri_dt_fac_min", c.dopri_dt_fac_min))
    c.dopri_dt_fac_max = float(o.get("dopri_dt_fac_max", c.dopri_dt_fac_max))
    c.dopri_atol = float(o.get("dopri_atol", c.dopri_atol))
    c.dopri_rtol = float(o.get("dopri_rtol", c.dopri_rtol))
    c.weights_brute = int(_bool(o.get("oracle_weights_brute", False)))
    c.setup_all = int(_bool(o.get("oracle_setup_all", False)))
    c.poly_degree = int(o.get("lagrbf_poly_degree", -1))
    c.local_eps = float(o.get("lagrbf_local_eps", 0.0))
    c.weights_corners = int(str(o.get("lagrbf_weights", "all")) == "corners")
    return c


class OracleCtx:
    """Oracle counterpart of the product's LagRBFCtx (same method names)."""

    def __init__(self, options):
        self.cfg = config_from_options(options)
        o = {k.lstrip("-"): v for k, v in dict(options).items()}
        blend = o.get("lagrbf_time_blend", False)
        self.blend = blend if isinstance(blend, bool) else str(blend).lower() in ("1", "true", "yes")
        bt = o.get("lagrbf_blend_times", (0.0, 1.0))
        if isinstance(bt, str):
            bt = [float(x) for x in bt.split(",")]
        self.blend_times = (float(bt[0]), float(bt[1]))
        self.dim = self.cfg.dim
        self.nfields = self.cfg.n_ifields
        self.nen = 1 << self.dim
        self.ctx = lib().orc_create(ctypes.byref(self.cfg))
        if not self.ctx:
            raise RuntimeError("orc_create failed (invalid options)")
        self._keep = []

    def __del__(self):
        self.destroy()

    def destroy(self):
        if getattr(self, "ctx", None):
            lib().orc_destroy(self.ctx)
            self.ctx = None

    @staticmethod
    def _chk(ierr, what):
        if ierr != 0:
            raise RuntimeError(f"{what} failed with error code {ierr}")

    def fill_swarm(self, coords, values=None, pidx=None, is_data=False):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        n = coords.shape[0]
        values = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        pidx = None if pidx is None else np.ascontiguousarray(pidx, dtype=np.int64)
        self._chk(lib().orc_fill_swarm(self.ctx, int(is_data), n, _dp(coords), _dp(values), _lp(pidx)), "orc_fill_swarm")

    def fill_data(self, coords, values, pidx=None):
        self.fill_swarm(coords, values, pidx, is_data=True)

    def fill_points(self, coords, values=None, pidx=None):
        self.fill_swarm(coords, values, pidx, is_data=False)

    def interpolate_c(self):
        self._chk(lib().orc_interpolation(self.ctx), "orc_interpolation")

    def reset_coefficients(self):
        self._chk(lib().orc_reset_coefficients(self.ctx), "orc_reset_coefficients")

    def output(self, coords=None, fvals=None, pidx=None, fieldname="fields", is_data=False):
        n = 0
        for a in (coords, fvals, pidx):
            if a is not None:
                n = a.shape[0]
        self._chk(lib().orc_output(self.ctx, int(is_data), n, _dp(coords), _dp(fvals), _lp(pidx),
                                   fieldname.encode()), "orc_output")

    def set_velocity_function(self, pyfunc):
        """pyfunc(X[n,dim]) -> V[n,dim]; wrapped as the reference's batch callback."""
        dim = self.dim

        def cb(n, d, X, V, user):
            x = np.ctypeslib.as_array(X, shape=(n, d))
            v = np.ctypeslib.as_array(V, shape=(n, d))
            v[:, :] = pyfunc(x)
            return 0
        cfun = VELOCITY_BATCH(cb)
        self._keep.append(cfun)
        self._chk(lib().orc_set_velocity_batch(self.ctx, cfun, None), "orc_set_velocity_batch")
        del dim

    def advect_interpolate_c(self, t0, t1):
        self._chk(lib().orc_advect_interpolate(self.ctx, float(t0), float(t1)), "orc_advect_interpolate")

    def advect_analytical_c(self, t0, t1):
        self._chk(lib().orc_advect_analytical(self.ctx, float(t0), float(t1)), "orc_advect_analytical")

    def advect_steps_c(self, nsteps, dt):
        self._chk(lib().orc_advect_steps(self.ctx, int(nsteps), float(dt)), "orc_advect_steps")

    def advect_adaptive_c(self, t_start, t_end, max_attempts=100000):
        """Per-particle adaptive DP45 (time blend from the options); returns
        {attempts, accepted, rejected, active}."""
        counts = np.zeros(4)
        self._chk(lib().orc_advect_adaptive(self.ctx, float(t_start), float(t_end), int(max_attempts), int(self.blend),
                                            self.blend_times[0], self.blend_times[1], _dp(counts)),
                  "orc_advect_adaptive")
        return dict(zip(("attempts", "accepted", "rejected", "active"), counts.tolist()))

    # ---- introspection ----
    def pu(self):
        ew = ctypes.c_double(); er = ctypes.c_double()
        P = lib().orc_get_pu(self.ctx, ctypes.byref(ew), ctypes.byref(er), None)
        nodes = np.zeros((P, self.dim))
        lib().orc_get_pu(self.ctx, None, None, _dp(nodes))
        return ew.value, er.value, nodes

    def cellids(self, is_data=False):
        n = lib().orc_get_cellids(self.ctx, int(is_data), None)
        out = np.zeros(n, dtype=np.int32)
        lib().orc_get_cellids(self.ctx, int(is_data), _ip(out))
        return out

    def cell_lists(self, N):
        C = lib().orc_get_cell_lists(self.ctx, None, None)
        cs = np.zeros(C + 1, dtype=np.int32); cn = np.zeros(max(N, 1), dtype=np.int32)
        lib().orc_get_cell_lists(self.ctx, _ip(cs), _ip(cn))
        return cs, cn[:N]

    def weights(self):
        n = lib().orc_get_weights(self.ctx, None)
        out = np.zeros((n, self.nen))
        lib().orc_get_weights(self.ctx, _dp(out))
        return out

    def patch_members(self, p):
        n = lib().orc_patch_members(self.ctx, p, None)
        out = np.zeros(max(n, 1), dtype=np.int32)
        lib().orc_patch_members(self.ctx, p, _ip(out))
        return out[:n]

    def patch_coeffs(self, p):
        n = lib().orc_patch_coeffs(self.ctx, p, None)
        out = np.zeros(max(n * self.nfields, 1))
        lib().orc_patch_coeffs(self.ctx, p, _dp(out))
        return out[: n * self.nfields].reshape(self.nfields, n).T.copy()   # [n, nf]

    def patch_poly(self, p):
        """(beta [m, nf], eps) of patch p: polynomial coefficients and shape parameter."""
        e = ctypes.c_double()
        m = lib().orc_patch_poly(self.ctx, p, None, ctypes.byref(e))
        out = np.zeros(max(m * self.nfields, 1))
        lib().orc_patch_poly(self.ctx, p, _dp(out), None)
        return out[: m * self.nfields].reshape(self.nfields, m).T.copy(), e.value

    def patch_flags(self, p):
        s = ctypes.c_int(); v = ctypes.c_int()
        lib().orc_patch_flags(self.ctx, p, ctypes.byref(s), ctypes.byref(v))
        return s.value, v.value

    def removed_count(self):
        return lib().orc_removed_count(self.ctx)

    def dt(self):
        return lib().orc_get_dt(self.ctx)


def kernel(ktype, r, eps):
    return lib().orc_kernel(int(ktype), float(r), float(eps))


def sytrf(A):
    A = np.array(A, dtype=np.float64, order="F")
    n = A.shape[0]
    ipiv = (ctypes.c_int * max(n, 1))()
    info = lib().orc_sytrf_lower(n, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, ipiv)
    return A, np.array(ipiv[:n], dtype=np.int32), info


def sytrs(LD, ipiv, B):
    n = LD.shape[0]
    B = np.array(B, dtype=np.float64, order="F")
    if B.ndim == 1:
        B = B.reshape(n, 1, order="F")
    ip = (ctypes.c_int * max(n, 1))(*[int(x) for x in ipiv])
    LDf = np.asfortranarray(LD)
    lib().orc_sytrs_lower(n, B.shape[1], LDf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, ip,
                          B.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return B


def potrf(A):
    A = np.array(A, dtype=np.float64, order="F")
    info = lib().orc_potrf_lower(A.shape[0], A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), A.shape[0])
    return A, info


def locate(dim, n, O, L, coords):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    nn = (ctypes.c_int * 3)(*list(n) + [0] * (3 - len(n)))
    OO = np.zeros(3); OO[:dim] = O
    LL = np.zeros(3); LL[:dim] = L
    out = np.zeros(coords.shape[0], dtype=np.int32)
    lib().orc_locate(dim, nn, _dp(OO), _dp(LL), coords.shape[0], _dp(coords), _ip(out))
    return out


def cell_sort(cellid, ncells):
    cellid = np.ascontiguousarray(cellid, dtype=np.int32)
    cs = np.zeros(ncells + 1, dtype=np.int32)
    cn = np.zeros(max(len(cellid), 1), dtype=np.int32)
    lib().orc_cell_sort(len(cellid), _ip(cellid), ncells, _ip(cs), _ip(cn))
    return cs, cn[: len(cellid)]


def poly_terms(dim, deg):
    return lib().orc_poly_terms(dim, deg)


def poly_eval(dim, deg, xi):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    out = np.zeros(10)
    lib().orc_poly_eval(dim, deg, _dp(xi), _dp(out))
    return out[: poly_terms(dim, deg)]


def patch_degree(dim, D, n):
    return lib().orc_patch_degree(dim, D, n)
