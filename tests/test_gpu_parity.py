"""GPU parity: liblagrbf.so (through its C ABI) against the FP64 CPU oracle on
identical seeded inputs.

Tolerances (stated per result):
  cell indices, per-cell node lists, patch membership   bit-exact
  FP64 interpolated values                               1e-12 relative (max-norm)
  RK positions after N steps                             1e-9 relative
"""
import numpy as np
import pytest

import lagrbf_amd as L
from tests import cases

pytestmark = pytest.mark.gpu

TOL_VAL = 1e-12
TOL_POS = 1e-9


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(opts):
    """GPU context and an oracle context built from the same (completed) options."""
    from oracle.oracle import OracleCtx
    g = L.LagRBFCtx(petsc_options=dict(opts))
    o = OracleCtx(g.options)
    return g, o


def interp_both(opts, X, V, P):
    g, o = pair(opts)
    nf = g.nfields
    res = []
    for c in (g, o):
        c.fill_data(X, V)
        c.fill_points(P)
        c.interpolate_c()
        vals = np.zeros((P.shape[0], nf)); xs = np.zeros_like(P); pid = np.zeros(P.shape[0], dtype=np.int64)
        c.output(xs, vals, pid)
        res.append((vals, xs, pid))
    return g, o, res


def test_notebook_known_answer_gpu():
    opts, X, V, P = cases.notebook_2d()
    print("case: notebook 2D", flush=True)
    g, o, res = interp_both(opts, X, V, P)
    (vg, xg, _), (vo, xo, _) = res
    an = np.exp(-np.sum(xg ** 2, axis=1) / (2 * 0.5 ** 2))
    d = np.abs(vg[:, 0] - an)
    assert f"{np.sqrt(np.mean(d ** 2)):.6e}" == f"{cases.NOTEBOOK_L2:.6e}"
    assert f"{d.max():.6e}" == f"{cases.NOTEBOOK_LINF:.6e}"
    assert rel_err(vg, vo) < TOL_VAL


# Shape factors that keep the patch matrices away from numerical singularity (with
# shape factor 1 the global kernels give cond ~1e18 here and no two FP64 solvers
# agree on a digit); the tolerance follows cond(A) * 1e-16.
@pytest.mark.parametrize("kernel,shape,tol", [(2, 1.0, TOL_VAL), (0, 1.0, 1e-11), (1, 1.0, TOL_VAL), (3, 6.0, 1e-11),
                                              (4, 10.0, 1e-9), (5, 10.0, 1e-11), (6, 10.0, 1e-11)])
def test_interp_2d_all_kernels(kernel, shape, tol):
    opts, X, V, P = cases.rotation_2d(N=2048, Np=3000, n=9, kernel=kernel, advection=False)
    opts["pu_mesh_shape_factor"] = shape
    print(f"case: rotation2d kernel={kernel} shape={shape}", flush=True)
    g, o, res = interp_both(opts, X, V, P)
    err = rel_err(res[0][0], res[1][0])
    print(f"kernel {kernel}: max rel err {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("kernel,shape", [(1, 1.0), (2, 1.0), (5, 10.0), (6, 10.0)])
def test_interp_2d_fp32_evaluation_kernels(kernel, shape):
    """The FP32 evaluation (mixed path) of every kernel type it covers besides r^3 and
    the per-patch Gaussian (tested with their polynomial blocks): Wendland C0 and C2,
    the inverse multiquadric (hardware reciprocal square root) and the Gaussian
    (hardware exp2), against the FP64 oracle.  The FP32 values carry the rounding of
    the FP32 coefficients and sums: 1e-4 relative bounds it where the patch sums cancel
    (the global kernels at shape factor 10)."""
    opts, X, V, P = cases.rotation_2d(N=2048, Np=3000, n=9, kernel=kernel, advection=False)
    opts["pu_mesh_shape_factor"] = shape
    opts["lagrbf_precision"] = "mixed"
    print(f"case: rotation2d FP32 evaluation kernel={kernel} shape={shape}", flush=True)
    g, o, res = interp_both(opts, X, V, P)
    assert np.all(np.isfinite(res[0][0]))
    err = rel_err(res[0][0], res[1][0])
    print(f"kernel {kernel} FP32 evaluation: max rel err {err:.3e}")
    assert err < 1e-4


@pytest.mark.parametrize("solve_method", [0, 1, 2])
def test_interp_3d(solve_method):
    opts, X, V, P = cases.fourier_3d(nlat=20, Np=2000, n=5, solve_method=solve_method)
    print(f"case: fourier3d solve={solve_method}", flush=True)
    g, o, res = interp_both(opts, X, V, P)
    assert rel_err(res[0][0], res[1][0]) < TOL_VAL


@pytest.mark.parametrize("dim,kernel,mode", [(3, 2, "cell"), (3, 2, "ref"), (2, 2, "cell"), (2, 0, "cell"),
                                             (3, 6, "cell"), (2, 1, "ref")])
def test_eval_kernels_vs_oracle(dim, kernel, mode):
    """Both evaluation kernels (cell-ordered and per-patch reference order) against the oracle."""
    if dim == 3:
        opts, X, V, P = cases.fourier_3d(nlat=24, Np=6000, n=6, kernel=kernel)
    else:
        opts, X, V, P = cases.rotation_2d(N=3000, Np=6000, n=11, kernel=kernel, advection=False)
    if kernel == 6:
        opts["pu_mesh_shape_factor"] = 10.0
    opts["lagrbf_eval"] = mode
    print(f"case: eval dim={dim} kernel={kernel} mode={mode}", flush=True)
    g, o, res = interp_both(opts, X, V, P)
    err = rel_err(res[0][0], res[1][0])
    print(f"max rel err {err:.3e}")
    assert err < (TOL_VAL if kernel != 0 else 1e-11)


def _coef_errors(g, o, P_nodes):
    errs = []
    for p in range(P_nodes):
        xo = o.patch_coeffs(p)
        if xo.size == 0:
            continue
        xg = g.debug_get("patch_coeffs", p, dtype=np.float64).reshape(xo.shape)
        errs.append(np.max(np.abs(xg - xo)) / np.max(np.abs(xo)))
    return np.array(errs)


@pytest.mark.parametrize("precision,factor,tol", [("fp64", "auto", 1e-12), ("fp64", "bk", 1e-12), ("mixed", "auto", 1e-10)])
def test_patch_coefficients_vs_oracle(precision, factor, tol):
    """Per-patch coefficients: FP64 Cholesky and FP64 Bunch-Kaufman within 1e-12,
    FP32 factor + FP64 refinement within 1e-10 of the oracle's FP64 dsytrf solve."""
    opts, X, V, P = cases.fourier_3d(nlat=24, Np=4000, n=6)
    opts["lagrbf_precision"] = precision
    opts["lagrbf_mixed_factor"] = "fp32"
    opts["lagrbf_factor"] = factor
    g, o = pair(opts)
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
    errs = _coef_errors(g, o, 6 ** 3)
    print(f"{precision}/{factor}: {errs.size} patches, max rel coefficient err {errs.max():.3e}, "
          f"fallback {g.info('fallback'):.0f}, refine iters max {g.info('refine_iters'):.0f}")
    assert errs.max() < tol
    if precision == "mixed":
        vg = np.zeros((P.shape[0], 3)); vo = np.zeros_like(vg)
        g.output(fvals=vg); o.output(fvals=vo)
        err = rel_err(vg, vo)
        print(f"mixed FP32 evaluation: max rel value err {err:.3e}")
        assert err < 1e-5      # FP32 evaluation (values), stated tolerance


def test_mixed_with_fp64_factor():
    """lagrbf_mixed_factor fp64: the mixed path factors in FP64 (no refinement) and
    keeps the FP32 evaluation: coefficients within 1e-12 of the oracle, values within
    the FP32 evaluation's 1e-5."""
    opts, X, V, P = cases.fourier_3d(nlat=24, Np=4000, n=6)
    print("case: mixed precision with FP64 factors", flush=True)
    with pytest.raises(RuntimeError, match="lagrbf_mixed_factor"):
        L.LagRBFCtx(petsc_options=dict(opts, lagrbf_mixed_factor="fp16"), precision="mixed")
    opts = dict(opts, lagrbf_precision="mixed", lagrbf_mixed_factor="fp64")
    g, o = pair(opts)
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
    errs = _coef_errors(g, o, 6 ** 3)
    print(f"mixed, FP64 factor: max rel coefficient err {errs.max():.3e}, refine iters {g.info('refine_iters'):.0f}")
    assert errs.max() < 1e-12
    assert g.info("refine_iters") == 0
    vg = np.zeros((P.shape[0], 3)); vo = np.zeros_like(vg)
    g.output(fvals=vg); o.output(fvals=vo)
    assert rel_err(vg, vo) < 1e-5


def test_mixed_rk4_positions():
    """Mixed path (FP32 factor + FP64 refinement, FP32 evaluation) through 50 RK4
    steps (the step count of the 3D workload) against the FP64 oracle.  The 1e-9
    position bar binds the FP64 path; the mixed path's bound follows from its FP32
    evaluation's value error e, measured on the initial swarm: positions drift by at
    most about N dt e over the run (e in velocity units), asserted with a factor 2 of
    margin relative to the position scale."""
    opts, X, V, P = cases.fourier_3d(nlat=24, Np=3000, n=6)
    opts["lagrbf_precision"] = "mixed"
    g, o = pair(opts)
    vals, out = [], []
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
        v = np.zeros((P.shape[0], 3)); c.output(fvals=v); vals.append(v)
        c.advect_steps_c(50, 1e-3)
        xs = np.zeros_like(P); c.output(coords=xs); out.append(xs)
    e = np.abs(vals[0] - vals[1]).max()
    err = rel_err(out[0], out[1])
    bound = 2 * 50 * 1e-3 * e / np.abs(out[1]).max()
    print(f"mixed RK4 50 steps: max rel position err {err:.3e} (bound {bound:.3e} from value err {e:.3e})")
    assert err < bound


def test_reorder_bit_identical():
    """The swarm stored in cell order (default) and in fill order give bit-identical
    positions, fields, pids and cell ids through RK4 steps that remove particles;
    both match the oracle."""
    opts, X, V, P = cases.fourier_3d(nlat=24, Np=4000, n=6)
    print("case: reorder on/off, RK4 25 steps dt=4e-3", flush=True)
    res = []
    for reorder in (True, False):
        o = dict(opts, lagrbf_reorder=reorder)
        g = L.LagRBFCtx(petsc_options=o)
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c(); g.advect_steps_c(25, 4e-3)
        xs = np.zeros_like(P); vals = np.zeros((P.shape[0], 3)); pid = np.zeros(P.shape[0], dtype=np.int64)
        g.output(xs, vals, pid)
        res.append((xs, vals, pid, g.debug_get("cellid_points"), g.debug_get("alive"), g.info("removed")))
    a, b = res
    assert a[5] == b[5] and a[5] > 0, (a[5], b[5])
    assert np.array_equal(a[4], b[4])
    for k in range(3):
        assert np.array_equal(a[k], b[k]), k
    live = a[4].astype(bool)
    assert np.array_equal(a[3][live], b[3][live])
    from oracle.oracle import OracleCtx
    oc = OracleCtx(L.LagRBFCtx(petsc_options=dict(opts)).options)
    oc.fill_data(X, V); oc.fill_points(P); oc.interpolate_c(); oc.advect_steps_c(25, 4e-3)
    xo = np.zeros_like(P); oc.output(coords=xo)
    assert oc.removed_count() == a[5]
    assert rel_err(a[0], xo) < TOL_POS
    assert np.array_equal(a[3][live], oc.cellids(False)[live])


def test_binning_bit_exact():
    opts, X, V, P = cases.fourier_3d(nlat=16, Np=3000, n=5)
    print("case: binning 3D", flush=True)
    g, o = pair(opts)
    for c in (g, o):
        c.fill_data(X, V)
        c.fill_points(P)
        c.interpolate_c()
    assert np.array_equal(g.debug_get("cellid_points"), o.cellids(False))
    assert np.array_equal(g.debug_get("cellid_data"), o.cellids(True))
    cs_o, cn_o = o.cell_lists(X.shape[0])
    assert np.array_equal(g.debug_get("cell_start"), cs_o)
    assert np.array_equal(g.debug_get("cell_nodes"), cn_o)
    P_nodes = int(np.prod(opts["pu_mesh_n"]))
    for p in range(P_nodes):
        m_g = g.debug_get("patch_members", p)
        m_o = o.patch_members(p)
        assert np.array_equal(m_g, m_o), p
    # weights (Shepard) to FP64 rounding
    wg = np.zeros((P.shape[0], 8)); wo = o.weights()
    g.output(fvals=wg, fieldname="weights")
    assert rel_err(wg, wo) < 1e-13


@pytest.mark.parametrize("per_cell", [300, 5000, 20000])
def test_cell_lists_large_cells_bit_exact(per_cell):
    """Cells far above 256 nodes (a cluster inside one cell of a 4^3 PU: the LDS
    bitonic sort up to 16,384 nodes per cell, the block-wide rank count beyond)
    next to ordinary cells: per-cell node lists in ascending node order (patch
    membership concatenates them), bit-exact against the oracle."""
    rng = np.random.default_rng(cases.SEED + 401)
    X = np.concatenate([rng.random((3000, 3)), 0.4 + 0.2 * rng.random((per_cell, 3))])
    X = np.ascontiguousarray(X[rng.permutation(X.shape[0])])
    V = np.zeros_like(X)
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[4] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=1.0, kernel_type=2, solve_method=2)
    print(f"case: cell lists, {per_cell} nodes in one cell", flush=True)
    g, o = pair(opts)
    for c in (g, o):
        c.fill_data(X, V)
    cs_o, cn_o = o.cell_lists(X.shape[0])
    assert np.diff(cs_o).max() >= per_cell
    assert np.array_equal(g.debug_get("cell_start"), cs_o)
    assert np.array_equal(g.debug_get("cell_nodes"), cn_o)


@pytest.mark.parametrize("dim,overlap", [(3, 0.9), (3, 1.0), (3, 1.7), (2, 1.0), (2, 3.0)])
def test_weights_vs_brute_force(dim, overlap):
    """Shepard weights from the support window against the oracle's sum over every
    PU node (swarm_interp.c:283-396): a node missing from the window would show."""
    if dim == 3:
        opts, X, V, P = cases.fourier_3d(nlat=8, Np=3000, n=7, overlap=overlap)
    else:
        opts, X, V, P = cases.rotation_2d(N=100, Np=3000, n=9, advection=False)
        opts["pu_mesh_overlap"] = overlap
    print(f"case: weights dim={dim} overlap={overlap}", flush=True)
    from oracle.oracle import OracleCtx
    g = L.LagRBFCtx(petsc_options=dict(opts))
    o = OracleCtx(dict(g.options, oracle_weights_brute=True))
    g.fill_points(P); o.fill_points(P)
    wg = np.zeros((P.shape[0], 1 << dim))
    g.output(fvals=wg, fieldname="weights")
    wo = o.weights()
    assert np.array_equal(wg == 0.0, wo == 0.0)
    assert rel_err(wg, wo) < 1e-13


def test_points_on_faces_and_edges():
    """x == O and x == L on every axis: located in the first / last cell."""
    opts, X, V, _ = cases.rotation_2d(N=500, Np=10, n=5, advection=False)
    P = np.array([[0.0, 0.0], [1.0, 1.0], [1.0, 0.0], [0.0, 1.0], [0.25, 0.75], [0.5, 0.5], [1.0, 0.5]])
    g, o = pair(opts)
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
    assert np.array_equal(g.debug_get("cellid_points"), o.cellids(False))


def test_out_of_box_rejected():
    opts, X, V, P = cases.rotation_2d(N=500, Np=10, n=5, advection=False)
    g = L.LagRBFCtx(petsc_options=dict(opts))
    bad = P.copy(); bad[3, 0] = 1.5
    with pytest.raises(RuntimeError, match="62"):
        g.fill_points(bad)


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_rk_fixed_steps(scheme):
    opts, X, V, P = cases.rotation_2d(N=2048, Np=2000, n=9, scheme=scheme)
    print(f"case: rk fixed steps scheme={scheme}", flush=True)
    g, o = pair(opts)
    out = []
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
        c.advect_steps_c(10, 0.02)
        xs = np.zeros_like(P); vals = np.zeros((P.shape[0], 2))
        c.output(xs, vals)
        out.append((xs, vals))
    assert rel_err(out[0][0], out[1][0]) < TOL_POS
    assert rel_err(out[0][1], out[1][1]) < 1e-10


@pytest.mark.parametrize("scheme", [2, 3])
def test_advect_interpolate_cfl_loop(scheme):
    """InternalForwardAdvectionLoop semantics: CFL dt, clipping to t_next, reset + clear."""
    opts, X, V, P = cases.rotation_2d(N=2048, Np=1500, n=9, scheme=scheme)
    opts["cfl_fac"] = 0.3
    print(f"case: advect_interpolate scheme={scheme}", flush=True)
    g, o = pair(opts)
    out = []
    for c in (g, o):
        c.fill_points(P)
        for step in range(2):   # AdvectionInterpolation.advect_step call sequence
            c.fill_data(X, V)
            c.interpolate_c()
            c.advect_interpolate_c(0.1 * step, 0.1 * (step + 1))
        xs = np.zeros_like(P)
        c.output(coords=xs)
        out.append(xs)
    assert rel_err(out[0], out[1]) < TOL_POS


def test_analytical_advection():
    opts, X, V, P = cases.rotation_2d(N=10, Np=2000, n=9, scheme=2)
    opts["cfl_fac"] = 0.3
    g, o = pair(opts)
    rot = lambda x: np.stack([-(x[:, 1] - 0.5), x[:, 0] - 0.5], axis=1)
    from lagrbf_amd.utils import numpy_velocity_callback
    g.fill_points(P); o.fill_points(P)
    g.set_advection_function(numpy_velocity_callback(rot))
    o.set_velocity_function(rot)
    g.advect_analytical_c(0.0, 0.5); o.advect_analytical_c(0.0, 0.5)
    xg = np.zeros_like(P); xo = np.zeros_like(P)
    g.output(coords=xg); o.output(coords=xo)
    assert rel_err(xg, xo) < TOL_POS
    # closure: rotation by 0.5 rad about (1/2, 1/2)
    c, s = np.cos(0.5), np.sin(0.5)
    ex = np.stack([0.5 + c * (P[:, 0] - 0.5) - s * (P[:, 1] - 0.5), 0.5 + s * (P[:, 0] - 0.5) + c * (P[:, 1] - 0.5)], axis=1)
    # RK4 truncation at CFL 0.3 (dt ~ 0.05): ~2e-6 closure error
    assert np.max(np.abs(xg - ex)) < 1e-5


from tests.golden import make_golden  # noqa: E402

_REG = make_golden.regression_cases()


@pytest.mark.parametrize("case", _REG, ids=[c[0] for c in _REG])
def test_gpu_vs_golden_regression(case):
    """The GPU library against the committed oracle vectors (tests/golden/oracle_regression.npz):
    binning bit-exact, weights 1e-13, values 1e-12 (1e-11 / 1e-9 for the worse-conditioned
    LINEAR, bump and multiquadric patches), positions 1e-9."""
    import os
    name, opts, X, V, P, action = case
    print(f"case: golden {name}", flush=True)
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_regression.npz"),
                allow_pickle=False)
    c = L.LagRBFCtx(petsc_options=dict(opts))
    c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
    if action[0] == "steps":
        c.advect_steps_c(action[1], action[2])
    xs = np.zeros_like(P); vals = np.zeros((P.shape[0], c.nfields)); w = np.zeros((P.shape[0], 1 << P.shape[1]))
    c.output(xs, vals)
    c.output(fvals=w, fieldname="weights")
    assert np.array_equal(c.debug_get("cellid_points"), g[f"{name}/cellid_points"])
    assert np.array_equal(c.debug_get("cellid_data"), g[f"{name}/cellid_data"])
    assert np.array_equal(c.debug_get("cell_start"), g[f"{name}/cell_start"])
    assert np.array_equal(c.debug_get("cell_nodes"), g[f"{name}/cell_nodes"])
    assert c.info("removed") == g[f"{name}/removed"][0]
    live = c.debug_get("alive").astype(bool)      # removed particles: zero output rows, stale oracle weights
    assert rel_err(w[live], g[f"{name}/weights"][live]) < 1e-13
    assert not np.any(w[~live])
    if action[0] == "steps":
        assert rel_err(xs, g[f"{name}/coords"]) < TOL_POS
        assert rel_err(vals, g[f"{name}/values"]) < 1e-10
    else:
        k = int(opts["kernel_type"])
        tol = {4: 1e-9, 0: 1e-11, 3: 1e-11}.get(k, TOL_VAL)
        assert rel_err(vals, g[f"{name}/values"]) < tol


def test_c1_rotation_full_size():
    """The 2D rotation configuration at full size: 4,096 nodes, 16x16 PU cells with
    weight radius 1.5 cell diagonals, Wendland C2, 10,000 particles in r < 0.4, RK4
    dt = 0.01 for 628 steps.  GPU vs oracle on the same inputs (positions 1e-9);
    the return-to-start error is reported (reference semantics: sum of the corner
    weights < 1 at this radius, so the rotation lags and particles drift inwards)."""
    opts, X, V, P = cases.rotation_2d(N=4096, Np=10000, n=17, scheme=2)
    print("case: C1 rotation 628 RK4 steps", flush=True)
    g, o = pair(opts)
    out = []
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
        c.advect_steps_c(628, 0.01)
        xs = np.zeros_like(P); c.output(coords=xs); out.append(xs)
    err = rel_err(out[0], out[1])
    ret = np.linalg.norm(out[0] - P, axis=1)
    print(f"C1: GPU vs oracle max rel position err {err:.3e}; return-to-start error max {ret.max():.4e}, "
          f"mean {ret.mean():.4e}; removed {g.info('removed'):.0f}")
    assert err < TOL_POS
    assert g.info("removed") == o.removed_count() == 0


def clustered_3d(n_uniform=6000, n_cluster=1500, sigma=0.03, Np=3000, n=6, seed=5):
    """Uniform nodes plus one Gaussian cluster (patches of several hundred to a few
    thousand members), the Fourier velocity, particles uniform in [0.1, 0.9]^3."""
    rng = np.random.default_rng(cases.SEED + 40 + seed)
    X = np.concatenate([rng.random((n_uniform, 3)),
                        np.clip(0.5 + sigma * rng.standard_normal((n_cluster, 3)), 0.0, 1.0)])
    V = cases.fourier_velocity(X, cases.fourier_modes())
    P = 0.1 + 0.8 * rng.random((Np, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[n] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=1.0, pu_mesh_shape_factor=1.0, kernel_type=2, solve_method=0,
                advection=True, advection_scheme=2, cfl_fac=1e6, dt=1e-3)
    return opts, np.ascontiguousarray(X), np.ascontiguousarray(V), P


def _wc2(r, eps):
    u = np.maximum(1.0 - eps * r, 0.0)
    return u ** 4 * (4.0 * eps * r + 1.0)


@pytest.mark.parametrize("precision,big_min,tol", [("fp64", 16, 1e-12), ("mixed", 16, 1e-10), ("fp64", 512, 1e-12),
                                                   ("mixed", 512, 1e-10)])
def test_large_patch_cholesky(precision, big_min, tol):
    """Uneven patch sizes (uniform nodes + a cluster: 40 to ~1900 members) through the
    large-patch Cholesky kernels (all patches above big_min).  Patches up to 512
    members: coefficients against the oracle's dsytrf solve (forward error, tol).
    The cluster patches are ill-conditioned (dense nodes inside the Wendland
    support): there the bar is the backward error, |A x - b| / (|A| |x|) < 1e-14
    with A assembled independently in numpy, and the forward error against the
    oracle is reported with cond(A)."""
    opts, X, V, P = clustered_3d()
    opts["lagrbf_precision"] = precision
    opts["lagrbf_mixed_factor"] = "fp32"
    opts["lagrbf_chol_big_min"] = big_min
    print(f"case: large patches {precision} big_min={big_min}", flush=True)
    g, o = pair(opts)
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
    eps = g.info("eps_r")
    sizes = np.array([int(g.debug_get("patch_flags", p)[2]) for p in range(6 ** 3)])
    small_err, big_bwd, big_fwd = 0.0, 0.0, 0.0
    for p in range(6 ** 3):
        xo = o.patch_coeffs(p)
        if xo.size == 0:
            continue
        xg = g.debug_get("patch_coeffs", p, dtype=np.float64).reshape(xo.shape)
        fwd = np.max(np.abs(xg - xo)) / np.max(np.abs(xo))
        if sizes[p] <= 512:
            small_err = max(small_err, fwd)
            continue
        m = g.debug_get("patch_members", p)
        xs = X[m]
        A = _wc2(np.sqrt(((xs[:, None, :] - xs[None, :, :]) ** 2).sum(-1)), eps)
        b = V[m]
        bwd = np.max(np.abs(A @ xg - b)) / (np.max(np.abs(A).sum(1)) * np.max(np.abs(xg)))
        big_bwd = max(big_bwd, bwd)
        big_fwd = max(big_fwd, fwd)
    big = int((sizes > 512).sum())
    pbig = int(np.argmax(sizes))
    mb = g.debug_get("patch_members", pbig)
    xs = X[mb]
    cond = np.linalg.cond(_wc2(np.sqrt(((xs[:, None, :] - xs[None, :, :]) ** 2).sum(-1)), eps))
    print(f"patch sizes {sizes.min()}..{sizes.max()}, {big} above 512; small patches: max rel coefficient err "
          f"{small_err:.3e}; large patches: backward err {big_bwd:.3e}, forward err {big_fwd:.3e} (cond of the largest "
          f"{cond:.2e}); fallback {g.info('fallback'):.0f}; refine iters max {g.info('refine_iters'):.0f}")
    assert big > 0
    if precision == "fp64":
        assert g.info("fallback") == 0     # FP32 factors of the dense cluster may fall back to FP64 (counted)
    assert small_err < tol
    assert big_bwd < 1e-14
    assert big_fwd < max(tol, 100 * cond * 1e-16)
    vg = np.zeros((P.shape[0], 3)); vo = np.zeros_like(vg)
    g.output(fvals=vg); o.output(fvals=vo)
    assert rel_err(vg, vo) < (max(TOL_VAL, 100 * cond * 1e-16) if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("scheme,precision", [(2, "fp64"), (2, "mixed"), (1, "fp64"), (0, "mixed")])
def test_graph_steps_match_eager(scheme, precision):
    """Fixed-step loop replayed from captured graphs (one per buffer parity) against
    the same loop run eagerly: positions and fields bit-identical after every call,
    with host reads and synchronisations between the calls' replays."""
    opts, X, V, P = cases.fourier_3d(nlat=20, Np=20000, n=9, scheme=scheme)
    print(f"case: graph replay fourier3d scheme={scheme} {precision}", flush=True)
    ctxs = []
    for graph in (False, True):
        o = dict(opts, lagrbf_graph=graph)
        g = L.LagRBFCtx(petsc_options=o, precision=precision)
        g.fill_data(X, V)
        g.fill_points(P)
        g.interpolate_c()
        ctxs.append(g)
    for call in range(4):
        for g in ctxs:
            g.advect_steps_c(3, 2e-3)
        out = []
        for g in ctxs:
            xs = np.zeros_like(P); vals = np.zeros((P.shape[0], 3))
            g.output(xs, vals)
            out.append((xs, vals, g.info("removed")))
        assert np.array_equal(out[0][0], out[1][0]), f"positions differ after call {call}"
        assert np.array_equal(out[0][1], out[1][1]), f"fields differ after call {call}"
        assert out[0][2] == out[1][2]
        assert ctxs[1].info("graph_steps") == 2 * (call + 1)
        assert ctxs[0].info("graph_steps") == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("Np", [600, 60000])
@pytest.mark.parametrize("case", ["3d", "2d", "3d-few-nodes"])
def test_eval_launch_shapes_identical(precision, Np, case):
    """Every evaluation launch shape the autotuner chooses between (one block per
    cell, per listed non-empty cell in XCD runs or round-robin, lane groups of
    32 / 16 / 8 over the ascending cell list, one or two table rows per wait and
    two occupancy floors in FP32, and in FP32 the accumulation on the matrix cores)
    gives the same values bit for bit, on a sparse and a dense swarm, in 3D and 2D,
    and with 0-9 nodes per cell (the node loops' short and empty tails); FP64
    against the oracle at 1e-12.  FP64 has neither the two-row loop nor the
    matrix-core accumulation: there the variants 6-9 run as 0, 1, 5 and 5, 12-13
    (runs of 4 and 16 listed cells dealt round-robin) as 10-11, and 14-17 as 1, 5,
    10 and 11."""
    if case == "3d":
        opts, X, V, P = cases.fourier_3d(nlat=20, Np=Np, n=9)
    elif case == "3d-few-nodes":
        opts, X, V, P = cases.fourier_3d(nlat=20, N=1500, Np=Np, n=9)
    else:
        opts, X, V, P = cases.rotation_2d(N=4096, Np=Np, n=17, advection=False)
    dim = X.shape[1]
    opts["advection"] = False
    print(f"case: launch shapes {case} Np={Np} {precision}", flush=True)
    vals = []
    nvar = 18
    for v in range(nvar):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_eval_variant=v), precision=precision)
        g.fill_data(X, V)
        g.fill_points(P)
        g.interpolate_c()
        out = np.zeros((P.shape[0], dim))
        g.output(fvals=out)
        used = g.info("eval_variant")
        g.destroy()
        vals.append(out)
        want = v if precision == "mixed" else {6: 0, 7: 1, 8: 5, 9: 5, 12: 10, 13: 11, 14: 1, 15: 5, 16: 10,
                                               17: 11}.get(v, v)
        assert used == want, (v, used)
    for v in range(1, nvar):
        assert np.array_equal(vals[0], vals[v]), f"variant {v} differs"
    if precision == "fp64":
        from oracle.oracle import OracleCtx
        o = OracleCtx(L.LagRBFCtx(petsc_options=dict(opts)).options)
        o.fill_data(X, V); o.fill_points(P); o.interpolate_c()
        vo = np.zeros_like(vals[0])
        o.output(None, vo, None)
        assert rel_err(vals[0], vo) < TOL_VAL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_six_field_evaluation(precision):
    """Six interpolated fields in 3D (two velocity snapshots): the FP32 cell table
    evaluates all six in one pass (its own kernel), FP64 falls back to the per-cell
    kernel.  Against the oracle (FP64 1e-12, FP32 evaluation 1e-5 relative), and every
    FP32 launch shape bit-identical."""
    opts, X, V, P = cases.fourier_3d(nlat=20, Np=30000, n=9)
    F = np.concatenate([V, 0.7 * np.roll(V, 1, axis=1)], axis=1)
    opts = dict(opts, n_interpolation_fields=6, advection=False)
    print(f"case: six fields fourier3d {precision}", flush=True)
    from oracle.oracle import OracleCtx
    o = OracleCtx(L.LagRBFCtx(petsc_options=dict(opts)).options)
    o.fill_data(X, F); o.fill_points(P); o.interpolate_c()
    vo = np.zeros((P.shape[0], 6))
    o.output(None, vo, None)
    vals = []
    for v in ((-1,) if precision == "fp64" else (0, 1, 5, 9, 10, 13)):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_eval_variant=v), precision=precision)
        g.fill_data(X, F)
        g.fill_points(P)
        g.interpolate_c()
        out = np.zeros((P.shape[0], 6))
        g.output(fvals=out)
        g.destroy()
        vals.append(out)
    err = rel_err(vals[0], vo)
    print(f"six fields {precision}: rel err vs oracle {err:.3e}", flush=True)
    assert err < (TOL_VAL if precision == "fp64" else 1e-5)
    for v in vals[1:]:
        assert np.array_equal(vals[0], v)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tuner_runoff_keeps_a_valid_shape(precision):
    """The autotuner's trial and run-off run to completion inside an RK loop (the
    run-off's decision prefers, among shapes within 2% of the fastest, the one that
    keeps neighbourhoods in fewer L2 caches): the shape it keeps is a valid one, the
    later steps run it, and the positions equal a run with that shape forced, bit
    for bit (every shape gives identical values)."""
    opts, X, V, P = cases.fourier_3d(nlat=20, Np=20000, n=9, scheme=2)
    opts.update(dt=1e-3, cfl_fac=1e6)
    print(f"case: tuner run-off over RK4 steps, {precision}", flush=True)
    out = {}
    for forced in (None, "kept"):
        o = dict(opts)
        if forced == "kept":
            o["lagrbf_eval_variant"] = kept
        g = L.LagRBFCtx(petsc_options=o, precision=precision)
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        g.advect_steps_c(12, 1e-3)   # the trial, a run-off of up to 18 shapes, then the kept one
        xs = np.zeros_like(P); g.output(coords=xs)
        used = int(g.info("eval_variant"))
        g.destroy()
        if forced is None:
            kept = used
        out[forced] = xs
        print(f"{precision} {forced}: shape {used}")
        assert 0 <= used < 18
    assert np.array_equal(out[None], out["kept"])
