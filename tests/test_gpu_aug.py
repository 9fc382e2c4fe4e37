"""GPU parity of the augmented patch systems (extensions: the r^3 kernel with a
linear or quadratic polynomial block, the Gaussian with per-patch shape parameter
eps = c / h_local, corner-normalised Shepard weights) through the C ABI, against the
FP64 CPU oracle on identical seeded inputs.

Tolerances (stated per result):
  FP64 interpolated values                      1e-12 relative (max-norm)
  FP64 coefficients c and beta                  1e-10 relative where the saddle-point
                                                matrix's condition number is ~1e5-1e6
                                                (3D, up to 300 members); 2e-9 at 1,000
                                                members (cond ~6e6) and 1e-8 for the 2D
                                                quadratic block (cond ~3e7): two backward
                                                stable FP64 solvers agree to ~cond * 1e-16
  FP32 factor + FP64 refinement coefficients    the same bounds (refinement against the
                                                FP64 residual reaches the FP64 floor; the
                                                mixed cases here factor in FP32,
                                                lagrbf_mixed_factor fp32)
  FP32-evaluation values (mixed path)           1e-4 relative on these coarse cases: the
                                                r^3 sums cancel by factors of 1e3-1e4 where
                                                a patch spans a sizeable part of a
                                                wavelength (relative error ~6e-8 x that)
  FP64 RK4 positions after N steps              1e-9 relative
"""
import numpy as np
import pytest

import lagrbf_amd as L
from tests import cases

pytestmark = pytest.mark.gpu

TOL_VAL = 1e-12
TOL_COEF = 1e-10
TOL_POS = 1e-9


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(opts):
    from oracle.oracle import OracleCtx
    g = L.LagRBFCtx(petsc_options=dict(opts))
    return g, OracleCtx(g.options)


def run_both(opts, X, V, P, steps=0, dt=1e-3):
    g, o = pair(opts)
    out = []
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
        if steps:
            c.advect_steps_c(steps, dt)
        vals = np.zeros((P.shape[0], V.shape[1])); xs = np.zeros_like(P)
        c.output(xs, vals)
        out.append((vals, xs))
    return g, o, out


def coef_errors(g, o, nP, nf):
    ec, eb = [], []
    for p in range(nP):
        xo = o.patch_coeffs(p)
        if xo.size == 0:
            continue
        bo, _ = o.patch_poly(p)
        xg = g.debug_get("patch_coeffs", p, dtype=np.float64).reshape(xo.shape)
        bg = g.debug_get("patch_beta", p, dtype=np.float64).reshape(bo.shape)
        scale = max(np.abs(xo).max(), np.abs(bo).max())
        ec.append(np.abs(xg - xo).max() / scale)
        eb.append(np.abs(bg - bo).max() / max(np.abs(bo).max(), 1e-300))
    return np.array(ec), np.array(eb)


def small_c2(nlat=16, Np=2000, n=6, deg=1, kernel=7, precision="fp64", **kw):
    """The C2 construction (jittered lattice, 16-mode Fourier velocity, r^3 +
    polynomial) at a size the oracle finishes in seconds; nlat 16 with 6^3 PU nodes
    gives patches of ~200-300 members, nlat 20 with 5^3 up to ~1,000 (the large-patch
    kernels)."""
    opts, X, V, P = cases.fourier_3d(nlat=nlat, Np=Np, n=n, kernel=kernel)
    opts.update(lagrbf_poly_degree=deg, lagrbf_precision=precision, lagrbf_mixed_factor="fp32", **kw)
    return opts, X, V, P


def clustered_small(local=3.0, deg=2, kernel=6, precision="fp64"):
    """Uniform nodes plus two Gaussian clusters: patches from ~10 to ~1,500 members,
    the Gaussian with eps = 3 / h_local and a quadratic block (the C3 construction)."""
    rng = np.random.default_rng(cases.SEED + 77)
    X = np.concatenate([rng.random((4000, 3)),
                        np.clip(0.35 + 0.03 * rng.standard_normal((1200, 3)), 0, 1),
                        np.clip(0.7 + 0.05 * rng.standard_normal((1200, 3)), 0, 1)])
    V = cases.fourier_velocity(X, cases.fourier_modes())
    P = rng.random((3000, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[7] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=1.0, pu_mesh_shape_factor=1.0, kernel_type=kernel, solve_method=0, advection=True,
                advection_scheme=1, cfl_fac=1e6, dt=1e-3, lagrbf_poly_degree=deg, lagrbf_local_eps=local,
                lagrbf_precision=precision, lagrbf_mixed_factor="fp32")
    return opts, np.ascontiguousarray(X), np.ascontiguousarray(V), P


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["c2-small", "c2-large-patches", "2d-quadratic"])
def test_phs3_values_and_coefficients(case, precision):
    if case == "c2-small":
        opts, X, V, P = small_c2(precision=precision)
        nP, tol_c = 6 ** 3, TOL_COEF
    elif case == "c2-large-patches":
        opts, X, V, P = small_c2(nlat=20, n=5, precision=precision)
        nP, tol_c = 5 ** 3, 2e-9
    else:
        opts, X, V, P = cases.rotation_2d(N=2048, Np=3000, n=9, kernel=7, advection=False)
        opts.update(lagrbf_poly_degree=2, lagrbf_precision=precision)
        nP, tol_c = 81, 1e-8
    print(f"case: r^3 {case} {precision}", flush=True)
    g, o, out = run_both(opts, X, V, P)
    ev = rel_err(out[0][0], out[1][0])
    ec, eb = coef_errors(g, o, nP, V.shape[1])
    print(f"{case}/{precision}: value err {ev:.3e}, coef err {ec.max():.3e}, beta err {eb.max():.3e}, "
          f"max patch {g.info('max_patch'):.0f}, fallback {g.info('fallback'):.0f}, "
          f"refine iters {g.info('refine_iters'):.0f}")
    assert ec.max() < tol_c and eb.max() < tol_c
    assert g.info("aug_fallback_bk") == 0
    assert ev < (TOL_VAL if precision == "fp64" else 1e-4)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gaussian_local_eps_quadratic(precision):
    opts, X, V, P = clustered_small(precision=precision)
    print(f"case: clustered Gaussian local eps quadratic {precision}", flush=True)
    g, o, out = run_both(opts, X, V, P)
    ev = rel_err(out[0][0], out[1][0])
    ec, eb = coef_errors(g, o, 7 ** 3, 3)
    for p in (0, 57, 171, 342):
        assert abs(g.debug_get("patch_eps", p, dtype=np.float64)[0] - o.patch_poly(p)[1]) <= 1e-14 * o.patch_poly(p)[1]
    print(f"local eps/{precision}: value err {ev:.3e}, coef err {ec.max():.3e}, beta err {eb.max():.3e}, "
          f"max patch {g.info('max_patch'):.0f}")
    assert ec.max() < TOL_COEF and eb.max() < TOL_COEF
    assert ev < (TOL_VAL if precision == "fp64" else 1e-5)


def test_bunch_kaufman_saddle_point():
    """-lagrbf_factor bk: the GPU's Bunch-Kaufman on the full (n+m) saddle-point
    matrix, the oracle's algorithm."""
    opts, X, V, P = small_c2(lagrbf_factor="bk")
    g, o, out = run_both(opts, X, V, P)
    ec, eb = coef_errors(g, o, 6 ** 3, 3)
    ev = rel_err(out[0][0], out[1][0])
    print(f"BK saddle point: value err {ev:.3e}, coef err {ec.max():.3e}, beta err {eb.max():.3e}")
    assert ev < TOL_VAL and ec.max() < TOL_COEF and eb.max() < TOL_COEF


def test_linear_reproduction_gpu():
    """Closed form: with a linear block a linear field is reproduced by every patch,
    so the value is f(x) times the sum of the used weights."""
    opts, X, _, P = small_c2(Np=3000)
    a = np.array([0.3, -1.1, 0.7, 2.0])
    V = np.stack([a[0] + X @ a[1:], -2 * a[0] + X @ a[:3], X[:, 0] * 0 + 1.0], axis=1)
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
    vals = np.zeros((P.shape[0], 3)); xs = np.zeros_like(P); w = np.zeros((P.shape[0], 8))
    g.output(xs, vals); g.output(fvals=w, fieldname="weights")
    ws = w.sum(axis=1)
    ex = np.stack([a[0] + xs @ a[1:], -2 * a[0] + xs @ a[:3], np.ones(len(xs))], axis=1) * ws[:, None]
    assert np.abs(vals - ex).max() < 1e-12 * np.abs(ex).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_phs3_launch_shapes_identical(precision):
    """Every evaluation launch shape (the matrix-core accumulation included) gives
    the same values bit for bit with the polynomial part."""
    opts, X, V, P = small_c2(Np=20000, precision=precision)
    opts["advection"] = False
    vals = []
    for v in range(18):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_eval_variant=v))
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        out = np.zeros((P.shape[0], 3)); g.output(fvals=out)
        g.destroy()
        vals.append(out)
    for v in range(1, 18):
        assert np.array_equal(vals[0], vals[v]), f"variant {v} differs"


@pytest.mark.parametrize("kernel", [7, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_particles_on_data_nodes(kernel, precision):
    """Particles placed exactly on data nodes, where r = 0: the FP32 evaluation forms
    r^2 as |p|^2 + |n|^2 - 2 p.n, which may round below zero there (taken by
    magnitude), so every value must stay finite and agree with the oracle to the
    path's bar (FP64 1e-12, FP32 evaluation 1e-4)."""
    opts, X, V, _ = small_c2(Np=10, kernel=kernel, precision=precision, deg=1 if kernel == 7 else -1)
    opts["advection"] = False
    P = np.ascontiguousarray(X[::3])
    print(f"case: {P.shape[0]} particles on data nodes, kernel {kernel}, {precision}", flush=True)
    g, o, out = run_both(opts, X, V, P)
    vg, vo = out[0][0], out[1][0]
    assert np.all(np.isfinite(vg))
    err = rel_err(vg, vo)
    print(f"on-node values: rel err {err:.3e}")
    assert err < (TOL_VAL if precision == "fp64" else 1e-4)


def test_phs3_rk4_positions_fp64():
    """FP64 r^3 + linear path through 50 RK4 steps (C2's step count) against the oracle."""
    opts, X, V, P = small_c2(Np=3000)
    g, o, out = run_both(opts, X, V, P, steps=50, dt=1e-3)
    err = rel_err(out[0][1], out[1][1])
    print(f"r^3 RK4 50 steps fp64: position err {err:.3e}, removed {g.info('removed'):.0f}/{o.removed_count()}")
    assert err < TOL_POS
    assert g.info("removed") == o.removed_count()


def test_phs3_rk4_positions_mixed():
    """Mixed path (FP32 factor + FP64 refinement, FP32 evaluation) through 50 RK4
    steps against the FP64 oracle.  Its bound follows from the FP32 evaluation's
    value error e (measured on the initial swarm): positions drift by at most about
    N dt e max|v| over the run, asserted with a factor 2 of margin relative to the
    position scale."""
    opts, X, V, P = small_c2(Np=3000, precision="mixed")
    g, o, out0 = run_both(opts, X, V, P)
    e = rel_err(out0[0][0], out0[1][0]) * np.abs(out0[1][0]).max()
    g, o, out = run_both(opts, X, V, P, steps=50, dt=1e-3)
    err = rel_err(out[0][1], out[1][1])
    bound = 2 * 50 * 1e-3 * e / np.abs(out[1][1]).max()
    print(f"r^3 RK4 50 steps mixed: position err {err:.3e} (bound {bound:.3e} from value err {e:.3e})")
    assert err < bound


def test_c1_rotation_phs3_full_size():
    """C1 as written: 4,096 nodes, 16x16 PU cells, weight radius 1.5 cell diagonals,
    r^3 + linear polynomial, 10,000 particles in r < 0.4, RK4 dt = 0.01, 628 steps.
    GPU vs oracle positions (1e-9), both weight normalisations; the return-to-start
    error and the error against the exact rotation by 6.28 rad are reported."""
    res = {}
    for wn in ("all", "corners"):
        opts, X, V, P = cases.rotation_2d(N=4096, Np=10000, n=17, kernel=7, scheme=2)
        opts.update(lagrbf_poly_degree=1, lagrbf_weights=wn)
        print(f"case: C1 r^3 + linear, weights {wn}, 628 RK4 steps", flush=True)
        g, o, out = run_both(opts, X, V, P, steps=628, dt=0.01)
        xg, xo = out[0][1], out[1][1]
        err = rel_err(xg, xo)
        th = 6.28
        c = np.array([0.5, 0.5])
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        exact = c + (P - c) @ R.T
        ret = np.linalg.norm(xg - P, axis=1)
        rot = np.linalg.norm(xg - exact, axis=1)
        res[wn] = (err, ret.max(), rot.max(), np.median(rot))
        print(f"C1 weights={wn}: GPU vs oracle {err:.3e}; return-to-start max {ret.max():.4e} mean {ret.mean():.4e}; "
              f"vs exact 6.28 rad max {rot.max():.4e} mean {rot.mean():.4e}")
        assert err < TOL_POS
    # corner-normalised weights reproduce the (linear) rotation field exactly, so only
    # RK4's truncation remains (~2e-8 at r = 0.4) for the particles whose corner
    # patches are set up; stage points entering cells of patches the lazy set-up has
    # not reached take a few biased stages (reference semantics, partition.c:452-469)
    assert res["corners"][3] < 1e-6 < res["all"][3]
