"""The tiled multi-workgroup Cholesky (chol_tiled.hip, patches too large for one
workgroup) against the FP64 oracle and against the one-workgroup large-patch
kernel: lowering -lagrbf_chol_big_min sends patches of a few hundred members
through it (many tile columns, partial last tiles, every polynomial case), and a
cluster of 3,000 nodes gives patches of thousands of members as the clustered
workload has them.

Tolerances: FP64 coefficients 1e-12 (Wendland C2) or 1e-10 (saddle-point
systems, cond ~1e4-1e6); FP32 factor + FP64 refinement 1e-10 (the mixed cases
set lagrbf_mixed_factor fp32)."""
import numpy as np
import pytest

import lagrbf_amd as L
from tests import cases
from tests.test_gpu_aug import clustered_small, coef_errors, rel_err, run_both, small_c2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["wendland", "phs3-linear", "gauss-local-quadratic"])
def test_tiled_factor_vs_oracle(case, precision):
    if case == "wendland":
        opts, X, V, P = cases.fourier_3d(nlat=24, Np=2000, n=6)
        opts.update(lagrbf_precision=precision, lagrbf_mixed_factor="fp32")
        nP, tol = 6 ** 3, (1e-12 if precision == "fp64" else 1e-10)
    elif case == "phs3-linear":
        opts, X, V, P = small_c2(precision=precision)
        nP, tol = 6 ** 3, 1e-10
    else:
        opts, X, V, P = clustered_small(precision=precision)
        nP, tol = 7 ** 3, 1e-10
    opts.update(lagrbf_chol_big_min=64, advection=False)
    print(f"case: tiled factor {case} {precision} (patches above 64 tiled)", flush=True)
    g, o, out = run_both(opts, X, V, P)
    if case == "wendland":
        ec = []
        for p in range(nP):
            xo = o.patch_coeffs(p)
            if xo.size:
                xg = g.debug_get("patch_coeffs", p, dtype=np.float64).reshape(xo.shape)
                ec.append(np.abs(xg - xo).max() / np.abs(xo).max())
        ec, eb = np.array(ec), np.zeros(1)
    else:
        ec, eb = coef_errors(g, o, nP, V.shape[1])
    ev = rel_err(out[0][0], out[1][0])
    print(f"tiled {case}/{precision}: coef err {ec.max():.3e}, beta err {eb.max():.3e}, value err {ev:.3e}, "
          f"max patch {g.info('max_patch'):.0f}, fallback {g.info('fallback'):.0f}, "
          f"refine iters {g.info('refine_iters'):.0f}")
    assert ec.max() < tol and eb.max() < tol
    assert g.info("fallback") == 0
    if precision == "fp64":
        assert ev < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiled_and_single_workgroup_agree(precision):
    """Both large-patch factors on the clustered Gaussian + quadratic case, every
    patch above 64 members: the interpolated values agree to the coefficients'
    rounding (FP64 1e-12; mixed: refined coefficients to 1e-10, FP32 tables 1e-6)."""
    opts, X, V, P = clustered_small(precision=precision)
    vals = {}
    for bf in ("tiled", "vector"):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_chol_big_min=64, lagrbf_big_factor=bf, advection=False))
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        v = np.zeros((P.shape[0], 3)); g.output(fvals=v)
        vals[bf] = v
        print(f"{bf}: max patch {g.info('max_patch'):.0f}, fallback {g.info('fallback'):.0f}")
        g.destroy()
    err = rel_err(vals["tiled"], vals["vector"])
    print(f"tiled vs one workgroup per patch ({precision}): {err:.3e}")
    assert err < (1e-12 if precision == "fp64" else 1e-6)


def _cluster_case(kernel, precision, deg=-1, local=0.0):
    """2,000 uniform nodes and 3,000 in a Gaussian cluster (sigma 0.06) about the
    central node of a 5^3 PU: its patch holds ~3,250 members, its neighbours up to
    ~1,700.  A narrower cluster inside a cell this large makes the patch-average
    h_local far wider than the cluster's spacing, the Gaussian nearly flat and the
    system singular to FP64 (cond 8e14 at sigma 0.02, 6e4 at 0.06)."""
    rng = np.random.default_rng(cases.SEED + 313)
    X = np.concatenate([rng.random((2000, 3)), np.clip(0.5 + 0.06 * rng.standard_normal((3000, 3)), 0, 1)])
    V = cases.fourier_velocity(X, cases.fourier_modes())
    P = rng.random((1500, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[5] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=1.0, pu_mesh_shape_factor=1.0, kernel_type=kernel, solve_method=0,
                lagrbf_precision=precision, lagrbf_mixed_factor="fp32", lagrbf_setup_all=True, oracle_setup_all=True)
    if deg >= 0:
        opts["lagrbf_poly_degree"] = deg
    if local > 0:
        opts["lagrbf_local_eps"] = local
    return opts, np.ascontiguousarray(X), np.ascontiguousarray(V), P


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiled_thousands_of_members(precision):
    """Patches of up to ~3,250 members (51 tile columns) with the Gaussian, eps = 3 / h_local
    and a quadratic block: coefficients and beta against the oracle's Bunch-Kaufman
    solve of the full saddle-point system.  The largest patch's saddle-point matrix
    has cond 4.7e5 at n = 3,254, so two backward-stable FP64 solvers agree to about
    n^(1/2) cond u ~ 3e-9 in the coefficients (bar 2e-9; measured 5.5e-10); the
    interpolated values to 1e-12.  The mixed path's FP32 factor breaks down on some
    of these patches, which are then refactored with the FP64 tiled Cholesky."""
    opts, X, V, P = _cluster_case(6, precision, deg=2, local=3.0)
    print(f"case: ~3,250-member patch, Gaussian local eps + quadratic, {precision}", flush=True)
    g, o, out = run_both(opts, X, V, P)
    ec, eb = coef_errors(g, o, 5 ** 3, 3)
    ev = rel_err(out[0][0], out[1][0])
    print(f"~3,250-member patch/{precision}: coef err {ec.max():.3e}, beta err {eb.max():.3e}, value err {ev:.3e}, "
          f"max patch {g.info('max_patch'):.0f}, fallback {g.info('fallback'):.0f}, "
          f"refine iters {g.info('refine_iters'):.0f}")
    assert g.info("max_patch") > 2500
    assert ec.max() < 2e-9 and eb.max() < 2e-9
    assert g.info("fallback") == g.info("big_refactor64")
    if precision == "fp64":
        assert ev < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiled_solve_matches_one_workgroup_solve(precision):
    """The large saddle-point patches' B22 solves over many workgroups (trsv_tiles.hip,
    tiles of 64, every patch above 64 members here) against the one-workgroup solve
    of the same factors: coefficients and beta to 1e-12 relative in FP64 (the sums
    run in another order); the mixed path's refinement reaches the FP64 solution
    either way (1e-10), with the same refinement-iteration count within one."""
    opts, X, V, P = clustered_small(precision=precision)
    opts.update(lagrbf_chol_big_min=64, advection=False)
    print(f"case: tiled vs one-workgroup B22 solve, {precision}", flush=True)
    res = {}
    for tiled in (True, False):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_tiled_solve=tiled))
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        cs = [g.debug_get("patch_coeffs", p, dtype=np.float64) for p in range(7 ** 3)]
        bs = [g.debug_get("patch_beta", p, dtype=np.float64) for p in range(7 ** 3)]
        res[tiled] = (cs, bs, g.info("refine_iters"), g.info("fallback"))
        g.destroy()
    ec = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(res[True][0], res[False][0]) if b.size)
    eb = max(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300) for a, b in zip(res[True][1], res[False][1]) if b.size)
    print(f"{precision}: coefficients {ec:.3e}, beta {eb:.3e}, refine iters {res[True][2]:.0f} / {res[False][2]:.0f}")
    tol = 1e-12 if precision == "fp64" else 1e-10
    assert ec < tol and eb < tol
    assert res[True][3] == res[False][3] == 0
    assert abs(res[True][2] - res[False][2]) <= 1
