"""The matrix-core Cholesky (chol_mfma.hip) against the FP64 oracle and against the
vector-ALU kernel: coefficients per patch (16-column panels), FP64
and FP32 factors (the latter refined in FP64; the mixed cases set
lagrbf_mixed_factor fp32), pure and saddle-point systems,
patches in every size bucket of the factor tuner.

Tolerances: FP64 coefficients 1e-12 (Wendland C2, cond ~1e2) or 1e-10 (r^3 +
linear block, cond ~1e5-1e6); FP32 factor + FP64 refinement 1e-10."""
import numpy as np
import pytest

import lagrbf_amd as L
from tests import cases

pytestmark = pytest.mark.gpu


def coef_err(g, o, nP):
    e = []
    for p in range(nP):
        xo = o.patch_coeffs(p)
        if xo.size == 0:
            continue
        xg = g.debug_get("patch_coeffs", p, dtype=np.float64).reshape(xo.shape)
        bo, _ = o.patch_poly(p)
        if bo.size:
            bg = g.debug_get("patch_beta", p, dtype=np.float64).reshape(bo.shape)
            xo = np.concatenate([xo.ravel(), bo.ravel()]); xg = np.concatenate([xg.ravel(), bg.ravel()])
        e.append(np.abs(xg - xo).max() / np.abs(xo).max())
    return np.array(e)


@pytest.mark.parametrize("nb", [16])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [2, 7])
def test_mfma_factor_coefficients(nb, precision, kernel):
    if kernel == 7:
        opts, X, V, P = cases.fourier_3d(nlat=16, Np=2000, n=6, kernel=7)
        opts["lagrbf_poly_degree"] = 1
        nP, tol = 6 ** 3, 1e-10
    else:
        opts, X, V, P = cases.fourier_3d(nlat=24, Np=2000, n=6)
        nP, tol = 6 ** 3, (1e-12 if precision == "fp64" else 1e-10)
    opts.update(lagrbf_factor_nb=nb, lagrbf_precision=precision, lagrbf_mixed_factor="fp32", advection=False)
    print(f"case: mfma factor nb={nb} {precision} kernel {kernel}", flush=True)
    from oracle.oracle import OracleCtx
    g = L.LagRBFCtx(petsc_options=dict(opts))
    o = OracleCtx(g.options)
    for c in (g, o):
        c.fill_data(X, V); c.fill_points(P); c.interpolate_c()
    e = coef_err(g, o, nP)
    print(f"nb={nb} {precision} k{kernel}: {e.size} patches, max rel coefficient err {e.max():.3e}, "
          f"fallback {g.info('fallback'):.0f}, refine iters {g.info('refine_iters'):.0f}")
    assert e.max() < tol
    assert g.info("fallback") == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_factor_kernels_agree_all_buckets(precision):
    """Patches from ~10 to ~500 members (every tuner bucket): the vector-ALU kernel,
    the matrix-core kernel at both panel widths and the tuner's choice give the
    same interpolated values to rounding (FP64 1e-12; mixed 1e-10 coefficients
    through refinement, so the FP32 tables agree to 1e-6)."""
    rng = np.random.default_rng(cases.SEED + 91)
    X = np.concatenate([rng.random((6000, 3)), np.clip(0.5 + 0.08 * rng.standard_normal((4000, 3)), 0, 1)])
    V = cases.fourier_velocity(X, cases.fourier_modes())
    P = rng.random((3000, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[8] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=1.0, kernel_type=2, solve_method=2, lagrbf_precision=precision,
                lagrbf_mixed_factor="fp32")
    vals = {}
    for nb in (0, 16, -1):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_factor_nb=nb))
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        out = np.zeros((P.shape[0], 3)); g.output(fvals=out)
        vals[nb] = out
        print(f"nb={nb}: max patch {g.info('max_patch'):.0f}, fallback {g.info('fallback'):.0f}, "
              f"trial ms {g.info('factor_trial_ms'):.2f}")
        g.destroy()
    tol = 1e-12 if precision == "fp64" else 1e-6
    for nb in (16, -1):
        err = np.abs(vals[nb] - vals[0]).max() / np.abs(vals[0]).max()
        print(f"nb={nb} vs vector kernel: {err:.3e}")
        assert err < tol


@pytest.mark.parametrize("nf", [1, 3, 5, 8])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lean_solve_matches_column_solve(precision, nf):
    """The lean one-workgroup solve (solve_llt.hip, matrix-core factors, the number
    of fields fixed at compile time and rounded up to 1, 2, 3, 4, 6 or 8) against
    the column-oriented solve of the same factors, patches of ~10 to ~500 members:
    FP64 coefficients to 1e-13 (the two solves sum in different orders); with FP32
    factors both refine to the FP64 solution (1e-10), and against the oracle."""
    from oracle.oracle import OracleCtx
    rng = np.random.default_rng(cases.SEED + 92)
    X = np.concatenate([rng.random((6000, 3)), np.clip(0.5 + 0.08 * rng.standard_normal((4000, 3)), 0, 1)])
    V0 = cases.fourier_velocity(X, cases.fourier_modes())
    V = np.ascontiguousarray(np.concatenate([V0, V0[:, ::-1] ** 2, V0], axis=1)[:, :nf])
    P = rng.random((2000, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[8] * 3, n_interpolation_fields=nf,
                pu_mesh_overlap=1.0, kernel_type=2, solve_method=0, lagrbf_precision=precision,
                lagrbf_mixed_factor="fp32", lagrbf_factor_nb=16, lagrbf_setup_all=True, oracle_setup_all=True)
    print(f"case: lean vs column solve, {precision}, nf={nf}", flush=True)
    coefs = {}
    for lean in (True, False):
        g = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_lean_solve=lean))
        g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
        coefs[lean] = [g.debug_get("patch_coeffs", p, dtype=np.float64) for p in range(8 ** 3)]
        if lean:
            o = OracleCtx(g.options)
            o.fill_data(X, V); o.fill_points(P); o.interpolate_c()
            ref = [o.patch_coeffs(p).ravel() for p in range(8 ** 3)]
            o.destroy()
        g.destroy()
    e_col = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(coefs[True], coefs[False]) if b.size)
    e_orc = np.array([np.abs(a - b).max() / np.abs(b).max() if b.size else 0.0 for a, b in zip(coefs[True], ref)])
    # the oracle solves by Bunch-Kaufman: two backward-stable FP64 solvers agree to
    # about cond(A) u, so the bar follows the worst patch's condition number
    pw = int(np.argmax(e_orc))
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V); g.fill_points(P); g.interpolate_c()
    m = g.debug_get("patch_members", pw)
    eps = g.info("eps_r")
    g.destroy()
    r = np.sqrt(((X[m][:, None, :] - X[m][None, :, :]) ** 2).sum(-1)) * eps
    A = np.where(r < 1, (1 - r) ** 4 * (4 * r + 1), 0.0)
    cond = np.linalg.cond(A)
    print(f"{precision} nf={nf}: lean vs column {e_col:.3e}, lean vs oracle {e_orc.max():.3e} "
          f"(patch {pw}, {len(m)} members, cond {cond:.2e})")
    assert e_col < (1e-13 if precision == "fp64" else 1e-10)
    assert e_orc.max() < max(1e-12, 10 * cond * 2.2e-16)
