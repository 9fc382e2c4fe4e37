"""The matrix-core Cholesky (chol_mfma.hip) against the FP64 oracle and against the
vector-ALU kernel: coefficients per patch (16-column panels), FP64
and FP32 factors (the latter refined in FP64), pure and saddle-point systems,
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
    opts.update(lagrbf_factor_nb=nb, lagrbf_precision=precision, advection=False)
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
                pu_mesh_overlap=1.0, kernel_type=2, solve_method=2, lagrbf_precision=precision)
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
