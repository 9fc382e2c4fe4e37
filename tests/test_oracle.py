"""Pins the FP64 CPU oracle against what the mathematics and the reference fix:
closed forms, LAPACK (scipy), brute force on tiny grids, the reference notebook's
recorded errors, and exact RK stability polynomials on linear flows."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from tests import cases


# ---------------------------------------------------------------- kernels
def test_kernel_closed_forms(oracle_mod):
    k = oracle_mod.kernel
    eps = 2.0
    # kernels_rbf.c:31-91 evaluated by hand at eps*r = 0.5 (u = 1 - eps*r = 0.5)
    assert k(0, 0.25, eps) == 0.5                               # linear: eps*r
    assert k(1, 0.25, eps) == 0.25                              # C0: u^2
    assert k(2, 0.25, eps) == 0.5 ** 4 * (5 - 4 * 0.5)          # C2: u^4 (5 - 4u) = 0.1875
    assert k(2, 0.0, eps) == 1.0
    assert abs(k(3, 0.25, eps) - np.exp(-1.0 / (0.5 * 1.5))) < 1e-16
    assert k(4, 0.25, eps) == np.sqrt(1.25)                     # MQ
    assert abs(k(5, 0.25, eps) - 1 / np.sqrt(1.25)) < 1e-16     # IMQ
    assert k(6, 0.25, eps) == np.exp(-0.25)                     # Gaussian
    for t in (1, 2, 3):                                         # compact support
        assert k(t, 0.5 + 1e-12, eps) == 0.0
    # Wendland C2 = (1-x)^4 (4x+1) in the x = eps*r form
    x = np.linspace(0, 1, 11)
    assert np.allclose([k(2, xi, 1.0) for xi in x], (1 - x) ** 4 * (4 * x + 1), rtol=0, atol=1e-15)


# ---------------------------------------------------------------- LAPACK
@pytest.mark.parametrize("n,kind,seed", [(5, "indef", 0), (17, "indef", 1), (40, "indef", 2), (33, "spd", 3),
                                         (25, "rbf_linear", 4), (30, "rbf_mq", 5)])
def test_sytrf_matches_lapack(oracle_mod, n, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "indef":
        A = rng.standard_normal((n, n)); A = A + A.T
    elif kind == "spd":
        B = rng.standard_normal((n, n)); A = B @ B.T + n * np.eye(n)
    else:
        X = rng.random((n, 2))
        r = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=-1)
        A = 3.0 * r if kind == "rbf_linear" else np.sqrt(1 + (3.0 * r) ** 2)
    LD, ipiv, info = oracle_mod.sytrf(A)
    ref, ref_ipiv, ref_info = lapack.dsytrf(A, lower=1)       # n < block size: dsytf2 path
    assert info == ref_info == 0
    assert np.array_equal(ipiv, ref_ipiv)
    tril = np.tril_indices(n)
    assert np.allclose(LD[tril], ref[tril], rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    b = rng.standard_normal((n, 3))
    x = oracle_mod.sytrs(LD, ipiv, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-9 * np.abs(x).max())


def test_potrf_matches_numpy(oracle_mod):
    rng = np.random.default_rng(7)
    B = rng.standard_normal((20, 20)); A = B @ B.T + 20 * np.eye(20)
    L, info = oracle_mod.potrf(A)
    assert info == 0
    assert np.allclose(np.tril(L), np.linalg.cholesky(A), rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- binning
def test_locate_brute_force(oracle_mod):
    rng = np.random.default_rng(11)
    n = [5, 4, 6]; O = [0.0, -1.0, 2.0]; Lm = [1.0, 1.0, 3.0]
    pts = np.column_stack([rng.uniform(O[d], Lm[d], 500) for d in range(3)])
    # exact node planes and the upper faces
    nodes = [np.linspace(O[d], Lm[d], n[d]) for d in range(3)]
    extra = np.array([[nodes[0][i], nodes[1][1], nodes[2][j]] for i in range(n[0]) for j in range(n[2])])
    pts = np.vstack([pts, extra, [[Lm[0], Lm[1], Lm[2]], [O[0], O[1], O[2]], [1.5, 0.0, 2.5], [0.5, np.nan, 2.5]]])
    got = oracle_mod.locate(3, n, O, Lm, pts)
    h = [(Lm[d] - O[d]) / (n[d] - 1) for d in range(3)]
    for p, c in zip(pts, got):
        inside = all(O[d] <= p[d] <= Lm[d] for d in range(3))
        if not inside:
            assert c == -1
            continue
        ijk = []
        for d in range(3):
            # brute force: the last cell whose lower node is <= x, the top face in the last cell
            lo = [i for i in range(n[d] - 1) if O[d] + i * h[d] <= p[d]]
            i = max(lo)
            if (p[d] - O[d]) / h[d] < i:   # floating tie: x just below node i
                i -= 1
            ijk.append(i)
        assert c == ijk[0] + (n[0] - 1) * (ijk[1] + (n[1] - 1) * ijk[2])


def test_cell_sort_stable(oracle_mod):
    rng = np.random.default_rng(5)
    cell = rng.integers(0, 37, 2000).astype(np.int32)
    cs, cn = oracle_mod.cell_sort(cell, 37)
    assert np.array_equal(cn, np.argsort(cell, kind="stable"))
    assert np.array_equal(cs, np.concatenate([[0], np.cumsum(np.bincount(cell, minlength=37))]))


def test_patch_membership_definition(oracle_mod):
    opts, X, V, P = cases.fourier_3d(nlat=12, Np=200, n=4)
    o = oracle_mod.OracleCtx(opts)
    o.fill_data(X, V); o.fill_points(P)
    cell = o.cellids(True)
    n = 4; nc = n - 1
    for p in range(n ** 3):
        a, b, c = p % n, (p // n) % n, p // (n * n)
        touching = sorted(i + nc * (j + nc * k) for k in (c - 1, c) for j in (b - 1, b) for i in (a - 1, a)
                          if 0 <= i < nc and 0 <= j < nc and 0 <= k < nc)
        expect = np.concatenate([np.nonzero(cell == q)[0] for q in touching]).astype(np.int32)
        assert np.array_equal(o.patch_members(p), expect)


# ---------------------------------------------------------------- weights
def test_weights_stencil_equals_brute_force(oracle_mod):
    opts, X, V, P = cases.rotation_2d(N=300, Np=800, n=9, advection=False)
    a = oracle_mod.OracleCtx(opts)
    b = oracle_mod.OracleCtx(dict(opts, oracle_weights_brute=True))
    a.fill_points(P); b.fill_points(P)
    assert np.array_equal(a.weights(), b.weights())     # omitted nodes add exact zeros


def test_weights_partition_of_unity_small_support(oracle_mod):
    # weight radius <= cell size: only the 2^d corners can be non-zero, so sum_k w_k = 1
    opts, X, V, P = cases.fourier_3d(nlat=8, Np=500, n=6, overlap=0.9)
    o = oracle_mod.OracleCtx(opts)
    o.fill_points(P)
    w = o.weights()
    s = w.sum(axis=1)
    # R = 0.15 < cell half-diagonal: points far from every corner get sum_w = 0 => w = 0
    eps_w, _, nodes = o.pu()
    far = np.array([np.min(np.linalg.norm(nodes - p, axis=1)) >= 1 / eps_w for p in P])
    assert far.any() and (~far).any()
    assert np.all(s[far] == 0.0)
    assert np.allclose(s[~far], 1.0, atol=1e-14)


def test_weights_hand_value(oracle_mod):
    # 2x2 nodes, one cell: w_k = W(|x - c_k|) / sum_k W, W = Wendland C2 with eps_w
    opts = dict(pu_mesh_dim=2, pu_mesh_n=[2, 2], pu_mesh_O=[0, 0], pu_mesh_L=[1, 1], pu_mesh_overlap=2.0)
    o = oracle_mod.OracleCtx(opts)
    eps_w, _, nodes = o.pu()
    assert eps_w == 2 ** 0.5 / (2.0 * 1.0) * 2 / 2 or abs(eps_w - (4 ** 0.5) / 2.0) < 1e-15
    x = np.array([[0.3, 0.6]])
    o.fill_points(x)
    c = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=float)       # Q1 order
    r = np.linalg.norm(c - x, axis=1) * eps_w
    W = np.where(r < 1, (1 - r) ** 4 * (4 * r + 1), 0.0)
    assert np.allclose(o.weights()[0], W / W.sum(), rtol=1e-14)


# ---------------------------------------------------------------- evaluation
def test_notebook_known_answer(oracle_mod):
    opts, X, V, P = cases.notebook_2d()
    o = oracle_mod.OracleCtx(dict(opts, advection=False))
    o.fill_data(X, V); o.fill_points(P); o.interpolate_c()
    xs = np.zeros_like(P); vals = np.zeros((P.shape[0], 1)); pid = np.zeros(P.shape[0], dtype=np.int64)
    o.output(xs, vals, pid)
    an = np.exp(-np.sum(xs ** 2, axis=1) / (2 * 0.5 ** 2))
    d = np.abs(vals[:, 0] - an)
    assert f"{np.sqrt(np.mean(d ** 2)):.6e}" == f"{cases.NOTEBOOK_L2:.6e}"      # interpolation_2d.ipynb:503
    assert f"{d.max():.6e}" == f"{cases.NOTEBOOK_LINF:.6e}"
    assert np.array_equal(pid, np.arange(1, P.shape[0] + 1))


@pytest.mark.parametrize("kernel", [0, 2, 4, 5, 6])
def test_single_cell_reproduces_data(oracle_mod, kernel):
    # one PU cell: every patch holds every node and the weights sum to 1, so the PU
    # interpolant is the global RBF interpolant and reproduces the data at the nodes
    rng = np.random.default_rng(9)
    X = rng.random((60, 2)); F = np.column_stack([np.sin(3 * X[:, 0]), X[:, 1] ** 2])
    opts = dict(pu_mesh_dim=2, pu_mesh_n=[2, 2], pu_mesh_O=[0, 0], pu_mesh_L=[1, 1], pu_mesh_overlap=3.0,
                pu_mesh_shape_factor=1.0 if kernel in (0, 2) else 3.0, kernel_type=kernel, solve_method=0,
                n_interpolation_fields=2)
    o = oracle_mod.OracleCtx(opts)
    o.fill_data(X, F); o.fill_points(X); o.interpolate_c()
    vals = np.zeros((60, 2)); o.output(fvals=vals)
    assert np.allclose(vals, F, rtol=0, atol=1e-8)


# ---------------------------------------------------------------- advection
def _rot(x):
    return np.stack([-(x[:, 1] - 0.5), x[:, 0] - 0.5], axis=1)


@pytest.mark.parametrize("scheme,poly", [(0, [1, 1]), (1, [1, 1, 1 / 2]), (2, [1, 1, 1 / 2, 1 / 6, 1 / 24]),
                                         (3, [1, 1, 1 / 2, 1 / 6, 1 / 24, 1 / 120, 1 / 600])])
def test_rk_stability_polynomial_on_rotation(oracle_mod, scheme, poly):
    """On x' = A (x - c), one step multiplies x - c by R(hA) exactly: the scheme's
    stability polynomial (DOPRI5: 1 + z + ... + z^5/120 + z^6/600)."""
    P = np.array([[0.7, 0.55], [0.3, 0.2], [0.5, 0.9]])
    opts = dict(pu_mesh_dim=2, pu_mesh_n=[3, 3], advection=True, advection_scheme=scheme, cfl_fac=1e9, dt=0.1)
    o = oracle_mod.OracleCtx(opts)
    o.fill_points(P)
    o.set_velocity_function(_rot)
    h = 0.1
    o.advect_analytical_c(0.0, h)
    xs = np.zeros_like(P); o.output(coords=xs)
    A = np.array([[0.0, -1.0], [1.0, 0.0]]) * h
    R = sum(c * np.linalg.matrix_power(A, k) for k, c in enumerate(poly))
    expect = 0.5 + (P - 0.5) @ R.T
    assert np.allclose(xs, expect, rtol=0, atol=2e-16 * 8)


def test_cfl_dt_closed_form(oracle_mod):
    P = np.array([[0.9, 0.5], [0.5, 0.6]])
    opts = dict(pu_mesh_dim=2, pu_mesh_n=[5, 5], pu_mesh_overlap=1.3, advection=True, advection_scheme=2, cfl_fac=0.4)
    o = oracle_mod.OracleCtx(opts)
    o.fill_points(P); o.set_velocity_function(_rot)
    o.advect_analytical_c(0.0, 10.0)
    eps_w, _, _ = o.pu()
    # the first step's dt: cfl * (1/eps_w) / (max |v| + 1e-16), max |v| = 0.4
    o2 = oracle_mod.OracleCtx(opts); o2.fill_points(P); o2.set_velocity_function(_rot)
    o2.advect_analytical_c(0.0, 1e-30 + 0.4 * (1 / eps_w) / (0.4 + 1e-16))
    assert abs(o2.dt() - 0.4 * (1.0 / eps_w) / (0.4 + 1e-16)) < 1e-15


def test_particles_leaving_are_removed(oracle_mod):
    P = np.array([[0.95, 0.5], [0.5, 0.5]])
    opts = dict(pu_mesh_dim=2, pu_mesh_n=[5, 5], advection=True, advection_scheme=0, cfl_fac=1e9, dt=0.2)
    o = oracle_mod.OracleCtx(opts)
    o.fill_points(P)
    o.set_velocity_function(lambda x: np.tile([1.0, 0.0], (x.shape[0], 1)))
    o.advect_analytical_c(0.0, 0.2)
    assert o.removed_count() == 1
    xs = np.zeros_like(P); pid = np.zeros(2, dtype=np.int64)
    o.output(coords=xs, pidx=pid)
    assert pid[0] == 0 and np.all(xs[0] == 0) and pid[1] == 2 and abs(xs[1, 0] - 0.7) < 1e-15


def test_rotation_revolution_c1_small(oracle_mod):
    """C1-like 2D rotation (coarser) with reference semantics: RK4 dt=0.01, 628 steps;
    error vs the exact rotation by 6.28 rad stays small but not zero (Σw_used < 1 bias)."""
    opts, X, V, P = cases.rotation_2d(N=1024, Np=50, n=9)
    o = oracle_mod.OracleCtx(opts)
    o.fill_data(X, V); o.fill_points(P); o.interpolate_c()
    o.advect_steps_c(100, 0.01)
    xs = np.zeros_like(P); o.output(coords=xs)
    c, s = np.cos(1.0), np.sin(1.0)
    ex = np.stack([0.5 + c * (P[:, 0] - 0.5) - s * (P[:, 1] - 0.5), 0.5 + s * (P[:, 0] - 0.5) + c * (P[:, 1] - 0.5)], axis=1)
    # radius about the centre is kept (the field is interpolated well in direction) ...
    ratio = np.linalg.norm(xs - 0.5, axis=1) / np.linalg.norm(P - 0.5, axis=1)
    assert np.max(np.abs(ratio - 1)) < 0.01
    # ... but with r_w ~ 2.1 h the corner weights sum to < 1 (normalised over all PU
    # nodes, swarm_interp.c:327-372), so the speed is low and the particles lag
    assert np.all(o.weights().sum(axis=1) < 0.999)
    assert 0.01 < np.max(np.abs(xs - ex)) < 0.2
