"""Pins the oracle's augmented patch systems (extensions with no reference
counterpart: the polyharmonic r^3 kernel, the polynomial block [A P; P^T 0] and the
per-patch shape parameter eps = c / h_local) by what the mathematics fixes:

* closed forms of phi = (eps r)^3, of the monomial basis and of h_local;
* polynomial reproduction: with a degree-k block every patch interpolant is exact on
  polynomials of degree <= k, so the PU value is f(x) * sum of the used weights;
* the interpolation condition: each corner patch of a node's cell holds the node, so
  at a data node the PU value is f_j * sum of the used weights, for any field;
* the saddle-point solution equals numpy's LAPACK solve of the same system built from
  its definition, and its residual is at rounding level.

A dropped polynomial term, a wrong sign, scale or transposed block, or a shape
parameter used differently in assembly and evaluation fails one of these."""
import numpy as np
import pytest

from tests import cases


def _ctx(oracle_mod, opts, X, V, P):
    o = oracle_mod.OracleCtx(opts)
    o.fill_data(X, V)
    o.fill_points(P)
    o.interpolate_c()
    vals = np.zeros((P.shape[0], V.shape[1]))
    xs = np.zeros_like(P)
    o.output(xs, vals)
    return o, vals, xs


def _opts(dim, n, kernel, deg, nf=1, local_eps=0.0, overlap=1.0):
    return dict(pu_mesh_dim=dim, pu_mesh_n=[n] * dim, pu_mesh_O=[0.0] * dim, pu_mesh_L=[1.0] * dim,
                n_interpolation_fields=nf, pu_mesh_overlap=overlap, pu_mesh_shape_factor=1.0, kernel_type=kernel,
                solve_method=2, lagrbf_poly_degree=deg, lagrbf_local_eps=local_eps)


def test_phs3_closed_form(oracle_mod):
    k = oracle_mod.kernel
    assert k(7, 0.25, 2.0) == 0.125
    assert k(7, 0.0, 5.0) == 0.0
    assert k(7, 3.0, 1.0) == 27.0


def test_monomials_and_degree_rule(oracle_mod):
    xi = np.array([2.0, 3.0, 5.0])
    assert np.array_equal(oracle_mod.poly_eval(3, 2, xi), [1, 2, 3, 5, 4, 6, 10, 9, 15, 25])
    assert np.array_equal(oracle_mod.poly_eval(2, 2, xi[:2]), [1, 2, 3, 4, 6, 9])
    assert np.array_equal(oracle_mod.poly_eval(3, 1, xi), [1, 2, 3, 5])
    assert [oracle_mod.poly_terms(3, d) for d in (-1, 0, 1, 2)] == [0, 1, 4, 10]
    assert [oracle_mod.poly_terms(2, d) for d in (-1, 0, 1, 2)] == [0, 1, 3, 6]
    # largest degree whose term count is at most max(1, n // 2)
    assert [oracle_mod.patch_degree(3, 2, n) for n in (0, 1, 3, 7, 8, 19, 20, 300)] == [-1, 0, 0, 0, 1, 1, 2, 2]
    assert oracle_mod.patch_degree(3, -1, 300) == -1
    assert oracle_mod.patch_degree(2, 1, 6) == 1 and oracle_mod.patch_degree(2, 1, 5) == 0


@pytest.mark.parametrize("dim,kernel,deg,local", [(2, 7, 1, 0.0), (3, 7, 1, 0.0), (2, 7, 2, 0.0), (3, 6, 2, 3.0),
                                                  (2, 2, 1, 0.0)])
def test_polynomial_reproduction(oracle_mod, dim, kernel, deg, local):
    rng = np.random.default_rng(11 + dim + deg)
    N = 1500 if dim == 2 else 4000
    X = rng.random((N, dim))
    P = rng.random((700, dim))
    c = rng.standard_normal(10)

    def field(Y):
        f = c[0] + Y @ c[1:1 + dim]
        if deg == 2:
            f = f + c[4] * Y[:, 0] ** 2 + c[5] * Y[:, 0] * Y[:, 1] + c[6] * Y[:, 1] ** 2
        return f

    V = np.stack([field(X), 2.0 - field(X)], axis=1)
    n = 6 if dim == 2 else 4
    o, vals, xs = _ctx(oracle_mod, _opts(dim, n, kernel, deg, nf=2, local_eps=local), X, V, P)
    w = o.weights().sum(axis=1)
    ex = np.stack([field(xs) * w, (2.0 - field(xs)) * w], axis=1)
    assert np.abs(vals - ex).max() < 1e-11 * max(1.0, np.abs(ex).max())


@pytest.mark.parametrize("kernel,deg,local", [(7, 1, 0.0), (7, 2, 0.0), (6, 2, 3.0), (2, 1, 0.0)])
def test_interpolation_condition_at_nodes(oracle_mod, kernel, deg, local):
    rng = np.random.default_rng(5 + deg)
    X = rng.random((3000, 3))
    V = np.sin(3 * X[:, :1]) * np.cos(2 * X[:, 1:2]) + X[:, 2:3] ** 3
    P = X[rng.permutation(3000)[:600]]          # evaluate at data nodes
    o, vals, xs = _ctx(oracle_mod, _opts(3, 5, kernel, deg, local_eps=local), X, V, P)
    w = o.weights().sum(axis=1)
    f = np.sin(3 * xs[:, :1]) * np.cos(2 * xs[:, 1:2]) + xs[:, 2:3] ** 3
    assert np.abs(vals[:, 0] - f[:, 0] * w).max() < 1e-9


def test_saddle_point_solution_against_numpy(oracle_mod):
    """The system assembled from its definition and solved by LAPACK (numpy)."""
    rng = np.random.default_rng(21)
    X = rng.random((2500, 3))
    V = np.cos(4 * X[:, :1] + X[:, 1:2])
    o, _, _ = _ctx(oracle_mod, _opts(3, 5, 7, 1), X, V, X[:50])
    eps_w = 1.0 / (1.0 / 125 ** (1.0 / 3.0))   # 1 / radius, radius = overlap * box / P^(1/d) (partition_of_unity.c:322-355)
    h = 0.25
    for p in (0, 31, 62):
        mem = o.patch_members(p)
        xm = X[mem]
        node = np.array([p % 5, (p // 5) % 5, p // 25]) * h
        r = np.linalg.norm(xm[:, None, :] - xm[None, :, :], axis=-1)
        beta, eps = o.patch_poly(p)
        assert abs(eps - eps_w) < 1e-15 * eps_w
        A = (eps * r) ** 3
        xi = (xm - node) / h
        Pm = np.hstack([np.ones((len(mem), 1)), xi])
        K = np.block([[A, Pm], [Pm.T, np.zeros((4, 4))]])
        rhs = np.concatenate([V[mem, 0], np.zeros(4)])
        sol = np.linalg.solve(K, rhs)
        cg = o.patch_coeffs(p)[:, 0]
        z = np.concatenate([cg, beta[:, 0]])
        assert np.abs(z - sol).max() <= 1e-10 * np.abs(sol).max()
        assert np.abs(K @ z - rhs).max() <= 1e-13 * np.abs(K).max() * np.abs(z).max()


def test_local_eps_closed_form(oracle_mod):
    """Exact lattice with k^3 nodes per cell: h_local = h / k for every patch."""
    n, k = 5, 3
    h = 1.0 / (n - 1)
    g = (np.arange((n - 1) * k) + 0.5) / ((n - 1) * k)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    V = X[:, :1].copy()
    o, _, _ = _ctx(oracle_mod, _opts(3, n, 6, 2, local_eps=3.0), X, V, X[:10])
    for p in (0, 2, 62, 124):                     # corner, edge, interior, corner
        _, eps = o.patch_poly(p)
        assert abs(eps - 3.0 * k / h) < 1e-12 * eps


def test_rotation_field_reproduced_c1(oracle_mod):
    """C1's solid-body rotation is linear: r^3 + linear block reproduces it exactly
    wherever the used weights sum to 1 (weight radius below one cell here)."""
    opts, X, V, P = cases.rotation_2d(N=4096, Np=2000, n=17, kernel=7, advection=False)
    opts.update(lagrbf_poly_degree=1, pu_mesh_overlap=0.9)
    o, vals, xs = _ctx(oracle_mod, opts, X, V, P)
    w = o.weights().sum(axis=1)
    assert np.abs(w - 1).max() < 1e-14
    ex = np.stack([-(xs[:, 1] - 0.5), xs[:, 0] - 0.5], axis=1)
    assert np.abs(vals - ex).max() < 1e-13


def test_corner_normalised_weights(oracle_mod):
    """-lagrbf_weights corners: the used weights sum to 1 at C1's radius (1.5 cell
    diagonals, where the reference's all-node normalisation leaves them below 1), so
    the rotation field is reproduced exactly by r^3 + linear."""
    opts, X, V, P = cases.rotation_2d(N=4096, Np=2000, n=17, kernel=7, advection=False)
    opts.update(lagrbf_poly_degree=1)
    o_all, _, _ = _ctx(oracle_mod, opts, X, V, P)
    assert o_all.weights().sum(axis=1).max() < 1 - 1e-3
    opts["lagrbf_weights"] = "corners"
    o, vals, xs = _ctx(oracle_mod, opts, X, V, P)
    w = o.weights()
    assert np.abs(w.sum(axis=1) - 1).max() < 1e-15
    assert np.allclose(w / w.sum(axis=1, keepdims=True), o_all.weights() / o_all.weights().sum(axis=1, keepdims=True),
                       rtol=1e-13, atol=0)
    ex = np.stack([-(xs[:, 1] - 0.5), xs[:, 0] - 0.5], axis=1)
    assert np.abs(vals - ex).max() < 1e-13
