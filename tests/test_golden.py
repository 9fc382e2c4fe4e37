"""The oracle against the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py): LAPACK dsytrf/dsytrs vectors, the reference
notebook's printed known answer, and the oracle's own regression vectors."""
import json
import os

import numpy as np
import pytest

from tests import cases
from tests.golden import make_golden

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("idx", range(len(make_golden.SYTRF_CASES)))
def test_oracle_sytrf_vs_lapack_fixture(oracle_mod, idx):
    g = load("lapack_sytrf.npz")
    A, b = g[f"A{idx}"], g[f"b{idx}"]
    n = A.shape[0]
    LD, ipiv, info = oracle_mod.sytrf(A)
    assert info == 0
    assert np.array_equal(ipiv, g[f"ipiv{idx}"])                    # same pivot sequence as dsytf2
    ref = g[f"ld{idx}"]
    tril = np.tril_indices(n)
    assert np.max(np.abs(LD[tril] - ref[tril])) <= 1e-12 * np.abs(ref).max()
    x = oracle_mod.sytrs(LD, ipiv, b)
    assert np.max(np.abs(x - g[f"x{idx}"])) <= 1e-10 * np.abs(g[f"x{idx}"]).max()


def test_fixture_matrices_regenerate():
    """The fixture's inputs are the seeded matrices make_golden describes."""
    g = load("lapack_sytrf.npz")
    for idx, (n, kind, seed) in enumerate(make_golden.SYTRF_CASES):
        A, b = make_golden.sytrf_matrix(n, kind, seed)
        assert np.array_equal(A, g[f"A{idx}"]) and np.array_equal(b, g[f"b{idx}"])


def test_oracle_vs_reference_notebook(oracle_mod):
    with open(os.path.join(GOLDEN, "reference_notebook.json")) as f:
        ref = json.load(f)
    opts, X, V, P = cases.notebook_2d()
    o = oracle_mod.OracleCtx(dict(opts, advection=False))
    o.fill_data(X, V); o.fill_points(P); o.interpolate_c()
    xs = np.zeros_like(P); vals = np.zeros((P.shape[0], 1))
    o.output(xs, vals)
    d = np.abs(vals[:, 0] - np.exp(-np.sum(xs ** 2, axis=1) / (2 * 0.5 ** 2)))
    assert f"{np.sqrt(np.mean(d ** 2)):.6e}" == f"{ref['L2']:.6e}"
    assert f"{d.max():.6e}" == f"{ref['Linf']:.6e}"


REG_CASES = make_golden.regression_cases()


@pytest.mark.parametrize("case", REG_CASES, ids=[c[0] for c in REG_CASES])
def test_oracle_regression(oracle_mod, case):
    name, opts, X, V, P, action = case
    g = load("oracle_regression.npz")
    res = make_golden.run_oracle(opts, X, V, P, action)
    for k, v in res.items():
        ref = g[f"{name}/{k}"]
        if v.dtype.kind in "iu":
            assert np.array_equal(v, ref), k
        else:
            assert np.allclose(v, ref, rtol=1e-14, atol=1e-14 * max(np.abs(ref).max(), 1e-300)), k
