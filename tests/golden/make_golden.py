"""Writes the golden fixtures of tests/golden/ (run from the repository root:
python tests/golden/make_golden.py).

  reference_notebook.json   the known answer the reference ships: L2 / Linf error of
                            the 2D Gaussian interpolation printed in
                            python_api/notebook/interpolation_2d.ipynb:503, with the
                            inputs' description (rebuilt by tests/cases.notebook_2d).
  lapack_sytrf.npz          LAPACK dsytrf / dsytrs (lower, Bunch-Kaufman) outputs from
                            scipy.linalg.lapack on seeded symmetric matrices (SPD,
                            indefinite with 2x2 pivots, RBF linear / multiquadric
                            matrices): the factorisation the reference reaches through
                            PETSc MatCholeskyFactor on SeqDense (partition_factor.c:37-55).
  oracle_regression.npz     oracle outputs on small seeded cases (binning, Shepard
                            weights, interpolation per kernel, RK1/2/4/DOPRI positions):
                            regression vectors that pin the oracle between changes and
                            give the GPU tests a check that does not rerun the oracle.

Every fixture records its seed; arrays are plain numpy (no pickles).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SYTRF_CASES = [(1, "spd", 1), (2, "indef", 2), (5, "indef", 3), (17, "spd", 4), (24, "indef", 5),
               (40, "rbf_linear", 6), (33, "rbf_mq", 7), (64, "indef", 8)]


def sytrf_matrix(n, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "indef":
        A = rng.standard_normal((n, n))
        A = A + A.T
    elif kind == "spd":
        B = rng.standard_normal((n, n))
        A = B @ B.T + n * np.eye(n)
    else:
        X = rng.random((n, 2))
        r = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=-1)
        A = 3.0 * r if kind == "rbf_linear" else np.sqrt(1 + (3.0 * r) ** 2)
    b = rng.standard_normal((n, 3))
    return A, b


def make_sytrf():
    from scipy.linalg import lapack
    out = {}
    for idx, (n, kind, seed) in enumerate(SYTRF_CASES):
        A, b = sytrf_matrix(n, kind, seed)
        ld, ipiv, info = lapack.dsytrf(A, lower=1)
        assert info == 0
        x, info2 = lapack.dsytrs(ld, ipiv, b, lower=1)
        assert info2 == 0
        out[f"A{idx}"] = A
        out[f"b{idx}"] = b
        out[f"ld{idx}"] = np.tril(ld)
        out[f"ipiv{idx}"] = ipiv.astype(np.int32)
        out[f"x{idx}"] = x
    out["cases"] = np.array([[n, ["spd", "indef", "rbf_linear", "rbf_mq"].index(k), s] for n, k, s in SYTRF_CASES],
                            dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "lapack_sytrf.npz"), **out)


def make_notebook():
    from tests import cases
    opts, X, V, P = cases.notebook_2d()
    d = {"source": "python_api/notebook/interpolation_2d.ipynb:503 (printed output of the reference)",
         "L2": cases.NOTEBOOK_L2, "Linf": cases.NOTEBOOK_LINF, "digits": 7,
         "inputs": "32x32 structured nodes on [-1,1]^2 (ravel order F), Gaussian sigma=0.5; 64^2 Halton "
                   "targets (bases 2,3, indices 1..4096) on [-1,1]^2; PU 4x4 on [-1.02,1.02]^2, overlap 1.1, "
                   "shape factor 1.0, Wendland C2, solve_method 2",
         "n_nodes": int(X.shape[0]), "n_targets": int(P.shape[0])}
    with open(os.path.join(HERE, "reference_notebook.json"), "w") as f:
        json.dump(d, f, indent=1)


REGRESSION_SEED = 20261203


def regression_cases():
    """(name, options, X, V, P, action) of the regression fixtures; action is
    ("interp",) or ("steps", nsteps, dt)."""
    from tests import cases
    out = []
    for k, sf in ((0, 1.0), (1, 1.0), (2, 1.0), (3, 6.0), (4, 10.0), (5, 10.0), (6, 10.0)):
        opts, X, V, P = cases.rotation_2d(N=600, Np=200, n=7, kernel=k, advection=False)
        opts["pu_mesh_shape_factor"] = sf
        out.append((f"interp2d_k{k}", opts, X, V, P, ("interp",)))
    opts, X, V, P = cases.fourier_3d(nlat=14, Np=300, n=5, solve_method=0)
    out.append(("interp3d_wc2", opts, X, V, P, ("interp",)))
    for scheme in (0, 1, 2, 3):
        opts, X, V, P = cases.rotation_2d(N=800, Np=300, n=7, scheme=scheme)
        out.append((f"rk2d_s{scheme}", opts, X, V, P, ("steps", 5, 0.02)))
    opts, X, V, P = cases.fourier_3d(nlat=14, Np=400, n=5, scheme=2)
    out.append(("rk3d_s2", opts, X, V, P, ("steps", 6, 5e-3)))
    return out


def run_oracle(opts, X, V, P, action):
    from oracle.oracle import OracleCtx
    o = OracleCtx(opts)
    o.fill_data(X, V)
    o.fill_points(P)
    o.interpolate_c()
    if action[0] == "steps":
        o.advect_steps_c(action[1], action[2])
    nf = o.nfields
    xs = np.zeros_like(P)
    vals = np.zeros((P.shape[0], nf))
    o.output(xs, vals)
    res = {"coords": xs, "values": vals, "weights": o.weights(), "cellid_points": o.cellids(False),
           "cellid_data": o.cellids(True)}
    cs, cn = o.cell_lists(X.shape[0])
    res["cell_start"], res["cell_nodes"] = cs, cn
    res["removed"] = np.array([o.removed_count()], dtype=np.int64)
    o.destroy()
    return res


def make_regression():
    out = {"seed": np.array([REGRESSION_SEED], dtype=np.int64)}
    for name, opts, X, V, P, action in regression_cases():
        res = run_oracle(opts, X, V, P, action)
        for k, v in res.items():
            out[f"{name}/{k}"] = v
    np.savez_compressed(os.path.join(HERE, "oracle_regression.npz"), **out)


if __name__ == "__main__":
    from oracle import oracle
    oracle.build()
    make_sytrf()
    make_notebook()
    make_regression()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
