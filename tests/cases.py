"""Seeded input generators shared by the tests, the smoke check and the benchmark.

Each case returns (options, data_coords, data_values, particle_coords).  Seeds are
20261203 + case index; numpy's PCG64 generator.
"""
import numpy as np

SEED = 20261203


def notebook_2d():
    """The reference notebook case (python_api/notebook/interpolation_2d.ipynb:96-125,
    238-244, 390-396): Gaussian sigma=0.5 on 32x32 structured nodes in [-1,1]^2,
    64^2 Halton targets, PU 4x4 on [-1.02,1.02]^2, Wendland C2, overlap 1.1,
    shape factor 1 (the notebook's shape_fac keyword is ignored), solve_method 2."""
    from lagrbf_amd.utils import structured_points, halton_points
    O = np.array([-1.0, -1.0]); E = np.array([1.0, 1.0])
    data = structured_points(O, E, [32, 32])
    g = np.exp(-np.sum(data ** 2, axis=1) / (2 * 0.5 ** 2))
    pts = halton_points(O, E, 64 ** 2, bases=(2, 3, 5))
    opts = dict(pu_mesh_dim=2, pu_mesh_O=O - 0.01 * (E - O), pu_mesh_L=E + 0.01 * (E - O), pu_mesh_n=[4, 4],
                n_interpolation_fields=1, pu_mesh_overlap=1.1, pu_mesh_shape_factor=1.0, kernel_type=2, solve_method=2)
    return opts, data, g[:, None].copy(), pts


NOTEBOOK_L2 = 6.585958e-05      # interpolation_2d.ipynb:503
NOTEBOOK_LINF = 3.538513e-04


def rotation_2d(N=4096, Np=10000, n=17, kernel=2, scheme=2, seed_off=0, solve_method=2, dt=0.01, advection=True):
    """2D solid-body rotation v = (-(y-1/2), x-1/2): N uniform nodes in [0,1]^2,
    Np particles uniform in the disk r < 0.4 about (1/2,1/2), PU n x n on [0,1]^2
    with weight radius 1.5 x cell diagonal."""
    rng = np.random.default_rng(SEED + 0 + seed_off)
    X = rng.random((N, 2))
    V = np.stack([-(X[:, 1] - 0.5), X[:, 0] - 0.5], axis=1)
    r = 0.4 * np.sqrt(rng.random(Np)); th = 2 * np.pi * rng.random(Np)
    P = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)], axis=1)
    overlap = 1.5 * np.sqrt(2.0) * n / (n - 1)
    opts = dict(pu_mesh_dim=2, pu_mesh_O=[0.0, 0.0], pu_mesh_L=[1.0, 1.0], pu_mesh_n=[n, n], n_interpolation_fields=2,
                pu_mesh_overlap=overlap, pu_mesh_shape_factor=1.0, kernel_type=kernel, solve_method=solve_method,
                advection=advection, advection_scheme=scheme, cfl_fac=1e6, dt=dt)
    return opts, X, V, P


def fourier_modes(nmodes=16, kmax=4, seed=SEED + 1):
    """Divergence-free Fourier field: v(x) = sum_m a_m sin(2 pi k_m.x + phi_m), a_m ⟂ k_m, |a_m| ~ N(0,1)/|k_m|^2."""
    rng = np.random.default_rng(seed)
    ks, amps, phis = [], [], []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if np.max(np.abs(k)) < 1:
            continue
        a = rng.standard_normal(3)
        a -= a.dot(k) / k.dot(k) * k
        a /= np.linalg.norm(a)
        a *= rng.standard_normal() / k.dot(k)
        ks.append(k); amps.append(a); phis.append(rng.random() * 2 * np.pi)
    return np.array(ks, dtype=np.float64), np.array(amps), np.array(phis)


def fourier_velocity(X, modes):
    ks, amps, phis = modes
    arg = 2 * np.pi * X @ ks.T + phis[None, :]
    return np.sin(arg) @ amps


def fourier_3d(nlat=102, N=None, Np=4096, n=32, kernel=2, scheme=2, solve_method=2, overlap=1.0, seed_off=0, dt=1e-3):
    """3D jittered lattice (nlat^3, +-0.3 spacing) sampling the Fourier field; the
    first N nodes of a seeded permutation; Np uniform particles; PU n^3 on [0,1]^3."""
    rng = np.random.default_rng(SEED + 1 + seed_off)
    g = (np.arange(nlat) + 0.5) / nlat
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    X = X + rng.uniform(-0.3, 0.3, size=X.shape) / nlat
    X = np.clip(X, 0.0, 1.0)
    if N is not None and N < X.shape[0]:
        X = X[rng.permutation(X.shape[0])[:N]]
    V = fourier_velocity(X, fourier_modes())
    P = rng.random((Np, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[n] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=overlap, pu_mesh_shape_factor=1.0, kernel_type=kernel, solve_method=solve_method,
                advection=True, advection_scheme=scheme, cfl_fac=1e6, dt=dt)
    return opts, np.ascontiguousarray(X), np.ascontiguousarray(V), P


def clustered_c3(n_uniform=1000000, n_clusters=64, per_cluster=15625, sigma=(0.03, 0.05), Np=8388608, n=49,
                 seed_off=0):
    """Uneven-batching configuration: uniform nodes in [0,1]^3 plus Gaussian clusters
    (sigma ~ U[sigma], centres ~ U[0.1,0.9]^3), the 16-mode Fourier velocity, PU
    n^3 nodes ((n-1)^3 cells), Wendland C2, Np uniform particles, RK2."""
    rng = np.random.default_rng(SEED + 2 + seed_off)
    parts = [rng.random((n_uniform, 3))]
    for _ in range(n_clusters):
        s = rng.uniform(*sigma)
        c = rng.uniform(0.1, 0.9, 3)
        parts.append(np.clip(c + s * rng.standard_normal((per_cluster, 3)), 0.0, 1.0))
    X = np.ascontiguousarray(np.concatenate(parts))
    V = fourier_velocity(X, fourier_modes())
    P = rng.random((Np, 3))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[1.0] * 3, pu_mesh_n=[n] * 3, n_interpolation_fields=3,
                pu_mesh_overlap=1.0, pu_mesh_shape_factor=1.0, kernel_type=2, solve_method=0,
                advection=True, advection_scheme=1, cfl_fac=1e6, dt=1e-3)
    return opts, X, np.ascontiguousarray(V), P
