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


def clustered_c3(n_uniform=1000000, n_clusters=64, per_cluster=15625, sigma=(0.01, 0.05), Np=8388608, n=49,
                 kernel=6, deg=2, local_eps=3.0, seed_off=0):
    """Uneven-batching configuration: uniform nodes in [0,1]^3 plus Gaussian clusters
    (sigma ~ U[sigma], centres ~ U[0.1,0.9]^3), the 16-mode Fourier velocity, PU
    n^3 nodes ((n-1)^3 cells), Np uniform particles, RK2.  As written: the Gaussian
    kernel with eps = local_eps / h_local per patch (h_local = (volume of the
    patch's cells / members)^(1/3)) and a quadratic polynomial block; kernel=2,
    deg=-1, local_eps=0 gives the Wendland C2 variant."""
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
                pu_mesh_overlap=1.0, pu_mesh_shape_factor=1.0, kernel_type=kernel, solve_method=0,
                advection=True, advection_scheme=1, cfl_fac=1e6, dt=1e-3)
    if deg >= 0:
        opts["lagrbf_poly_degree"] = deg
    if local_eps > 0:
        opts["lagrbf_local_eps"] = local_eps
    return opts, X, np.ascontiguousarray(V), P


def slab_velocity(X, xs, modes, U0=0.5, w=0.3, noise=0.05):
    """Nondimensional corner-flow-like sinking slab in the x-z plane of a 4:1:1 box
    plus small divergence-free Fourier noise: stream function
    psi(x, z) = U0 sin(pi z) (x - xs) exp(-((x - xs)/w)^2), v = (dpsi/dz, 0, -dpsi/dx)
    (downwelling U0 sin(pi z) at x = xs, no normal flow through z = 0, 1)."""
    u = X[:, 0] - xs
    g = u * np.exp(-(u / w) ** 2)
    dg = (1.0 - 2.0 * (u / w) ** 2) * np.exp(-(u / w) ** 2)
    z = X[:, 2]
    V = np.zeros_like(X)
    V[:, 0] = U0 * np.pi * np.cos(np.pi * z) * g
    V[:, 2] = -U0 * np.sin(np.pi * z) * dg
    return V + noise * fourier_velocity(X / np.array([4.0, 1.0, 1.0]), modes)


def slab_c4(m=126, N=8000000, Np=16777216, pu=38, seed_off=0):
    """Adaptive DP45 configuration: a 4m x m x m lattice in [0,4] x [0,1] x [0,1]
    (nodes at (i+1/2)h, jittered +-0.3 h), the first N of a seeded permutation;
    velocity snapshots A (slab at x = 1.5, noise seed 1) and B (slab at x = 2.0,
    noise seed 2), blended linearly between t = 0 and 1 (fields 0-2 and 3-5); PU
    (4 pu + 1) x (pu + 1)^2 nodes, weight radius 0.97 cell (overlap scaled by the
    PU node count, as the reference's radius = overlap / P^(1/d) assumes a unit
    box); Wendland C2; Np particles uniform in the box."""
    rng = np.random.default_rng(SEED + 3 + seed_off)
    h = 1.0 / m
    gx = (np.arange(4 * m) + 0.5) * h
    g = (np.arange(m) + 0.5) * h
    X = np.stack(np.meshgrid(gx, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    X = X + rng.uniform(-0.3, 0.3, size=X.shape) * h
    X = np.clip(X, 0.0, [4.0, 1.0, 1.0])
    if N is not None and N < X.shape[0]:
        X = X[rng.permutation(X.shape[0])[:N]]
    X = np.ascontiguousarray(X)
    VA = slab_velocity(X, 1.5, fourier_modes(seed=SEED + 11))
    VB = slab_velocity(X, 2.0, fourier_modes(seed=SEED + 12))
    P = rng.random((Np, 3)) * np.array([4.0, 1.0, 1.0])
    npu = np.array([4 * pu + 1, pu + 1, pu + 1])
    hc = 1.0 / pu
    overlap = 0.97 * hc * float(np.prod(npu)) ** (1.0 / 3.0)
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[0.0] * 3, pu_mesh_L=[4.0, 1.0, 1.0], pu_mesh_n=npu.tolist(),
                n_interpolation_fields=6, pu_mesh_overlap=overlap, pu_mesh_shape_factor=1.0, kernel_type=2,
                solve_method=0, advection=True, advection_scheme=3, dt=1e-3, dopri_rtol=1e-6, dopri_atol=1e-9,
                lagrbf_time_blend=True, lagrbf_blend_times="0.0,1.0", lagrbf_setup_all=True, oracle_setup_all=True)
    return opts, X, np.ascontiguousarray(np.concatenate([VA, VB], axis=1)), P
