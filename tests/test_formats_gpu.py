"""File paths that need a context on the GPU: SwarmOutput / SwarmLoadFromFiles
round trips, data ingestion from PETSc-binary and StagYY files into the node
swarm (against the in-memory fill and the oracle), and the advection drivers'
.vtp/PVD output read back by restart_from_pvd.

Tolerances: file contents and reloaded swarms bit-exact; interpolation of
ingested data 1e-12 relative against the in-memory path run on the same
library (bit-exact expected), 1e-10 relative against the oracle."""
import os

import numpy as np
import pytest

import lagrbf_amd as L
from lagrbf_amd import fileio, petsc_io
from tests import cases
from tests.stag_files import write_stag, yin_axes

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _interp_ctx(opts, X, V, P, pid=None):
    g = L.LagRBFCtx(petsc_options=dict(opts))
    g.fill_data(X, V)
    g.fill_points(P, pidx=pid)
    g.interpolate_c()
    return g


def test_swarm_output_files_and_reload(tmp_path):
    opts, X, V, P = cases.rotation_2d(N=2048, Np=3000, n=9, advection=False)
    print("case: swarm output rotation2d 3000 particles", flush=True)
    rng = np.random.default_rng(11)
    pid = rng.permutation(P.shape[0]).astype(np.int64) + 1
    g = _interp_ctx(opts, X, V, P, pid)
    n = P.shape[0]
    vals = np.zeros((n, 2)); w = np.zeros((n, 4))
    g.output(fvals=vals, pidx=np.zeros(n, dtype=np.int64))
    g.output(fvals=w, fieldname="weights")
    root = str(tmp_path / "out")
    g.swarm_output(root, 3, 0.5, ["fields", "DMSwarm_pid", "weights"])
    g.swarm_output(root, 4, 0.75, ["fields"], direction=1, permute=True)
    base = os.path.join(root, "step3", "step000003.swarm_interp_")
    xs = petsc_io.PetscVecLoad(base + "DMSwarmPIC_coor.pbvec")
    np.testing.assert_array_equal(xs, P)                       # live particles in fill order
    fp = petsc_io.PetscISLoadFromFile(base + "DMSwarm_pid.pbis")
    np.testing.assert_array_equal(fp, pid)
    np.testing.assert_array_equal(petsc_io.PetscVecLoad(base + "fields.pbvec"), vals[pid - 1])
    np.testing.assert_array_equal(petsc_io.PetscVecLoad(base + "weights.pbvec"), w[pid - 1])
    # permute: the row-major field read as if component-major (swarm_output.c:28-36)
    perm = petsc_io.PetscVecLoadFromFile(os.path.join(root, "step4", "step000004_forward.swarm_interp_fields.pbvec"))
    np.testing.assert_array_equal(perm.reshape(n, 2), vals[pid - 1].reshape(-1).reshape(2, n).T)
    pvd = open(os.path.join(root, "timeseries_X.pvd")).read()
    assert '  <DataSet timestep="5.000000e-01" file="step3/step000003_coords.pbvec"/>\n' in pvd
    assert '  <DataSet timestep="7.500000e-01" file="step4/step000004_coords.pbvec"/>\n' in pvd
    # SwarmLoadFromFiles into a new context: same swarm, same interpolation
    h = L.LagRBFCtx(petsc_options=dict(opts))
    h.fill_data(X, V)
    h.swarm_load(os.path.join(root, "step3"), 3, ["fields", "DMSwarm_pid"])
    h.interpolate_c()
    xs2, v2, p2 = np.zeros_like(P), np.zeros((n, 2)), np.zeros(n, dtype=np.int64)
    h.output(xs2, v2, p2)
    xs1 = np.zeros_like(P)
    g.output(coords=xs1)
    np.testing.assert_array_equal(xs2, xs1)
    np.testing.assert_array_equal(p2, np.arange(1, n + 1))
    np.testing.assert_array_equal(v2, vals)


def test_load_data_petsc_files_matches_fill(tmp_path):
    opts, X, V, P = cases.fourier_3d(nlat=16, Np=2000, n=5)
    opts["advection"] = False
    print("case: PETSc-binary data load fourier3d", flush=True)
    d = tmp_path / "data"
    d.mkdir()
    petsc_io.PetscVecWrite(str(d / "run_xyz00007.pbvec"), X)
    petsc_io.PetscVecWrite(str(d / "run_vel00007.pbvec"), V)
    ref = _interp_ctx(opts, X, V, P)
    want = np.zeros((P.shape[0], 3))
    ref.output(fvals=want)
    o = dict(opts, data_format=1, swarm_data_fname_nfiles=1, swarm_data_fname_fields="vel", swarm_data_fname_prefix="run",
             swarm_data_fname_separator="_", swarm_data_dirname_base=str(d), swarm_data_fname_coordinates="xyz")
    g = L.LagRBFCtx(petsc_options=o)
    g.load_data(7)
    g.fill_points(P)
    g.interpolate_c()
    got = np.zeros_like(want)
    g.output(fvals=got)
    assert rel_err(got, want) < 1e-12
    np.testing.assert_array_equal(got, want)


def test_stag_dataset_interpolation(tmp_path):
    from oracle.oracle import OracleCtx
    print("case: StagYY vp+T files -> node swarm, 3D shell", flush=True)
    n0, n1, n2 = 8, 16, 4
    x, y, z = yin_axes(n0, n1, n2)
    lat = (0.25 * np.pi - x.astype(np.float64))[None, None, :]
    lon = (y.astype(np.float64) - 0.75 * np.pi)[None, :, None]
    r = (0.55 + z.astype(np.float64))[:, None, None]
    vp = np.zeros((2, n2, n1, n0, 4), dtype=np.float32)
    vp[..., 0] = (np.sin(2 * lon) * np.cos(lat) * np.ones_like(r))[None]
    vp[..., 1] = (np.cos(lon) * r * np.ones_like(lat))[None]
    vp[..., 2] = (r * np.sin(lat) * np.ones_like(lon))[None]
    vp[..., 3] = (r ** 2 * np.ones_like(lat) * np.ones_like(lon))[None]
    T = (np.cos(3 * lat) * r * np.ones_like(lon)).astype(np.float32)[None, ..., None].repeat(2, axis=0)
    fv, ft = str(tmp_path / "vp"), str(tmp_path / "t")
    write_stag(fv, vp, x, y, z, npar=(2, 2, 1, 1))
    write_stag(ft, T, x, y, z)
    ds = fileio.StagBinSwarmSetFromFiles([ft, fv], [0, 1], 5)
    Xn, Fn = ds.arrays()
    rng = np.random.default_rng(12)
    u = rng.standard_normal((1500, 3))
    P = u / np.linalg.norm(u, axis=1)[:, None] * rng.uniform(0.6, 0.95, (1500, 1))
    opts = dict(pu_mesh_dim=3, pu_mesh_O=[-1.05] * 3, pu_mesh_L=[1.05] * 3, pu_mesh_n=[6] * 3,
                n_interpolation_fields=5, pu_mesh_overlap=1.0, pu_mesh_shape_factor=1.0, kernel_type=2,
                solve_method=2, advection=False)
    out = []
    for use_ds in (True, False):
        g = L.LagRBFCtx(petsc_options=dict(opts))
        if use_ds:
            g.fill_from_dataset(ds, is_data=True)
        else:
            g.fill_data(Xn, Fn)
        g.fill_points(P)
        g.interpolate_c()
        v = np.zeros((P.shape[0], 5))
        g.output(fvals=v)
        out.append(v)
    ds.free()
    np.testing.assert_array_equal(out[0], out[1])
    o = OracleCtx(L.LagRBFCtx(petsc_options=dict(opts)).options)
    o.fill_data(Xn, Fn)
    o.fill_points(P)
    o.interpolate_c()
    vo = np.zeros((P.shape[0], 5))
    o.output(None, vo, None)
    err = rel_err(out[0], vo)
    print(f"stag dataset: max rel err vs oracle {err:.3e}", flush=True)
    assert err < 1e-10


def test_advection_driver_output_and_restart(tmp_path):
    opts, X, V, P = cases.rotation_2d(N=10, Np=800, n=9, scheme=2)
    print("case: AdvectionAnalytical vtp/pvd output + restart_from_pvd", flush=True)
    from lagrbf_amd.utils import PointContainer, numpy_velocity_callback
    rot = lambda x: np.stack([-(x[:, 1] - 0.5), x[:, 0] - 0.5], axis=1)
    a = L.AdvectionAnalytical(petsc_options=dict(opts), output_path=str(tmp_path), time_start=0.0, time_end=0.2,
                              nsteps=3, advection_cfl=0.3)
    a.set_advection_function(numpy_velocity_callback(rot))
    n = P.shape[0]
    out = PointContainer(coords=np.zeros((n, 2)), values=np.zeros((n, 2)), pidx=np.zeros(n, dtype=np.int64))
    a.advect(PointContainer(coords=P), out, None, prefix="_r")
    assert sorted(os.listdir(tmp_path)) == ["step00001_r.vtp", "step00002_r.vtp", "timeseries-advection.pvd"]
    ts = L.timeseries_process(str(tmp_path / "timeseries-advection.pvd"))
    assert ts["time"] == [f"{0.1:.6e}", f"{0.2:.6e}"] and ts["step"] == ["1", "2"]
    pts, data = L.read_vtp(str(tmp_path / "step00002_r.vtp"))
    np.testing.assert_array_equal(pts[:, :2], out.coords)
    np.testing.assert_array_equal(data["velocity"], out.values)
    np.testing.assert_array_equal(data["pidx"], out.pidx)
    b = L.AdvectionAnalytical(petsc_options=dict(opts), output_path=str(tmp_path), time_start=0.0, time_end=0.2,
                              nsteps=3, write_output=False)
    pc = L.restart_from_pvd(None, b, "timeseries-advection.pvd")
    assert b.start_idx == 1
    np.testing.assert_array_equal(pc.coords, out.coords)
    np.testing.assert_array_equal(pc.pidx, out.pidx)
