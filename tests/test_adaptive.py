"""Per-particle adaptive Dormand-Prince 5(4) with a time-blended velocity
(lagrbf_advect_adaptive; no reference counterpart, the reference's DOPRI takes
one dt for the swarm).

CPU tests pin the oracle's restatement by closed forms: a blend of v and -v
between t = 0 and 1 moves every particle along its streamline by the parameter
s(t) = t - t^2, which returns to 0 at t = 1, so particles come back to their
start; without a blend the rotation field's flow is a rotation; particles are
independent (alone or together, bit for bit); a tighter tolerance converges.
GPU tests compare the library with the oracle per particle: at least 99% of the
trajectories agree to 1e-11 relative (rounding of the velocities only); a step
decision that flips near the acceptance threshold (the velocities differ in the
last bits, so an error estimate at 1.0 can fall either side) moves that
particle by at most the local tolerance, so every position is within rtol
relative; accepted and rejected counts within 1%."""
import numpy as np
import pytest

from tests import cases


def _rotation_case(blend, rtol=1e-11, n=400, seed=3):
    opts, X, V, _ = cases.rotation_2d(N=2048, Np=10, n=9, scheme=3)
    rng = np.random.default_rng(seed)
    r = 0.3 * np.sqrt(rng.random(n)); th = 2 * np.pi * rng.random(n)
    P = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)], axis=1)
    # weight radius below the cell size: the corner weights sum to 1 and the
    # interpolated velocity is continuous (wider radii drop weight, as the reference does)
    o = dict(opts, pu_mesh_overlap=1.0, dopri_rtol=rtol, dopri_atol=rtol * 1e-3, dt=0.01, oracle_setup_all=True,
             lagrbf_setup_all=True)
    if blend:
        o.update(n_interpolation_fields=4, lagrbf_time_blend=True, lagrbf_blend_times="0.0,1.0")
        F = np.concatenate([V, -V], axis=1)
    else:
        F = V
    return o, X, F, P


def _oracle_run(o, X, F, P, t0=0.0, t1=1.0):
    from oracle.oracle import OracleCtx
    c = OracleCtx(o)
    c.fill_data(X, F)
    c.fill_points(P)
    c.interpolate_c()
    counts = c.advect_adaptive_c(t0, t1)
    xs = np.zeros_like(P)
    c.output(coords=xs)
    c.destroy()
    return xs, counts


def test_oracle_blend_returns_to_start():
    o, X, F, P = _rotation_case(blend=True)
    xs, counts = _oracle_run(o, X, F, P)
    assert counts["active"] == 0 and counts["accepted"] >= P.shape[0]
    err = np.max(np.abs(xs - P))
    assert err < 1e-7, err


def test_oracle_flow_of_the_interpolant_matches_dop853():
    """Without a blend the trajectories are those of the interpolated field: an
    independent integrator (scipy's DOP853 at rtol 1e-12, velocity from the
    oracle's interpolation at the stage points) agrees to the tolerance."""
    from scipy.integrate import solve_ivp
    from oracle.oracle import OracleCtx
    o, X, F, P = _rotation_case(blend=False, rtol=1e-10, n=16)
    xs, counts = _oracle_run(o, X, F, P, 0.0, 0.8)
    ev = OracleCtx(o)
    ev.fill_data(X, F)

    def rhs(t, y):
        pts = y.reshape(-1, 2)
        ev.fill_points(pts)
        ev.interpolate_c()
        v = np.zeros_like(pts)
        ev.output(fvals=v)
        return v.reshape(-1)

    sol = solve_ivp(rhs, (0.0, 0.8), P.reshape(-1), method="DOP853", rtol=1e-12, atol=1e-14)
    ev.destroy()
    ref = sol.y[:, -1].reshape(-1, 2)
    assert np.max(np.abs(xs - ref)) < 1e-8
    # and the field is close to the rotation it samples (RBF without polynomial, 2048 nodes)
    c, s = np.cos(0.8), np.sin(0.8)
    ex = np.stack([0.5 + c * (P[:, 0] - 0.5) - s * (P[:, 1] - 0.5), 0.5 + s * (P[:, 0] - 0.5) + c * (P[:, 1] - 0.5)], 1)
    assert np.max(np.abs(xs - ex)) < 0.02
    # the global error follows the tolerance
    o2 = dict(o, dopri_rtol=1e-13, dopri_atol=1e-16)
    xs2, counts2 = _oracle_run(o2, X, F, P, 0.0, 0.8)
    assert counts2["accepted"] > counts["accepted"]
    assert np.max(np.abs(xs - xs2)) < 1e-8


def test_oracle_particles_independent():
    o, X, F, P = _rotation_case(blend=True, n=60)
    xs, _ = _oracle_run(o, X, F, P)
    for i in (0, 17, 59):
        one, _ = _oracle_run(o, X, F, P[i:i + 1])
        assert np.array_equal(one[0], xs[i])


def _wall_case(dt):
    """The rotation field with four particles: three on circles about the centre that
    stay inside the unit square (radii 0.49, 0.485, 0.3, two of them next to a
    wall) and one whose circle crosses the top wall (radius 0.59); first step
    size dt."""
    opts, X, F, _ = _rotation_case(blend=False, rtol=1e-9, n=4)
    P = np.array([[0.99, 0.5], [0.5, 0.015], [0.97, 0.9], [0.2, 0.5]])
    return dict(opts, dt=dt), X, F, P


def test_oracle_wall_overshoot_rejected_not_removed():
    """A first step of 0.5 rad sends the trial points of the near-wall particles
    outside the square: those attempts are rejected (the step size shrinks) instead
    of removing the particles, so they end where a run started with steps of 0.01
    ends, on the rotated positions up to the interpolant's error; the particle whose
    path leaves through the wall is removed, and the loop ends (a forced rejection
    below a step size of 1e-10 removes a particle sitting at the wall)."""
    from oracle.oracle import OracleCtx
    out = {}
    for dt in (0.5, 0.01):
        o, X, F, P = _wall_case(dt)
        c = OracleCtx(o)
        c.fill_data(X, F); c.fill_points(P); c.interpolate_c()
        counts = c.advect_adaptive_c(0.0, 1.0)
        xs = np.zeros_like(P); c.output(coords=xs)
        out[dt] = (xs, counts, c.removed_count())
        c.destroy()
        assert counts["active"] == 0
        assert out[dt][2] == 1
    keep = [0, 1, 3]
    assert out[0.5][1]["rejected"] > 0
    assert np.max(np.abs(out[0.5][0][keep] - out[0.01][0][keep])) < 1e-7
    th = 1.0
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    ex = 0.5 + (P - 0.5) @ R.T
    assert np.max(np.abs(out[0.5][0][keep] - ex[keep])) < 0.02


@pytest.mark.gpu
def test_adaptive_wall_overshoot_matches_oracle():
    """The library applies the same rule: out-of-domain trial points reject the
    attempt, the particle crossing the wall is removed.  The step sequences are the
    oracle's up to accept/reject decisions that the interpolated velocity's rounding
    flips (an error norm within rounding of 1): counts within 5%, and positions to
    1e-7, the bar two step sequences of this integration (first steps 0.5 and 0.01)
    meet in the oracle (rtol 1e-9)."""
    import lagrbf_amd as L
    o, X, F, P = _wall_case(0.5)
    print("case: adaptive DP45, trial points beyond the wall, first step 0.5", flush=True)
    g = L.LagRBFCtx(petsc_options=dict(o))
    g.fill_data(X, F); g.fill_points(P); g.interpolate_c()
    cg = g.advect_adaptive_c(0.0, 1.0)
    xg = np.zeros_like(P); g.output(coords=xg)
    xo, co = _oracle_run(o, X, F, P, 0.0, 1.0)
    print(f"wall case: gpu {cg}, removed {g.info('removed'):.0f}; oracle {co}", flush=True)
    assert cg["active"] == 0 and g.info("removed") == 1
    assert cg["rejected"] > 0
    assert abs(cg["accepted"] - co["accepted"]) <= 0.05 * co["accepted"]
    assert abs(cg["rejected"] - co["rejected"]) <= 0.05 * co["rejected"]
    keep = [0, 1, 3]
    assert np.all(xg[2] == 0)
    assert np.max(np.abs(xg[keep] - xo[keep])) < 1e-7


def _fourier_blend_case(Np=3000):
    opts, X, V, P = cases.fourier_3d(nlat=18, Np=Np, n=6, scheme=3)
    rng = np.random.default_rng(9)
    P = 0.2 + 0.6 * rng.random((Np, 3))
    VB = np.roll(V, 1, axis=1) * 0.7
    o = dict(opts, n_interpolation_fields=6, lagrbf_time_blend=True, lagrbf_blend_times="0.0,1.0", dt=1e-3,
             dopri_rtol=1e-7, dopri_atol=1e-10, lagrbf_setup_all=True, oracle_setup_all=True)
    return o, X, np.concatenate([V, VB], axis=1), P


@pytest.mark.gpu
def test_adaptive_gpu_matches_oracle():
    import lagrbf_amd as L
    o, X, F, P = _fourier_blend_case()
    print("case: adaptive DP45 + blend, fourier3d 3000 particles", flush=True)
    g = L.LagRBFCtx(petsc_options=dict(o))
    g.fill_data(X, F)
    g.fill_points(P)
    g.interpolate_c()
    cg = g.advect_adaptive_c(0.0, 0.05)
    xg = np.zeros_like(P)
    g.output(coords=xg)
    xo, co = _oracle_run(o, X, F, P, 0.0, 0.05)
    per = np.max(np.abs(xg - xo), axis=1) / np.max(np.abs(xo))
    same = float(np.mean(per < 1e-11))
    print(f"adaptive gpu vs oracle: rel max {per.max():.3e}, {same * 100:.2f}% of particles within 1e-11, "
          f"counts gpu {cg} oracle {co}", flush=True)
    assert cg["active"] == 0 and co["active"] == 0
    assert abs(cg["accepted"] - co["accepted"]) <= 0.01 * co["accepted"]
    assert abs(cg["rejected"] - co["rejected"]) <= 0.01 * max(co["rejected"], 100)
    assert same >= 0.99
    assert per.max() < o["dopri_rtol"]


@pytest.mark.gpu
def test_adaptive_gpu_mixed_and_blend_return():
    import lagrbf_amd as L
    print("case: adaptive DP45 blend v/-v return to start, fp64 and mixed", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        o, X, F, P = _rotation_case(blend=True, n=2000, rtol=1e-11 if prec == "fp64" else 1e-8)
        g = L.LagRBFCtx(petsc_options=dict(o), precision=prec)
        g.fill_data(X, F)
        g.fill_points(P)
        g.interpolate_c()
        c = g.advect_adaptive_c(0.0, 1.0)
        xs = np.zeros_like(P)
        g.output(coords=xs)
        out[prec] = xs
        assert c["active"] == 0
    assert np.max(np.abs(out["fp64"] - P)) < 1e-7
    # FP32 evaluation: ~1e-6 relative velocities
    assert np.max(np.abs(out["mixed"] - P)) < 1e-4


@pytest.mark.gpu
def test_slab_configuration_matches_oracle():
    """The adaptive DP45 configuration (4:1:1 box, slab snapshots blended in time)
    at a small size: library against the oracle, FP64."""
    import lagrbf_amd as L
    o, X, F, P = cases.slab_c4(m=16, N=None, Np=3000, pu=5)
    print("case: slab 4:1:1, 16,384 nodes, 3000 particles, t = 0 .. 0.2", flush=True)
    g = L.LagRBFCtx(petsc_options=dict(o))
    g.fill_data(X, F)
    g.fill_points(P)
    g.interpolate_c()
    cg = g.advect_adaptive_c(0.0, 0.2)
    xg = np.zeros_like(P)
    g.output(coords=xg)
    xo, co = _oracle_run(o, X, F, P, 0.0, 0.2)
    live = np.any(xo != 0, axis=1)
    rel = np.max(np.abs(xg[live] - xo[live])) / np.max(np.abs(xo[live]))
    print(f"slab: rel {rel:.3e}, gpu {cg} oracle {co}, removed {g.info('removed'):.0f}", flush=True)
    assert cg["active"] == 0 and co["active"] == 0
    assert abs(cg["accepted"] - co["accepted"]) <= 0.01 * co["accepted"]
    assert np.array_equal(np.any(xg != 0, axis=1), live)
    assert rel < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_blend_fused_matches_two_passes(precision):
    """Both snapshots evaluated in one pass (6 fields, the FP32 default) against one
    evaluation per snapshot: the same sums per field, so the same trajectories and
    step counts, bit for bit (FP64 runs one pass per snapshot either way)."""
    import lagrbf_amd as L
    o, X, F, P = _fourier_blend_case(Np=4000)
    print(f"case: adaptive blend fused vs two passes, {precision}", flush=True)
    res = []
    for fused in (1, 0):
        g = L.LagRBFCtx(petsc_options=dict(o, lagrbf_blend_fused=fused), precision=precision)
        g.fill_data(X, F)
        g.fill_points(P)
        g.interpolate_c()
        cnt = g.advect_adaptive_c(0.0, 0.05)
        xs = np.zeros_like(P)
        g.output(coords=xs)
        res.append((xs, cnt))
        g.destroy()
    assert res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])
