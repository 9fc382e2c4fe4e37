"""Throughput benchmark: particle-RK4 steps/s on one MI355X.

Workload (the 3D plain-vs-mixed configuration): 1,048,576 data nodes (first 2^20
of a seeded permutation of a 102^3 lattice jittered by +-0.3 spacing) sampling a
divergence-free 16-mode Fourier velocity; PU grid 32^3 nodes on [0,1]^3, r^3
kernel with a linear polynomial block (saddle-point patch systems), overlap 1,
shape factor 1; 4,194,304 particles uniform in the box; RK4 with dt = 1e-3.  A
step is one RK4 step of the whole swarm (four PU evaluations, Shepard weights,
re-location, lazy patch set-up check).  The same workload with the reference's
Wendland C2 kernel (no polynomial) is reported as a secondary key, and the 2D
rotation configuration (r^3 + linear, 628 RK4 steps) with its closure errors.

Both precision paths run on the same inputs, one after the other: FP32 factor +
FP64 refinement with FP32 evaluation ("mixed", the headline `value`) and plain
FP64.  Each timed region holds exactly --steps steps with the factored patches
and the particles resident in HBM; set-up (binning, assembly, factorisation,
first interpolation) and --warmup steps run before it.  The line also carries
the mixed path's error against the FP64 path (patch coefficients, positions).

Roofline of the dominant kernel (PU evaluation): algorithmic FLOPs per launch
(per particle n_nb * F_k + sum_k 2 nf n_pk + 2^d 2 nf + polynomial terms 2^d (2 nf m
+ m d); F_k = 10 for r^3 in 3D) over the kernel's average duration, timed with HIP
events on the library's stream.  r^3 has no compact support, so every counted
(particle, node) pair is computed; the Wendland key's count is nominal (the
kernel skips node rows outside the support).  HBM traffic per
launch comes from rocprofv3 FETCH_SIZE / WRITE_SIZE runs stored beside this
file in profiles/pu_evaluation_c2_phs3_traffic.json (null when absent, and for the
Wendland key, which has no traffic profile).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "mixed": 157.3e12}     # MI355X dense FP64 / FP32, FLOP/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pu_evaluation_c2_phs3_traffic.json")
WORKLOAD = ("3D plain-vs-mixed: 2^20 jittered-lattice nodes, 16-mode Fourier field, 32^3 PU, r^3 kernel + linear "
            "polynomial, 4,194,304 particles, RK4 dt=1e-3")
FK = {7: 10, 2: 16}      # flops per (particle, node) pair, distance included: r^3, Wendland C2 (3D)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def eval_flops(cell_start, nc, n_p, cellid, alive, nf=3, Fk=10, m=4, d=3):
    """Algorithmic FLOPs of one PU evaluation over the swarm (3D):
    per particle n_nb * F_k + sum_k 2 nf n_pk + 2^d 2 nf + 2^d (2 nf m + m d), with
    n_nb the population of the 3x3x3 cell neighbourhood, n_pk the sizes of the 8
    corner patches and m the polynomial terms per patch (0 without a block)."""
    cnt = np.diff(cell_start).reshape(nc, nc, nc)          # [k][j][i]
    pad = np.pad(cnt, 1)
    nb = sum(pad[1 + dk:1 + dk + nc, 1 + dj:1 + dj + nc, 1 + di:1 + di + nc]
             for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)).ravel()
    n = nc + 1
    npk = n_p.reshape(n, n, n)
    corner_sum = sum(npk[dk:dk + nc, dj:dj + nc, di:di + nc] for dk in (0, 1) for dj in (0, 1) for di in (0, 1)).ravel()
    c = cellid[alive.astype(bool)]
    return float(np.sum(nb[c]) * Fk + np.sum(corner_sum[c]) * 2 * nf + c.size * 8 * (2 * nf + 2 * nf * m + m * d))


def build_workload(Np, nlat=102, N=1 << 20, n=32, kernel=7):
    """C2: r^3 + linear polynomial (kernel 7); kernel 2 gives the Wendland C2 variant."""
    from tests import cases
    opts, X, V, P = cases.fourier_3d(nlat=nlat, N=N, Np=Np, n=n, kernel=kernel, scheme=2, solve_method=0, overlap=1.0)
    if kernel == 7:
        opts["lagrbf_poly_degree"] = 1
    return opts, X, V, P


def patch_sizes(cs, nc):
    """Patch sizes (members) from the per-cell node counts: patch p holds the <= 8 cells around node p."""
    cnt = np.diff(cs).reshape(nc, nc, nc)
    n = nc + 1
    pc = np.zeros((n, n, n), dtype=np.int64)
    padc = np.pad(cnt, 1)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                pc += padc[dk:dk + n, dj:dj + n, di:di + n]
    return pc.ravel()


def load_traffic(precision, shape, path=TRAFFIC_FILE):
    """HBM bytes per evaluation launch of `precision` in launch shape `shape`, from
    the committed counter runs (one per shape); None when that shape was not
    measured."""
    if path is None:
        return None
    try:
        with open(path) as f:
            t = json.load(f)[precision]["shapes"][str(int(shape))]
        return float(t["fetch_bytes"]) + float(t["write_bytes"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def copy_calibration(mib):
    """A plain copy of known size (reads and writes mib MiB) for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE: torch.neg over float64, launched three times."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 x torch.neg over {n * 8} bytes")
    del a, b


def run_path(L, precision, opts, X, V, P, steps, warmup, dt, sample, traffic_file=TRAFFIC_FILE):
    """Set up one context, run warmup + timed steps; returns the result record,
    a sample of patch coefficients and the final positions (rows by pid)."""
    ctx = L.LagRBFCtx(petsc_options=dict(opts), precision=precision)
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ts = time.perf_counter()
    ctx.interpolate_c()
    t_setup = time.perf_counter() - ts
    log(f"[{precision}] set-up + first interpolation: {t_setup:.2f}s, max patch {ctx.info('max_patch'):.0f}, "
        f"fallback {ctx.info('fallback'):.0f}, refine iters max {ctx.info('refine_iters'):.0f}")
    ctx.advect_steps_c(warmup, dt)
    removed0 = ctx.info("removed")
    ev0, n_ev0 = ctx.info("eval_ms"), ctx.info("eval_count")
    t1 = time.perf_counter()
    ctx.advect_steps_c(steps, dt)
    t2 = time.perf_counter()
    ev1, n_ev1 = ctx.info("eval_ms"), ctx.info("eval_count")
    wall = t2 - t1
    removed = ctx.info("removed")
    # particle-steps of the live swarm: particles alive at the end count every step,
    # the ones removed inside the timed region none (a lower bound)
    live_steps = (P.shape[0] - removed) * steps
    log(f"[{precision}] timed {steps} steps: {wall:.3f}s ({wall / max(steps, 1) * 1e3:.2f} ms/step), removed {removed:.0f}")

    cs = ctx.debug_get("cell_start")
    alive = ctx.debug_get("alive")
    cellid = ctx.debug_get("cellid_points")
    nc = opts["pu_mesh_n"][0] - 1
    kt = int(opts["kernel_type"])
    m = 4 if int(opts.get("lagrbf_poly_degree", -1)) == 1 else 0
    flops = eval_flops(cs, nc, patch_sizes(cs, nc), cellid, alive, Fk=FK[kt], m=m)
    # evaluations are timed once their launch shape is settled (the autotuner's
    # trials and run-off are not): with too few warmup steps none may be
    n_launch = int(n_ev1 - n_ev0)
    avg_ms = achieved = None
    if n_launch > 0:
        avg_ms = (ev1 - ev0) / n_launch
        achieved = flops / (avg_ms * 1e-3)
    else:
        log(f"[{precision}] no settled evaluation in the timed steps: raise --warmup for the roofline figures")
    traffic = load_traffic(precision, ctx.info("eval_variant"), traffic_file)
    coefs = {}
    for p in sample:
        c = ctx.debug_get("patch_coeffs", int(p), dtype=np.float64)
        if c.size:
            # c and, with a polynomial block, beta
            coefs[int(p)] = np.concatenate([c, ctx.debug_get("patch_beta", int(p), dtype=np.float64)])
    pos = np.zeros_like(P)
    ctx.output(coords=pos)
    rec = {
        "value": live_steps / wall, "ms_per_step": wall / max(steps, 1) * 1e3,
        "end_to_end": live_steps / (wall + t_setup), "setup_s": t_setup, "setup_phases_ms": setup_phases(ctx),
        "removed_particles": removed,
        "removed_in_timed_steps": removed - removed0,
        "fallback_patches": ctx.info("fallback"), "refine_iters_max": ctx.info("refine_iters"),
        "roofline": {"bound": "valu", "kernel": "pu_evaluation (k_eval_blk)",
                     "achieved": achieved / 1e12 if achieved else None, "peak": PEAK[precision] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK[precision] if achieved else None, "traffic": traffic,
                     "flops_per_launch": flops, "avg_launch_ms": avg_ms, "launches_timed": n_launch,
                     "launch_shape": int(ctx.info("eval_variant")),
                     "flops_note": "exact: every (particle, node) pair is computed" if kt == 7 else
                     "nominal: node rows outside the compact support are skipped"},
    }
    ctx.destroy()
    return rec, coefs, pos


def run_graph(L, precision, opts, X, V, P, steps, warmup, dt):
    """The same timed steps with every step after each call's first replayed from a
    captured HIP graph (option lagrbf_graph): steps/s and the graph share."""
    ctx = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_graph=True), precision=precision)
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ctx.interpolate_c()
    ctx.advect_steps_c(warmup, dt)
    g0 = ctx.info("graph_steps")
    t1 = time.perf_counter()
    ctx.advect_steps_c(steps, dt)
    wall = time.perf_counter() - t1
    rec = {"value": P.shape[0] * steps / wall, "ms_per_step": wall / max(steps, 1) * 1e3,
           "graph_steps": int(ctx.info("graph_steps") - g0), "steps": steps}
    log(f"[{precision}] graph replay: {steps} steps {wall:.3f}s ({rec['ms_per_step']:.2f} ms/step), "
        f"{rec['graph_steps']} replayed")
    ctx.destroy()
    return rec


def c1_bench(L, precision, weights):
    """2D rotation configuration as written: 4,096 uniform nodes, 16x16 PU cells with
    weight radius 1.5 cell diagonals, r^3 kernel + linear polynomial, 10,000 particles
    uniform in the disk r < 0.4 about (1/2, 1/2), RK4 dt = 0.01, 628 steps.
    weights "all": the reference's Shepard normalisation over every PU node (the used
    corner weights sum to less than 1 at this radius, swarm_interp.c:327-372);
    "corners": normalised over the used corners.  Reports the return-to-start error
    and the error against the exact rotation by 6.28 rad (628 x 0.01 < 2 pi)."""
    from tests import cases
    opts, X, V, P = cases.rotation_2d(N=4096, Np=10000, n=17, kernel=7, scheme=2)
    opts.update(lagrbf_poly_degree=1, lagrbf_weights=weights)
    ctx = L.LagRBFCtx(petsc_options=dict(opts), precision=precision)
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ctx.interpolate_c()
    t = time.perf_counter()
    ctx.advect_steps_c(628, 0.01)
    wall = time.perf_counter() - t
    xs = np.zeros_like(P)
    ctx.output(coords=xs)
    removed = ctx.info("removed")
    ctx.destroy()
    th = 6.28
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    exact = 0.5 + (P - 0.5) @ R.T
    ret = np.linalg.norm(xs - P, axis=1)
    rot = np.linalg.norm(xs - exact, axis=1)
    res = {"precision": precision, "weights": weights, "wall_s": wall, "steps_per_s": 628 / wall,
           "particle_rk4_steps_per_s": P.shape[0] * 628 / wall, "removed": removed,
           "return_to_start_max": float(ret.max()), "return_to_start_mean": float(ret.mean()),
           "vs_exact_rotation_max": float(rot.max()), "vs_exact_rotation_mean": float(rot.mean()),
           "vs_exact_rotation_median": float(np.median(rot))}
    log(f"c1 [{precision}, weights {weights}]: {wall:.2f}s, return-to-start max {ret.max():.4e} mean {ret.mean():.4e}, "
        f"vs exact 6.28 rad max {rot.max():.4e} mean {rot.mean():.4e}")
    return res


def cpu_baseline(opts, X, V, P, steps, dt, nsample, threads):
    """The oracle (C restatement of the reference, OpenMP) on a bounded sample:
    the first nsample particles of the workload inside [0.25,0.5]^3, same nodes,
    same PU, steady-state RK4 steps after set-up.  Reference-faithful mode
    (Shepard sums over every PU node, as swarm_interp.c:283-396), and the
    cutoff-equivalent mode (stencil of PU nodes inside the weight support)."""
    from oracle.oracle import OracleCtx
    sel = np.all((P >= 0.25) & (P <= 0.5), axis=1)
    Ps = P[sel][:nsample]
    rates = {}
    for mode, extra in (("faithful", {"oracle_weights_brute": True}), ("cutoff", {})):
        o = OracleCtx(dict(opts, **extra))
        o.fill_data(X, V)
        o.fill_points(Ps)
        o.interpolate_c()
        t0 = time.perf_counter()
        o.advect_steps_c(steps, dt)
        rates[mode] = (Ps.shape[0] * steps / (time.perf_counter() - t0), time.perf_counter() - t0)
        o.destroy()
    v, t = rates["faithful"]
    return {"value": v, "unit": "particle-RK4-steps/s", "cores": threads, "kind": "port",
            "sample": f"{Ps.shape[0]} particles of the workload inside [0.25,0.5]^3, {steps} RK4 steps after set-up "
                      f"(same 2^20 nodes and 32^3 PU), brute-force Shepard sums as the reference, {t:.1f} s",
            "cutoff_value": rates["cutoff"][0]}


def serve_datasets(seed=20261203 + 4):
    """The serving configuration's four datasets: 2D 65,536 uniform nodes under the
    solid-body rotation (PU 45^2, Wendland C2); 3D jittered lattices with the 16-mode
    Fourier field of 2^20 nodes (PU 32^3), 2,000,000 nodes (PU 40^3) and 4,000,000
    nodes (PU 50^3), built as the 3D configuration is: r^3 kernel with a linear
    polynomial block; about 34 nodes per PU cell."""
    from tests import cases
    rng = np.random.default_rng(seed)
    X = rng.random((65536, 2))
    V = np.stack([-(X[:, 1] - 0.5), X[:, 0] - 0.5], axis=1)
    out = [(dict(pu_mesh_dim=2, pu_mesh_O=[0.0, 0.0], pu_mesh_L=[1.0, 1.0], pu_mesh_n=[45, 45], pu_mesh_overlap=1.0,
                 pu_mesh_shape_factor=1.0, kernel_type=2), X, V, 1e-2)]
    for nlat, N, n in ((102, 1 << 20, 32), (128, 2000000, 40), (160, 4000000, 50)):
        o, X, V, _ = cases.fourier_3d(nlat=nlat, N=N, Np=1, n=n, overlap=1.0, kernel=7)
        keep = ("pu_mesh_dim", "pu_mesh_O", "pu_mesh_L", "pu_mesh_n", "pu_mesh_overlap", "pu_mesh_shape_factor",
                "kernel_type")
        out.append((dict({k: o[k] for k in keep}, lagrbf_poly_degree=1), X, V, 1e-3))
    return out


def serve_bench(L, precision, dsets, nreq, rate, seed=20261203 + 5, budget=200e9):
    """Poisson arrivals at `rate` req/s of `nreq` requests: particle count log-uniform
    in [1e3, 1e6], dataset uniform, RK4 or DOPRI with equal odds, 10..200 steps.
    Latency = completion - scheduled arrival; throughput = particle-steps / wall."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    ts = time.perf_counter()
    srv = L.Server(hbm_budget=budget)
    for o, X, V, _ in dsets:
        srv.add_dataset(dict(o, lagrbf_precision=precision), X, V)
    t_setup = time.perf_counter() - ts
    reqs = []
    for k in range(nreq):
        ds = int(rng.integers(len(dsets)))
        dim = dsets[ds][1].shape[1]
        n = int(10 ** rng.uniform(3, 6))
        scheme = "rk4" if rng.random() < 0.5 else "dopri"
        nsteps = int(rng.integers(10, 201))
        P = 0.05 + 0.9 * rng.random((n, dim))
        reqs.append((ds, scheme, nsteps, dsets[ds][3], P))
    offered = sum(r[4].shape[0] * r[2] for r in reqs) / arrivals[-1]
    t0 = srv.info("now")
    w0 = time.perf_counter()
    i, pending, lat, work = 0, {}, np.zeros(nreq), 0.0
    ndone = 0
    while i < nreq or pending:
        now = time.perf_counter() - w0
        while i < nreq and arrivals[i] <= now:
            ds, scheme, nsteps, dt, P = reqs[i]
            pending[srv.submit(ds, scheme, nsteps, dt, P)] = (t0 + arrivals[i], P.shape[0] * nsteps, i)
            i += 1
        if not pending:
            time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - w0)))
            continue
        srv.step()
        for rid in list(pending):
            done, _, t = srv.poll(rid, want_result=False)
            if done:
                arr, w, k = pending.pop(rid)
                lat[k] = t[2] - arr
                work += w
                ndone += 1
                if nreq >= 500 and ndone % 100 == 0:
                    log(f"serve [{precision}] {ndone}/{nreq} requests done at {time.perf_counter() - w0:.0f} s")
                srv.release(rid)
    wall = time.perf_counter() - w0
    h = nreq // 2
    res = {"precision": precision, "requests": nreq, "rate_hz": rate, "wall_s": wall, "setup_s": t_setup,
           "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "p50_ms_first_half": float(np.percentile(lat[:h], 50) * 1e3),
           "p50_ms_second_half": float(np.percentile(lat[h:], 50) * 1e3),
           "offered_particle_steps_per_s": float(offered),
           "particle_steps_per_s": work / wall, "ticks": srv.info("ticks"), "max_active": srv.info("max_active"),
           "host_enqueue_ms": srv.info("enqueue_ms"), "host_tick_ms": srv.info("tick_ms")}
    srv.close()
    return res


def c3_bench(L, precision, steps, warmup):
    """Uneven batching as written: 1,000,000 uniform nodes + 64 Gaussian clusters of
    15,625 (sigma in [0.01, 0.05]), PU 49^3, Gaussian kernel with eps = 3 / h_local
    per patch and a quadratic polynomial block, 8,388,608 particles, RK2 dt=1e-3.
    Reports the patch-size histogram, the fallbacks and the distribution of
    refinement iterations."""
    from tests import cases
    opts, X, V, P = cases.clustered_c3()
    ctx = L.LagRBFCtx(petsc_options=dict(opts), precision=precision)
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ts = time.perf_counter()
    ctx.interpolate_c()
    t_setup = time.perf_counter() - ts
    ctx.advect_steps_c(warmup, 1e-3)
    ev0, n0 = ctx.info("eval_ms"), ctx.info("eval_count")
    t1 = time.perf_counter()
    ctx.advect_steps_c(steps, 1e-3)
    wall = time.perf_counter() - t1
    ev1, n1 = ctx.info("eval_ms"), ctx.info("eval_count")
    removed = ctx.info("removed")
    cs = ctx.debug_get("cell_start")
    nc = opts["pu_mesh_n"][0] - 1
    sizes = patch_sizes(cs, nc)
    edges = [0, 16, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 1 << 30]
    hist = {f"{a}-{b - 1}" if b < (1 << 30) else f">={a}": int(((sizes >= a) & (sizes < b)).sum())
            for a, b in zip(edges[:-1], edges[1:])}
    iters = {}
    if precision == "mixed":
        for i in range(64):
            v = ctx.info(f"refine_hist_{i}")
            if v:
                iters[str(i)] = int(v)
    res = {"precision": precision, "nodes": int(X.shape[0]), "particles": int(P.shape[0]), "setup_s": t_setup,
           "setup_phases_ms": setup_phases(ctx), "patches_setup": ctx.info("patches_setup"),
           "patch_size": {"min": int(sizes.min()), "p50": float(np.median(sizes)), "p99": float(np.percentile(sizes, 99)),
                          "max": int(sizes.max()), "above_512": int((sizes > 512).sum()), "histogram": hist},
           "fallback_patches": ctx.info("fallback"), "refactored_fp64_tiled": ctx.info("big_refactor64"),
           "refactored_bunch_kaufman": ctx.info("aug_fallback_bk"), "refine_iters_max": ctx.info("refine_iters"),
           "refine_iters_histogram": iters,
           "particle_rk2_steps_per_s": (P.shape[0] - removed) * steps / wall, "ms_per_step": wall / steps * 1e3,
           "eval_ms": (ev1 - ev0) / max(n1 - n0, 1), "removed": removed}
    ctx.destroy()
    if precision == "mixed":
        res["with_fp64_factor"] = setup_with_fp64_factor(L, opts, X, V, P)
    log(f"c3 [{precision}]: set-up {t_setup:.2f}s, patch sizes {res['patch_size']}, fallback {res['fallback_patches']:.0f} "
        f"(FP64 tiled {res['refactored_fp64_tiled']:.0f}, Bunch-Kaufman {res['refactored_bunch_kaufman']:.0f}), "
        f"refinement iterations {iters}, {res['particle_rk2_steps_per_s']:.3e} particle-RK2-steps/s "
        f"({res['ms_per_step']:.1f} ms/step)")
    return res


def setup_phases(ctx):
    """Patch set-up by phase, ms: pool growth (host; after another context freed tens
    of GB this waits on the driver), factors, solves, and their total on the stream."""
    return {k: round(ctx.info(f"setup_{k}_ms"), 2) for k in ("alloc", "factor", "solve")} | {
        "total": round(ctx.info("setup_ms"), 2)}


def setup_with_fp64_factor(L, opts, X, V, P):
    """Set-up of the mixed path with lagrbf_mixed_factor fp64 (FP64 factors, no
    refinement, FP32 evaluation) on the same nodes and particles: patch set-up phases
    in ms (pool growth, factor, solve) and wall seconds."""
    ctx = L.LagRBFCtx(petsc_options=dict(opts, lagrbf_mixed_factor="fp64"), precision="mixed")
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ts = time.perf_counter()
    ctx.interpolate_c()
    t = time.perf_counter() - ts
    iters = {str(i): int(ctx.info(f"refine_hist_{i}")) for i in range(64) if ctx.info(f"refine_hist_{i}")}
    res = {"setup_s": t, "setup_phases_ms": setup_phases(ctx), "refine_iters_histogram": iters,
           "fallback_patches": ctx.info("fallback")}
    ctx.destroy()
    return res


def c4_sizes(scale):
    """C4 sizes scaled by `scale` (nodes, particles and PU cells together)."""
    if scale >= 1.0:
        return dict(m=126, N=8000000, Np=16777216, pu=38)
    m = max(8, int(round(126 * scale ** (1 / 3))))
    return dict(m=m, N=None, Np=max(1024, int(16777216 * scale)), pu=max(4, int(round(38 * scale ** (1 / 3)))))


def c4_bench(L, precision, t_end, m=126, N=8000000, Np=16777216, pu=38):
    """Adaptive DP45 geodynamics-like flow: 8,000,000 perturbed-lattice nodes in a
    4:1:1 box, two slab-plus-noise velocity snapshots blended linearly in time,
    16,777,216 particles, per-particle Dormand-Prince 5(4) (rtol 1e-6, atol 1e-9)
    from t = 0 to t_end."""
    from tests import cases
    t0 = time.perf_counter()
    opts, X, V, P = cases.slab_c4(m=m, N=N, Np=Np, pu=pu)
    log(f"c4 workload built: {X.shape[0]} nodes, {P.shape[0]} particles ({time.perf_counter() - t0:.1f}s)")
    ctx = L.LagRBFCtx(petsc_options=dict(opts), precision=precision)
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ts = time.perf_counter()
    ctx.interpolate_c()
    t_setup = time.perf_counter() - ts
    ev0, n0 = ctx.info("eval_ms"), ctx.info("eval_count")
    t1 = time.perf_counter()
    c = ctx.advect_adaptive_c(0.0, t_end)
    wall = time.perf_counter() - t1
    ev1, n1 = ctx.info("eval_ms"), ctx.info("eval_count")
    res = {"precision": precision, "nodes": int(X.shape[0]), "particles": int(P.shape[0]), "t_end": t_end,
           "setup_s": t_setup, "setup_phases_ms": setup_phases(ctx), "wall_s": wall, "attempts": c["attempts"],
           "accepted_steps": c["accepted"],
           "rejected_steps": c["rejected"], "unfinished": c["active"], "removed": ctx.info("removed"),
           "particle_steps_per_s": c["accepted"] / wall, "eval_ms": (ev1 - ev0) / max(n1 - n0, 1),
           "evals": int(n1 - n0)}
    ctx.destroy()
    if precision == "mixed":
        res["with_fp64_factor"] = setup_with_fp64_factor(L, opts, X, V, P)
    log(f"c4 [{precision}]: set-up {t_setup:.2f}s, {c['attempts']:.0f} attempts, {c['accepted']:.3e} accepted / "
        f"{c['rejected']:.3e} rejected steps in {wall:.2f}s = {res['particle_steps_per_s']:.3e} particle-steps/s, "
        f"eval {res['eval_ms']:.2f} ms, removed {res['removed']:.0f}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=4194304)
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--cpu-particles", type=int, default=65536)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--c3-only", action="store_true", help="the uneven-batching configuration only")
    ap.add_argument("--c3-precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-c3", action="store_true", help="skip the uneven-batching measurement")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay measurement of the fixed-step loop")
    ap.add_argument("--no-wendland", action="store_true", help="skip the Wendland C2 variant of the workload")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2D rotation configuration")
    ap.add_argument("--c4-only", action="store_true", help="the adaptive DP45 configuration only")
    ap.add_argument("--no-c4", action="store_true", help="skip the adaptive DP45 measurement")
    ap.add_argument("--c4-t-end", type=float, default=1.0,
                    help="end time of the adaptive DP45 run (1.0: the whole interval of the time blend)")
    ap.add_argument("--c4-scale", type=float, default=1.0, help="fraction of the C4 node and particle counts")
    ap.add_argument("--eval-variant", type=int, default=-1,
                    help="fix the evaluation launch shape (0..13, for profiling one kernel); -1: autotuned")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="launch a plain copy of MIB MiB first (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only (one GPU's share of the swarm); "
                                   "--gpus must be 1"}), flush=True)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import lagrbf_amd as L
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    if args.c3_only:
        precs = ("mixed", "fp64") if args.c3_precision == "both" else (args.c3_precision,)
        out = {p: c3_bench(L, p, args.steps, args.warmup) for p in precs}
        print(json.dumps({"c3": out}), flush=True)
        return
    if args.c4_only:
        out = {p: c4_bench(L, p, args.c4_t_end, **c4_sizes(args.c4_scale)) for p in ("mixed", "fp64")}
        print(json.dumps({"c4": out}), flush=True)
        return
    if args.serve_only:
        dsets = serve_datasets()
        precs = ("mixed", "fp64") if args.serve_precision == "both" else (args.serve_precision,)
        out = {p: serve_bench(L, p, dsets, args.serve_requests, args.serve_rate) for p in precs}
        print(json.dumps({"serve": out}), flush=True)
        return
    dt = 1e-3
    t0 = time.perf_counter()
    opts, X, V, P = build_workload(args.particles)
    if args.eval_variant >= 0:
        opts["lagrbf_eval_variant"] = args.eval_variant
    n_nodes, n_particles = int(X.shape[0]), int(P.shape[0])
    log(f"workload built: {n_nodes} nodes, {n_particles} particles ({time.perf_counter() - t0:.1f}s)")
    n_pu = int(np.prod(opts["pu_mesh_n"]))
    sample = np.linspace(0, n_pu - 1, 512).astype(np.int64)
    paths = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    res, coefs, pos = {}, {}, {}
    for prec in paths:
        res[prec], coefs[prec], pos[prec] = run_path(L, prec, opts, X, V, P, args.steps, args.warmup, dt, sample)
        if not args.no_graph:
            res[prec]["graph"] = run_graph(L, prec, opts, X, V, P, args.steps, args.warmup, dt)
    main_path = paths[0]
    r = res[main_path]
    # the same workload with the reference's Wendland C2 kernel (no polynomial block)
    wendland = None
    if not args.no_wendland:
        opts_w, _, _, _ = build_workload(1, kernel=2)
        if args.eval_variant >= 0:
            opts_w["lagrbf_eval_variant"] = args.eval_variant
        wendland = {}
        for prec in paths:
            rw, _, _ = run_path(L, prec, opts_w, X, V, P, args.steps, args.warmup, dt, [],
                                traffic_file=None)
            wendland[prec] = {k: rw[k] for k in ("value", "ms_per_step", "setup_s", "roofline", "removed_particles")}
        wendland["config"] = WORKLOAD.replace("r^3 kernel + linear polynomial", "Wendland C2 (no polynomial)")
    c1 = None
    if not args.no_c1:
        c1 = [c1_bench(L, prec, wn) for prec in paths for wn in ("all", "corners")]
    cmp = None
    if len(paths) == 2:
        common = [p for p in coefs["mixed"] if p in coefs["fp64"]]
        cerr = max((np.max(np.abs(coefs["mixed"][p] - coefs["fp64"][p])) / max(np.max(np.abs(coefs["fp64"][p])), 1e-300)
                    for p in common), default=float("nan"))
        live = np.any(pos["fp64"] != 0, axis=1) & np.any(pos["mixed"] != 0, axis=1)
        perr = float(np.max(np.abs(pos["mixed"][live] - pos["fp64"][live])) / np.max(np.abs(pos["fp64"][live])))
        cmp = {"speedup_mixed_over_fp64": res["mixed"]["value"] / res["fp64"]["value"],
               "coef_rel_err_max": float(cerr), "patches_compared": len(common),
               "position_rel_err_max": perr, "position_steps": args.warmup + args.steps}
        log(f"mixed vs fp64: speedup {cmp['speedup_mixed_over_fp64']:.3f}, coefficient rel err {cerr:.3e} "
            f"({len(common)} patches), position rel err {perr:.3e}")
    cpu = None
    if rank == 0 and not args.no_cpu:
        thr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"cpu baseline ({thr} threads) ...")
        cpu = cpu_baseline(opts, X, V, P, args.cpu_steps, dt, args.cpu_particles, thr)
        log(f"cpu baseline: {cpu['value']:.3e} {cpu['unit']} ({cpu['sample']}); cutoff mode {cpu['cutoff_value']:.3e}")
    serve, c3 = None, None
    del X, V, P
    if not args.no_serve:
        try:
            dsets = serve_datasets()
            serve = {}
            for p in ("mixed", "fp64"):
                serve[p] = serve_bench(L, p, dsets, args.serve_requests, args.serve_rate)
                log(f"serve [{p}]: {args.serve_requests} requests at {args.serve_rate:.0f}/s: p50 {serve[p]['p50_ms']:.1f} ms, "
                    f"p99 {serve[p]['p99_ms']:.1f} ms, {serve[p]['particle_steps_per_s']:.3e} particle-steps/s, "
                    f"wall {serve[p]['wall_s']:.1f} s, set-up {serve[p]['setup_s']:.1f} s")
            serve["config"] = ("4 device-resident datasets (2D 65,536 nodes; 3D 2^20, 2,000,000, 4,000,000 nodes), "
                               f"{args.serve_requests} requests, Poisson arrivals at {args.serve_rate:.0f}/s, particles "
                               "log-uniform in [1e3, 1e6], RK4 or DOPRI, 10-200 steps, HBM budget 200 GB; 3D datasets r^3 + linear "
                               "polynomial, 2D Wendland C2")
            del dsets
        except Exception as e:       # the headline measurement above stands on its own
            serve = {"error": str(e)}
            log(f"serve failed: {e}")
    if not args.no_c3:
        try:
            c3 = {p: c3_bench(L, p, 5, 1) for p in ("mixed", "fp64")}
            c3["config"] = ("uneven batching: 1,000,000 uniform + 64 Gaussian clusters x 15,625 nodes (sigma 0.01-0.05), "
                            "PU 49^3, Gaussian eps = 3/h_local + quadratic polynomial, 8,388,608 particles, RK2 dt=1e-3, "
                            "5 steps")
        except Exception as e:
            c3 = {"error": str(e)}
            log(f"c3 failed: {e}")
    c4 = None
    if not args.no_c4:
        try:
            c4 = {p: c4_bench(L, p, args.c4_t_end, **c4_sizes(args.c4_scale)) for p in ("mixed", "fp64")}
            c4["config"] = ("adaptive DP45: 8,000,000 perturbed-lattice nodes in a 4:1:1 box, slab + Fourier-noise "
                            "snapshots blended linearly in t, 16,777,216 particles, per-particle Dormand-Prince 5(4) "
                            f"rtol 1e-6 atol 1e-9, t = 0 .. {args.c4_t_end}")
        except Exception as e:
            c4 = {"error": str(e)}
            log(f"c4 failed: {e}")
    out = {
        "metric": "particle-RK4 steps/s, FP32+refine vs FP64 @1e-10 rel err; serve p99 ms",
        "value": r["value"], "unit": "particle-RK4-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if main_path == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": WORKLOAD, "nodes": n_nodes, "particles": n_particles, "pu_n": 32,
                   "kernel": "phs3+linear", "precision": main_path},
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
        "setup_s": r["setup_s"], "end_to_end": r["end_to_end"], "removed_particles": r["removed_particles"],
        "paths": {p: {k: v for k, v in res[p].items()} for p in paths},
        "mixed_vs_fp64": cmp,
        "wendland_c2": wendland,
        "rotation_2d": c1,
        "serve": serve,
        "uneven_batching": c3,
        "adaptive_dp45": c4,
    }
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
