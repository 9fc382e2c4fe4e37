"""Throughput benchmark: particle-RK4 steps/s on one MI355X.

Workload (the 3D plain-vs-mixed configuration): 1,048,576 data nodes (first 2^20
of a seeded permutation of a 102^3 lattice jittered by +-0.3 spacing) sampling a
divergence-free 16-mode Fourier velocity; PU grid 32^3 nodes on [0,1]^3, Wendland
C2 kernel, overlap 1, shape factor 1; 4,194,304 particles uniform in the box;
RK4 with dt = 1e-3.  A step is one RK4 step of the whole swarm (four PU
evaluations, weights, re-location, lazy patch set-up check).

The timed region holds exactly --steps steps with the factored patches and the
particles resident in HBM; set-up (binning, assembly, factorisation, first
interpolation) and --warmup steps run before it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12}      # MI355X dense vector FP64 / FP32, FLOP/s
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def eval_flops(cell_start, nc, n_p, cellid, alive, nf=3, Fk=16):
    """Algorithmic FLOPs of one PU evaluation over the swarm (Wendland C2, 3D):
    per particle n_nb * F_k + sum_k 2 nf n_pk + 2^d 2 nf, with n_nb the population of
    the 3x3x3 cell neighbourhood and n_pk the sizes of the 8 corner patches."""
    cnt = np.diff(cell_start).reshape(nc, nc, nc)          # [k][j][i]
    pad = np.pad(cnt, 1)
    nb = sum(pad[1 + dk:1 + dk + nc, 1 + dj:1 + dj + nc, 1 + di:1 + di + nc]
             for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)).ravel()
    n = nc + 1
    npk = n_p.reshape(n, n, n)
    corner_sum = sum(npk[dk:dk + nc, dj:dj + nc, di:di + nc] for dk in (0, 1) for dj in (0, 1) for di in (0, 1)).ravel()
    c = cellid[alive.astype(bool)]
    return float(np.sum(nb[c]) * Fk + np.sum(corner_sum[c]) * 2 * nf + c.size * 8 * 2 * nf)


def build_workload(Np, nlat=102, N=1 << 20, n=32):
    from tests import cases
    opts, X, V, P = cases.fourier_3d(nlat=nlat, N=N, Np=Np, n=n, kernel=2, scheme=2, solve_method=0, overlap=1.0)
    return opts, X, V, P


def cpu_baseline(opts, X, V, P, steps, dt, threads):
    """The oracle (C restatement, OpenMP) on a bounded sample: particles of the
    workload that start inside [0.25,0.5]^3, same nodes, same PU; steady-state RK4
    steps after set-up."""
    from oracle.oracle import OracleCtx
    sel = np.all((P >= 0.25) & (P <= 0.5), axis=1)
    Ps = P[sel][:65536]
    o = OracleCtx(opts)
    o.fill_data(X, V)
    o.fill_points(Ps)
    o.interpolate_c()
    t0 = time.perf_counter()
    o.advect_steps_c(steps, dt)
    t = time.perf_counter() - t0
    return {"value": Ps.shape[0] * steps / t, "unit": "particle-RK4-steps/s", "cores": threads, "kind": "port",
            "sample": f"{Ps.shape[0]} particles of the workload inside [0.25,0.5]^3, {steps} RK4 steps after set-up "
                      f"(same 2^20 nodes and 32^3 PU), {t:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=4194304)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import lagrbf_amd as L
    dt = 1e-3
    t0 = time.perf_counter()
    opts, X, V, P = build_workload(args.particles)
    log(f"workload built: {X.shape[0]} nodes, {P.shape[0]} particles ({time.perf_counter() - t0:.1f}s)")
    ctx = L.LagRBFCtx(petsc_options=dict(opts), precision=args.precision)
    ctx.fill_data(X, V)
    ctx.fill_points(P)
    ts = time.perf_counter()
    ctx.interpolate_c()
    t_setup = time.perf_counter() - ts
    log(f"set-up + first interpolation: {t_setup:.2f}s, max patch {ctx.info('max_patch'):.0f}")
    ctx.advect_steps_c(args.warmup, dt)
    log(f"warmup {args.warmup} steps done")
    ev0, n_ev0 = ctx.info("eval_ms"), ctx.info("eval_count")
    t1 = time.perf_counter()
    ctx.advect_steps_c(args.steps, dt)
    t2 = time.perf_counter()
    ev1, n_ev1 = ctx.info("eval_ms"), ctx.info("eval_count")
    wall = t2 - t1
    removed = ctx.info("removed")
    log(f"timed {args.steps} steps: {wall:.3f}s, removed {removed:.0f}")
    value = P.shape[0] * args.steps / wall

    # roofline of the dominant kernel (PU evaluation), algorithmic FLOPs per launch
    cs = ctx.debug_get("cell_start")
    alive = ctx.debug_get("alive")
    cellid = ctx.debug_get("cellid_points")
    nc = opts["pu_mesh_n"][0] - 1
    cnt = np.diff(cs).reshape(nc, nc, nc)
    n = nc + 1
    pc = np.zeros((n, n, n), dtype=np.int64)
    padc = np.pad(cnt, 1)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                pc += padc[dk:dk + n, dj:dj + n, di:di + n]
    flops = eval_flops(cs, nc, pc.ravel(), cellid, alive)
    n_launch = max(n_ev1 - n_ev0, 1)
    avg_ms = (ev1 - ev0) / n_launch
    achieved = flops / (avg_ms * 1e-3)
    peak = PEAK["f64" if args.precision == "fp64" else "f32"]
    cpu = None
    if rank == 0 and not args.no_cpu:
        import multiprocessing
        thr = int(os.environ.get("OMP_NUM_THREADS", multiprocessing.cpu_count()))
        log("cpu baseline ...")
        cpu = cpu_baseline(opts, X, V, P, args.cpu_steps, dt, thr)
        log(f"cpu baseline: {cpu['value']:.3e} {cpu['unit']} ({cpu['sample']})")
    out = {
        "metric": "particle-RK4 steps/s, FP32+refine vs FP64 @1e-10 rel err; serve p99 ms",
        "value": value, "unit": "particle-RK4-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64-refine", "data": "synthetic",
        "config": {"workload": "3D plain-vs-mixed: 2^20 jittered-lattice nodes, 16-mode Fourier field, 32^3 PU, "
                               "Wendland C2, 4,194,304 particles, RK4 dt=1e-3",
                   "nodes": int(X.shape[0]), "particles": int(P.shape[0]), "pu_n": 32, "kernel": "wendland_c2",
                   "precision": args.precision},
        "roofline": {"bound": "valu", "kernel": "pu_evaluation", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "flops_per_launch": flops, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
        "setup_s": t_setup, "removed_particles": removed,
    }
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
