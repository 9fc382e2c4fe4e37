/*
 * lagrbf.h -- C ABI of the MI355X-native LagRBF hot path (liblagrbf.so).
 *
 * Drop-in for the symbols the reference's Python layer binds through ctypes
 * (python_api/lagrbfpy/interface/register_cfunc.py:44-195).  Types follow the
 * reference's 32-bit-index PETSc build: PetscInt = int32_t, PetscReal = double,
 * PetscInt64 = int64_t, PetscBool is passed as a C bool (ctypes.c_bool).
 * Every function returns 0 on success or a PETSc-style error code (62 wrong
 * argument, 63 out of range, 73 wrong state, 76 device/library failure,
 * 85 null argument); the message is printed to stderr and kept for
 * lagrbf_last_error().
 *
 * Configuration does not cross the ABI in the reference: lagrbf_create() reads
 * the PETSc options database.  Here the same option names (without PETSc) are
 * given through lagrbf_set_option(), a file (lagrbf_options_insert_file) or the
 * LAGRBF_OPTIONS environment variable ("-name value -name value ...").
 *
 * All computation runs on the current HIP device.  Host arrays passed in are
 * copied; host output arrays are caller-allocated and fully written.
 */
#ifndef LAGRBF_H
#define LAGRBF_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- reference ABI ---------------- */

/* Replaces lagrbf_create, src/interface/interface_ctx.c:93-104.
 * Builds the PU grid, patches and both swarms from the options database.
 * Returns NULL on error (bad options, no HIP device). */
void *lagrbf_create(void);

/* Replaces lagrbf_destroy, src/interface/interface_ctx.c:118-123. */
void lagrbf_destroy(void *ctx);

/* Replaces lagrbf_fill_swarm, src/interface/interface_utils.c:403-430.
 * is_data: fill the data (node) swarm, else the interpolation (particle) swarm.
 * coords [npoints*dim] AoS, values [npoints*n_interpolation_fields] or NULL,
 * pidx [npoints] (1-based unique ids) or NULL (ids 1..npoints assigned).
 * Points outside the PU box [O,L] are rejected with error 62. */
int lagrbf_fill_swarm(void *ctx, bool is_data, int32_t npoints, double *coords,
                      double *values, int64_t *pidx);

/* Replaces lagrbf_interpolation, src/interface/interface_interp.c:210-221:
 * lazy patch set-up + solve (partition.c:417-478) then PU evaluation of "fields". */
int lagrbf_interpolation(void *ctx);

/* Replaces lagrbf_output, src/interface/interface_utils.c:362-384.
 * Writes rows by pid-1 (AssembleArray_Allreduce semantics); rows of removed
 * particles stay zero.  fieldname: "fields", "velocity" or "weights". */
int lagrbf_output(void *ctx, bool is_data, int32_t npoints, double *coords, double *fields,
                  int64_t *pidx, const char *fieldname);

/* Replaces lagrbf_avect_analytical (sic), src/interface/interface_adv.c:75-88:
 * ForwardAnalyticalAdvection (time_loop.c:176-221) with the registered callback. */
int lagrbf_avect_analytical(void *ctx, double time, double time_next);

/* Replaces lagrbf_advect_interpolate, src/interface/interface_adv.c:90-106:
 * InternalForwardAdvectionLoop (time_loop.c:39-100), CFL-driven dt; at exit the
 * coefficients are reset and the data swarm is cleared, as in the reference. */
int lagrbf_advect_interpolate(void *ctx, double time, double time_next);

/* Replaces lagrbf_advection_function, src/interface/interface_adv.c:35-53. */
int lagrbf_advection_function(void *ctx, int (*velocity_fn)(double *, double *, void *), void *fn_ctx);

/* Replaces lagrbf_advection_function_batch, src/interface/interface_adv.c:55-73. */
int lagrbf_advection_function_batch(void *ctx,
                                    int (*velocity_fn_batch)(int32_t, int32_t, double *, double *, void *),
                                    void *fn_ctx);

/* Replaces lagrbf_reset_coefficients, src/interface/interface_utils.c:447-458
 * (LagRBF_PartitionsReset, partition.c:494-524). */
int lagrbf_reset_coefficients(void *ctx);

/* Replaces lagrbf_interpolate_driver, src/interface/interface_interp.c:167-194:
 * one-shot interpolation with options read from options_file (may be NULL). */
void lagrbf_interpolate_driver(int32_t ndata_local, double *data_coords, double *data_values,
                               int32_t npoints_local, double *points_coords, double *points_values,
                               int32_t npoints_global, char *options_file);

/* ---------------- extensions (no reference counterpart) ---------------- */

/* Options database replacing PETSc.Options (context.py:119-201 writes these names). */
int lagrbf_set_option(const char *name, const char *value);
int lagrbf_clear_options(void);
int lagrbf_options_insert_file(const char *path);

/* Fixed-step form of the internal advection loop: nsteps steps of size dt with
 * the per-step sequence of time_loop.c:80-90, no CFL, no end-of-call reset.
 * Used for fixed-dt workloads and the throughput benchmark.  With the option
 * lagrbf_graph, steps after the call's first one are captured as HIP graphs
 * (one per particle-buffer parity) and replayed, when the step's launches are
 * fixed: RK1/2/4 with interpolated velocity, every patch set up and solved. */
int lagrbf_advect_steps(void *ctx, int32_t nsteps, double dt);

/* Per-particle adaptive Dormand-Prince 5(4) from t_start to t_end (no reference
 * counterpart; the reference's DOPRI uses one dt for the swarm).  Each particle
 * has its own time and step size (initial -dt), accepts or rejects each attempt
 * with the DOPRI error norm and controller (-dopri_atol, -dopri_rtol,
 * -dopri_dt_fac*), and stops at t_end; at most max_attempts attempts.  With
 * -lagrbf_time_blend (n_interpolation_fields = 2 x dim) the velocity at stage
 * time t is (1 - tau) v_A + tau v_B, tau = (t - tA) / (tB - tA), v_A the first dim
 * fields and v_B the next (-lagrbf_blend_times tA,tB, default 0,1).  Counters:
 * "dp_attempts", "dp_accepted", "dp_rejected", "dp_active" (unfinished).  A trial
 * point of stages 2-6 outside the domain rejects the particle's attempt (back to
 * its step's start, step size times dopri_dt_fac_min), twice per particle at most;
 * the third, or a step's end point outside the domain, removes it ("removed"). */
int lagrbf_advect_adaptive(void *ctx, double t_start, double t_end, int32_t max_attempts);

/* Runtime counters: "removed" (particles that left the domain), "fallback"
 * (patches refactored in FP64 after FP32 refinement stalled), "refine_iters",
 * "patches_setup", "dt" (current advection dt), "eval_ms" and "eval_count"
 * (evaluation kernel time and launches timed, outside graphs), "eval_variant"
 * (launch shape 0..17 of the last cell-table evaluation, -1 before one),
 * "graph_ms" and "graph_steps" (replayed steps and their time), "setup_ms" and its
 * phases "setup_alloc_ms" (pool growth), "setup_factor_ms", "setup_solve_ms",
 * "n_particles", "n_nodes", "max_patch", "eps_w", "eps_r". */
int lagrbf_get_info(void *ctx, const char *key, double *value);

/* Introspection for tests: copies device state into `out` (capacity `cap`
 * elements) and returns the element count (negative error code on failure).
 * what: "cellid_points", "cellid_data", "cell_start", "cell_nodes", "alive"
 * (int32), "patch_members" (int32, patch `index`), "patch_coeffs" (double,
 * row-major [n][nf]), "patch_flags" (int32 setup, solved, n). */
int64_t lagrbf_debug_get(void *ctx, const char *what, int32_t index, void *out, int64_t cap);

/* Copy the last error message (NUL-terminated, truncated to len). */
int lagrbf_last_error(char *buf, int32_t len);

/* No counterpart in the reference: an entry point that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The variant of
 * lagrbf_fill_swarm(is_data = false) + lagrbf_interpolation + lagrbf_output
 * ("fields") for points already in device memory: interpolates the context's
 * fields at n points (coords: device [n * dim] AoS) into values (device
 * [n * n_interpolation_fields] AoS, rows in point order; points outside the PU box
 * and points with an empty 3^d neighbourhood give zero rows), with the same
 * kernels and the same values bit for bit.  Every work array lives in a static
 * device array of 40 MiB per GPU (held until the process ends): at most 262,144
 * points per call, 262,144 PU cells and 4 fields (error 60 beyond).  The context
 * must have every patch set up and solved and its evaluation table built
 * (-lagrbf_setup_all, then one lagrbf_interpolation; error 73 otherwise); its own
 * particle swarm is untouched.  Returns after the values are written.  Calls that
 * overlap on one GPU share its scratch and race. */
int lagrbf_interpolate_device_noalloc(void *ctx, int32_t n, const double *coords, double *values);

/* ---------------- serving (no reference counterpart) ----------------
 * Continuous batching of advection requests against device-resident datasets.
 * A dataset is filled once and all its patches are set up and solved (factors
 * are then released; the evaluation table stays).  Requests on a dataset and
 * scheme share one particle batch: each tick admits queued requests while the
 * HBM budget allows, advances every active request by one step (own dt; DOPRI
 * step size controlled per request) and retires the finished ones.  A
 * request's result does not depend on the other requests in the batch. */

/* New server; hbm_budget_bytes <= 0 selects 200e9. */
void *lagrbf_serve_create(double hbm_budget_bytes);
void lagrbf_serve_destroy(void *srv);

/* Dataset from an options string ("-pu_mesh_dim 3 -pu_mesh_n 32,32,32 ...", the
 * names lagrbf_set_option accepts; n_interpolation_fields must equal the
 * dimension: values are the velocity) and n nodes (coords [n*dim], values
 * [n*dim]).  -lagrbf_serve_tune_min_particles m (default 4194304): batch
 * evaluations of fewer than m particles run one fixed launch shape instead of
 * the autotuner's trials.  Returns the dataset id (>= 0) or a negative error
 * code. */
int32_t lagrbf_serve_add_dataset(void *srv, const char *options, int32_t n, double *coords, double *values);

/* Queue a request: n particles (coords [n*dim], inside the dataset's box), scheme
 * 0..3 (RK1, RK2, RK4, DOPRI), nsteps steps of size dt (DOPRI: initial dt).
 * Returns the request id (>= 0) or a negative error code. */
int64_t lagrbf_serve_submit(void *srv, int32_t dataset, int32_t scheme, int32_t nsteps, double dt, int32_t n,
                            const double *coords);

/* One tick (admission, one step of every active request, retirement).  Returns
 * the number of queued + active requests, or a negative error code. */
int32_t lagrbf_serve_step(void *srv);

/* 1 when the request is done (final positions copied to out [n*dim], rows of
 * particles that left the box are zero), 0 while pending, negative on error.
 * times (may be NULL): submit, admit, done, in seconds since lagrbf_serve_create
 * (-1 while not reached). */
int32_t lagrbf_serve_poll(void *srv, int64_t id, double *out, double *times);

/* Drop a finished request's host copy. */
int lagrbf_serve_release(void *srv, int64_t id);

/* "completed", "particle_steps", "ticks", "reserved" (bytes), "queued",
 * "max_active", "now", "datasets". */
int lagrbf_serve_info(void *srv, const char *key, double *value);

#ifdef __cplusplus
}
#endif
#endif
