/*
 * lagrbf_oracle.h -- FP64 CPU restatement of the LagRBF partition-of-unity RBF
 * interpolation + Lagrangian advection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (lagrbf_amd/liblagrbf.so): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links, loads or
 * calls anything in oracle/.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the LagRBF source tree).  Numerics: strict IEEE double, no fast-math, no FMA
 * contraction (built with -ffp-contract=off).
 *
 * Pinning (what checks each function, see tests/test_oracle.py and
 * tests/test_golden.py):
 *   orc_kernel          closed forms + hand values                       (pinned)
 *   augmented systems   (extension: r^3 kernel, polynomial block, local eps)
 *                       polynomial reproduction of linear/quadratic fields, the
 *                       interpolation condition at every node, KKT residual and
 *                       numpy/LAPACK solve of the same system, closed-form h_local
 *                                                                        (pinned)
 *   orc_sytrf/sytrs     LAPACK dsytrf/dsytrs via scipy + committed vectors (pinned)
 *   orc_potrf/potrs     LAPACK dpotrf via scipy                          (pinned)
 *   orc_locate          brute-force interval search on tiny grids        (pinned)
 *   orc_cell_sort       numpy stable argsort                             (pinned)
 *   patch membership    definition by brute force (cells touching node)  (pinned)
 *   weights             brute-force O(P) sum == stencil sum, sum==1 case (pinned)
 *   evaluation          single-patch reproduction at nodes, notebook KAT (pinned)
 *   RK1/RK2/RK4/DOPRI   analytic rotation, tableau order conditions      (pinned)
 *   CFL dt / DOPRI dt   closed forms                                     (pinned)
 *   PETSc-defined parts (DMLocatePoints tie rule at x==L, DMSwarmSort
 *   within-cell order, DMDAGetElements Q1 corner order, DMSwarmMigrate removal)
 *   follow the documented choice: parity unpinned against PETSc itself.
 */
#ifndef LAGRBF_ORACLE_H
#define LAGRBF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KernelType, include/kernels/kernels_rbf.h:34-42 */
enum { ORC_KERNEL_LINEAR = 0, ORC_KERNEL_WENDLAND_C0, ORC_KERNEL_WENDLAND_C2,
       ORC_KERNEL_BUMP, ORC_KERNEL_MQ, ORC_KERNEL_IMQ, ORC_KERNEL_GAUSSIAN,
       /* extension, no reference counterpart: polyharmonic spline phi = (eps r)^3 */
       ORC_KERNEL_PHS3 };
/* ASolveMethod, include/lagrangian_rbf.h:91-95 */
enum { ORC_A_FACTOR = 0, ORC_A_KSP, ORC_A_KSP_ASMB };
/* AdvectionScheme, include/advection/advection_ctx.h:34-39 */
enum { ORC_RK1 = 0, ORC_RK2, ORC_RK4, ORC_DOPRI };

typedef struct {
  int    dim;             /* -pu_mesh_dim */
  int    n[3];            /* -pu_mesh_n   */
  double O[3], L[3];      /* -pu_mesh_O / -pu_mesh_L (L is the max corner) */
  double overlap;         /* -pu_mesh_overlap */
  double shape_factor;    /* -pu_mesh_shape_factor */
  int    kernel_type;     /* -kernel_type */
  int    solve_method;    /* -solve_method */
  int    n_ifields;       /* -n_interpolation_fields */
  double rmin, rmax;      /* -rmin / -rmax */
  double cfl_fac;         /* -cfl_fac */
  int    advection;       /* -advection */
  int    scheme;          /* -advection_scheme */
  double dt;              /* -dt */
  double dopri_dt_fac, dopri_dt_fac_min, dopri_dt_fac_max, dopri_atol, dopri_rtol;
  int    weights_brute;   /* 1: O(Np*P) Shepard sum as swarm_interp.c:335-360 */
  int    nthreads;        /* OpenMP threads (0 = runtime default) */
  int    setup_all;       /* 1: every patch is set up (serving datasets), not only the
                             corners of particle cells */
  /* Extensions (no reference counterpart; the reference solves pure RBF systems,
   * python_api/lagrbfpy/interpolation.py:43-45): */
  int    poly_degree;     /* -lagrbf_poly_degree: -1 none, 0 constant, 1 linear, 2 quadratic
                             augmentation of every patch system ([A P; P^T 0]) */
  double local_eps;       /* -lagrbf_local_eps c: shape parameter of patch p is c / h_local(p),
                             h_local = (patch cells' volume / members)^(1/d); 0: global eps_r */
  int    weights_corners; /* -lagrbf_weights corners: Shepard weights normalised over the cell's
                             2^d corners instead of every PU node (swarm_interp.c:327-372) */
} orc_config;

typedef struct orc_ctx orc_ctx;
typedef int (*orc_velocity_batch_fn)(int32_t n, int32_t dim, double *X, double *V, void *user);

void     orc_config_default(orc_config *cfg);

/* ---- pure building blocks ---- */
double   orc_kernel(int ktype, double r, double eps);
/* number of polynomial terms of degree <= deg in dim dimensions (0 for deg < 0) */
int      orc_poly_terms(int dim, int deg);
/* monomials 1, x, y, (z), xx, xy, (xz), yy, (yz), (zz) of degree <= deg at xi */
void     orc_poly_eval(int dim, int deg, const double *xi, double *out);
/* degree a patch of n members takes under -lagrbf_poly_degree D: the largest
 * k <= D with terms(k) <= max(1, n / 2); -1 for D < 0 or n == 0 */
int      orc_patch_degree(int dim, int D, int n);
int      orc_sytrf_lower(int n, double *A, int lda, int *ipiv);
void     orc_sytrs_lower(int n, int nrhs, const double *A, int lda, const int *ipiv, double *B, int ldb);
int      orc_potrf_lower(int n, double *A, int lda);
void     orc_potrs_lower(int n, int nrhs, const double *A, int lda, double *B, int ldb);
int      orc_locate(int dim, const int *n, const double *O, const double *L,
                    int npts, const double *coords, int32_t *cellid);
void     orc_cell_sort(int N, const int32_t *cellid, int ncells, int32_t *cell_start, int32_t *cell_nodes);

/* ---- stateful path mirroring the lagrbf_* C ABI ---- */
orc_ctx *orc_create(const orc_config *cfg);
void     orc_destroy(orc_ctx *ctx);
int      orc_fill_swarm(orc_ctx *ctx, int is_data, int32_t n, const double *coords,
                        const double *values, const int64_t *pidx);
int      orc_interpolation(orc_ctx *ctx);
int      orc_output(orc_ctx *ctx, int is_data, int32_t n, double *coords, double *fields,
                    int64_t *pidx, const char *fieldname);
int      orc_reset_coefficients(orc_ctx *ctx);
int      orc_set_velocity_batch(orc_ctx *ctx, orc_velocity_batch_fn fn, void *user);
int      orc_advect_interpolate(orc_ctx *ctx, double t, double t_next);
int      orc_advect_analytical(orc_ctx *ctx, double t, double t_next);
int      orc_advect_steps(orc_ctx *ctx, int32_t nsteps, double dt);
int      orc_advect_adaptive(orc_ctx *ctx, double t0, double t1, int max_attempts, int blend, double tA, double tB,
                             double *counts);

/* ---- introspection (tests) ---- */
int      orc_get_pu(const orc_ctx *ctx, double *eps_w, double *eps_r, double *pu_coords);
int      orc_get_cellids(const orc_ctx *ctx, int is_data, int32_t *out);
int      orc_get_cell_lists(const orc_ctx *ctx, int32_t *cell_start, int32_t *cell_nodes);
int      orc_get_weights(const orc_ctx *ctx, double *out);
int      orc_patch_size(const orc_ctx *ctx, int p);
int      orc_patch_members(const orc_ctx *ctx, int p, int32_t *members);
int      orc_patch_coeffs(const orc_ctx *ctx, int p, double *X);
int      orc_patch_flags(const orc_ctx *ctx, int p, int *setup, int *solved);
/* polynomial part of a patch: returns its term count m and writes beta
 * column-major [nf][m] (NULL: count only); eps (NULL ok): the patch's shape parameter */
int      orc_patch_poly(const orc_ctx *ctx, int p, double *beta, double *eps);
int64_t  orc_removed_count(const orc_ctx *ctx);
double   orc_get_dt(const orc_ctx *ctx);
int      orc_get_alive(const orc_ctx *ctx, int32_t *alive);

#ifdef __cplusplus
}
#endif
#endif
