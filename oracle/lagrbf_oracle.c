/*
 * lagrbf_oracle.c -- FP64 CPU restatement of LagRBF's PU-RBF interpolation and
 * Lagrangian advection hot path (single MPI rank semantics).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the timed CPU
 * baseline ("port").  Nothing in the product links or loads this file.
 *
 * Reference paths below are relative to the LagRBF source tree.
 */
#include "lagrbf_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* PETSc error codes the reference returns for the same conditions */
#define ERR_MEM        55
#define ERR_SUP        56
#define ERR_ARG_SIZ    60
#define ERR_ARG_WRONG  62
#define ERR_ARG_RANGE  63
#define ERR_WRONGSTATE 73
#define ERR_LIB        76
#define ERR_ARG_NULL   85

typedef struct {
  int      setup, solved;
  int      n;          /* ndata */
  int32_t *members;    /* data_idx, partition.c:337-371 */
  double  *A;          /* factored n x n (column-major, lower used) */
  int     *ipiv;
  double  *X;          /* coefficients, column-major [nf][n] as coeffs[f*ndata+nd] */
  int      nf;
  /* extensions: polynomial block of m terms (degree deg), its coefficients
   * column-major [nf][m], and the patch's shape parameter */
  int      m, deg;
  double  *beta;
  double   eps;
} orc_patch;

struct orc_ctx {
  orc_config cfg;
  int dim, nen, P, C, nc[3];
  double h[3], eps_w, eps_r;
  double *pu;                      /* P*dim node coordinates */
  int32_t (*corner)[8];            /* per cell: Q1 node list */
  /* data swarm */
  int N; double *dx; double *dval; int64_t *dpid; int32_t *dcell;
  int32_t *cell_start, *cell_nodes;
  int data_filled;
  /* particle swarm */
  int Np; double *px; int32_t *pcell; int32_t *alive; double *w; double *fields;
  double *vel; double *x0; double *vadv[7]; int64_t *ppid; int order;
  int points_filled;
  int64_t removed;
  /* patches */
  orc_patch *patch;
  /* advection */
  double adv_dt, dopri_err;
  orc_velocity_batch_fn vfn; void *vuser; int analytical;
};

void orc_config_default(orc_config *c)
{
  memset(c, 0, sizeof(*c));
  c->dim = 3; c->n[0] = c->n[1] = c->n[2] = 4;
  c->L[0] = c->L[1] = c->L[2] = 1.0;
  c->overlap = 1.0; c->shape_factor = 1.0;
  c->kernel_type = ORC_KERNEL_LINEAR;     /* lagrangian_rbf.c:47 */
  c->solve_method = ORC_A_FACTOR;         /* lagrangian_rbf.c:63 */
  c->rmin = 0.0; c->rmax = DBL_MAX;       /* lagrangian_rbf.c:56-57 */
  c->cfl_fac = 0.4;                       /* lagrangian_rbf.c:58 */
  c->scheme = ORC_DOPRI;                  /* lagrangian_rbf.c:186 */
  c->dt = 1e-10;                          /* advection_ctx.c: AdvectionCtxCreate */
  c->dopri_dt_fac = 0.9; c->dopri_dt_fac_min = 1e-2; c->dopri_dt_fac_max = 1e2;  /* dormand_prince.c:318-322 */
  c->dopri_atol = 1e-3; c->dopri_rtol = 1e-3;
  c->poly_degree = -1;                    /* extension: no augmentation, as the reference */
  c->local_eps = 0.0;
}

/* ------------------------------------------------------------------------ */
/* kernels_rbf.c:31-91                                                        */
double orc_kernel(int ktype, double r, double eps)
{
  double u;
  switch (ktype) {
  case ORC_KERNEL_LINEAR:     return eps * r;
  case ORC_KERNEL_WENDLAND_C0: u = 1.0 - eps * r; if (u < 0.0) return 0.0; return u * u;
  case ORC_KERNEL_WENDLAND_C2: u = 1.0 - eps * r; if (u < 0.0) return 0.0; return pow(u, 4) * (5.0 - 4.0 * u);
  case ORC_KERNEL_BUMP:       u = 1.0 - eps * r; if (u < 0.0) return 0.0; return exp(-1.0 / (u * (2.0 - u)));
  case ORC_KERNEL_MQ:         u = eps * r; return sqrt(1.0 + u * u);
  case ORC_KERNEL_IMQ:        u = eps * r; return 1.0 / sqrt(1.0 + u * u);
  case ORC_KERNEL_GAUSSIAN:   u = eps * r; return exp(-u * u);
  case ORC_KERNEL_PHS3:       u = eps * r; return u * u * u;   /* extension */
  default: return NAN;
  }
}

/* ------------------------------------------------------------------------ */
/* Polynomial augmentation (extension; the reference has none,
 * python_api/lagrbfpy/interpolation.py:43-45).  The patch system is the saddle
 * point [A P; P^T 0] [c; beta] = [f; 0] with P_ik = p_k(xi_i), xi = (x - x_p) / h
 * the member shifted to the patch's PU node x_p and scaled by the largest cell
 * width h.  The interpolant of patch p is s_p(x) = sum_j c_j phi(|x - x_j|) +
 * sum_k beta_k p_k(xi(x)).                                                    */
int orc_poly_terms(int dim, int deg)
{
  if (deg < 0) return 0;
  if (deg == 0) return 1;
  if (deg == 1) return dim + 1;
  return (dim + 1) * (dim + 2) / 2;
}

void orc_poly_eval(int dim, int deg, const double *xi, double *out)
{
  int a, b, k = 0;
  if (deg < 0) return;
  out[k++] = 1.0;
  if (deg >= 1) for (a = 0; a < dim; a++) out[k++] = xi[a];
  if (deg >= 2)
    for (a = 0; a < dim; a++)
      for (b = a; b < dim; b++) out[k++] = xi[a] * xi[b];
}

int orc_patch_degree(int dim, int D, int n)
{
  int k, lim = n / 2 > 1 ? n / 2 : 1;
  if (D < 0 || n <= 0) return -1;
  for (k = D; k >= 0; k--) if (orc_poly_terms(dim, k) <= lim) return k;
  return -1;
}

/* ------------------------------------------------------------------------ */
/* Bunch-Kaufman LDL^T, lower storage: the LAPACK dsytf2 algorithm that
 * MatCholeskyFactor_SeqDense reaches through dsytrf when no SPD flag is set
 * (partition_factor.c:37-55, partition_io.c:114-141, petsc_dense.c:64-69).
 * Column-major, 0-based; ipiv follows LAPACK's 1-based sign convention.      */
static int idamax(int n, const double *x, int inc)
{
  int i, im = 0; double m = -1.0;
  for (i = 0; i < n; i++) { double a = fabs(x[(size_t)i * inc]); if (a > m) { m = a; im = i; } }
  return im;
}
#define AE(i, j) A[(size_t)(j) * lda + (i)]
int orc_sytrf_lower(int n, double *A, int lda, int *ipiv)
{
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  int info = 0, k = 0;
  while (k < n) {
    int kstep = 1, kp, imax = k, jmax, i, j;
    double absakk = fabs(AE(k, k)), colmax = 0.0, rowmax;
    if (k < n - 1) { imax = k + 1 + idamax(n - k - 1, &AE(k + 1, k), 1); colmax = fabs(AE(imax, k)); }
    if (fmax(absakk, colmax) == 0.0 || isnan(absakk)) {
      if (info == 0) info = k + 1;
      kp = k;
    } else {
      if (absakk >= alpha * colmax) {
        kp = k;
      } else {
        jmax = k + idamax(imax - k, &AE(imax, k), lda);
        rowmax = fabs(AE(imax, jmax));
        if (imax < n - 1) {
          jmax = imax + 1 + idamax(n - imax - 1, &AE(imax + 1, imax), 1);
          rowmax = fmax(rowmax, fabs(AE(jmax, imax)));
        }
        if (absakk >= alpha * colmax * (colmax / rowmax)) kp = k;
        else if (fabs(AE(imax, imax)) >= alpha * rowmax) kp = imax;
        else { kp = imax; kstep = 2; }
      }
      {
        int kk = k + kstep - 1;
        if (kp != kk) {
          double t;
          for (i = kp + 1; i < n; i++) { t = AE(i, kk); AE(i, kk) = AE(i, kp); AE(i, kp) = t; }
          for (j = kk + 1; j < kp; j++) { t = AE(j, kk); AE(j, kk) = AE(kp, j); AE(kp, j) = t; }
          t = AE(kk, kk); AE(kk, kk) = AE(kp, kp); AE(kp, kp) = t;
          if (kstep == 2) { t = AE(k + 1, k); AE(k + 1, k) = AE(kp, k); AE(kp, k) = t; }
        }
      }
      if (kstep == 1) {
        if (k < n - 1) {
          double d11 = 1.0 / AE(k, k);
          for (j = k + 1; j < n; j++) {
            double t = -d11 * AE(j, k);
            for (i = j; i < n; i++) AE(i, j) += AE(i, k) * t;
          }
          for (i = k + 1; i < n; i++) AE(i, k) *= d11;
        }
      } else {
        if (k < n - 2) {
          double d21 = AE(k + 1, k);
          double d11 = AE(k + 1, k + 1) / d21;
          double d22 = AE(k, k) / d21;
          double t = 1.0 / (d11 * d22 - 1.0);
          d21 = t / d21;
          for (j = k + 2; j < n; j++) {
            double wk = d21 * (d11 * AE(j, k) - AE(j, k + 1));
            double wkp1 = d21 * (d22 * AE(j, k + 1) - AE(j, k));
            for (i = j; i < n; i++) AE(i, j) = AE(i, j) - AE(i, k) * wk - AE(i, k + 1) * wkp1;
            AE(j, k) = wk; AE(j, k + 1) = wkp1;
          }
        }
      }
    }
    if (kstep == 1) ipiv[k] = kp + 1;
    else { ipiv[k] = -(kp + 1); ipiv[k + 1] = -(kp + 1); }
    k += kstep;
  }
  return info;
}

/* LAPACK dsytrs, lower (petsc_dense.c:66) */
void orc_sytrs_lower(int n, int nrhs, const double *A, int lda, const int *ipiv, double *B, int ldb)
{
  int k, i, r;
#define BE(i, j) B[(size_t)(j) * ldb + (i)]
  k = 0;
  while (k < n) {
    if (ipiv[k] > 0) {
      int kp = ipiv[k] - 1;
      for (r = 0; r < nrhs; r++) {
        if (kp != k) { double t = BE(k, r); BE(k, r) = BE(kp, r); BE(kp, r) = t; }
        for (i = k + 1; i < n; i++) BE(i, r) -= AE(i, k) * BE(k, r);
        BE(k, r) *= 1.0 / AE(k, k);
      }
      k += 1;
    } else {
      int kp = -ipiv[k] - 1;
      double akm1k = AE(k + 1, k), akm1 = AE(k, k) / akm1k, ak = AE(k + 1, k + 1) / akm1k;
      double denom = akm1 * ak - 1.0;
      for (r = 0; r < nrhs; r++) {
        double bkm1, bk;
        if (kp != k + 1) { double t = BE(k + 1, r); BE(k + 1, r) = BE(kp, r); BE(kp, r) = t; }
        for (i = k + 2; i < n; i++) BE(i, r) -= AE(i, k) * BE(k, r);
        for (i = k + 2; i < n; i++) BE(i, r) -= AE(i, k + 1) * BE(k + 1, r);
        bkm1 = BE(k, r) / akm1k; bk = BE(k + 1, r) / akm1k;
        BE(k, r) = (ak * bkm1 - bk) / denom;
        BE(k + 1, r) = (akm1 * bk - bkm1) / denom;
      }
      k += 2;
    }
  }
  k = n - 1;
  while (k >= 0) {
    if (ipiv[k] > 0) {
      int kp = ipiv[k] - 1;
      for (r = 0; r < nrhs; r++) {
        double s = 0.0;
        for (i = k + 1; i < n; i++) s += BE(i, r) * AE(i, k);
        BE(k, r) -= s;
        if (kp != k) { double t = BE(k, r); BE(k, r) = BE(kp, r); BE(kp, r) = t; }
      }
      k -= 1;
    } else {
      int kp = -ipiv[k] - 1;
      for (r = 0; r < nrhs; r++) {
        double s = 0.0, s1 = 0.0;
        for (i = k + 1; i < n; i++) s += BE(i, r) * AE(i, k);
        for (i = k + 1; i < n; i++) s1 += BE(i, r) * AE(i, k - 1);
        BE(k, r) -= s; BE(k - 1, r) -= s1;
        if (kp != k) { double t = BE(k, r); BE(k, r) = BE(kp, r); BE(kp, r) = t; }
      }
      k -= 2;
    }
  }
#undef BE
}

/* LAPACK dpotrf/dpotrs, lower (petsc_dense.c:48-53, the SPD branch) */
int orc_potrf_lower(int n, double *A, int lda)
{
  int j, k, i;
  for (j = 0; j < n; j++) {
    double d = AE(j, j);
    for (k = 0; k < j; k++) d -= AE(j, k) * AE(j, k);
    if (!(d > 0.0)) return j + 1;
    d = sqrt(d);
    AE(j, j) = d;
    for (i = j + 1; i < n; i++) {
      double s = AE(i, j);
      for (k = 0; k < j; k++) s -= AE(i, k) * AE(j, k);
      AE(i, j) = s / d;
    }
  }
  return 0;
}
void orc_potrs_lower(int n, int nrhs, const double *A, int lda, double *B, int ldb)
{
  int r, i, k;
  for (r = 0; r < nrhs; r++) {
    double *b = B + (size_t)r * ldb;
    for (i = 0; i < n; i++) { double s = b[i]; for (k = 0; k < i; k++) s -= AE(i, k) * b[k]; b[i] = s / AE(i, i); }
    for (i = n - 1; i >= 0; i--) { double s = b[i]; for (k = i + 1; k < n; k++) s -= AE(k, i) * b[k]; b[i] = s / AE(i, i); }
  }
}
#undef AE

/* ------------------------------------------------------------------------ */
/* Particle -> cell location: DMLocatePoints on the uniform Q1 DMDA
 * (swarm_utils.c:995-1030).  Cell index is lexicographic, i fastest.  A point
 * with O <= x <= L on every axis is found; x == L belongs to the last cell;
 * anything else (incl. NaN) is DMLOCATEPOINT_POINT_NOT_FOUND = -1.           */
static int32_t locate_point(int dim, const int *n, const double *O, const double *L, const double *h, const double *x)
{
  int d; int32_t cell = 0, stride = 1;
  for (d = 0; d < dim; d++) {
    double xd = x[d], t; int i, nc = n[d] - 1;
    if (!(xd >= O[d] && xd <= L[d])) return -1;
    t = (xd - O[d]) / h[d];
    i = (int)floor(t);
    if (i >= nc) i = nc - 1;
    if (i < 0) i = 0;
    cell += i * stride;
    stride *= nc;
  }
  return cell;
}

int orc_locate(int dim, const int *n, const double *O, const double *L, int npts, const double *coords, int32_t *cellid)
{
  double h[3]; int d, p, nf = 0;
  for (d = 0; d < dim; d++) h[d] = (L[d] - O[d]) / (double)(n[d] - 1);
  for (p = 0; p < npts; p++) { cellid[p] = locate_point(dim, n, O, L, h, coords + (size_t)p * dim); if (cellid[p] < 0) nf++; }
  return nf;
}

/* DMSwarmSortGetAccess / GetPointsPerCell (cell_info.c:468-501, partition.c:350-368):
 * per-cell node lists, ascending node index inside a cell (stable counting sort). */
void orc_cell_sort(int N, const int32_t *cellid, int ncells, int32_t *cell_start, int32_t *cell_nodes)
{
  int i, c;
  int32_t *fill = (int32_t *)calloc((size_t)ncells + 1, sizeof(int32_t));
  for (c = 0; c <= ncells; c++) cell_start[c] = 0;
  for (i = 0; i < N; i++) if (cellid[i] >= 0) cell_start[cellid[i] + 1]++;
  for (c = 0; c < ncells; c++) cell_start[c + 1] += cell_start[c];
  for (i = 0; i < N; i++) { c = cellid[i]; if (c >= 0) cell_nodes[cell_start[c] + fill[c]++] = i; }
  free(fill);
}

/* ------------------------------------------------------------------------ */
static void cell_ijk(const orc_ctx *ctx, int32_t cell, int *ijk)
{
  ijk[0] = cell % ctx->nc[0];
  ijk[1] = (cell / ctx->nc[0]) % ctx->nc[1];
  ijk[2] = ctx->dim == 3 ? cell / (ctx->nc[0] * ctx->nc[1]) : 0;
}

/* DMDAGetElements Q1 connectivity of the single-rank DMDA (used as elidx[nen*cell+k]) */
static void build_corners(orc_ctx *ctx)
{
  int c;
  const int n0 = ctx->cfg.n[0], n1 = ctx->cfg.n[1];
  ctx->corner = (int32_t (*)[8])calloc((size_t)ctx->C, sizeof(int32_t[8]));
  for (c = 0; c < ctx->C; c++) {
    int ijk[3], i, j, k; cell_ijk(ctx, c, ijk); i = ijk[0]; j = ijk[1]; k = ijk[2];
    if (ctx->dim == 2) {
      ctx->corner[c][0] = i + j * n0;           ctx->corner[c][1] = (i + 1) + j * n0;
      ctx->corner[c][2] = (i + 1) + (j + 1) * n0; ctx->corner[c][3] = i + (j + 1) * n0;
    } else {
      int s = n0 * n1;
      ctx->corner[c][0] = i + j * n0 + k * s;           ctx->corner[c][1] = (i + 1) + j * n0 + k * s;
      ctx->corner[c][2] = (i + 1) + (j + 1) * n0 + k * s; ctx->corner[c][3] = i + (j + 1) * n0 + k * s;
      ctx->corner[c][4] = i + j * n0 + (k + 1) * s;     ctx->corner[c][5] = (i + 1) + j * n0 + (k + 1) * s;
      ctx->corner[c][6] = (i + 1) + (j + 1) * n0 + (k + 1) * s; ctx->corner[c][7] = i + (j + 1) * n0 + (k + 1) * s;
    }
  }
}

orc_ctx *orc_create(const orc_config *cfg)
{
  orc_ctx *ctx; int d, p; double box = 0.0, ov, radius;
  if (!cfg || cfg->dim < 2 || cfg->dim > 3) return NULL;
  if (cfg->shape_factor <= 0.0) return NULL;
  if (cfg->kernel_type < 0 || cfg->kernel_type > ORC_KERNEL_PHS3) return NULL;
  if (cfg->poly_degree < -1 || cfg->poly_degree > 2) return NULL;
  if (!(cfg->local_eps >= 0.0)) return NULL;
  if (cfg->solve_method < 0 || cfg->solve_method > 2) return NULL;
  for (d = 0; d < cfg->dim; d++) if (cfg->n[d] < 2 || !(cfg->L[d] > cfg->O[d])) return NULL;
  ctx = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  ctx->cfg = *cfg;
  ctx->dim = cfg->dim; ctx->nen = 1 << cfg->dim;
  ctx->P = 1; ctx->C = 1;
  for (d = 0; d < ctx->dim; d++) {
    ctx->P *= cfg->n[d]; ctx->nc[d] = cfg->n[d] - 1; ctx->C *= ctx->nc[d];
    ctx->h[d] = (cfg->L[d] - cfg->O[d]) / (double)(cfg->n[d] - 1);
    if (cfg->L[d] - cfg->O[d] > box) box = cfg->L[d] - cfg->O[d];
  }
  /* partition_of_unity.c:322-355 */
  ov = cfg->overlap * box;
  radius = ov / pow((double)ctx->P, 1.0 / ctx->dim);
  ctx->eps_w = 1.0 / radius;
  ctx->eps_r = cfg->shape_factor * ctx->eps_w;
  /* DMDASetUniformCoordinates: x_i = O + i*hx, partition_of_unity.c:143 */
  ctx->pu = (double *)malloc(sizeof(double) * (size_t)ctx->P * ctx->dim);
  for (p = 0; p < ctx->P; p++) {
    int i = p % cfg->n[0], j = (p / cfg->n[0]) % cfg->n[1], k = ctx->dim == 3 ? p / (cfg->n[0] * cfg->n[1]) : 0;
    ctx->pu[(size_t)p * ctx->dim + 0] = cfg->O[0] + i * ctx->h[0];
    ctx->pu[(size_t)p * ctx->dim + 1] = cfg->O[1] + j * ctx->h[1];
    if (ctx->dim == 3) ctx->pu[(size_t)p * ctx->dim + 2] = cfg->O[2] + k * ctx->h[2];
  }
  build_corners(ctx);
  ctx->patch = (orc_patch *)calloc((size_t)ctx->P, sizeof(orc_patch));
  ctx->order = 0;
  if (cfg->advection) {
    static const int ord[4] = {1, 2, 4, 7};   /* Advection_*Create */
    if (cfg->scheme < 0 || cfg->scheme > 3) { orc_destroy(ctx); return NULL; }
    ctx->order = ord[cfg->scheme];
    ctx->adv_dt = cfg->dt;
  }
  ctx->cell_start = (int32_t *)calloc((size_t)ctx->C + 1, sizeof(int32_t));
  return ctx;
}

static void patch_clear(orc_patch *p)
{
  free(p->members); free(p->A); free(p->ipiv); free(p->X); free(p->beta);
  memset(p, 0, sizeof(*p));
}

static void free_particles(orc_ctx *ctx)
{
  int s;
  free(ctx->px); free(ctx->pcell); free(ctx->alive); free(ctx->w); free(ctx->fields);
  free(ctx->vel); free(ctx->x0); free(ctx->ppid);
  for (s = 0; s < 7; s++) { free(ctx->vadv[s]); ctx->vadv[s] = NULL; }
  ctx->px = NULL; ctx->pcell = NULL; ctx->alive = NULL; ctx->w = NULL; ctx->fields = NULL;
  ctx->vel = NULL; ctx->x0 = NULL; ctx->ppid = NULL; ctx->Np = 0;
}

static void free_data(orc_ctx *ctx)
{
  free(ctx->dx); free(ctx->dval); free(ctx->dpid); free(ctx->dcell); free(ctx->cell_nodes);
  ctx->dx = NULL; ctx->dval = NULL; ctx->dpid = NULL; ctx->dcell = NULL; ctx->cell_nodes = NULL;
  ctx->N = 0;
  memset(ctx->cell_start, 0, sizeof(int32_t) * ((size_t)ctx->C + 1));
}

void orc_destroy(orc_ctx *ctx)
{
  int p;
  if (!ctx) return;
  for (p = 0; p < ctx->P; p++) patch_clear(&ctx->patch[p]);
  free_particles(ctx); free_data(ctx);
  free(ctx->patch); free(ctx->pu); free(ctx->corner); free(ctx->cell_start);
  free(ctx);
}

/* ------------------------------------------------------------------------ */
/* Shepard weights, swarm_interp.c:283-380 (scalar order).  The sum runs over
 * every PU node whose Wendland-C2 weight can be non-zero, in ascending node
 * index; the omitted nodes contribute exact zeros, so the stencil sum equals the
 * brute-force O(P) sum bit for bit.                                          */
static void weights_point(const orc_ctx *ctx, const double *x, int32_t cell, double *w)
{
  const int dim = ctx->dim, nen = ctx->nen;
  const int32_t *cn = ctx->corner[cell];
  double sum = 0.0; int k;
  for (k = 0; k < nen; k++) w[k] = 0.0;
  if (ctx->cfg.weights_brute) {
    int p;
    for (p = 0; p < ctx->P; p++) {
      const double *c = ctx->pu + (size_t)p * dim; double dist = 0.0, wt; int d;
      for (d = 0; d < dim; d++) dist += (x[d] - c[d]) * (x[d] - c[d]);
      dist = sqrt(dist);
      wt = orc_kernel(ORC_KERNEL_WENDLAND_C2, dist, ctx->eps_w);
      sum += wt;
      for (k = 0; k < nen; k++) if (cn[k] == p) { w[k] = wt; break; }
    }
  } else {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, d, i, j, kk;
    double R = 1.0 / ctx->eps_w;
    for (d = 0; d < dim; d++) {
      double a = floor((x[d] - ctx->cfg.O[d] - R) / ctx->h[d]) - 1.0;
      double b = ceil((x[d] - ctx->cfg.O[d] + R) / ctx->h[d]) + 1.0;
      lo[d] = a < 0.0 ? 0 : (a > ctx->cfg.n[d] - 1 ? ctx->cfg.n[d] - 1 : (int)a);
      hi[d] = b < 0.0 ? 0 : (b > ctx->cfg.n[d] - 1 ? ctx->cfg.n[d] - 1 : (int)b);
    }
    for (kk = lo[2]; kk <= hi[2]; kk++)
      for (j = lo[1]; j <= hi[1]; j++)
        for (i = lo[0]; i <= hi[0]; i++) {
          int p = i + j * ctx->cfg.n[0] + kk * ctx->cfg.n[0] * ctx->cfg.n[1];
          const double *c = ctx->pu + (size_t)p * dim; double dist = 0.0, wt;
          for (d = 0; d < dim; d++) dist += (x[d] - c[d]) * (x[d] - c[d]);
          dist = sqrt(dist);
          wt = orc_kernel(ORC_KERNEL_WENDLAND_C2, dist, ctx->eps_w);
          sum += wt;
          for (k = 0; k < nen; k++) if (cn[k] == p) { w[k] = wt; break; }
        }
  }
  if (ctx->cfg.weights_corners) {   /* extension: normalised over the used corners */
    sum = 0.0;
    for (k = 0; k < nen; k++) sum += w[k];
  }
  for (k = 0; k < nen; k++) w[k] = sum > 0.0 ? w[k] / sum : 0.0;
}

static void set_weights(orc_ctx *ctx)
{
  int p;
#pragma omp parallel for schedule(static)
  for (p = 0; p < ctx->Np; p++)
    if (ctx->alive[p]) weights_point(ctx, ctx->px + (size_t)p * ctx->dim, ctx->pcell[p], ctx->w + (size_t)p * ctx->nen);
}

/* DMSwarmMigrate(swi, PETSC_TRUE) on one rank: re-locate, remove points that
 * left every cell (runge_kutta_4.c:78 etc.).  Removal = the point is dropped
 * from every later computation and from the pid-keyed output.               */
static void migrate(orc_ctx *ctx)
{
  int p; int64_t rem = 0;
#pragma omp parallel for schedule(static) reduction(+:rem)
  for (p = 0; p < ctx->Np; p++) {
    if (!ctx->alive[p]) continue;
    ctx->pcell[p] = locate_point(ctx->dim, ctx->cfg.n, ctx->cfg.O, ctx->cfg.L, ctx->h, ctx->px + (size_t)p * ctx->dim);
    if (ctx->pcell[p] < 0) { ctx->alive[p] = 0; rem++; }
  }
  ctx->removed += rem;
}

/* ------------------------------------------------------------------------ */
/* Swarm filling, interface_utils.c:68-165 */
static int check_bbox(const orc_ctx *ctx, int n, const double *coords)
{
  int p, d;
  for (p = 0; p < n; p++)
    for (d = 0; d < ctx->dim; d++) {
      double x = coords[(size_t)p * ctx->dim + d];
      /* swarm_utils.c:862-870; NaN is rejected as well */
      if (!(x >= ctx->cfg.O[d] && x <= ctx->cfg.L[d])) return ERR_ARG_WRONG;
    }
  return 0;
}

int orc_fill_swarm(orc_ctx *ctx, int is_data, int32_t n, const double *coords, const double *values, const int64_t *pidx)
{
  int i, err; const int dim = ctx->dim, nf = ctx->cfg.n_ifields;
  if (n < 0) return ERR_ARG_RANGE;
  if (n > 0 && !coords) return ERR_ARG_NULL;
  if ((err = check_bbox(ctx, n, coords))) return err;
  if (is_data) {
    free_data(ctx);
    ctx->N = n;
    ctx->dx = (double *)malloc(sizeof(double) * ((size_t)n * dim + 1));
    ctx->dval = (double *)calloc((size_t)n * (nf > 0 ? nf : 1) + 1, sizeof(double));
    ctx->dpid = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    ctx->dcell = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    ctx->cell_nodes = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    memcpy(ctx->dx, coords, sizeof(double) * (size_t)n * dim);
    if (values && nf > 0) memcpy(ctx->dval, values, sizeof(double) * (size_t)n * nf);
    for (i = 0; i < n; i++) ctx->dpid[i] = pidx ? pidx[i] : (int64_t)i + 1;
    orc_locate(dim, ctx->cfg.n, ctx->cfg.O, ctx->cfg.L, n, ctx->dx, ctx->dcell);
    orc_cell_sort(n, ctx->dcell, ctx->C, ctx->cell_start, ctx->cell_nodes);
    ctx->data_filled = 1;
  } else {
    int s;
    free_particles(ctx);
    ctx->Np = n;
    ctx->px = (double *)malloc(sizeof(double) * ((size_t)n * dim + 1));
    ctx->pcell = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    ctx->alive = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    ctx->w = (double *)calloc((size_t)n * ctx->nen + 1, sizeof(double));
    ctx->fields = (double *)calloc((size_t)n * (nf > 0 ? nf : 1) + 1, sizeof(double));
    ctx->ppid = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(ctx->px, coords, sizeof(double) * (size_t)n * dim);
    if (values && nf > 0) memcpy(ctx->fields, values, sizeof(double) * (size_t)n * nf);
    for (i = 0; i < n; i++) { ctx->ppid[i] = pidx ? pidx[i] : (int64_t)i + 1; ctx->alive[i] = 1; }
    if (ctx->cfg.advection) {
      ctx->vel = (double *)calloc((size_t)n * dim + 1, sizeof(double));
      ctx->x0 = (double *)calloc((size_t)n * dim + 1, sizeof(double));
      for (s = 0; s < ctx->order; s++) ctx->vadv[s] = (double *)calloc((size_t)n * dim + 1, sizeof(double));
    }
    orc_locate(dim, ctx->cfg.n, ctx->cfg.O, ctx->cfg.L, n, ctx->px, ctx->pcell);
    set_weights(ctx);
    ctx->points_filled = 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Patch membership: partition.c:54-87 (inverse map, ascending cell index) and
 * partition.c:337-371 (concatenate each cell's node list).                  */
static int patch_fill_members(orc_ctx *ctx, int p, int32_t *out)
{
  int a[3] = {0, 0, 0}, cnt = 0, i, j, k;
  const int *n = ctx->cfg.n;
  a[0] = p % n[0]; a[1] = (p / n[0]) % n[1]; a[2] = ctx->dim == 3 ? p / (n[0] * n[1]) : 0;
  for (k = (ctx->dim == 3 ? a[2] - 1 : 0); k <= (ctx->dim == 3 ? a[2] : 0); k++) {
    if (ctx->dim == 3 && (k < 0 || k >= ctx->nc[2])) continue;
    for (j = a[1] - 1; j <= a[1]; j++) {
      if (j < 0 || j >= ctx->nc[1]) continue;
      for (i = a[0] - 1; i <= a[0]; i++) {
        int c, s;
        if (i < 0 || i >= ctx->nc[0]) continue;
        c = i + j * ctx->nc[0] + k * ctx->nc[0] * ctx->nc[1];
        for (s = ctx->cell_start[c]; s < ctx->cell_start[c + 1]; s++) { if (out) out[cnt] = ctx->cell_nodes[s]; cnt++; }
      }
    }
  }
  return cnt;
}

/* Number of grid cells touching PU node p (the cells whose nodes form patch p). */
static int patch_ncells(const orc_ctx *ctx, int p)
{
  const int *n = ctx->cfg.n; int a[3], d, cnt = 1;
  a[0] = p % n[0]; a[1] = (p / n[0]) % n[1]; a[2] = ctx->dim == 3 ? p / (n[0] * n[1]) : 0;
  for (d = 0; d < ctx->dim; d++) cnt *= (a[d] > 0 ? 1 : 0) + (a[d] < ctx->nc[d] ? 1 : 0);
  return cnt;
}

/* Shape parameter of patch p: eps_r, or with -lagrbf_local_eps c (extension)
 * c / h_local, h_local = (cells of p x cell volume / members)^(1/d).         */
static double patch_eps(const orc_ctx *ctx, int p, int n)
{
  double vol = 1.0; int d;
  if (!(ctx->cfg.local_eps > 0.0) || n <= 0) return ctx->eps_r;
  for (d = 0; d < ctx->dim; d++) vol *= ctx->h[d];
  vol *= (double)patch_ncells(ctx, p);
  return ctx->cfg.local_eps / pow(vol / (double)n, 1.0 / ctx->dim);
}

/* polynomial scale: 1 / largest cell width */
static double poly_scale(const orc_ctx *ctx)
{
  double hm = ctx->h[0]; int d;
  for (d = 1; d < ctx->dim; d++) if (ctx->h[d] > hm) hm = ctx->h[d];
  return 1.0 / hm;
}

/* PartitionSetUpRBFMatrix_*: assemble (partition_solve.c:47-82) and factor
 * (partition_factor.c:37-55; PETSc SeqDense Cholesky without SPD flag = dsytrf).
 * With a polynomial block (extension) the same dsytrf factors the (n+m) x (n+m)
 * saddle-point matrix [A P; P^T 0]. */
static int patch_setup(orc_ctx *ctx, int p)
{
  orc_patch *pt = &ctx->patch[p]; int n, i, j, d, nt, m;
  const int dim = ctx->dim;
  n = patch_fill_members(ctx, p, NULL);
  pt->n = n;
  pt->members = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
  patch_fill_members(ctx, p, pt->members);
  pt->nf = ctx->cfg.n_ifields;
  pt->eps = patch_eps(ctx, p, n);
  pt->deg = orc_patch_degree(dim, ctx->cfg.poly_degree, n);
  pt->m = m = orc_poly_terms(dim, pt->deg);
  nt = n + m;
  pt->X = (double *)calloc((size_t)n * (pt->nf > 0 ? pt->nf : 1) + 1, sizeof(double));
  pt->beta = (double *)calloc((size_t)m * (pt->nf > 0 ? pt->nf : 1) + 1, sizeof(double));
  pt->A = (double *)malloc(sizeof(double) * ((size_t)nt * nt + 1));
  pt->ipiv = (int *)malloc(sizeof(int) * ((size_t)nt + 1));
  for (i = 0; i < n; i++)
    for (j = i; j < n; j++) {
      double dist = 0.0, v;
      const double *xi = ctx->dx + (size_t)pt->members[i] * dim, *xj = ctx->dx + (size_t)pt->members[j] * dim;
      for (d = 0; d < dim; d++) dist += (xi[d] - xj[d]) * (xi[d] - xj[d]);
      v = orc_kernel(ctx->cfg.kernel_type, sqrt(dist), pt->eps);
      pt->A[(size_t)j * nt + i] = v; pt->A[(size_t)i * nt + j] = v;
    }
  if (m > 0) {
    const double ps = poly_scale(ctx); const double *c = ctx->pu + (size_t)p * dim;
    for (i = 0; i < n; i++) {
      double xi[3], pv[10];
      const double *x = ctx->dx + (size_t)pt->members[i] * dim;
      for (d = 0; d < dim; d++) xi[d] = (x[d] - c[d]) * ps;
      orc_poly_eval(dim, pt->deg, xi, pv);
      for (j = 0; j < m; j++) { pt->A[(size_t)(n + j) * nt + i] = pv[j]; pt->A[(size_t)i * nt + n + j] = pv[j]; }
    }
    for (i = 0; i < m; i++) for (j = 0; j < m; j++) pt->A[(size_t)(n + j) * nt + n + i] = 0.0;
  }
  if (nt > 0) orc_sytrf_lower(nt, pt->A, nt, pt->ipiv);
  pt->setup = 1;
  return 0;
}

/* PartitionSolveLocalCoefficients*: RHS gather (partition_factor.c:57-90) + solve */
static void patch_solve(orc_ctx *ctx, int p)
{
  orc_patch *pt = &ctx->patch[p]; int i, f; const int nf = pt->nf, n = pt->n, m = pt->m, nt = n + m;
  double *B;
  if (n == 0 || nf == 0) { pt->solved = 1; return; }
  if (m == 0) {
    for (i = 0; i < n; i++)
      for (f = 0; f < nf; f++) pt->X[(size_t)f * n + i] = ctx->dval[(size_t)pt->members[i] * nf + f];
    orc_sytrs_lower(n, nf, pt->A, n, pt->ipiv, pt->X, n);
    pt->solved = 1;
    return;
  }
  /* saddle point: right-hand side [f; 0], solution [c; beta] */
  B = (double *)calloc((size_t)nt * nf, sizeof(double));
  for (i = 0; i < n; i++)
    for (f = 0; f < nf; f++) B[(size_t)f * nt + i] = ctx->dval[(size_t)pt->members[i] * nf + f];
  orc_sytrs_lower(nt, nf, pt->A, nt, pt->ipiv, B, nt);
  for (f = 0; f < nf; f++) {
    for (i = 0; i < n; i++) pt->X[(size_t)f * n + i] = B[(size_t)f * nt + i];
    for (i = 0; i < m; i++) pt->beta[(size_t)f * m + i] = B[(size_t)f * nt + n + i];
  }
  free(B);
  pt->solved = 1;
}

/* LagRBF_PartitionsSetUp, partition.c:417-478: only the corner patches of the
 * cells holding a particle are set up (once) and solved (once per reset).   */
static int partitions_setup(orc_ctx *ctx)
{
  int p, k, cnt = 0;
  char *touch = (char *)calloc((size_t)ctx->P, 1);
  int *list = (int *)malloc(sizeof(int) * ((size_t)ctx->P + 1));
  for (p = 0; p < ctx->Np; p++) {
    if (!ctx->alive[p]) continue;
    for (k = 0; k < ctx->nen; k++) touch[ctx->corner[ctx->pcell[p]][k]] = 1;
  }
  if (ctx->cfg.setup_all) memset(touch, 1, (size_t)ctx->P);
  for (p = 0; p < ctx->P; p++) if (touch[p]) list[cnt++] = p;
#pragma omp parallel for schedule(dynamic, 1)
  for (k = 0; k < cnt; k++) {
    int q = list[k];
    if (!ctx->patch[q].setup) patch_setup(ctx, q);
    if (ctx->patch[q].setup && !ctx->patch[q].solved && ctx->patch[q].n > 0) patch_solve(ctx, q);
  }
  free(touch); free(list);
  return 0;
}

/* LagRBF_PartitionsReset, partition.c:494-524 */
int orc_reset_coefficients(orc_ctx *ctx)
{
  int p;
  for (p = 0; p < ctx->P; p++) {
    if (ctx->cfg.solve_method == ORC_A_KSP_ASMB) patch_clear(&ctx->patch[p]);
    else ctx->patch[p].solved = 0;
  }
  return 0;
}

/* PointwiseInterpolation_Scalar, interpolation.c:39-154.  values has nfields
 * columns (nfields <= n_ifields; the first nfields coefficient columns are used). */
static void rbf_interpolation(orc_ctx *ctx, double *values, int nfields)
{
  int np;
  const int dim = ctx->dim, nen = ctx->nen;
#pragma omp parallel for schedule(dynamic, 64)
  for (np = 0; np < ctx->Np; np++) {
    int32_t cell; int ijk[3], d, k, nb_count = 0, f;
    const double *x;
    double interp[64];
    if (!ctx->alive[np]) continue;
    cell = ctx->pcell[np];
    x = ctx->px + (size_t)np * dim;
    /* 3^d neighbourhood population (cell_info.c:468-501); empty => stale values */
    cell_ijk(ctx, cell, ijk);
    {
      int di, dj, dk;
      for (dk = (dim == 3 ? -1 : 0); dk <= (dim == 3 ? 1 : 0); dk++)
        for (dj = -1; dj <= 1; dj++)
          for (di = -1; di <= 1; di++) {
            int i = ijk[0] + di, j = ijk[1] + dj, kk = ijk[2] + dk, c;
            if (i < 0 || i >= ctx->nc[0] || j < 0 || j >= ctx->nc[1]) continue;
            if (dim == 3 && (kk < 0 || kk >= ctx->nc[2])) continue;
            c = i + j * ctx->nc[0] + kk * ctx->nc[0] * ctx->nc[1];
            nb_count += ctx->cell_start[c + 1] - ctx->cell_start[c];
          }
    }
    if (nb_count == 0) continue;
    for (f = 0; f < nfields; f++) values[(size_t)np * nfields + f] = 0.0;
    for (k = 0; k < nen; k++) {
      const orc_patch *pt = &ctx->patch[ctx->corner[cell][k]];
      double wgt; int nd;
      if (pt->n == 0) continue;
      for (f = 0; f < nfields; f++) interp[f] = 0.0;
      for (nd = 0; nd < pt->n; nd++) {
        const double *xj = ctx->dx + (size_t)pt->members[nd] * dim; double dist = 0.0, rbf;
        for (d = 0; d < dim; d++) { double dx = x[d] - xj[d]; dist += dx * dx; }
        rbf = orc_kernel(ctx->cfg.kernel_type, sqrt(dist), pt->eps);
        for (f = 0; f < nfields; f++) interp[f] += rbf * pt->X[(size_t)f * pt->n + nd];
      }
      if (pt->m > 0) {   /* polynomial part (extension) */
        double xi[3], pv[10]; int q;
        const double *c = ctx->pu + (size_t)ctx->corner[cell][k] * dim, ps = poly_scale(ctx);
        for (d = 0; d < dim; d++) xi[d] = (x[d] - c[d]) * ps;
        orc_poly_eval(dim, pt->deg, xi, pv);
        for (f = 0; f < nfields; f++)
          for (q = 0; q < pt->m; q++) interp[f] += pv[q] * pt->beta[(size_t)f * pt->m + q];
      }
      wgt = ctx->w[(size_t)np * nen + k];
      for (f = 0; f < nfields; f++) values[(size_t)np * nfields + f] += interp[f] * wgt;
    }
  }
}

/* Interface_Interpolation, interface_interp.c:83-96 */
int orc_interpolation(orc_ctx *ctx)
{
  if (!ctx->data_filled || !ctx->points_filled) return ERR_ARG_NULL;
  if (ctx->cfg.n_ifields > 64) return ERR_SUP;
  partitions_setup(ctx);
  rbf_interpolation(ctx, ctx->fields, ctx->cfg.n_ifields);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Advection (advection.c, runge_kutta_{1,2,4}.c, dormand_prince.c, time_loop.c) */
static void project_point(double *x, int dim, double rmin, double rmax)
{
  /* AdvectionProjectOntoSphereBoundary_Pointwise, advection.c:472-491 */
  double pos = 0.0, nrm[3]; int d;
  for (d = 0; d < dim; d++) pos += x[d] * x[d];
  if (pos < rmax * rmax && pos > rmin * rmin) return;
  pos = sqrt(pos);
  for (d = 0; d < dim; d++) nrm[d] = x[d] / pos;
  if (pos > rmax) for (d = 0; d < dim; d++) x[d] = rmax * nrm[d];
  else for (d = 0; d < dim; d++) x[d] = rmin * nrm[d];
}

static int eval_velocity(orc_ctx *ctx)
{
  /* AdvectionEvaluateVelocity, advection.c:391-414 */
  if (!ctx->analytical) { rbf_interpolation(ctx, ctx->vel, ctx->dim); return 0; }
  else {
    int p, m = 0, d, err; const int dim = ctx->dim;
    double *X, *V;
    if (!ctx->vfn) return ERR_ARG_NULL;
    X = (double *)malloc(sizeof(double) * ((size_t)ctx->Np * dim + 1));
    V = (double *)malloc(sizeof(double) * ((size_t)ctx->Np * dim + 1));
    for (p = 0; p < ctx->Np; p++) if (ctx->alive[p]) { for (d = 0; d < dim; d++) X[(size_t)m * dim + d] = ctx->px[(size_t)p * dim + d]; m++; }
    err = ctx->vfn(m, dim, X, V, ctx->vuser);
    m = 0;
    for (p = 0; p < ctx->Np; p++) if (ctx->alive[p]) { for (d = 0; d < dim; d++) ctx->vel[(size_t)p * dim + d] = V[(size_t)m * dim + d]; m++; }
    free(X); free(V);
    return err;
  }
}

static int init_step_velocity(orc_ctx *ctx)
{
  /* AdvectionInitializeTimestepVelocity, advection.c:353-373 */
  if (ctx->analytical) return eval_velocity(ctx);
  else {
    int p, d; const int dim = ctx->dim, nf = ctx->cfg.n_ifields;
    for (p = 0; p < ctx->Np; p++)
      if (ctx->alive[p]) for (d = 0; d < dim; d++) ctx->vel[(size_t)p * dim + d] = ctx->fields[(size_t)p * nf + d];
    return 0;
  }
}

static double cfl_dt(const orc_ctx *ctx)
{
  /* AdvectionComputeTimestep_CFL, advection.c:430-470 */
  double dt = DBL_MAX, radius = 1.0 / ctx->eps_w; int p;
  for (p = 0; p < ctx->Np; p++) {
    double nv = 0.0, dl; int d;
    if (!ctx->alive[p]) continue;
    for (d = 0; d < ctx->dim; d++) nv += ctx->vel[(size_t)p * ctx->dim + d] * ctx->vel[(size_t)p * ctx->dim + d];
    nv = sqrt(nv);
    dl = ctx->cfg.cfl_fac * radius / (nv + 1e-16);
    if (dl < dt) dt = dl;
  }
  return dt;
}

static void after_move(orc_ctx *ctx) { migrate(ctx); set_weights(ctx); }

static int stage_eval(orc_ctx *ctx) { after_move(ctx); return eval_velocity(ctx); }

static int advect_rk1(orc_ctx *ctx, double dt)
{
  int p, d; const int dim = ctx->dim;
  for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
    for (d = 0; d < dim; d++) ctx->px[(size_t)p * dim + d] = ctx->px[(size_t)p * dim + d] + dt * ctx->vel[(size_t)p * dim + d];
  after_move(ctx);
  return 0;
}

static int advect_rk2(orc_ctx *ctx, double dt)
{
  int p, d, err; const int dim = ctx->dim; size_t i;
  for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
    for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->x0[i] = ctx->px[i]; ctx->px[i] = ctx->x0[i] + 0.5 * dt * ctx->vel[i]; }
  if ((err = stage_eval(ctx))) return err;
  for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
    for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->px[i] = ctx->x0[i] + dt * ctx->vel[i]; }
  after_move(ctx);
  return 0;
}

static int advect_rk4(orc_ctx *ctx, double dt)
{
  static const double c[3] = {0.5, 0.5, 1.0}, b[4] = {1.0, 2.0, 2.0, 1.0};
  int p, d, s, err; const int dim = ctx->dim; size_t i;
  for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
    for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->x0[i] = ctx->px[i]; ctx->vadv[0][i] = ctx->vel[i]; }
  for (s = 0; s < 3; s++) {
    const double sdt = c[s] * dt;
    for (p = 0; p < ctx->Np; p++) if (ctx->alive[p]) {
      for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->vel[i] = ctx->vadv[s][i]; ctx->px[i] = ctx->x0[i] + sdt * ctx->vel[i]; }
      project_point(ctx->px + (size_t)p * dim, dim, ctx->cfg.rmin, ctx->cfg.rmax);
    }
    if ((err = stage_eval(ctx))) return err;
    for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
      for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->vadv[s + 1][i] = ctx->vel[i]; }
  }
  for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
    for (d = 0; d < dim; d++) {
      double sum = 0.0; i = (size_t)p * dim + d;
      for (s = 0; s < 4; s++) sum += b[s] * ctx->vadv[s][i];
      ctx->px[i] = ctx->x0[i] + dt / 6.0 * sum;
      project_point(ctx->px + (size_t)p * dim, dim, ctx->cfg.rmin, ctx->cfg.rmax);  /* inside the d loop, runge_kutta_4.c:134-141 */
    }
  after_move(ctx);
  return 0;
}

static int advect_dopri(orc_ctx *ctx)
{
  static const double A[6][6] = {
    {1.0 / 5.0, 0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
    {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
  static const double Bs[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
  const int dim = ctx->dim; const double dt = ctx->adv_dt;
  int p, d, s, j, err; size_t i; double max_err = -DBL_MAX, fac_max;
  for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
    for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->x0[i] = ctx->px[i]; ctx->vadv[0][i] = ctx->vel[i]; }
  for (s = 0; s < 6; s++) {
    const int stage = s + 1;
    for (p = 0; p < ctx->Np; p++) if (ctx->alive[p]) {
      for (d = 0; d < dim; d++) {
        i = (size_t)p * dim + d; ctx->vel[i] = 0.0;
        for (j = 0; j < stage; j++) ctx->vel[i] += A[s][j] * ctx->vadv[j][i];
      }
      for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->px[i] = ctx->x0[i] + dt * ctx->vel[i]; }
      project_point(ctx->px + (size_t)p * dim, dim, ctx->cfg.rmin, ctx->cfg.rmax);
    }
    if ((err = stage_eval(ctx))) return err;
    for (p = 0; p < ctx->Np; p++) if (ctx->alive[p])
      for (d = 0; d < dim; d++) { i = (size_t)p * dim + d; ctx->vadv[stage][i] = ctx->vel[i]; }
  }
  /* AdvectionEvaluateError_DOPRI, dormand_prince.c:37-102 */
  for (p = 0; p < ctx->Np; p++) {
    double vlow[3] = {0, 0, 0}, E = 0.0;
    if (!ctx->alive[p]) continue;
    for (j = 0; j < 7; j++) for (d = 0; d < dim; d++) vlow[d] += Bs[j] * ctx->vadv[j][(size_t)p * dim + d];
    for (d = 0; d < dim; d++) {
      i = (size_t)p * dim + d;
      double mv = fmax(fabs(ctx->x0[i]), fabs(ctx->px[i]));
      double tol = ctx->cfg.dopri_atol + ctx->cfg.dopri_rtol * mv;
      double xl = ctx->x0[i] + dt * vlow[d];
      double e = ctx->px[i] - xl;
      E += (e * e) / (tol * tol);
    }
    E = sqrt(1.0 / dim * E);
    if (E > max_err) max_err = E;
  }
  ctx->dopri_err = max_err;
  /* AdvectionComputeAdaptiveTimeStep_DOPRI, dormand_prince.c:179-214 (always accepted) */
  after_move(ctx);
  fac_max = max_err <= 1.0 ? ctx->cfg.dopri_dt_fac_max : 1.0;
  ctx->adv_dt = dt * fmin(fac_max, fmax(ctx->cfg.dopri_dt_fac_min, ctx->cfg.dopri_dt_fac * pow(1.0 / max_err, 1.0 / 5.0)));
  return 0;
}

static int advect_swarm(orc_ctx *ctx, double dt)
{
  switch (ctx->cfg.scheme) {
  case ORC_RK1: return advect_rk1(ctx, dt);
  case ORC_RK2: return advect_rk2(ctx, dt);
  case ORC_RK4: return advect_rk4(ctx, dt);
  case ORC_DOPRI: return advect_dopri(ctx);
  default: return ERR_ARG_RANGE;
  }
}

/* Per-particle adaptive Dormand-Prince 5(4) from t0 to t1, optionally with the
 * velocity blended in time between two snapshots: fields 0..d-1 at tA and
 * d..2d-1 at tB, v = (1 - tau) vA + tau vB, tau = (t - tA) / (tB - tA) at each
 * particle's stage time.  An extension with no reference counterpart (the
 * reference steps the swarm with one dt, dormand_prince.c:216-310); the tableau,
 * error norm (dormand_prince.c:37-102) and controller (dormand_prince.c:179-214)
 * are applied per particle, and an attempt with error norm > 1 is rejected.
 * Particles are advanced attempt by attempt in lockstep, each with its own t and
 * dt; one that reaches t1 stops.  counts: attempts, accepted, rejected, active. */
static void adaptive_eval(orc_ctx *ctx, double *k, double *vals, int blend, const double *tp, const double *dtp,
                          double cst, double tA, double tB)
{
  const int dim = ctx->dim, nv = blend ? 2 * dim : dim;
  int p, d;
  rbf_interpolation(ctx, vals, nv);
  for (p = 0; p < ctx->Np; p++) {
    if (!ctx->alive[p]) continue;
    if (blend) {
      const double t = tp[p] + cst * dtp[p];
      const double tau = (t - tA) / (tB - tA);
      for (d = 0; d < dim; d++)
        k[(size_t)p * dim + d] = (1.0 - tau) * vals[(size_t)p * nv + d] + tau * vals[(size_t)p * nv + dim + d];
    } else {
      for (d = 0; d < dim; d++) k[(size_t)p * dim + d] = vals[(size_t)p * nv + d];
    }
  }
}

int orc_advect_adaptive(orc_ctx *ctx, double t0, double t1, int max_attempts, int blend, double tA, double tB,
                        double *counts)
{
  static const double A[6][6] = {
    {1.0 / 5.0, 0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
    {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
  static const double Cst[6] = {1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
  static const double Bs[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
  const int dim = ctx->dim, Np = ctx->Np, nf = ctx->cfg.n_ifields;
  const size_t n = (size_t)Np * dim + 1;
  double *k[7], *x0, *tp, *dtp, *vals, dt0;
  int32_t *fin, *susp;
  int a, s, j, p, d;
  if (!ctx->data_filled || !ctx->points_filled) return ERR_ARG_NULL;
  if (!(t1 > t0) || max_attempts < 0 || !(ctx->cfg.dt > 0.0)) return ERR_ARG_RANGE;
  if ((blend && nf != 2 * dim) || (!blend && nf < dim)) return ERR_ARG_SIZ;
  for (j = 0; j < 7; j++) k[j] = (double *)calloc(n, sizeof(double));
  x0 = (double *)calloc(n, sizeof(double));
  tp = (double *)calloc((size_t)Np + 1, sizeof(double));
  dtp = (double *)calloc((size_t)Np + 1, sizeof(double));
  fin = (int32_t *)calloc((size_t)Np + 1, sizeof(int32_t));
  susp = (int32_t *)calloc((size_t)Np + 1, sizeof(int32_t));
  vals = (double *)calloc((size_t)Np * (blend ? 2 * dim : dim) + 1, sizeof(double));
  for (j = 0; j < 4; j++) counts[j] = 0.0;
  partitions_setup(ctx);
  dt0 = fmin(ctx->cfg.dt, t1 - t0);
  for (p = 0; p < Np; p++) {
    tp[p] = t0; dtp[p] = dt0;
    if (ctx->alive[p]) for (d = 0; d < dim; d++) x0[(size_t)p * dim + d] = ctx->px[(size_t)p * dim + d];
  }
  adaptive_eval(ctx, k[0], vals, blend, tp, dtp, 0.0, tA, tB);
  for (a = 0; a < max_attempts; a++) {
    double acc = 0, rej = 0, act = 0;
    for (s = 0; s < 6; s++) {
      for (p = 0; p < Np; p++) {
        double *x;
        if (!ctx->alive[p]) continue;
        x = ctx->px + (size_t)p * dim;
        for (d = 0; d < dim; d++) {
          double v = 0.0;
          for (j = 0; j <= s; j++) v += A[s][j] * k[j][(size_t)p * dim + d];
          x[d] = x0[(size_t)p * dim + d] + dtp[p] * v;
        }
        project_point(x, dim, ctx->cfg.rmin, ctx->cfg.rmax);
        /* a trial point of stages 2-6 outside the domain suspends the particle for
         * the rest of the attempt, which is then rejected (below); only the step's
         * end point (stage 7) leaving the domain removes it */
        if (s < 5 && locate_point(dim, ctx->cfg.n, ctx->cfg.O, ctx->cfg.L, ctx->h, x) < 0) {
          susp[p] |= 1;
          ctx->alive[p] = 0;
        }
      }
      after_move(ctx);
      adaptive_eval(ctx, k[s + 1], vals, blend, tp, dtp, Cst[s], tA, tB);
    }
    for (p = 0; p < Np; p++) {
      double E = 0.0, t, g, inner, f, fmx, dn, rest; int ok;
      if (susp[p] & 1) {
        /* forced rejection: back to x0, the step size shrunk by fac_min; the third
         * of the particle's integration (or a step size below 1e-10 max(1, |t1|))
         * means the path leaves the domain: removed.  susp: bit 0 this attempt, the
         * rest the count */
        const int nrej = (susp[p] >> 1) + 1;
        for (d = 0; d < dim; d++) ctx->px[(size_t)p * dim + d] = x0[(size_t)p * dim + d];
        susp[p] = nrej << 1;
        rej += 1;
        dn = dtp[p] * fmin(1.0, ctx->cfg.dopri_dt_fac_min);
        if (nrej > 2 || dn <= 1e-10 * fmax(1.0, fabs(t1))) { susp[p] = 0; ctx->removed += 1; continue; }
        ctx->alive[p] = 1;
        rest = t1 - tp[p];
        if (dn > rest) dn = rest;
        dtp[p] = dn;
        act += 1;
        continue;
      }
      if (!ctx->alive[p]) continue;
      for (d = 0; d < dim; d++) {
        const size_t i = (size_t)p * dim + d;
        double vl = 0.0, mv, tol, xl, e;
        for (j = 0; j < 7; j++) vl += Bs[j] * k[j][i];
        mv = fmax(fabs(x0[i]), fabs(ctx->px[i]));
        tol = ctx->cfg.dopri_atol + ctx->cfg.dopri_rtol * mv;
        xl = x0[i] + dtp[p] * vl;
        e = ctx->px[i] - xl;
        E += (e * e) / (tol * tol);
      }
      E = sqrt(1.0 / dim * E);
      ok = E <= 1.0;
      t = tp[p];
      if (ok) {
        for (d = 0; d < dim; d++) {
          const size_t i = (size_t)p * dim + d;
          x0[i] = ctx->px[i];
          k[0][i] = k[6][i];
        }
        t = t + dtp[p];
        tp[p] = t;
        acc += 1;
      } else {
        for (d = 0; d < dim; d++) ctx->px[(size_t)p * dim + d] = x0[(size_t)p * dim + d];
        rej += 1;
      }
      fmx = ok ? ctx->cfg.dopri_dt_fac_max : 1.0;
      g = ctx->cfg.dopri_dt_fac * pow(1.0 / E, 1.0 / 5.0);
      inner = (ctx->cfg.dopri_dt_fac_min < g) ? g : ctx->cfg.dopri_dt_fac_min;
      f = (fmx < inner) ? fmx : inner;
      dn = dtp[p] * f;
      rest = t1 - t;
      if (rest <= 1e-12 * fmax(1.0, fabs(t1))) {
        ctx->alive[p] = 0;
        fin[p] = 1;
      } else {
        if (dn > rest) dn = rest;
        dtp[p] = dn;
        act += 1;
      }
    }
    counts[0] += 1; counts[1] += acc; counts[2] += rej; counts[3] = act;
    if (act == 0) break;
    partitions_setup(ctx);
  }
  for (p = 0; p < Np; p++) if (fin[p]) ctx->alive[p] = 1;
  for (j = 0; j < 7; j++) free(k[j]);
  free(x0); free(tp); free(dtp); free(fin); free(susp); free(vals);
  return 0;
}

int orc_set_velocity_batch(orc_ctx *ctx, orc_velocity_batch_fn fn, void *user)
{
  if (!ctx->cfg.advection) return ERR_ARG_NULL;
  ctx->vfn = fn; ctx->vuser = user; ctx->analytical = 1;
  return 0;
}

/* InternalForwardAdvectionLoop, time_loop.c:39-100 */
int orc_advect_interpolate(orc_ctx *ctx, double t, double t_next)
{
  int err;
  if (!ctx->cfg.advection || !ctx->points_filled) return ERR_ARG_NULL;
  if (!ctx->analytical && ctx->cfg.n_ifields < ctx->dim) return ERR_ARG_SIZ;
  if ((err = init_step_velocity(ctx))) return err;
  while (t < t_next) {
    double dt;
    if (ctx->cfg.scheme == ORC_DOPRI) { double dc = cfl_dt(ctx); dt = fmin(ctx->adv_dt, dc); }
    else dt = cfl_dt(ctx);
    if (t + dt > t_next) dt = t_next - t;
    ctx->adv_dt = dt;
    if ((err = advect_swarm(ctx, dt))) return err;
    partitions_setup(ctx);
    rbf_interpolation(ctx, ctx->fields, ctx->cfg.n_ifields);
    if ((err = init_step_velocity(ctx))) return err;
    t += dt;
  }
  orc_reset_coefficients(ctx);
  free_data(ctx);   /* CellInfoReset + DMSwarmSetLocalSizes(swarm_data,0,-1), time_loop.c:94-98 */
  ctx->data_filled = 0;
  return 0;
}

/* ForwardAnalyticalAdvection, time_loop.c:176-221 */
int orc_advect_analytical(orc_ctx *ctx, double t, double t_next)
{
  int err;
  if (!ctx->cfg.advection || !ctx->points_filled || !ctx->analytical) return ERR_ARG_NULL;
  if ((err = init_step_velocity(ctx))) return err;
  while (t < t_next) {
    double dt;
    if (ctx->cfg.scheme == ORC_DOPRI) { double dc = cfl_dt(ctx); dt = fmin(ctx->adv_dt, dc); }
    else dt = cfl_dt(ctx);
    if (t + dt > t_next) dt = t_next - t;
    ctx->adv_dt = dt;
    if ((err = advect_swarm(ctx, dt))) return err;
    if ((err = init_step_velocity(ctx))) return err;
    t += dt;
  }
  return 0;
}

/* Fixed-step variant of the internal loop (no CFL, no end-of-call reset): the
 * same per-step sequence as time_loop.c:80-90 with dt given by the caller.  */
int orc_advect_steps(orc_ctx *ctx, int32_t nsteps, double dt)
{
  int err, s;
  if (!ctx->cfg.advection || !ctx->points_filled) return ERR_ARG_NULL;
  if (!ctx->analytical && ctx->cfg.n_ifields < ctx->dim) return ERR_ARG_SIZ;
  if ((err = init_step_velocity(ctx))) return err;
  for (s = 0; s < nsteps; s++) {
    ctx->adv_dt = dt;
    if ((err = advect_swarm(ctx, dt))) return err;
    if (!ctx->analytical) {
      partitions_setup(ctx);
      rbf_interpolation(ctx, ctx->fields, ctx->cfg.n_ifields);
    }
    if ((err = init_step_velocity(ctx))) return err;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* lagrbf_output / AssembleArray_Allreduce, interface_utils.c:188-384 */
int orc_output(orc_ctx *ctx, int is_data, int32_t n, double *coords, double *fields, int64_t *pidx, const char *fieldname)
{
  const int dim = ctx->dim; int p, f, nfl = 0; int Ns = is_data ? ctx->N : ctx->Np;
  const int64_t *pid = is_data ? ctx->dpid : ctx->ppid;
  const double *src = NULL, *xs = is_data ? ctx->dx : ctx->px;
  const int32_t *alive = is_data ? NULL : ctx->alive;
  if (fields) {
    if (!fieldname || !strcmp(fieldname, "fields")) { src = is_data ? ctx->dval : ctx->fields; nfl = ctx->cfg.n_ifields; }
    else if (!strcmp(fieldname, "velocity") && !is_data) { src = ctx->vel; nfl = dim; }
    else if (!strcmp(fieldname, "weights") && !is_data) { src = ctx->w; nfl = ctx->nen; }
    else return ERR_ARG_WRONG;
    if (!src) return ERR_WRONGSTATE;
    memset(fields, 0, sizeof(double) * (size_t)n * nfl);
  }
  if (coords) memset(coords, 0, sizeof(double) * (size_t)n * dim);
  if (pidx) memset(pidx, 0, sizeof(int64_t) * (size_t)n);
  for (p = 0; p < Ns; p++) {
    int64_t g = pid[p] - 1;
    if (alive && !alive[p]) continue;
    if (g < 0 || g >= n) return ERR_ARG_RANGE;
    if (coords) for (f = 0; f < dim; f++) coords[g * dim + f] = xs[(size_t)p * dim + f];
    if (fields) for (f = 0; f < nfl; f++) fields[g * nfl + f] = src[(size_t)p * nfl + f];
    if (pidx) pidx[g] = pid[p];
  }
  return 0;
}

/* ---- introspection ---- */
int orc_get_pu(const orc_ctx *ctx, double *eps_w, double *eps_r, double *pu)
{
  if (eps_w) *eps_w = ctx->eps_w;
  if (eps_r) *eps_r = ctx->eps_r;
  if (pu) memcpy(pu, ctx->pu, sizeof(double) * (size_t)ctx->P * ctx->dim);
  return ctx->P;
}
int orc_get_cellids(const orc_ctx *ctx, int is_data, int32_t *out)
{
  int n = is_data ? ctx->N : ctx->Np;
  if (out) memcpy(out, is_data ? ctx->dcell : ctx->pcell, sizeof(int32_t) * (size_t)n);
  return n;
}
int orc_get_cell_lists(const orc_ctx *ctx, int32_t *cs, int32_t *cn)
{
  if (cs) memcpy(cs, ctx->cell_start, sizeof(int32_t) * ((size_t)ctx->C + 1));
  if (cn && ctx->N) memcpy(cn, ctx->cell_nodes, sizeof(int32_t) * (size_t)ctx->N);
  return ctx->C;
}
int orc_get_weights(const orc_ctx *ctx, double *out)
{
  if (out) memcpy(out, ctx->w, sizeof(double) * (size_t)ctx->Np * ctx->nen);
  return ctx->Np;
}
int orc_patch_size(const orc_ctx *ctx, int p) { return (p < 0 || p >= ctx->P) ? -1 : ctx->patch[p].n; }
int orc_patch_members(const orc_ctx *ctx, int p, int32_t *m)
{
  if (p < 0 || p >= ctx->P) return -1;
  if (ctx->patch[p].setup) { if (m) memcpy(m, ctx->patch[p].members, sizeof(int32_t) * (size_t)ctx->patch[p].n); return ctx->patch[p].n; }
  return patch_fill_members((orc_ctx *)ctx, p, m);
}
int orc_patch_coeffs(const orc_ctx *ctx, int p, double *X)
{
  const orc_patch *pt;
  if (p < 0 || p >= ctx->P) return -1;
  pt = &ctx->patch[p];
  if (X && pt->X) memcpy(X, pt->X, sizeof(double) * (size_t)pt->n * pt->nf);
  return pt->n;
}
int orc_patch_poly(const orc_ctx *ctx, int p, double *beta, double *eps)
{
  const orc_patch *pt;
  if (p < 0 || p >= ctx->P) return -1;
  pt = &ctx->patch[p];
  if (eps) *eps = pt->setup ? pt->eps : patch_eps(ctx, p, patch_fill_members((orc_ctx *)ctx, p, NULL));
  if (beta && pt->beta) memcpy(beta, pt->beta, sizeof(double) * (size_t)pt->m * pt->nf);
  return pt->m;
}
int orc_patch_flags(const orc_ctx *ctx, int p, int *setup, int *solved)
{
  if (p < 0 || p >= ctx->P) return -1;
  if (setup) *setup = ctx->patch[p].setup;
  if (solved) *solved = ctx->patch[p].solved;
  return 0;
}
int64_t orc_removed_count(const orc_ctx *ctx) { return ctx->removed; }
double  orc_get_dt(const orc_ctx *ctx) { return ctx->adv_dt; }
int orc_get_alive(const orc_ctx *ctx, int32_t *a)
{
  if (a) memcpy(a, ctx->alive, sizeof(int32_t) * (size_t)ctx->Np);
  return ctx->Np;
}
