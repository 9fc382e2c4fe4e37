// Host-side launchers of the gfx950 kernels (defined in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "common.hpp"

namespace lrb {

// Combination x = x0 + scale * sum_j a_j v_j used by every RK stage.
struct StageArgs {
  const double* v[7];
  double a[7];
  int nv;
  double scale;
  int project;        // 1: sphere projection after the update
  int proj_in_loop;   // 1: projection inside the per-axis loop (RK4 combine)
  int write_vel;      // 1: velocity field receives the stage combination
  double rmin, rmax;
  // per-request step sizes (serving): scale_i = (dt_req[req[i]] * mul) / div
  const double* dt_req;
  const int32_t* req;
  double mul, div;
  // per-particle adaptive steps: a stage point outside the domain sets suspend[i]
  // and takes the particle out of the attempt's remaining stages (the controller
  // rejects the attempt and restores it) instead of removing it; null: removal
  int32_t* suspend;
};

// Read-only patch tables used by evaluation.
struct PatchView {
  const int32_t* n;        // [P] patch size (frozen at set-up)
  const uint8_t* active;   // [P] setup && n > 0
  const int64_t* mem_off;  // [P] offset into member / coefficient pools
  const int32_t* members;  // member pool
  const double* coef;      // coefficient pool, row-major [n][nf_total]
  int nf_total;
};

// Extensions of the patch systems (no reference counterpart): a polynomial block
// ([A P; P^T 0], monomials of the member coordinates shifted to the patch's PU node
// and scaled by ps) and a shape parameter per patch.  Every pointer may be null:
// m == nullptr means no polynomial block, eps == nullptr the grid's eps_r.
struct AugView {
  const int32_t* m;        // [P] polynomial terms of each patch
  const int32_t* deg;      // [P] their degree
  const double* eps;       // [P] shape parameter of each patch
  const int64_t* aux_off;  // [P] offset of the patch's null-space data in aux
  double* aux;             // per patch: xs [3][n] (member - PU node), V [m][n] (Householder
                           // vectors of P), X [m][n], B1 [m][n] (first m columns of Q^T A Q),
                           // R [m][m], T [m][m] (Q = I - V T V^T)
  double* beta;            // polynomial coefficients, patch p at beta + p * bstride, [m][nf]
  int bstride;
  int mmax;                // terms of the requested degree (the largest m)
  double ps;               // polynomial scale: 1 / largest cell width
  // [P] panel width of the matrix-core factor of each patch (its solve is blocked,
  // blocked_solve.hpp), 0: a factor of the vector-ALU kernels (column solve)
  const uint8_t* fnb;
};
__host__ __device__ inline size_t aug_aux_size(int n, int m) {
  return m > 0 ? (size_t)3 * n + (size_t)3 * m * n + (size_t)2 * m * m : 0;
}
// pivots of a Bunch-Kaufman factor (n + m of them: with a polynomial block the
// member-indexed pools give each patch n + POLY_MAX slots)
__host__ __device__ inline int64_t bk_piv_base(const AugView&, int, int64_t mem_off) { return mem_off; }

// ---- binning ----
void launch_locate_aos(const Grid& g, const double* x, int n, int32_t* cell, hipStream_t s);
void launch_locate_particles(const Grid& g, double* px, int64_t ld, int Np, int32_t* cell, int32_t* alive,
                             unsigned long long* removed, hipStream_t s);
void launch_cell_lists(const int32_t* cell, int n, int C, int32_t* cell_start, int32_t* cell_nodes,
                       int32_t* scratch /* >= C+1 + scan blocks */, hipStream_t s);
void launch_patch_count(const Grid& g, const int32_t* cell_start, int32_t* pcount, hipStream_t s);
void launch_patch_gather(const Grid& g, const int32_t* list, int nlist, const int32_t* cell_start,
                         const int32_t* cell_nodes, const int64_t* mem_off, int32_t* mem_pool, hipStream_t s);
void launch_touch(const Grid& g, const int32_t* cell, const int32_t* alive, int Np, uint8_t* touched, hipStream_t s);

// ---- patch factor / solve ----
// With a polynomial block the Bunch-Kaufman factor covers the (n+m) x (n+m)
// saddle-point matrix; the solve writes c to coef and beta to a.beta.
void launch_factor_bk(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                      const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool,
                      const double* xdata, double* fac, int32_t* piv, int32_t* info, const AugView& a, hipStream_t s);
void launch_solve_bk(const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off,
                     const int64_t* fac_off, const int32_t* mem_pool, const double* fac, const int32_t* piv,
                     const double* values, int nf, double* coef, const AugView& a, hipStream_t s);

// blocked Cholesky (positive definite kernels), FP64 or FP32 factor; the FP32
// solve refines FP64 coefficients against the FP64 matrix (patch_chol.hip).
// big: patches above chol_max_n() (no n-sized LDS); their solve needs a workspace
// ws of 2 nf doubles per member, indexed like the member pool.
// With a polynomial block (a.m) the factor is that of B22 = (Q^T A Q)[m:, m:], the
// saddle-point matrix on the null space of P^T (k_aug_prep computes Q first); the
// factor dimension is n - m.
bool chol_kernel_supported(int ktype);
int chol_max_n();
// nb: 0 the vector-ALU kernels below, 16 the matrix-core kernel of chol_mfma.hip
// (16-column panels, factor dimensions up to nmax <= chol_mfma_max_n())
void launch_factor_chol(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                        const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                        void* fac, int32_t* status, bool big, const AugView& a, int nb, int nmax, hipStream_t s);
void launch_solve_chol(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                       const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                       const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                       int32_t* status, int32_t* iters, double* ws, int nmax /* largest listed patch */,
                       const AugView& a, hipStream_t s);
// (solve_llt.hip) the same solve for patches of up to solve_llt_max_n() members whose
// factor holds inverse diagonal blocks (matrix-core factor, panel width 16)
void launch_solve_llt(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                      const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                      const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                      int32_t* status, int32_t* iters, int nmax, const double* peps, hipStream_t s);
int solve_llt_max_n();
// (aug_lean.hip) the saddle-point solve of k_solve_aug for patches of up to
// solve_aug_lean_max_n() members with a matrix-core B22 factor; FP32 factors (with
// refinement) for the kernels solve_aug_lean_supported() names
void launch_solve_aug_lean(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const void* fac,
                           const double* values, int nf, double* coef, int max_iter, double tol, int32_t* status,
                           int32_t* iters, int nmax, const AugView& a, hipStream_t s);
int solve_aug_lean_max_n();
bool solve_aug_lean_supported(int ktype, bool fp32);
// (chol_mfma.hip) the same factor with the panel products on the matrix cores,
// panels of 16 columns, factor dimension up to nmax <= chol_mfma_max_n()
int chol_mfma_max_n();
void launch_factor_chol_mfma(const Grid& g, int ktype, bool fp32, int nmax, const int32_t* list, int nlist,
                             const int32_t* pn, const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool,
                             const double* xd, void* fac, int32_t* status, const AugView& a, hipStream_t s);
// (chol_tiled.hip) Cholesky of large patches over many workgroups, tiles of
// TILE_NB columns: tile_tasks lists, for each tile column k, the (patch, row tile
// s >= k) pairs of every patch with more than k tiles (off[k] .. off[k + 1]); the
// factor slot is laid out as by launch_factor_chol_mfma (blocked solves apply)
constexpr int TILE_NB = 64;
void tile_tasks(const std::vector<int32_t>& list, const std::vector<int32_t>& nn, std::vector<int2>& tasks,
                std::vector<int64_t>& off);
void launch_factor_tiled(const Grid& g, int ktype, bool fp32, const int2* tasks, const std::vector<int64_t>& off,
                         const int32_t* pn, const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool,
                         const double* xd, void* fac, int32_t* status, const AugView& a, hipStream_t s);
// (aug.hip) null-space data of the listed patches with a polynomial block
void launch_aug_prep(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                     const int64_t* mem_off, const int32_t* mem_pool, const double* xd, const AugView& a, hipStream_t s);
// the same for large patches: the O(n^2) product U = A V spread over many
// workgroups, one block of aug_resid_rows() rows per task (patch, row block)
void launch_aug_prep_tasks(const Grid& g, int ktype, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int32_t* mem_pool, const double* xd, const AugView& a,
                           const int2* tasks, int ntasks, hipStream_t s);
// solve of the saddle-point systems: c and beta from the Cholesky factor of B22
// (FP64 direct, or FP32 factor with FP64 refinement of the full saddle-point
// residual); ws: (3 + 3 nf) doubles per member for patches larger than the LDS
// arrays (nmax: largest listed patch)
void launch_solve_aug(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                      const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                      const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                      int32_t* status, int32_t* iters, double* ws, int nmax, const AugView& a, hipStream_t s);
size_t aug_solve_ws_per_member(int nf);
// The same solve one refinement iteration per launch (step_it = 0, 1, ...; FP32
// factors of large patches, vectors in the workspace ws): status[p] is 2 while the
// patch iterates, dstate[p] its last correction; iterations >= 1 read the residual
// that launch_aug_resid_big wrote, one block of aug_resid_rows() rows per task
// (patch, row block).  step_it = -1: the whole solve in one launch.
void launch_solve_aug_step(const Grid& g, int ktype, bool fp32, const int32_t* list, int nlist, const int32_t* pn,
                           const int64_t* mem_off, const int64_t* fac_off, const int32_t* mem_pool, const double* xd,
                           const void* fac, const double* values, int nf, double* coef, int max_iter, double tol,
                           int32_t* status, int32_t* iters, double* ws, int nmax, const AugView& a, int step_it,
                           double* dstate, int phase, double* hstate, hipStream_t s);
// (trsv_tiles.hip) multi-workgroup B22 solves of the large saddle-point patches
// between phase 1 and phase 2 of launch_solve_aug_step: task lists per launch from
// trsv_tasks (the patches and their factor dimensions), uploaded once
int trsv_tile();
void trsv_tasks(const std::vector<int32_t>& l, const std::vector<int>& nn, std::vector<int2>& tasks,
                std::vector<int>& off);
void launch_trsv_tiles(bool fp32, const int2* dtasks, const std::vector<int>& off, const int32_t* pn,
                       const AugView& a, const int64_t* mem_off, const int64_t* fac_off, const void* fac, double* ws,
                       int nf, hipStream_t s);
int aug_resid_rows();
void launch_aug_resid_big(const Grid& g, int ktype, const int2* tasks, int ntasks, const int32_t* pn,
                          const int64_t* mem_off, const int32_t* mem_pool, const double* values, int nf,
                          const int32_t* status, double* ws, const AugView& a, hipStream_t s);
// polynomial table of the evaluation: [P][mm][nft], beta of fields f0.. of the
// active patches, zero elsewhere (mm = terms of the requested degree)
void launch_build_ptab(const Grid& g, const uint8_t* active, const AugView& a, int mm, int nft, int nf, int f0,
                       double* ptab, hipStream_t s);

// ---- weights / evaluation ----
void launch_weights(const Grid& g, const double* pux, const double* puy, const double* puz, const double* px,
                    int64_t ld, int Np, const int32_t* cell, const int32_t* alive, double* w, hipStream_t s);
void launch_eval_ref(const Grid& g, int ktype, const double* px, int64_t ld, int Np, const int32_t* cell,
                     const int32_t* alive, const double* w, const PatchView& pv, const double* xdata,
                     const int32_t* cell_start, int nfields, double* out, const AugView& a, hipStream_t s);

// cell-ordered evaluation: particles grouped by cell (pstart[c]..pstart[c+1]), dead
// particles in the tail bucket C, so perm covers the whole swarm
void launch_sort_particles(const int32_t* cell, const int32_t* alive, int Np, int C,
                           int32_t* pstart /* C+2 */, int32_t* perm,
                           int32_t* scratch /* sort_particles_scratch(C, Np) */, hipStream_t s);
size_t sort_particles_scratch(int C, int Np);
// gather of SoA particle arrays (rows of ld elements) into the order of perm
struct ReorderSet {
  const double* src[6];
  double* dst[6];
  int rows[6];
  int nd;
  const int32_t* isrc[4];
  int32_t* idst[4];
  int ni;
};
// hipErrorInvalidValue (nothing launched) for a null source or target
hipError_t launch_reorder(const ReorderSet& r, const int32_t* perm, int n, int64_t ld, hipStream_t s);
void launch_iota(int32_t* a, int n, hipStream_t s);
void launch_sorted_coords(const double* x, int n, int dim, const int32_t* cell_nodes, double* xs, hipStream_t s);
// (eval_cell.hip) fp32: coef/zeros/xs are float (cell-relative coordinates)
bool eval_f32_supported(int ktype);
void launch_eval_cell(const Grid& g, int ktype, bool fp32, const double* px, int64_t ld, const int32_t* perm,
                      const int32_t* pstart, const double* w, const uint8_t* active, const int64_t* mem_off,
                      const void* coef, int nft, const void* zeros, const void* xs, int N, const int32_t* cell_start,
                      int nfields, double* out, int block, hipStream_t s);
// (eval_blk.hip) cell-major node table: per node row [coords, pad, coefficients in
// the 2^d corner patches of its cell]; evaluation for nft <= 4 fields
bool eval_blk_supported(int nft, int dim, bool f32);
size_t blk_row(int dim, int nft);
// zero rows after the N node rows of the evaluation table (the node loops read up
// to two rows past a cell's last node)
constexpr int BLK_PAD_ROWS = 2;
void launch_build_blk(const Grid& g, bool fp32, const double* x, int N, const int32_t* cn, const int32_t* dcell,
                      const int32_t* cs, const uint8_t* active, const int64_t* mem_off, const double* coef, int nft,
                      int cstride, int f0, double cscale, void* blk, hipStream_t s);   // fields f0 .. f0+nft-1 of
                                                                                       // cstride, times cscale
// cell_list / list_count (device): evaluate only the listed non-empty cells, one
// block per cell (group 0: contiguous runs of cells per XCD; group -K: runs of K
// cells dealt round-robin over the XCDs) or 64/group cells per wave in lane groups
// of group = 8, 16, 32 (needs nft == dim); nullptr: one block per cell.  Every
// variant performs the same arithmetic per particle.
// Events the next launch_eval_blk dispatch records its start and end in (nullptr: none).
void set_eval_timing_events(hipEvent_t start, hipEvent_t stop);
// Extensions: ev.ptab (polynomial table, launch_build_ptab) adds each corner
// patch's polynomial part, in a pass over the particles (cell ids and alive flags
// in pcell / palive, np of them) after the evaluation kernel; ev.peps with ktype
// K_GAUSS_LOCAL gives the Gaussian the shape parameter of each corner patch.
struct EvalExt {
  const double* ptab = nullptr;
  int mm = 0;              // polynomial terms per patch in ptab
  int deg = -1;            // their degree
  double ps = 1.0;         // polynomial scale
  const double* peps = nullptr;
  const int32_t* pcell = nullptr;
  const int32_t* palive = nullptr;
};
void launch_eval_blk(const Grid& g, int ktype, bool fp32, const double* px, int64_t ld, const int32_t* perm,
                     const int32_t* pstart, const double* w, const void* blk, int nft, const int32_t* cs, int nfields,
                     double* out, int block, int np, const int32_t* cell_list, const int32_t* list_count, int group,
                     int pipe, int waves, const EvalExt& ev, hipStream_t s);
// ascending non-empty cells; list needs nonempty_list_capacity(C) entries (the tail is scratch)
void launch_nonempty_cells(const int32_t* pstart, int C, int32_t* list, int32_t* count, hipStream_t s);
size_t nonempty_list_capacity(int C);
void launch_sorted_coords32(const Grid& g, const double* x, int n, const int32_t* cell_nodes, const int32_t* cell,
                            float* xs, hipStream_t s);
void launch_coef_to_f32(const int32_t* list, int nlist, const int32_t* pn, const int64_t* mem_off, int nf,
                        const double* coef, float* coef32, hipStream_t s);

// ---- RK stage loop ----
void launch_stage(const Grid& g, const double* pux, const double* puy, const double* puz, double* px,
                  const double* x0, double* vel, int64_t ld, int Np, int32_t* cell, int32_t* alive,
                  unsigned long long* removed, double* w, const StageArgs& a, hipStream_t s);
void launch_save(double* px, double* x0, const double* vel, double* v0, int64_t ld, int Np, int dim,
                 const int32_t* alive, hipStream_t s);
void launch_copy_rows(const double* src, int64_t lds, double* dst, int64_t ldd, int nrows, int Np,
                      const int32_t* alive, hipStream_t s);
void launch_cfl(const double* vel, int64_t ld, int Np, int dim, const int32_t* alive, double c,
                unsigned long long* out_bits, hipStream_t s);
void launch_dopri_err(const double* px, const double* x0, const double* const* vadv, int64_t ld, int Np,
                      int dim, const int32_t* alive, double dt, double atol, double rtol,
                      unsigned long long* max_bits, unsigned int* count, hipStream_t s);

// ---- serving: per-request reductions and retirement ----
void launch_dopri_err_req(const double* px, const double* x0, const double* const* vadv, int64_t ld, int Np, int dim,
                          const int32_t* alive, const int32_t* req, const double* dt_req, double atol, double rtol,
                          unsigned long long* maxb_req, int nslots, hipStream_t s);
// per-particle adaptive DP45 with time-blended velocity (adaptive.hip)
void launch_blend(double* out, const double* va, const double* vb, int64_t ld, int Np, int dim, const int32_t* alive,
                  const double* tp, const double* dtp, double cst, double tA, double tB, hipStream_t s);
// susp: particles suspended by an out-of-domain trial point (StageArgs::suspend):
// rejected, restored to x0 and alive again with step size times fac_min, or
// removed (counted in removed) once that step size falls below 1e-10 max(1, |t_end|)
void launch_dp45_control(const Grid& g, int32_t* cell, double* px, double* x0, const double* const* k, double* k1,
                         int64_t ld, int Np, int dim,
                         int32_t* alive, int32_t* fin, int32_t* susp, unsigned long long* removed, double* tp,
                         double* dtp, double t_end, double atol, double rtol, double fac, double fac_min,
                         double fac_max, unsigned long long* counts, hipStream_t s);
void launch_dp45_init(int Np, const int32_t* alive, double* tp, double* dtp, int32_t* fin, int32_t* self, double t0,
                      double dt0, hipStream_t s);
void launch_dp45_finish(int Np, int32_t* alive, const int32_t* fin, hipStream_t s);
void launch_retire(const double* px, int64_t ld, int Np, int dim, int32_t* alive, const int32_t* req,
                   const int32_t* porig, const uint8_t* done, double* const* out_req, hipStream_t s);
void launch_fill_i32(int32_t* a, int n, int32_t v, hipStream_t s);

// ---- layout conversion ----
void launch_aos_to_soa(const double* aos, int n, int bs, double* soa, int64_t ld, hipStream_t s);
void launch_soa_to_aos(const double* soa, int64_t ld, int n, int bs, double* aos, hipStream_t s);

}  // namespace lrb
