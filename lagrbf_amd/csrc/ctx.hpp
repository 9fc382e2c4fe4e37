// Device state of a LagRBF context (shared by context.hip and serve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "options.hpp"
#include "tuner.hpp"

namespace lrb {

// PETSc error codes returned for the same conditions as the reference
enum { E_MEM = 55, E_SUP = 56, E_SIZ = 60, E_WRONG = 62, E_RANGE = 63, E_STATE = 73, E_LIB = 76, E_NULL = 85 };

int set_err(int code, const char* fn, const char* fmt, ...);

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return set_err(E_LIB, __func__, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define CHK(call)              \
  do {                         \
    int r_ = (call);           \
    if (r_) return r_;         \
  } while (0)

// (devmem.cpp) device allocations through the process-wide cache of large blocks:
// a released block of 64 MiB or more is kept for the next reservation on its
// device instead of returned to the driver (LAGRBF_DEVICE_CACHE=0: no cache)
hipError_t dev_malloc(void** p, size_t bytes);
void dev_free(void* p);
size_t dev_cached_bytes();
void dev_cache_release();

// Device buffer with grow-only capacity.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  ~DBuf() { release(); }
  void release() { if (p) dev_free(p); p = nullptr; cap = 0; }
  hipError_t reserve(size_t n) {
    if (n <= cap && p) return hipSuccess;
    release();
    size_t m = n ? n : 1;
    hipError_t e = dev_malloc((void**)&p, m * sizeof(T));
    if (e == hipSuccess) cap = m;
    return e;
  }
  // grow keeping the first `keep` elements
  hipError_t grow(size_t n, size_t keep, hipStream_t s) {
    if (n <= cap && p) return hipSuccess;
    T* q = nullptr;
    hipError_t e = dev_malloc((void**)&q, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) return e;
    if (p && keep) {
      e = hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) { dev_free(q); return e; }
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) { dev_free(q); return e; }
    }
    release();
    p = q;
    cap = n ? n : 1;
    return hipSuccess;
  }
};

struct Ctx {
  Grid g{};
  int ktype = K_LINEAR, solve_method = 0, nf = 0;
  double rmin = 0.0, rmax = DBL_MAX, cfl_fac = 0.4;
  bool advection = false;
  int scheme = 3, order = 0;
  double adv_dt = 1e-10;
  double dopri_fac = 0.9, dopri_fac_min = 1e-2, dopri_fac_max = 1e2, dopri_atol = 1e-3, dopri_rtol = 1e-3;
  double dopri_err = 0.0;
  std::string velocity_name = "velocity";
  bool pvd_written = false;                // this context has written its PVD collection
  // time blend of two snapshots (fields 0..dim-1 at tA, dim..2dim-1 at tB)
  bool setup_all = false;                  // every patch at the first set-up, not only the touched ones
  bool blend = false;
  int blend_fused = -1;                    // both snapshots in one evaluation: 1, 0, -1 = FP32 only
  double blend_t[2] = {0.0, 1.0};
  // per-particle adaptive DP45 state and counters
  DBuf<double> ptime, pdt, vtmp;
  DBuf<int32_t> pfin, pself, psusp, psusp_alt;   // psusp: out-of-domain trial points (adaptive DP45)
  DBuf<double> ptime_alt, pdt_alt;     // reorder targets of the adaptive state
  bool cell_ordered = false;              // swarm stored in cell order since the last fill
  int bucket_np = -1;                     // swarm size for the tuner's bucket while Np is a prefix
  DBuf<int32_t> pfin_alt;
  double dp_attempts = 0, dp_accepted = 0, dp_rejected = 0, dp_active = 0;
  hipStream_t s = nullptr;
  int device = 0;

  DBuf<double> pux, puy, puz;
  // data swarm
  int N = 0;
  bool data_filled = false;
  DBuf<double> dx, dval;
  DBuf<int32_t> dcell, cell_start, cell_nodes, scratch, pcount;
  std::vector<int32_t> h_pcount;
  DBuf<double> xs_sorted, zeros;           // cell-sorted node coordinates (SoA), zero coefficients
  DBuf<float> xs32, zeros32, coef32_pool;  // mixed path: cell-relative FP32 coordinates, FP32 coefficients
  // cell-major node table of the evaluation (coordinates + corner-patch coefficients per node)
  DBuf<double> blk64;
  DBuf<float> blk32;
  bool blk_valid = false, blk_f32 = false;
  // tables of one field group (fields g*dim .. g*dim+dim-1) for the time-blended
  // velocity: rebuilt when table_gen moved on since they were built
  DBuf<double> blkg64[2];
  DBuf<float> blkg32[2];
  int blkg_gen[2] = {-1, -1};
  bool blkg_f32[2] = {false, false};
  int table_gen = 0;
  EvalTuner tuner;                         // evaluation launch shape per particle density (this data)
  int eval_variant = -1;                   // forced launch shape (tests), -1: tuned
  int serve_tune_min = 0;                  // serving: smaller batch evaluations skip the tuner
  int last_variant = -1;                   // launch shape of the last cell-table evaluation
  std::vector<int32_t> h_cell_start_prev, h_cell_nodes_prev;   // cell lists the frozen patches were built from
  bool memb_current = true;                // every set-up patch's members == current cell lists
  int eval_mode = 0;                       // 0 auto (cell-ordered when possible), 1 cell-ordered, 2 reference order
  std::vector<int64_t> h_dpid;
  // particle swarm
  int Np = 0;
  int64_t ld = 0;
  bool points_filled = false;
  DBuf<double> px, w, fields, vel, x0, vadv[7];
  DBuf<int32_t> pcell, alive;
  DBuf<int32_t> perm, pstart, pscratch;    // particles grouped by cell
  DBuf<int32_t> cell_list, list_count;     // non-empty particle cells (listed evaluation shapes)
  bool sorted_valid = false;
  // the swarm is stored in cell order at the start of every advection step;
  // porig[s] = index (fill order) of the particle held in slot s
  bool reorder = true;
  DBuf<int32_t> porig, porig_alt, pcell_alt, alive_alt;
  DBuf<double> px_alt, vel_alt, fields_alt, w_alt;
  std::vector<int64_t> h_ppid;
  DBuf<unsigned long long> counters;   // [0] removed, [1] reduction bits, [2] count
  // patches
  std::vector<uint8_t> h_setup, h_solved, h_active;
  std::vector<int32_t> h_n;
  std::vector<int64_t> h_mem_off, h_fac_off;
  size_t mem_used = 0, fac_used = 0;
  int setup_N = -1;
  DBuf<int32_t> p_n, mem_pool, piv_pool, list, info;
  DBuf<uint8_t> p_active, touched;
  DBuf<int64_t> p_mem_off, p_fac_off;
  DBuf<double> fac_pool, coef_pool;
  DBuf<float> fac32_pool;
  size_t fac32_used = 0;
  std::vector<uint8_t> h_method;           // per patch: M_BK64, M_CHOL64, M_CHOL32, M_CHOL64B, M_CHOL32B
  DBuf<double> solve_ws;                   // solve workspace of the large patches (2 nf per member)
  DBuf<int32_t> pstatus, piters;
  int precision = 0;                       // 0 FP64, 1 FP32 factor + FP64 refinement
  // mixed: FP32 factor and FP64 refinement (the default); true: factor in FP64 (no
  // refinement) with the FP32 evaluation kept
  bool mixed_factor64 = false;
  int factor_mode = 0;                     // 0 auto (Cholesky for PD kernels), 1 Bunch-Kaufman always
  bool factors_released = false;           // serving datasets keep coefficients only
  int chol_big_min = 512;                  // patches above this size use the large-patch Cholesky kernels
  int big_factor = 1;                      // large patches: 1 tiled over many workgroups (matrix cores), 0 one
                                           // workgroup per patch (vector ALU)
  DBuf<int2> tile_tasks;                   // (patch, row tile) pairs of the tiled factor, step by step
  std::vector<int2> h_tile_tasks;
  std::vector<int64_t> h_tile_off;         // steps: tasks h_tile_off[k] .. h_tile_off[k + 1]
  DBuf<int2> resid_tasks;                  // (patch, row block) pairs of the stepped refinement's residual
  DBuf<double> refine_dstate;              // [P] last correction of a stepped refinement
  bool stepped_refine = true;              // FP32 factors of large saddle-point patches: refine step by step
  bool lean_solve = true;                  // small Cholesky patches with matrix-core factors: solve_llt.hip
  bool tiled_solve = true;                 // large saddle-point patches: B22 solves over many workgroups
  DBuf<int2> trsv_dtasks;                  // their task lists (trsv_tiles.hip)
  DBuf<double> aug_hstate;                 // h1 and u1 between the two phases of those solves
  int refine_max_iter = 10;
  double refine_tol = 1e-12;
  double fallback = 0, refine_iters_max = 0, refine_iters_sum = 0, refined_patches = 0;
  // analytical velocity
  int (*vfn)(double*, double*, void*) = nullptr;
  int (*vfn_batch)(int32_t, int32_t, double*, double*, void*) = nullptr;
  void* vuser = nullptr;
  bool analytical = false;
  // stats
  double removed = 0, patches_setup = 0, setup_ms = 0, eval_ms = 0, eval_count = 0;
  // evaluation timing: event pairs resolved at the next synchronisation point
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
  std::vector<hipEvent_t> ev_pool;
  // graph capture of the fixed-step loop (lagrbf_advect_steps): one executable
  // graph per buffer parity (the reorder swaps the particle buffers every step)
  bool use_graph = false, capturing = false;
  int parity = 0;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  double graph_dt = 0.0, graph_steps = 0, graph_ms = 0;
  int max_patch = 0;
  // extensions (no reference counterpart): polynomial block of every patch system
  // and a shape parameter per patch
  int poly_deg = -1;                       // -lagrbf_poly_degree (-1: pure RBF systems, as the reference)
  int poly_mm = 0;                         // terms of poly_deg
  double local_eps = 0.0;                  // -lagrbf_local_eps c: eps_p = c / h_local(p), 0: eps_r
  std::vector<int32_t> h_m, h_deg;         // per patch: polynomial terms and degree
  std::vector<double> h_eps;               // per patch: shape parameter
  std::vector<int64_t> h_aux_off;
  size_t aux_used = 0;
  DBuf<int32_t> p_m, p_deg;
  DBuf<double> p_eps, aux_pool, beta_pool;
  DBuf<int64_t> p_aux_off;
  DBuf<double> ptab, ptabg[2];             // polynomial tables of the evaluation (main, blend groups)
  int ptab_gen = -1, ptabg_gen[2] = {-1, -1};
  double aug_fallback_bk = 0;              // saddle-point patches refactored with Bunch-Kaufman
  double big_refactor64 = 0;               // large FP32-factored patches refactored with the FP64 tiled Cholesky
  std::vector<double> refine_hist = std::vector<double>(64, 0.0);   // refined patches by iteration count
  // factor kernel of the small-patch Cholesky per size bucket: -1 tune (default),
  // 0 vector ALU, 16 / 32 matrix cores with panels of that width (-lagrbf_factor_nb)
  int factor_nb = -1;
  std::vector<uint8_t> h_fnb;              // per patch: panel width of its matrix-core factor, 0 otherwise
  DBuf<uint8_t> p_fnb;
  double factor_trial_ms = 0;              // time spent in the factor tuner's trials
  // patch set-up by phase: pool growth (host wall time, the stream idle), factors
  // and solves (stream time between events)
  double setup_alloc_ms = 0, setup_factor_ms = 0, setup_solve_ms = 0;
};

// factor methods: Bunch-Kaufman FP64; Cholesky FP64 / FP32 (n <= chol_max_n());
// the same for larger patches (kernels without n-sized LDS); the saddle-point
// systems' null-space Cholesky (B22, FP64 / FP32, small / large)
enum { M_BK64 = 0, M_CHOL64 = 1, M_CHOL32 = 2, M_CHOL64B = 3, M_CHOL32B = 4,
       M_AUG64 = 5, M_AUG32 = 6, M_AUG64B = 7, M_AUG32B = 8, N_METHODS = 9 };
inline bool method_fp32(int m) { return m == M_CHOL32 || m == M_CHOL32B || m == M_AUG32 || m == M_AUG32B; }
inline bool method_big(int m) { return m == M_CHOL64B || m == M_CHOL32B || m == M_AUG64B || m == M_AUG32B; }
inline bool method_aug(int m) { return m >= M_AUG64 && m <= M_AUG32B; }

template <typename T>
inline void swap_buf(DBuf<T>& a, DBuf<T>& b) {
  std::swap(a.p, b.p);
  std::swap(a.cap, b.cap);
}

// context.hip
Ctx* create_from_view(const OptView& o);
OptView current_options();
int fill_data(Ctx& c, int n, const double* coords, const double* values, const int64_t* pidx);
// all_patches: set up every patch (not only the corners of particle cells)
int partitions_setup(Ctx& c, bool all_patches = false);
int ensure_blk(Ctx& c, bool f32);
// kernel of the evaluation (K_GAUSS_LOCAL for the Gaussian with per-patch shape)
int eval_ktype(const Ctx& c);
// FP32 evaluation table for this context's kernel
bool eval_f32(const Ctx& c);
// polynomial table and per-patch shape of the evaluation (group -1: main table)
EvalExt eval_ext(const Ctx& c, int group);
// (noalloc.hip) device address of the no-allocation entry point's static scratch on
// `device`, resolved once per device (called at context creation)
int noalloc_init_device(int device);
// patch views of the extensions (with_poly false: shape parameter only)
AugView aug_view(Ctx& c, bool with_poly);
// free the factor pools once every patch is solved (coefficients stay)
void release_factors(Ctx& c);

}  // namespace lrb
