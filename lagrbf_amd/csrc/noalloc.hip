// Interpolation at device-resident points without any allocation (no reference
// counterpart): for callers that cannot call hipMalloc (code under graph capture,
// tight serving loops, harnesses that forbid it).  Every work array lives in one
// file-scope device array, `scratch`, of fixed size; each GPU holds its own copy
// until the process ends.  Its device address is taken once per device with
// hipGetSymbolAddress at context creation and used everywhere host code needs
// the array (kernel arguments, fills): the array's name in host code is only a
// host-side stand-in.
//
// The kernels are those of lagrbf_fill_swarm (particles) + lagrbf_interpolation:
// location, Shepard weights, grouping by cell, the cell-table evaluation of the
// context's resident factored data.  Every call on a GPU shares that GPU's
// scratch (the previous call's contents stay; what must start at zero is cleared
// with fill kernels): calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "ctx.hpp"

namespace lrb {

// limits of the scratch layout
constexpr int NOALLOC_MAX_POINTS = 1 << 18;   // points per call
constexpr int NOALLOC_MAX_CELLS = 1 << 18;    // PU cells of the context
constexpr int NOALLOC_MAX_FIELDS = 4;         // interpolated fields
constexpr size_t NA_LD = NOALLOC_MAX_POINTS;  // leading dimension of the SoA rows

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Byte offsets of the work arrays (dimension 3 and 8 corners at most).
struct NaLayout {
  size_t px, cell, alive, w, perm, out, pstart, sort, list, count, removed, end;
  NaLayout() {
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t at = o; o += align256(bytes); return at; };
    px = take(3 * NA_LD * sizeof(double));
    cell = take(NA_LD * sizeof(int32_t));
    alive = take(NA_LD * sizeof(int32_t));
    w = take(8 * NA_LD * sizeof(double));
    perm = take(NA_LD * sizeof(int32_t));
    out = take((size_t)NOALLOC_MAX_FIELDS * NA_LD * sizeof(double));
    pstart = take(((size_t)NOALLOC_MAX_CELLS + 2) * sizeof(int32_t));
    sort = take(sort_particles_scratch(NOALLOC_MAX_CELLS, NOALLOC_MAX_POINTS) * sizeof(int32_t));
    list = take(nonempty_list_capacity(NOALLOC_MAX_CELLS) * sizeof(int32_t));
    count = take(sizeof(int32_t));
    removed = take(2 * sizeof(unsigned long long));
    end = o;
  }
};

// 256 KiB points x (3 + 8 + 4) doubles + cell lists: about 36 MB per GPU
constexpr size_t NOALLOC_SCRATCH_BYTES = 40u << 20;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};

// Device address of this GPU's scratch, resolved once per device with that device
// current (called from lagrbf_create).
int noalloc_init_device(int device) {
  if (device < 0 || device >= 64) return set_err(E_RANGE, "lagrbf_create", "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  static_assert(sizeof(scratch) >= 1, "scratch");
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_err(E_LIB, "lagrbf_create", "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return 0;
}

__global__ void k_fill_f64(double* __restrict__ a, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace lrb

using namespace lrb;

extern "C" int lagrbf_interpolate_device_noalloc(void* ctx, int32_t n, const double* coords, double* values) {
  const char* fn = "lagrbf_interpolate_device_noalloc";
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err(E_NULL, fn, "LagRBFCtx is NULL");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(E_LIB, fn, "hipSetDevice failed");
  static const NaLayout lay;
  static_assert(NOALLOC_SCRATCH_BYTES >= 36u << 20, "scratch size");
  if (lay.end > NOALLOC_SCRATCH_BYTES) return set_err(E_LIB, fn, "scratch layout exceeds the array");
  if (n < 0) return set_err(E_RANGE, fn, "negative point count %d", n);
  if (n > NOALLOC_MAX_POINTS)
    return set_err(E_SIZ, fn, "%d points exceed the scratch maximum of %d per call", n, NOALLOC_MAX_POINTS);
  if (c->g.C > NOALLOC_MAX_CELLS)
    return set_err(E_SIZ, fn, "%d PU cells exceed the scratch maximum of %d", c->g.C, NOALLOC_MAX_CELLS);
  if (c->nf < 1 || c->nf > NOALLOC_MAX_FIELDS)
    return set_err(E_SIZ, fn, "%d interpolated fields (the scratch holds 1 to %d)", c->nf, NOALLOC_MAX_FIELDS);
  if (n > 0 && (!coords || !values)) return set_err(E_NULL, fn, "coords/values is NULL");
  const bool f32 = eval_f32(*c);
  // nothing here may allocate: the patches, their coefficients and the evaluation
  // table must already be resident
  bool complete = true;
  for (int p = 0; p < c->g.P && complete; p++) complete = c->h_setup[p] && c->h_solved[p];
  if (!c->data_filled || !complete || !c->blk_valid || c->blk_f32 != f32 || !c->memb_current ||
      !eval_blk_supported(c->nf, c->g.dim, f32))
    return set_err(E_STATE, fn, "needs every patch set up and solved and the evaluation table built "
                   "(-lagrbf_setup_all and one lagrbf_interpolation on the current data)");
  void* base = g_scratch_ptr[c->device];
  if (!base) return set_err(E_STATE, fn, "scratch of device %d not resolved (context not created here)", c->device);
  if (n == 0) return 0;
  unsigned char* b = static_cast<unsigned char*>(base);
  double* px = reinterpret_cast<double*>(b + lay.px);
  int32_t* cell = reinterpret_cast<int32_t*>(b + lay.cell);
  int32_t* alive = reinterpret_cast<int32_t*>(b + lay.alive);
  double* w = reinterpret_cast<double*>(b + lay.w);
  int32_t* perm = reinterpret_cast<int32_t*>(b + lay.perm);
  double* out = reinterpret_cast<double*>(b + lay.out);
  int32_t* pstart = reinterpret_cast<int32_t*>(b + lay.pstart);
  int32_t* sortw = reinterpret_cast<int32_t*>(b + lay.sort);
  int32_t* list = reinterpret_cast<int32_t*>(b + lay.list);
  int32_t* count = reinterpret_cast<int32_t*>(b + lay.count);
  unsigned long long* removed = reinterpret_cast<unsigned long long*>(b + lay.removed);
  const int64_t ld = (int64_t)NA_LD;
  const int dim = c->g.dim, nf = c->nf;
  hipStream_t s = c->s;
  // what must start at zero: outputs (an empty 3^d neighbourhood leaves its
  // particle's values untouched), weights, the removal counter; points alive
  const size_t nz = (size_t)nf * ld;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, out, nz, 0.0);
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)(((size_t)c->g.nen * ld + 255) / 256)), dim3(256), 0, s, w,
                     (size_t)c->g.nen * ld, 0.0);
  launch_fill_i32(reinterpret_cast<int32_t*>(removed), 2, 0, s);
  launch_fill_i32(alive, n, 1, s);
  // lagrbf_fill_swarm: SoA coordinates, location (points outside the box are
  // dropped and output as zero rows), Shepard weights
  launch_aos_to_soa(coords, n, dim, px, ld, s);
  launch_locate_particles(c->g, px, ld, n, cell, alive, removed, s);
  launch_weights(c->g, c->pux.p, c->puy.p, c->puz.p, px, ld, n, cell, alive, w, s);
  // lagrbf_interpolation: grouping by cell and the cell-table evaluation
  launch_sort_particles(cell, alive, n, c->g.C, pstart, perm, sortw, s);
  int v = c->eval_variant >= 0 ? c->eval_variant : EvalTuner::DEFAULT_VARIANT;
  if (!EvalTuner::valid(v, true, f32)) v = EvalTuner::DEFAULT_VARIANT;
  bool listed;
  int group, pipe, waves;
  eval_variant(v, listed, group, pipe, waves);
  if (listed) launch_nonempty_cells(pstart, c->g.C, list, count, s);
  const int per_cell = (n + c->g.C - 1) / std::max(c->g.C, 1);
  const int block = std::min(256, std::max(64, (per_cell + 63) / 64 * 64));
  set_eval_timing_events(nullptr, nullptr);
  EvalExt ev = eval_ext(*c, -1);
  ev.pcell = cell;
  ev.palive = alive;
  launch_eval_blk(c->g, eval_ktype(*c), f32, px, ld, perm, pstart, w, f32 ? (const void*)c->blk32.p : (const void*)c->blk64.p,
                  nf, c->cell_start.p, nf, out, block, n, listed ? list : nullptr, listed ? count : nullptr, group, pipe,
                  waves, ev, s);
  // lagrbf_output("fields"): AoS rows in point order
  launch_soa_to_aos(out, ld, n, nf, values, s);
  const hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return set_err(E_LIB, fn, "%s", hipGetErrorString(e));
  return 0;
}
