// Serving: continuous batching of advection requests against device-resident
// datasets.  No reference counterpart (the reference advects one swarm per call).
//
// A dataset is a context whose data nodes are filled and whose patches are ALL
// set up and solved up front (the evaluation table is built, factors released):
// it is read-only afterwards and shared by every request on it.
//
// Requests on one (dataset, scheme) share one batch: one particle swarm in SoA
// with per-particle request slot and index-within-request, advanced by the same
// kernels as a context's swarm, one RK step per tick for every active request.
// Step sizes are per request (stage scale = dt_req * c / d, the products and
// quotients of the single-swarm path), DOPRI error maxima are per request, so a
// particle's arithmetic depends only on its own position and its request: a
// request's result is the same alone and in any mix.
//
// Each tick: (1) admit queued requests while the HBM budget allows (append their
// particles, locate, weights); (2) per batch, on its own stream: group by cell,
// store in cell order (finished and removed particles drop to the tail and are
// cut), velocity at the start (after admissions), one RK step, velocity at the
// end; (3) after every batch: count steps, DOPRI step-size control, retire the
// finished requests (positions scattered to their output rows).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

#include "../../include/lagrbf.h"
#include "ctx.hpp"

namespace lrb {

namespace {

constexpr int RSLOTS = 1024;   // request slots per batch

struct ServeReq {
  int64_t id = 0;
  int ds = 0, scheme = 2, nsteps = 0, n = 0, slot = -1;
  double dt = 0.0;
  std::vector<double> coords, result;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int steps_done = 0;
  size_t bytes = 0;
  double* d_out = nullptr;
  bool done = false;
};

struct Batch {
  Ctx* ds = nullptr;
  int scheme = 2;
  hipStream_t s = nullptr;
  int Np = 0;
  int64_t ld = 0;
  DBuf<double> px, x0, vel, w, vadv[7], stage;
  DBuf<int32_t> cell, alive, req, porig, perm, pstart, pscratch, cell_list, list_count;
  DBuf<double> px_b, vel_b, w_b;
  DBuf<int32_t> cell_b, alive_b, req_b, porig_b;
  DBuf<unsigned long long> counters, err_req;
  DBuf<double> dt_req;
  DBuf<uint8_t> done_req;
  DBuf<double*> out_req;
  std::vector<ServeReq*> slots = std::vector<ServeReq*>(RSLOTS, nullptr);
  std::vector<double> h_dt = std::vector<double>(RSLOTS, 0.0);
  bool appended = false;
  int32_t* h_count = nullptr;               // pinned: alive count after the tick's regrouping
  unsigned long long* h_err = nullptr;      // pinned: DOPRI per-request maxima
  ~Batch() {
    if (h_count) (void)hipHostFree(h_count);
    if (h_err) (void)hipHostFree(h_err);
    if (s) (void)hipStreamDestroy(s);
  }
  int order() const { return scheme == 3 ? 7 : (scheme == 2 ? 4 : (scheme == 1 ? 1 : 0)); }
};

struct Server {
  double budget = 200e9, reserved = 0;
  std::vector<std::unique_ptr<Ctx>> ds;
  std::map<std::pair<int, int>, std::unique_ptr<Batch>> batches;
  std::deque<ServeReq*> queue;
  std::map<int64_t, std::unique_ptr<ServeReq>> reqs;
  int64_t next_id = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double completed = 0, particle_steps = 0, ticks = 0, max_active = 0;
  double enqueue_ms = 0, tick_ms = 0;       // host time: enqueueing the batches' ticks; whole ticks
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// Bytes a request holds while active: its share of the batch's SoA arrays (both
// reorder buffers, v_adv stages of the scheme) plus its output rows.
size_t request_bytes(int dim, int nen, int scheme, int n) {
  const int stages = scheme == 3 ? 7 : (scheme == 2 ? 4 : (scheme == 1 ? 1 : 0));
  const size_t dbl = (size_t)dim * (2 + 2 + 1 + stages) + (size_t)2 * nen;   // px, vel (x2), x0, v_adv, w (x2)
  const size_t ints = 9;                                                         // cell, alive, req, porig (x2), perm
  return (size_t)n * (8 * dbl + 4 * ints + 8 * dim);
}

template <typename T>
int grow_rows(DBuf<T>& b, int rows, int64_t old_ld, int64_t new_ld, int n, hipStream_t s) {
  if (rows <= 0) return 0;
  const char* fn = "batch_reserve";
  T* q = nullptr;
  hipError_t e = dev_malloc((void**)&q, sizeof(T) * (size_t)new_ld * rows);
  if (e != hipSuccess)
    return set_err(E_MEM, fn, "%d rows of %lld elements: %s", rows, (long long)new_ld, hipGetErrorString(e));
  // one copy per row: a pitched copy is refused when its rows span several of the
  // device cache's physical chunks (the range is contiguous only virtually)
  for (int r = 0; b.p && n > 0 && r < rows; r++) {
    e = hipMemcpyAsync(q + (size_t)r * new_ld, b.p + (size_t)r * old_ld, sizeof(T) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) {
      dev_free(q);
      return set_err(E_LIB, fn, "copying row %d of %d (%d elements, leading dimension %lld -> %lld): %s", r, rows, n,
                     (long long)old_ld, (long long)new_ld, hipGetErrorString(e));
    }
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    dev_free(q);
    return set_err(E_LIB, fn, "stream synchronisation before the row copy's release: %s", hipGetErrorString(e));
  }
  b.release();
  b.p = q;
  b.cap = (size_t)new_ld * rows;
  return 0;
}

int batch_reserve(Batch& b, int need) {
  if (need <= b.ld) return 0;
  const int64_t nl = std::max<int64_t>({(int64_t)need, 2 * b.ld, 4096}) / 64 * 64 + 64;
  const int dim = b.ds->g.dim, nen = b.ds->g.nen;
  const int64_t ol = b.ld;
  const int n = b.Np;
  CHK(grow_rows(b.px, dim, ol, nl, n, b.s));
  CHK(grow_rows(b.vel, dim, ol, nl, n, b.s));
  CHK(grow_rows(b.w, nen, ol, nl, n, b.s));
  CHK(grow_rows(b.cell, 1, ol, nl, n, b.s));
  CHK(grow_rows(b.alive, 1, ol, nl, n, b.s));
  CHK(grow_rows(b.req, 1, ol, nl, n, b.s));
  CHK(grow_rows(b.porig, 1, ol, nl, n, b.s));
  // per-step scratch: contents need not survive
  HIPCHK(b.x0.reserve((size_t)nl * dim));
  for (int j = 0; j < b.order(); j++) HIPCHK(b.vadv[j].reserve((size_t)nl * dim));
  HIPCHK(b.px_b.reserve((size_t)nl * dim)); HIPCHK(b.vel_b.reserve((size_t)nl * dim));
  HIPCHK(b.w_b.reserve((size_t)nl * nen));
  HIPCHK(b.cell_b.reserve(nl)); HIPCHK(b.alive_b.reserve(nl)); HIPCHK(b.req_b.reserve(nl)); HIPCHK(b.porig_b.reserve(nl));
  HIPCHK(b.perm.reserve(nl));
  HIPCHK(b.pscratch.reserve(sort_particles_scratch(b.ds->g.C, (int)nl)));
  b.ld = nl;
  return 0;
}

Batch* get_batch(Server& sv, int ds, int scheme) {
  auto key = std::make_pair(ds, scheme);
  auto it = sv.batches.find(key);
  if (it != sv.batches.end()) return it->second.get();
  auto b = std::make_unique<Batch>();
  b->ds = sv.ds[ds].get();
  b->scheme = scheme;
  if (hipSetDevice(b->ds->device) != hipSuccess || hipStreamCreate(&b->s) != hipSuccess) return nullptr;
  const int C = b->ds->g.C;
  if (b->pstart.reserve((size_t)C + 2) != hipSuccess || b->cell_list.reserve(nonempty_list_capacity(C)) != hipSuccess ||
      b->list_count.reserve(1) != hipSuccess ||
      b->pscratch.reserve(sort_particles_scratch(C, 1)) != hipSuccess ||
      b->counters.reserve(4) != hipSuccess || b->err_req.reserve(RSLOTS) != hipSuccess ||
      b->dt_req.reserve(RSLOTS) != hipSuccess || b->done_req.reserve(RSLOTS) != hipSuccess ||
      b->out_req.reserve(RSLOTS) != hipSuccess ||
      hipHostMalloc((void**)&b->h_count, sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc((void**)&b->h_err, sizeof(unsigned long long) * RSLOTS) != hipSuccess)
    return nullptr;
  (void)hipMemsetAsync(b->done_req.p, 0, RSLOTS, b->s);
  (void)hipMemsetAsync(b->counters.p, 0, 4 * sizeof(unsigned long long), b->s);
  Batch* raw = b.get();
  sv.batches[key] = std::move(b);
  return raw;
}

// Append a request's particles at the batch tail: located, weighted, marked with
// its slot.  Velocities are computed for the whole batch at the tick's start.
int admit(Server& sv, Batch& b, ServeReq& r) {
  const char* fn = "lagrbf_serve";
  Ctx& d = *b.ds;
  int slot = -1;
  for (int k = 0; k < RSLOTS; k++) if (!b.slots[k]) { slot = k; break; }
  if (slot < 0) return set_err(E_STATE, fn, "no free request slot in the batch");
  const int dim = d.g.dim, off = b.Np, n = r.n;
  CHK(batch_reserve(b, off + n));
  HIPCHK(b.stage.reserve((size_t)n * dim));
  // stream-ordered allocation: no device-wide synchronisation when it is freed
  HIPCHK(hipMallocAsync((void**)&r.d_out, sizeof(double) * (size_t)n * dim, b.s));
  HIPCHK(hipMemsetAsync(r.d_out, 0, sizeof(double) * (size_t)n * dim, b.s));
  HIPCHK(hipMemcpyAsync(b.stage.p, r.coords.data(), sizeof(double) * (size_t)n * dim, hipMemcpyHostToDevice, b.s));
  launch_aos_to_soa(b.stage.p, n, dim, b.px.p + off, b.ld, b.s);
  launch_fill_i32(b.req.p + off, n, slot, b.s);
  launch_iota(b.porig.p + off, n, b.s);
  launch_fill_i32(b.alive.p + off, n, 1, b.s);
  launch_locate_particles(d.g, b.px.p + off, b.ld, n, b.cell.p + off, b.alive.p + off, b.counters.p, b.s);
  launch_weights(d.g, d.pux.p, d.puy.p, d.puz.p, b.px.p + off, b.ld, n, b.cell.p + off, b.alive.p + off, b.w.p + off, b.s);
  b.h_dt[slot] = r.dt;
  HIPCHK(hipMemcpyAsync(b.dt_req.p + slot, &b.h_dt[slot], sizeof(double), hipMemcpyHostToDevice, b.s));
  HIPCHK(hipMemcpyAsync(b.out_req.p + slot, &r.d_out, sizeof(double*), hipMemcpyHostToDevice, b.s));
  HIPCHK(hipStreamSynchronize(b.s));   // host sources above are short-lived
  b.slots[slot] = &r;
  r.slot = slot;
  r.t_admit = sv.now();
  b.Np = off + n;
  b.appended = true;
  return 0;
}

// untuned batches: FP64 listed cells in contiguous XCD runs; FP32 listed cells in
// runs of 4 with the accumulation on the matrix cores (the fastest FP32 shape at
// serving densities: 1.46e8 against 1.16e8 particle-steps/s for the FP32 form of the
// FP64 choice, and 1.30e8 with every batch density tuned)
static constexpr int SMALL_BATCH_VARIANT = 1;
static constexpr int SMALL_BATCH_VARIANT_F32 = 16;
static constexpr int SERVE_TUNE_MIN_DEFAULT = 1 << 22;

// Evaluation launch shape chosen by the dataset's autotuner for this batch's
// particle density (the first batch of a density times every shape).
void eval_batch(Batch& b, double* out) {
  Ctx& d = *b.ds;
  const bool f32 = eval_f32(d);
  EvalExt ev = eval_ext(d, -1);
  ev.pcell = b.cell.p;     // the batch's particles, for the polynomial pass
  ev.palive = b.alive.p;
  const int per_cell = (b.Np + d.g.C - 1) / std::max(d.g.C, 1);
  const int block = std::min(256, std::max(64, (per_cell + 63) / 64 * 64));
  auto launch = [&](int v) {
    bool listed;
    int grp, pipe, waves;
    eval_variant(v, listed, grp, pipe, waves);
    launch_eval_blk(d.g, eval_ktype(d), f32, b.px.p, b.ld, b.perm.p, b.pstart.p, b.w.p,
                    f32 ? (const void*)d.blk32.p : (const void*)d.blk64.p, d.nf, d.cell_start.p, d.g.dim, out, block,
                    b.Np, listed ? b.cell_list.p : nullptr, listed ? b.list_count.p : nullptr, grp, pipe, waves, ev, b.s);
  };
  // the batch is stored in cell order at every tick
  const int bucket = density_bucket(b.Np, d.g.C);
  // A batch below serve_tune_min particles runs a fixed listed-cell shape untimed:
  // batch sizes wander over many density buckets as requests come and go, and a
  // trial of every shape there costs more than the fastest shape saves on so few
  // particles.
  const bool small = d.tuner.forced < 0 && b.Np < d.serve_tune_min;
  const int small_v = f32 ? SMALL_BATCH_VARIANT_F32 : SMALL_BATCH_VARIANT;
  const int planned = small ? small_v : d.tuner.planned(bucket);
  // the non-empty cell list, for the shapes that use it (and the tuner's trials)
  if (planned < 0 || eval_variant_listed(planned))
    launch_nonempty_cells(b.pstart.p, d.g.C, b.cell_list.p, b.list_count.p, b.s);
  if (small)
    launch(small_v);
  else
    d.tuner.run(bucket, d.nf == d.g.dim, f32, true, true, b.s, launch);
}

void group(Batch& b) {
  launch_sort_particles(b.cell.p, b.alive.p, b.Np, b.ds->g.C, b.pstart.p, b.perm.p, b.pscratch.p, b.s);
}

StageArgs serve_stage_args(const Batch& b, int nv, const double* const* v, const double* a, double mul, double div,
                           int project, int proj_in_loop, int write_vel) {
  StageArgs s{};
  s.nv = nv;
  for (int j = 0; j < nv; j++) { s.v[j] = v[j]; s.a[j] = a[j]; }
  s.scale = 0.0;
  s.project = project; s.proj_in_loop = proj_in_loop; s.write_vel = write_vel;
  s.rmin = b.ds->rmin; s.rmax = b.ds->rmax;
  s.dt_req = b.dt_req.p; s.req = b.req.p; s.mul = mul; s.div = div;
  return s;
}

void stage(Batch& b, const StageArgs& a) {
  Ctx& d = *b.ds;
  launch_stage(d.g, d.pux.p, d.puy.p, d.puz.p, b.px.p, b.x0.p, b.vel.p, b.ld, b.Np, b.cell.p, b.alive.p, b.counters.p,
               b.w.p, a, b.s);
}

// One tick of a batch, enqueued on its stream (no host synchronisation).
int batch_tick(Batch& b) {
  Ctx& d = *b.ds;
  const int dim = d.g.dim, C = d.g.C;
  // cell order (dead particles at the tail, counted for the cut after the tick)
  group(b);
  ReorderSet r{};
  r.src[0] = b.px.p; r.dst[0] = b.px_b.p; r.rows[0] = dim;
  r.src[1] = b.vel.p; r.dst[1] = b.vel_b.p; r.rows[1] = dim;
  r.src[2] = b.w.p; r.dst[2] = b.w_b.p; r.rows[2] = d.g.nen;
  r.nd = 3;
  r.isrc[0] = b.cell.p; r.idst[0] = b.cell_b.p;
  r.isrc[1] = b.alive.p; r.idst[1] = b.alive_b.p;
  r.isrc[2] = b.req.p; r.idst[2] = b.req_b.p;
  r.isrc[3] = b.porig.p; r.idst[3] = b.porig_b.p;
  r.ni = 4;
  HIPCHK(launch_reorder(r, b.perm.p, b.Np, b.ld, b.s));
  swap_buf(b.px, b.px_b); swap_buf(b.vel, b.vel_b); swap_buf(b.w, b.w_b);
  swap_buf(b.cell, b.cell_b); swap_buf(b.alive, b.alive_b); swap_buf(b.req, b.req_b); swap_buf(b.porig, b.porig_b);
  launch_iota(b.perm.p, b.Np, b.s);
  HIPCHK(hipMemcpyAsync(b.h_count, b.pstart.p + C, sizeof(int32_t), hipMemcpyDeviceToHost, b.s));
  if (b.appended) {
    // velocity at the start for the admitted particles (the others recompute the
    // same values: a particle's velocity depends on its position only)
    eval_batch(b, b.vel.p);
    b.appended = false;
  }
  switch (b.scheme) {
    case 0: {   // runge_kutta_1.c:37-75 (x0 = the current position)
      launch_save(b.px.p, b.x0.p, b.vel.p, nullptr, b.ld, b.Np, dim, b.alive.p, b.s);
      const double* v[1] = {b.vel.p};
      const double a[1] = {1.0};
      stage(b, serve_stage_args(b, 1, v, a, 1.0, 1.0, 0, 0, 0));
      break;
    }
    case 1: {   // runge_kutta_2.c:38-91
      launch_save(b.px.p, b.x0.p, b.vel.p, nullptr, b.ld, b.Np, dim, b.alive.p, b.s);
      const double* v[1] = {b.vel.p};
      const double a[1] = {1.0};
      stage(b, serve_stage_args(b, 1, v, a, 0.5, 1.0, 0, 0, 0));
      group(b);
      eval_batch(b, b.vel.p);
      stage(b, serve_stage_args(b, 1, v, a, 1.0, 1.0, 0, 0, 0));
      break;
    }
    case 2: {   // runge_kutta_4.c:45-156
      static const double cc[3] = {0.5, 0.5, 1.0};
      static const double bw[4] = {1.0, 2.0, 2.0, 1.0};
      launch_save(b.px.p, b.x0.p, b.vel.p, b.vadv[0].p, b.ld, b.Np, dim, b.alive.p, b.s);
      for (int s = 0; s < 3; s++) {
        const double* v[1] = {b.vadv[s].p};
        const double a[1] = {1.0};
        stage(b, serve_stage_args(b, 1, v, a, cc[s], 1.0, 1, 0, 1));
        group(b);
        eval_batch(b, b.vadv[s + 1].p);
      }
      const double* v[4] = {b.vadv[0].p, b.vadv[1].p, b.vadv[2].p, b.vadv[3].p};
      stage(b, serve_stage_args(b, 4, v, bw, 1.0, 6.0, 1, 1, 0));
      break;
    }
    default: {  // dormand_prince.c:216-310, step size per request
      static const double A[6][6] = {
          {1.0 / 5.0, 0, 0, 0, 0, 0},
          {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
          {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
          {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
          {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
          {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
      launch_save(b.px.p, b.x0.p, b.vel.p, b.vadv[0].p, b.ld, b.Np, dim, b.alive.p, b.s);
      for (int s = 0; s < 6; s++) {
        const double* v[7];
        for (int j = 0; j <= s; j++) v[j] = b.vadv[j].p;
        stage(b, serve_stage_args(b, s + 1, v, A[s], 1.0, 1.0, 1, 0, 1));
        group(b);
        eval_batch(b, b.vadv[s + 1].p);
      }
      const double* vv[7];
      for (int j = 0; j < 7; j++) vv[j] = b.vadv[j].p;
      HIPCHK(hipMemsetAsync(b.err_req.p, 0, sizeof(unsigned long long) * RSLOTS, b.s));
      launch_dopri_err_req(b.px.p, b.x0.p, vv, b.ld, b.Np, dim, b.alive.p, b.req.p, b.dt_req.p, d.dopri_atol,
                           d.dopri_rtol, b.err_req.p, RSLOTS, b.s);
      HIPCHK(hipMemcpyAsync(b.h_err, b.err_req.p, sizeof(unsigned long long) * RSLOTS, hipMemcpyDeviceToHost, b.s));
      break;
    }
  }
  // velocity at the end of the step (the next step's start)
  group(b);
  eval_batch(b, b.vel.p);
  return 0;
}

// After the batch's stream has drained: cut the tail, count the step, control the
// DOPRI step sizes, retire finished requests.
int batch_finish(Server& sv, Batch& b) {
  Ctx& d = *b.ds;
  const int dim = d.g.dim;
  // the tick's device-to-host copies (alive count, DOPRI maxima) must have landed
  HIPCHK(hipStreamSynchronize(b.s));
  b.Np = *b.h_count;
  std::vector<uint8_t> done(RSLOTS, 0);
  bool any_done = false, dt_changed = false;
  for (int k = 0; k < RSLOTS; k++) {
    ServeReq* r = b.slots[k];
    if (!r) continue;
    r->steps_done++;
    sv.particle_steps += r->n;
    if (b.scheme == 3) {
      // AdvectionComputeAdaptiveTimeStep_DOPRI, dormand_prince.c:179-214, per request
      double max_err;
      const unsigned long long bits = b.h_err[k];
      std::memcpy(&max_err, &bits, sizeof(double));
      if (bits == 0ull) max_err = 0.0;
      const double fac_max = max_err <= 1.0 ? d.dopri_fac_max : 1.0;
      const double gfac = d.dopri_fac * std::pow(1.0 / max_err, 1.0 / 5.0);
      const double inner = (d.dopri_fac_min < gfac) ? gfac : d.dopri_fac_min;
      const double f = (fac_max < inner) ? fac_max : inner;
      b.h_dt[k] = b.h_dt[k] * f;
      dt_changed = true;
    }
    if (r->steps_done >= r->nsteps) { done[k] = 1; any_done = true; }
  }
  if (dt_changed) HIPCHK(hipMemcpyAsync(b.dt_req.p, b.h_dt.data(), sizeof(double) * RSLOTS, hipMemcpyHostToDevice, b.s));
  if (any_done) {
    HIPCHK(hipMemcpyAsync(b.done_req.p, done.data(), RSLOTS, hipMemcpyHostToDevice, b.s));
    launch_retire(b.px.p, b.ld, b.Np, dim, b.alive.p, b.req.p, b.porig.p, b.done_req.p, b.out_req.p, b.s);
    HIPCHK(hipMemsetAsync(b.done_req.p, 0, RSLOTS, b.s));
    for (int k = 0; k < RSLOTS; k++) {
      if (!done[k]) continue;
      ServeReq* r = b.slots[k];
      r->result.resize((size_t)r->n * dim);
      HIPCHK(hipMemcpyAsync(r->result.data(), r->d_out, sizeof(double) * r->result.size(), hipMemcpyDeviceToHost, b.s));
    }
    HIPCHK(hipStreamSynchronize(b.s));
    const double t = sv.now();
    for (int k = 0; k < RSLOTS; k++) {
      if (!done[k]) continue;
      ServeReq* r = b.slots[k];
      (void)hipFreeAsync(r->d_out, b.s);
      r->d_out = nullptr;
      r->done = true;
      r->t_done = t;
      sv.reserved -= (double)r->bytes;
      sv.completed += 1;
      b.slots[k] = nullptr;
    }
  } else {
    HIPCHK(hipStreamSynchronize(b.s));   // step-size upload done before the host array changes
  }
  return 0;
}

int serve_tick(Server& sv) {
  const auto t0 = std::chrono::steady_clock::now();
  // (1) admission, first come first served, while the HBM budget allows
  while (!sv.queue.empty()) {
    ServeReq* r = sv.queue.front();
    if (sv.reserved + (double)r->bytes > sv.budget && sv.reserved > 0) break;
    Batch* b = get_batch(sv, r->ds, r->scheme);
    if (!b) return set_err(E_LIB, "lagrbf_serve_step", "cannot create a batch");
    int busy = 0;
    for (auto* q : b->slots) busy += q != nullptr;
    if (busy >= RSLOTS) break;
    CHK(admit(sv, *b, *r));
    sv.reserved += (double)r->bytes;
    sv.queue.pop_front();
  }
  // (2) one step of every active batch, each on its own stream
  const auto t1 = std::chrono::steady_clock::now();
  std::vector<Batch*> act;
  for (auto& kv : sv.batches) {
    Batch& b = *kv.second;
    bool any = false;
    for (auto* q : b.slots) any = any || q != nullptr;
    if (!any) continue;
    CHK(batch_tick(b));
    act.push_back(&b);
  }
  sv.enqueue_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
  // (3) bookkeeping
  double active = 0;
  for (Batch* b : act) {
    CHK(batch_finish(sv, *b));
    for (auto* q : b->slots) active += q != nullptr;
  }
  sv.ticks += 1;
  sv.max_active = std::max(sv.max_active, active);
  sv.tick_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace

}  // namespace lrb

using namespace lrb;

extern "C" {

void* lagrbf_serve_create(double hbm_budget_bytes) {
  auto* sv = new Server();
  if (hbm_budget_bytes > 0) sv->budget = hbm_budget_bytes;
  return sv;
}

void lagrbf_serve_destroy(void* srv) {
  auto* sv = static_cast<Server*>(srv);
  if (!sv) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : sv->reqs)
    if (kv.second->d_out) (void)hipFree(kv.second->d_out);
  delete sv;
}

int32_t lagrbf_serve_add_dataset(void* srv, const char* options, int32_t n, double* coords, double* values) {
  const char* fn = "lagrbf_serve_add_dataset";
  auto* sv = static_cast<Server*>(srv);
  if (!sv || !options || !coords || !values) return -set_err(E_NULL, fn, "NULL argument");
  Options db;
  db.insert_string(options);
  OptView v;
  v.db = db.snapshot();
  Ctx* c = create_from_view(v);
  if (!c) return -E_WRONG;
  std::unique_ptr<Ctx> own(c);
  bool bad = false;
  c->serve_tune_min = SERVE_TUNE_MIN_DEFAULT;
  v.get_int("lagrbf_serve_tune_min_particles", c->serve_tune_min, bad);
  if (bad || c->serve_tune_min < 0)
    return -set_err(E_WRONG, fn, "-lagrbf_serve_tune_min_particles must be a non-negative integer");
  if (c->nf != c->g.dim || !eval_blk_supported(c->nf, c->g.dim, false))
    return -set_err(E_WRONG, fn, "a serving dataset carries the velocity: n_interpolation_fields must equal the dimension");
  if (hipSetDevice(c->device) != hipSuccess) return -set_err(E_LIB, fn, "hipSetDevice failed");
  int e = fill_data(*c, n, coords, values, nullptr);
  if (!e) e = partitions_setup(*c, true);
  if (!e) e = ensure_blk(*c, eval_f32(*c));
  if (e) return -e;
  release_factors(*c);
  sv->ds.push_back(std::move(own));
  return (int32_t)sv->ds.size() - 1;
}

int64_t lagrbf_serve_submit(void* srv, int32_t dataset, int32_t scheme, int32_t nsteps, double dt, int32_t n,
                            const double* coords) {
  const char* fn = "lagrbf_serve_submit";
  auto* sv = static_cast<Server*>(srv);
  if (!sv || !coords) return -set_err(E_NULL, fn, "NULL argument");
  if (dataset < 0 || dataset >= (int32_t)sv->ds.size()) return -set_err(E_RANGE, fn, "dataset %d out of range", dataset);
  if (scheme < 0 || scheme > 3) return -set_err(E_RANGE, fn, "scheme %d out of range", scheme);
  if (n <= 0 || nsteps <= 0 || !(dt > 0.0)) return -set_err(E_RANGE, fn, "n, nsteps and dt must be positive");
  Ctx& d = *sv->ds[dataset];
  const int dim = d.g.dim;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < dim; k++) {
      const double x = coords[(size_t)i * dim + k];
      if (!(x >= d.g.O[k] && x <= d.g.L[k]))
        return -set_err(E_WRONG, fn, "Point %d, coordinate %d (%g) outside the dataset's box", i, k, x);
    }
  auto r = std::make_unique<ServeReq>();
  r->id = sv->next_id++;
  r->ds = dataset; r->scheme = scheme; r->nsteps = nsteps; r->dt = dt; r->n = n;
  r->coords.assign(coords, coords + (size_t)n * dim);
  r->bytes = request_bytes(dim, d.g.nen, scheme, n);
  r->t_submit = sv->now();
  sv->queue.push_back(r.get());
  const int64_t id = r->id;
  sv->reqs[id] = std::move(r);
  return id;
}

int32_t lagrbf_serve_step(void* srv) {
  auto* sv = static_cast<Server*>(srv);
  if (!sv) return -set_err(E_NULL, "lagrbf_serve_step", "NULL server");
  int e = serve_tick(*sv);
  if (e) return -e;
  int32_t pending = (int32_t)sv->queue.size();
  for (auto& kv : sv->batches)
    for (auto* q : kv.second->slots) pending += q != nullptr;
  return pending;
}

int32_t lagrbf_serve_poll(void* srv, int64_t id, double* out, double* times) {
  auto* sv = static_cast<Server*>(srv);
  if (!sv) return -set_err(E_NULL, "lagrbf_serve_poll", "NULL server");
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) return -set_err(E_RANGE, "lagrbf_serve_poll", "unknown request %lld", (long long)id);
  ServeReq& r = *it->second;
  if (times) { times[0] = r.t_submit; times[1] = r.t_admit; times[2] = r.t_done; }
  if (!r.done) return 0;
  if (out) std::memcpy(out, r.result.data(), sizeof(double) * r.result.size());
  return 1;
}

int lagrbf_serve_release(void* srv, int64_t id) {
  auto* sv = static_cast<Server*>(srv);
  if (!sv) return set_err(E_NULL, "lagrbf_serve_release", "NULL server");
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end() || !it->second->done) return set_err(E_STATE, "lagrbf_serve_release", "request not finished");
  sv->reqs.erase(it);
  return 0;
}

int lagrbf_serve_info(void* srv, const char* key, double* value) {
  auto* sv = static_cast<Server*>(srv);
  if (!sv || !key || !value) return set_err(E_NULL, "lagrbf_serve_info", "NULL argument");
  const std::string k(key);
  if (k == "completed") *value = sv->completed;
  else if (k == "particle_steps") *value = sv->particle_steps;
  else if (k == "ticks") *value = sv->ticks;
  else if (k == "reserved") *value = sv->reserved;
  else if (k == "queued") *value = (double)sv->queue.size();
  else if (k == "max_active") *value = sv->max_active;
  else if (k == "enqueue_ms") *value = sv->enqueue_ms;
  else if (k == "tick_ms") *value = sv->tick_ms;
  else if (k == "now") *value = sv->now();
  else if (k == "datasets") *value = (double)sv->ds.size();
  else return set_err(E_WRONG, "lagrbf_serve_info", "unknown key '%s'", key);
  return 0;
}

}  // extern "C"
