// Evaluation autotuner: the PU evaluation has several launch shapes with the same
// per-particle arithmetic (one block per cell; one block per listed non-empty
// cell, in contiguous XCD runs or dealt round-robin singly or in runs of 4 or 16
// cells; lane groups of 32, 16 or 8 particles with several cells per wave; one or
// two table rows per wait).
// Which is fastest depends on the particles per cell and the dataset, so the
// first evaluation in each density bucket (particles per cell, half-octave
// steps) times every valid shape on the live launch and keeps the fastest.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <vector>

namespace lrb {

constexpr int N_EVAL_VARIANTS = 18;

// variant -> (evaluate listed cells only, lane group size; 0 = one block per cell,
// -K = one block per listed cell with runs of K consecutive listed cells dealt
// round-robin over the XCDs instead of one contiguous run per XCD: better balanced
// when cells differ widely in work, and with K > 1 neighbouring cells still share
// an L2; rows per wait of the FP32 node loop, 1 or 2: variants 6-8 are 0, 1 and 5
// with two; FP32 occupancy floor in waves per SIMD, 0 = the default: variant 9 is
// 8 with 6 waves, which leaves the loop its registers without spilling; variants
// 10-11 are 5 and 12-13 are 9 in runs of 4 and 16 cells; pipe 3: the accumulation
// on the matrix cores (FP32, up to 4 fields), variants 14-17 are 1, 5, 10 and 11
// with it)
inline void eval_variant(int v, bool& listed, int& group, int& pipe, int& waves) {
  static const int G[N_EVAL_VARIANTS] = {0, 0, 32, 16, 8, -1, 0, 0, -1, -1, -4, -16, -4, -16, 0, -1, -4, -16};
  static const int L[N_EVAL_VARIANTS] = {0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static const int PIPE[N_EVAL_VARIANTS] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 1, 1, 2, 2, 3, 3, 3, 3};
  static const int WAVES[N_EVAL_VARIANTS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 6, 0, 0, 6, 6, 0, 0, 0, 0};
  listed = L[v] != 0;
  group = G[v];
  pipe = PIPE[v];
  waves = WAVES[v];
}

// the same shape with one row per wait (the FP64 evaluation has no two-row loop)
inline int eval_variant_pipe1(int v) {
  static const int P1[N_EVAL_VARIANTS] = {0, 1, 2, 3, 4, 5, 0, 1, 5, 5, 10, 11, 10, 11, 1, 5, 10, 11};
  return P1[v];
}

// How far a shape spreads a cell's neighbourhood over the XCDs' L2 caches: 0 keeps
// neighbouring cells on one XCD (one block per cell in contiguous ranges, listed
// cells in contiguous runs or in runs of 4), 1 runs of 16 and the lane-group shapes,
// 2 single cells dealt round-robin.  Measured HBM fetch per FP64 evaluation at the
// r^3 + linear workload of 2^20 nodes: 1.5, 3.7 and 5.6 GB.
inline int eval_variant_locality(int v) {
  bool listed;
  int group, pipe, waves;
  eval_variant(v, listed, group, pipe, waves);
  if (group == 0 || group == -4) return 0;
  if (group == -1) return 2;
  return 1;
}

inline bool eval_variant_listed(int v) {
  bool listed;
  int group, pipe, waves;
  eval_variant(v, listed, group, pipe, waves);
  return listed;
}

inline int density_bucket(int64_t np, int ncells) {
  const double ppc = (double)np / (double)(ncells > 0 ? ncells : 1);
  return (int)std::floor(2.0 * std::log2(ppc + 0.0625));
}

struct EvalTuner {
  // Per density bucket:
  //  - a swarm still in fill order (the first evaluation, before an RK step stores
  //    the swarm in cell order) runs the default shape untimed: timings there
  //    favour other shapes than the steady state;
  //  - the first evaluation in cell order times every valid shape back to back on
  //    the live launch (lane groups only below 64 particles per cell, where they
  //    can win);
  //  - back-to-back launches do not see the caches and clocks of the steady state
  //    (an RK stage runs between two evaluations), so the shapes within 5% of the
  //    best trial time get a run-off: the bucket's next evaluations use one of them
  //    each, timed with stream events, and the fastest of those stays.
  // All shapes write identical values, so none of this changes results.
  static constexpr float RUNOFF_MARGIN = 1.05f;
  // run-off times within this factor of the fastest are a tie, which the shape that
  // moves fewer bytes wins: concurrent work (other serving batches, the set-up's
  // solves) shares HBM and L2 with the evaluation
  static constexpr float TIE_MARGIN = 1.02f;
  static constexpr int DEFAULT_VARIANT = 5;
  static constexpr int GROUPED_MAX_BUCKET = 12;   // bucket of 64 particles per cell
  struct Runoff {
    std::vector<int> cand;                 // shapes in the run-off
    std::vector<hipEvent_t> ev;            // start/stop per launched candidate
    size_t next = 0;
  };
  std::map<int, int> choice;    // density bucket -> variant
  std::map<int, bool> final_;   // density bucket -> choice stays
  std::map<int, Runoff> runoff; // density bucket -> run-off in progress
  std::map<int, float> best_ms;
  int forced = -1;              // a fixed variant (option lagrbf_eval_variant), -1: tune
  bool untuned = false;         // run DEFAULT_VARIANT without trials (unless forced)

  EvalTuner() = default;
  EvalTuner(const EvalTuner&) = delete;
  EvalTuner& operator=(const EvalTuner&) = delete;
  EvalTuner& operator=(EvalTuner&& o) noexcept {
    release();
    choice = std::move(o.choice);
    final_ = std::move(o.final_);
    runoff = std::move(o.runoff);
    best_ms = std::move(o.best_ms);
    forced = o.forced;
    untuned = o.untuned;
    o.runoff.clear();
    return *this;
  }
  ~EvalTuner() { release(); }

  // events of unfinished run-offs
  void release() {
    for (auto& kv : runoff)
      for (hipEvent_t e : kv.second.ev) (void)hipEventDestroy(e);
    runoff.clear();
  }

  // variant run() will launch without timing candidates, or -1 when it will tune
  int planned(int bucket) const {
    if (forced >= 0) return forced;
    if (untuned) return DEFAULT_VARIANT;
    auto it = choice.find(bucket);
    auto f = final_.find(bucket);
    return it == choice.end() || f == final_.end() || !f->second ? -1 : it->second;
  }

  // valid(v): whether variant v applies (lane groups need the velocity fields
  // only; two rows per wait exist for FP32 only)
  static bool valid(int v, bool grouped_ok, bool pipe2_ok) {
    bool listed;
    int group, pipe, waves;
    eval_variant(v, listed, group, pipe, waves);
    return !(group > 0 && !grouped_ok) && !(pipe >= 2 && !pipe2_ok);
  }

  // the variant a forced choice stands for where it does not apply
  static int fallback(int v, bool grouped_ok, bool pipe2_ok) {
    if (!pipe2_ok) v = eval_variant_pipe1(v);
    bool listed;
    int group, pipe, waves;
    eval_variant(v, listed, group, pipe, waves);
    return (group > 0 && !grouped_ok) ? 1 : v;
  }

  // launch(v) enqueues the evaluation with variant v on stream s; returns the
  // variant used.  cell_ordered: the swarm is stored in cell order; runoff: a
  // run-off may follow the trial (callers about to capture the step in a graph
  // take the trial's choice at once).
  template <typename F>
  int run(int bucket, bool grouped_ok, bool pipe2_ok, bool cell_ordered, bool allow_runoff, hipStream_t s,
          F&& launch) {
    if (forced >= 0) {
      const int v = fallback(forced, grouped_ok, pipe2_ok);
      launch(v);
      return v;
    }
    if (untuned) {
      launch(DEFAULT_VARIANT);
      return DEFAULT_VARIANT;
    }
    const int planned_v = planned(bucket);
    if (planned_v >= 0) {
      launch(planned_v);
      return planned_v;
    }
    auto ro = runoff.find(bucket);
    if (ro != runoff.end()) return runoff_step(bucket, ro->second, s, launch);
    if (!cell_ordered) {
      launch(DEFAULT_VARIANT);
      return DEFAULT_VARIANT;
    }
    const bool grouped = grouped_ok && bucket < GROUPED_MAX_BUCKET;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      launch(DEFAULT_VARIANT);
      return DEFAULT_VARIANT;
    }
    launch(DEFAULT_VARIANT);   // warm caches and code objects before timing
    float tmin[N_EVAL_VARIANTS];
    float best = 1e30f;
    int bv = DEFAULT_VARIANT;
    for (int v = 0; v < N_EVAL_VARIANTS; v++) {
      tmin[v] = 1e30f;
      if (!valid(v, grouped, pipe2_ok)) continue;
      (void)hipEventRecord(e0, s);
      launch(v);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      tmin[v] = ms;
      if (ms < best) { best = ms; bv = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    choice[bucket] = bv;
    best_ms[bucket] = best;
    Runoff r;
    if (allow_runoff)
      for (int v = 0; v < N_EVAL_VARIANTS; v++)
        if (tmin[v] <= best * RUNOFF_MARGIN) r.cand.push_back(v);
    final_[bucket] = r.cand.size() <= 1;
    if (r.cand.size() > 1) runoff[bucket] = std::move(r);
    return bv;
  }

 private:

  template <typename F>
  int runoff_step(int bucket, Runoff& r, hipStream_t s, F&& launch) {
    const int v = r.cand[r.next];
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      if (a) (void)hipEventDestroy(a);
      final_[bucket] = true;   // keep the trial's choice
      for (hipEvent_t e : r.ev) (void)hipEventDestroy(e);
      runoff.erase(bucket);
      const int c = choice[bucket];
      launch(c);
      return c;
    }
    (void)hipEventRecord(a, s);
    launch(v);
    (void)hipEventRecord(b, s);
    r.ev.push_back(a);
    r.ev.push_back(b);
    r.next += 1;
    if (r.next == r.cand.size()) {
      // every candidate ran once in the steady state: the fastest stays, ties going
      // to the shape that keeps neighbourhoods in fewer L2 caches
      std::vector<float> t(r.cand.size(), 1e30f);
      float best = 1e30f;
      for (size_t k = 0; k < r.cand.size(); k++) {
        (void)hipEventSynchronize(r.ev[2 * k + 1]);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.ev[2 * k], r.ev[2 * k + 1]) == hipSuccess) t[k] = ms;
        best = t[k] < best ? t[k] : best;
      }
      int bk = -1;
      for (size_t k = 0; k < r.cand.size(); k++) {
        if (!(t[k] <= best * TIE_MARGIN)) continue;
        if (bk < 0) { bk = (int)k; continue; }
        const int lk = eval_variant_locality(r.cand[k]), lb = eval_variant_locality(r.cand[bk]);
        if (lk < lb || (lk == lb && t[k] < t[bk])) bk = (int)k;
      }
      // the winner is read before the run-off (which r refers to) is erased
      const int bv = bk >= 0 ? r.cand[bk] : choice[bucket];
      const float bt = bk >= 0 ? t[bk] : best;
      (void)hipGetLastError();
      for (hipEvent_t e : r.ev) (void)hipEventDestroy(e);
      runoff.erase(bucket);
      choice[bucket] = bv;
      best_ms[bucket] = bt;
      final_[bucket] = true;
    }
    return v;
  }
};

}  // namespace lrb
