// Evaluation autotuner: the PU evaluation has several launch shapes with the same
// per-particle arithmetic (one block per cell; one block per listed non-empty
// cell, in contiguous XCD runs or round-robin; lane groups of 32, 16 or 8
// particles with several cells per wave; one or two table rows per wait).
// Which is fastest depends on the particles per cell and the dataset, so the
// first evaluation in each density bucket (particles per cell, half-octave
// steps) times every valid shape on the live launch and keeps the fastest.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>

namespace lrb {

constexpr int N_EVAL_VARIANTS = 9;

// variant -> (evaluate listed cells only, lane group size; 0 = one block per cell,
// -1 = one block per listed cell with the blocks dealt round-robin over the XCDs
// instead of in contiguous runs: better balanced when cells differ widely in work;
// rows per wait of the FP32 node loop, 1 or 2: variants 6-8 are 0, 1 and 5 with two)
inline void eval_variant(int v, bool& listed, int& group, int& pipe) {
  static const int G[N_EVAL_VARIANTS] = {0, 0, 32, 16, 8, -1, 0, 0, -1};
  static const int L[N_EVAL_VARIANTS] = {0, 1, 1, 1, 1, 1, 0, 1, 1};
  listed = L[v] != 0;
  group = G[v];
  pipe = v >= 6 ? 2 : 1;
}

// the same shape with one row per wait (the FP64 evaluation has no two-row loop)
inline int eval_variant_pipe1(int v) { return v == 6 ? 0 : v == 7 ? 1 : v == 8 ? 5 : v; }

inline bool eval_variant_listed(int v) {
  bool listed;
  int group, pipe;
  eval_variant(v, listed, group, pipe);
  return listed;
}

inline int density_bucket(int64_t np, int ncells) {
  const double ppc = (double)np / (double)(ncells > 0 ? ncells : 1);
  return (int)std::floor(2.0 * std::log2(ppc + 0.0625));
}

struct EvalTuner {
  // A choice timed on a swarm still in fill order (the first evaluation, before an
  // RK step stores the swarm in cell order) favours other shapes than the steady
  // state: it is provisional, and the bucket's next evaluation times the
  // candidates again.  A choice timed on a swarm in cell order stays.
  std::map<int, int> choice;   // density bucket -> variant
  std::map<int, bool> final_;  // density bucket -> choice timed on a swarm in cell order
  std::map<int, float> best_ms;
  int forced = -1;             // a fixed variant (option lagrbf_eval_variant), -1: tune

  // variant run() will launch without timing candidates, or -1 when it will tune
  int planned(int bucket) const {
    if (forced >= 0) return forced;
    auto it = choice.find(bucket);
    auto f = final_.find(bucket);
    return it == choice.end() || f == final_.end() || !f->second ? -1 : it->second;
  }

  // launch(v) enqueues the evaluation with variant v on stream s.  Returns the
  // variant used.  Candidates run back to back on the same inputs; each writes
  // the same values, so the output of the last one stands.
  // valid(v): whether variant v applies (lane groups need the velocity fields
  // only; two rows per wait exist for FP32 only)
  static bool valid(int v, bool grouped_ok, bool pipe2_ok) {
    bool listed;
    int group, pipe;
    eval_variant(v, listed, group, pipe);
    return !(group > 0 && !grouped_ok) && !(pipe == 2 && !pipe2_ok);
  }

  // the variant a forced choice stands for where it does not apply
  static int fallback(int v, bool grouped_ok, bool pipe2_ok) {
    if (!pipe2_ok) v = eval_variant_pipe1(v);
    bool listed;
    int group, pipe;
    eval_variant(v, listed, group, pipe);
    return (group > 0 && !grouped_ok) ? 1 : v;
  }

  // cell_ordered: the swarm is stored in cell order (a choice made now stays)
  template <typename F>
  int run(int bucket, bool grouped_ok, bool pipe2_ok, bool cell_ordered, hipStream_t s, F&& launch) {
    if (forced >= 0) {
      const int v = fallback(forced, grouped_ok, pipe2_ok);
      launch(v);
      return v;
    }
    const int planned_v = planned(bucket);
    if (planned_v >= 0) {
      launch(planned_v);
      return planned_v;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      launch(0);
      return 0;
    }
    launch(0);   // warm caches and code objects before timing
    float best = 1e30f;
    int bv = 0;
    for (int v = 0; v < N_EVAL_VARIANTS; v++) {
      if (!valid(v, grouped_ok, pipe2_ok)) continue;
      (void)hipEventRecord(e0, s);
      launch(v);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bv = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the second round stays whatever the order (positions drift between reorders)
    const bool again = choice.count(bucket) > 0;
    choice[bucket] = bv;
    best_ms[bucket] = best;
    final_[bucket] = cell_ordered || again;
    return bv;
  }
};

}  // namespace lrb
