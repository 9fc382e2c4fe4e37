// Evaluation autotuner: the PU evaluation has several launch shapes with the same
// per-particle arithmetic (one block per cell; one block per listed non-empty
// cell; lane groups of 32, 16 or 8 particles with several cells per wave).
// Which is fastest depends on the particles per cell and the dataset, so the
// first evaluation in each density bucket (particles per cell, half-octave
// steps) times every valid shape on the live launch and keeps the fastest.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>

namespace lrb {

constexpr int N_EVAL_VARIANTS = 6;

// variant -> (evaluate listed cells only, lane group size; 0 = one block per cell,
// -1 = one block per listed cell with the blocks dealt round-robin over the XCDs
// instead of in contiguous runs: better balanced when cells differ widely in work)
inline void eval_variant(int v, bool& listed, int& group) {
  static const int G[N_EVAL_VARIANTS] = {0, 0, 32, 16, 8, -1};
  listed = v > 0;
  group = G[v];
}

inline int density_bucket(int64_t np, int ncells) {
  const double ppc = (double)np / (double)(ncells > 0 ? ncells : 1);
  return (int)std::floor(2.0 * std::log2(ppc + 0.0625));
}

struct EvalTuner {
  std::map<int, int> choice;   // density bucket -> variant
  std::map<int, float> best_ms;
  int forced = -1;             // a fixed variant (option lagrbf_eval_variant), -1: tune

  // variant run() will launch without timing candidates, or -1 when it will tune
  int planned(int bucket) const {
    if (forced >= 0) return forced;
    auto it = choice.find(bucket);
    return it == choice.end() ? -1 : it->second;
  }

  // launch(v) enqueues the evaluation with variant v on stream s.  Returns the
  // variant used.  Candidates run back to back on the same inputs; each writes
  // the same values, so the output of the last one stands.
  template <typename F>
  int run(int bucket, bool grouped_ok, hipStream_t s, F&& launch) {
    if (forced >= 0) {
      bool listed;
      int group;
      eval_variant(forced, listed, group);
      const int v = (group > 0 && !grouped_ok) ? 1 : forced;
      launch(v);
      return v;
    }
    auto it = choice.find(bucket);
    if (it != choice.end()) {
      launch(it->second);
      return it->second;
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
      bool listed;
      int group;
      eval_variant(v, listed, group);
      if (group > 0 && !grouped_ok) continue;
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
    choice[bucket] = bv;
    best_ms[bucket] = best;
    return bv;
  }
};

}  // namespace lrb
