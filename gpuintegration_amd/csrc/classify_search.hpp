// classify_search.hpp -- the heuristic classifier's threshold search
// (Heuristic_classifier::classify, heuristic_classifier.cuh:362-451, with
// get_larger_threshold_results :327-346 and evaluate_error_budget :305-325)
// as a replay over measured candidates, shared by the host and the device.
//
// The search walks over states (t, low, high): raising moves t halfway to
// high (low = t), lowering moves it halfway to low (high = t), and every
// decision depends only on the count of errors above t and their sum.  A
// probe pass (k_multi_threshold) measures a state together with every state
// kTreeDepth moves away; search_replay then runs the reference's loop from
// the start, reading measured states from the cache, until it finishes or
// meets a state not measured yet.  Run to completion this way, the result
// is that of the sequential loop.  The device runs the replay in one
// workgroup (k_search), so the host enqueues probe passes without reading
// anything back; the host runs the same function when it searches itself.
// Plain IEEE double without contraction on both sides, so both make the
// same decisions.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pagani.h"
#include "phases.hpp"

namespace pagani {
namespace internal {

constexpr int kRaiseLimit = 20;  // :335
constexpr int kTreeDepth = 3;
constexpr int kTreeNodes = (2 << kTreeDepth) - 1;
static_assert(kTreeNodes < phases::kMaxThresholds, "one pass measures the whole tree");
// measured states kept by one device search (passes of kTreeNodes each);
// a search that needs more is finished by the host
constexpr int kSearchCache = 16 * kTreeNodes;
// loop steps of one device replay before it gives the search to the host
constexpr int kSearchSteps = 100000;
// probe rounds the host queues between two reads of the outcome: the
// headline's classifications take 3.6 passes on average
constexpr int kSearchRounds = 4;

struct SearchInputs {
  std::uint64_t n;
  double mn, mx;  // smallest and largest error
  double iter_err, iter_fin_err, tot_fin_err;
  double target_error;  // |latest estimate| * epsrel
  double max_budget, max_active;
};

struct SearchStep {
  int done;      // res and last_thr hold the result
  int overflow;  // the replay hit kSearchSteps
  double t, low, high;  // !done: the state to measure next
  double last_thr;      // the last applied threshold
  pagani_class_res res;
};

// The reference's loop.  find(t, cnt, dot) returns false for a state not
// measured yet.
template <class Find>
__host__ __device__ inline void search_replay(const SearchInputs& in, Find&& find, SearchStep& out)
{
#pragma clang fp contract(off)
  pagani_class_res res{};
  res.range_low = in.mn;
  res.range_high = in.mx;
  res.threshold = in.iter_err / static_cast<double>(in.n);
  double max_budget = in.max_budget, max_active = in.max_active;
  std::uint64_t n_inc = 0, n_dec = 0;
  double last_dot = 0.0, last_thr = res.threshold;
  out.done = 0;
  out.overflow = 0;
  int steps = 0;
  do {
    if (++steps > kSearchSteps) {
      out.overflow = 1;
      return;
    }
    if (!res.pass_mem && n_inc <= 20) {
      // get_larger_threshold_results with apply_threshold
      res.pass_mem = 0;
      for (int k = 0; k < kRaiseLimit && !res.pass_mem; ++k) {
        std::uint64_t cnt = 0;
        double dot = 0.0;
        if (!find(res.threshold, cnt, dot)) {
          out.t = res.threshold;
          out.low = res.range_low;
          out.high = res.range_high;
          return;
        }
        last_dot = dot;
        last_thr = res.threshold;
        res.threshold_passes++;
        res.num_active = cnt;
        res.percent_mem_active = static_cast<double>(static_cast<int>(cnt)) /
                                 static_cast<double>(static_cast<int>(in.n));
        res.pass_mem = res.percent_mem_active <= max_active;
        if (!res.pass_mem) {
          res.range_low = res.threshold;
          const double diff = fabs(res.range_high - res.threshold);
          res.threshold += diff * .5;
        }
      }
      n_inc++;
    }
    if (res.pass_mem) {
      // evaluate_error_budget
      const double extra = in.iter_err - last_dot - in.iter_fin_err;
      const double budget = in.target_error - in.tot_fin_err;
      res.pass_errorest_budget = extra <= max_budget * budget;
      res.finished_errorest = extra;
      if (!res.pass_errorest_budget && n_dec <= 20) {
        res.range_high = res.threshold;
        const double diff = fabs(res.range_low - res.threshold);
        res.threshold -= diff * .5;
        res.pass_mem = 0;
        n_dec++;
      }
    }
    const bool exhausted = n_dec >= 20 || n_inc >= 20;
    if (exhausted && max_budget < .7) {
      max_budget += 0.1;
      n_dec = 0;
      n_inc = 0;
      res.range_low = in.mn;
      res.range_high = in.mx;
      res.threshold = in.iter_err / static_cast<double>(in.n);
    } else if (exhausted && max_budget >= .7 && max_active <= .7) {
      max_active += .1;
    } else if (exhausted) {
      break;
    }
  } while (!res.pass_mem || !res.pass_errorest_budget);
  res.max_budget_perc_to_cover = max_budget;
  res.max_active_perc = max_active;
  out.res = res;
  out.last_thr = last_thr;
  out.done = 1;
}

// The states one probe pass measures with (t, low, high): the state and
// every state kTreeDepth moves away, breadth first (node[]), and the pass's
// candidates sorted ascending (NaN measured as +inf, which counts the same:
// err > NaN holds for no error), padded with +inf: slot i of S holds node
// perm[i].
__host__ __device__ inline void search_tree(double t, double low, double high, double* node,
                                            phases::ThresholdSet& S, int* perm)
{
#pragma clang fp contract(off)
  double lo[kTreeNodes], hi[kTreeNodes];
  int k = 0;
  node[k] = t;
  lo[k] = low;
  hi[k] = high;
  ++k;
  for (int i = 0; k < kTreeNodes; ++i) {
    const double x = node[i];
    node[k] = x + fabs(hi[i] - x) * .5;  // raise
    lo[k] = x;
    hi[k] = hi[i];
    ++k;
    node[k] = x - fabs(lo[i] - x) * .5;  // lower
    lo[k] = lo[i];
    hi[k] = x;
    ++k;
  }
  auto key = [&](int i) { return node[i] != node[i] ? HUGE_VAL : node[i]; };
  for (int i = 0; i < kTreeNodes; ++i) perm[i] = i;
  for (int i = 1; i < kTreeNodes; ++i) {  // insertion sort, stable
    const int p = perm[i];
    int j = i - 1;
    while (j >= 0 && key(perm[j]) > key(p)) {
      perm[j + 1] = perm[j];
      --j;
    }
    perm[j + 1] = p;
  }
  S.k = kTreeNodes;
  for (int i = 0; i < phases::kMaxThresholds; ++i) S.t[i] = i < kTreeNodes ? key(perm[i]) : HUGE_VAL;
}

// Device search state, one per workspace (device memory).
struct SearchState {
  SearchInputs in;
  int pending;  // a probe pass of S is queued (its results go to the cache)
  int len;      // cache entries
  std::uint64_t passes;
  SearchStep step;
  phases::ThresholdSet S;
  int perm[kTreeNodes];
  double node[kTreeNodes];
  double ct[kSearchCache], cdot[kSearchCache];
  std::uint64_t ccnt[kSearchCache];
};

}  // namespace internal
}  // namespace pagani
