// devloop_decide.hpp -- the per-iteration decisions of the device-resident
// loop.
//
// Workspace::integrate (Workspace.cuh:362-494) decides, after each
// iteration's reductions, whether the estimate is accurate enough, which
// regions finished, whether the classifier must run and what the next list
// holds.  While the region list is small and the memory model far from its
// budget, those decisions need nothing but the iteration's sums, so they run
// on the device, in one thread of the workgroup that finishes the
// iteration's reductions last (k_step_dev), before the filter + split of the
// same iteration; the host queues whole rounds of iterations and reads the
// state back once per round.  Everything the device cannot decide alone --
// the error-budget fix (Workspace.cuh:121-146), a memory model past a tenth
// of its budget (where the heuristic classifier may run, :74-118), a next
// list beyond the device loop's capacity -- stops the loop with kDevHost
// before any state changes, and the host finishes that iteration exactly as
// it finishes its own.
//
// Plain IEEE double without contraction: the same operations in the same
// order as the host loop, so both make identical decisions.
#pragma once

#include <hip/hip_runtime.h>

#include "ws_internal.hpp"

namespace pagani {
namespace internal {

__device__ __forceinline__ void push_trace(DevLoop* s, std::uint64_t n, std::uint64_t m,
                                           double est, double err)
{
  if (s->ntrace < kDevTrace) s->trace[s->ntrace++] = TraceRec{n, m, est, err};
}

// One iteration's decisions from its sums (estimates and errors of all
// regions and of the active ones) and its count m of active regions.  Runs
// in one thread while it.run is set.
__device__ __forceinline__ void decide_step(DevLoop* s, double iter_est, double iter_err,
                                            double act_est, double act_err, std::uint64_t m)
{
#pragma clang fp contract(off)
  const std::uint64_t n = s->it.n;
  s->evaluated += n;
  s->iterations += 1;
  if (n > s->max_n) s->max_n = n;

  // accuracy_reached, PaganiUtils.cuh:387-394
  const double tot_est = fabs(s->cum_est + iter_est);
  const double tot_err = s->cum_err + iter_err;
  if (tot_err / tot_est <= s->epsrel || tot_err <= s->epsabs) {
    if (s->user_regions) s->cum_iters++;
    push_trace(s, n, 0, s->cum_est + iter_est, tot_err);
    s->cum_est += iter_est;
    s->cum_err += iter_err;
    s->cum_status = 0;
    s->cum_nregions += n;
    s->status = kDevDone;
    s->it.run = 0;
    s->cp.run = 0;
    return;
  }

  // compute_finished_estimates, PaganiUtils.cuh:368-385
  const double fin_est = iter_est - act_est;
  const double fin_err = iter_err - act_err;
  const bool overflow = s->cum_err + fin_err > fabs(s->cum_est + iter_est) * s->epsrel;
  const double ratio = mem_ratio(n, s->d, s->mem_total, s->mem_model);
  if (overflow || ratio > .1 || 2 * m > s->n_cap) {
    s->status = kDevHost;
    s->it.run = 0;
    s->cp.run = 0;
    return;
  }
  if (s->user_regions) s->cum_iters++;
  // Heuristic_classifier::store_estimate
  s->hist[0] = s->hist[1];
  s->hist[1] = s->hist[2];
  s->hist[2] = s->cum_est + iter_est;
  s->iters_collected++;
  // heuristic_classify: with the ratio at most .1, criteria_met is false
  // (heuristic_classifier.cuh:349-360) and the split is not necessary
  // (ratio <= 1), so nothing is classified and nothing terminates.
  push_trace(s, n, m, s->cum_est + iter_est, tot_err);
  s->cum_est += fin_est;
  s->cum_err += fin_err;
  s->cum_nregions += n - m;
  if (m == 0) {
    s->it.n = 0;
    s->it.run = 0;
    s->cp.run = 0;
    s->status = kDevDoneEmpty;
  } else {
    s->cp.n = n;
    s->cp.m = m;
    s->cp.cur = s->it.cur;
    s->cp.has_parents = s->it.has_parents;
    s->cp.split = s->it.split;
    s->cp.run = 1;
    // the kept regions become the next list's parents: 2m children in
    // split form in the other pair
    s->it.n = 2 * m;
    s->it.cur ^= 1u;
    s->it.has_parents = 1;
    s->it.split = 1;
    s->it.slices = static_cast<std::uint32_t>(
        choose_slices_n(2 * m, s->num_cus, s->slices_override, s->generic,
                        s->has_table ? s->slice_table : nullptr));
    // the round's launches have launch_slices waves per workgroup and
    // grids for round_cap regions: an iteration that needs another slice
    // count (idle waves would take the place of working ones) or a larger
    // grid waits for the next round
    if (static_cast<std::int32_t>(s->it.slices) != s->launch_slices || 2 * m > s->round_cap)
      s->it.run = 0;
  }
  if (!s->user_regions) s->cum_iters++;
  s->iter++;
  if (m > 0 && s->iter >= 700) {
    s->status = kDevLimit;
    s->it.run = 0;
  }
}

}  // namespace internal
}  // namespace pagani
