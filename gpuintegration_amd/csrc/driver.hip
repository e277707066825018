// driver.hip -- the PAGANI iteration loop and the C ABI (include/pagani.h).
//
// Host restatement of Workspace::integrate (cuda/pagani/quad/GPUquad/
// Workspace.cuh:74-494) over the device phases of phases.hip and an
// evaluation callback.  Per iteration: evaluate -> refine+classify (with the
// four region sums and the per-chunk active counts) -> one finalize kernel
// -> ONE device-to-host copy of a 64-byte scalar record -> the host decides
// (convergence, budget fix, heuristic classification) -> fused filter+split.
// Buffers are owned by the workspace and grow geometrically; the steady
// state allocates nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pagani.h"
#include "pagani/eval_kernel.hpp"
#include "pagani_ext.h"
#include "phases.hpp"
#include "classify_search.hpp"
#include "rule.hpp"

#include "ws_internal.hpp"

// kTunedSlices[family][ndim][precision][size class]: the launch shapes
// gpuintegration_amd/tune.py measured (gpuintegration_amd/tuning/
// eval_slices.json), written out as a C table by the build; family 0 is an
// evaluation callback, precision 0 FP64 and 1 FP32; 0 = no measurement.
#include "tuned_slices.inc"

namespace pagani {
namespace internal {

// ---------------------------------------------------------------------------
// small utilities

void harvest_compact(pagani_ws* ws)
{
  if (ws->timing && ws->compact_pending) {
    ws->stats.compact_ms += elapsed(ws, EV_CMP0, EV_CMP1);
    ws->compact_pending = false;
  }
}

int pool_handover(pagani_ws* from, pagani_ws* to)
{
  if (!from || !to || from == to) return PAGANI_EBADARG;
  if (from->susp.held || to->susp.held)
    return fail(to, PAGANI_EBADARG, "a suspended run holds the region pool");
  if (from->device != to->device) return fail(to, PAGANI_EBADARG, "workspaces on two devices");
  if (hipSetDevice(from->device) != hipSuccess) return fail(to, PAGANI_EHIP, "hipSetDevice");
  HIP_TRY(from, hipStreamSynchronize(from->stream));
  HIP_TRY(to, hipStreamSynchronize(to->stream));
  harvest_compact(from);
  harvest_compact(to);
  for_each_pool_pair(from, to, [](auto& a, auto& b) {
    if (a.cap > b.cap) std::swap(a, b);
  });
  for (pagani_ws* w : {from, to}) {
    w->live_n = 0;
    w->live_evaluated = false;
    w->snap_n = 0;
  }
  return PAGANI_OK;
}

// ---------------------------------------------------------------------------
// evaluation launch data

// The tuned table of the workspace's running integrand, or NULL (the rule).
const std::uint8_t* tuned_table(const pagani_ws* ws)
{
  if (!ws->tuned || ws->tune_family < 0 || ws->tune_family > PAGANI_GENZ_DISCONTINUOUS)
    return nullptr;
  const std::uint8_t* t =
      kTunedSlices[ws->tune_family][ws->ndim][ws->tune_precision == PAGANI_FP32 ? 1 : 0];
  for (int b = 0; b < kSliceBuckets; ++b)
    if (t[b]) return t;
  return nullptr;
}

int choose_slices(const pagani_ws* ws, std::size_t n)
{
  return choose_slices_n(n, ws->num_cus, ws->slices_override,
                         ws->rule.feval - ws->rule.set_begin[3], tuned_table(ws));
}

void set_slices(pagani_eval_consts& C, int S);

void make_consts(const pagani_ws* ws, const double* lows, const double* highs, std::size_t n,
                 pagani_eval_consts& C, const ph::Levels* lv = nullptr)
{
  std::memset(&C, 0, sizeof(C));
  C.abi_version = PAGANI_EVAL_ABI_VERSION;
  if (lv && lv->on) {
    C.levels = 1;
    for (int j = 0; j < PAGANI_MAX_DIM; ++j) C.h0[j] = lv->h0[j];
  }
  const int d = ws->ndim;
  const pagani::RuleTables& R = ws->rule;
  std::memcpy(C.w, R.w, sizeof(C.w));
  std::memcpy(C.scale, R.scale, sizeof(C.scale));
  std::memcpy(C.norm, R.norm, sizeof(C.norm));
  double jac = 1.0;
  for (int j = 0; j < d; ++j) {
    C.lows[j] = lows ? lows[j] : 0.0;
    C.range[j] = (highs ? highs[j] : 1.0) - C.lows[j];
    jac = jac * C.range[j];
  }
  C.jac = jac;
  C.lambda1 = R.lambda[0];
  C.lambda2 = R.lambda[1];
  C.lambda3 = R.lambda[2];
  C.lambda4 = R.lambda[3];
  C.lambda5 = R.lambda[4];
  const double q = R.lambda[1] / R.lambda[0];
  C.ratio = q * q;
  C.gtab = ws->d_gtab;
  C.ndim = d;
  C.feval = R.feval;
  for (int s = 0; s <= PAGANI_NSETS; ++s) C.set_begin[s] = R.set_begin[s];
  set_slices(C, choose_slices(ws, n));
}

// The slice count of a launch and its split of the generic points (the
// split itself is pagani::slice_range; slice_begin records it).
void set_slices(pagani_eval_consts& C, int S)
{
  C.slices = S;
  for (int i = 0; i <= PAGANI_MAX_SLICES; ++i) C.slice_begin[i] = C.feval;
  for (int w = 0; w < S; ++w) {
    int pa, pb;
    pagani::slice_range(S, w, C.set_begin[3], C.feval, pa, pb);
    C.slice_begin[w] = pa;
    C.slice_begin[w + 1] = pb;
  }
}

// ---------------------------------------------------------------------------
// Heuristic classifier, heuristic_classifier.cuh:89-452.  The threshold
// search (classify_search.hpp) runs on the device: the host queues rounds of
// probe passes and replays and reads the outcome once; a search too long for
// the device's cache is finished here with the same replay.

// A state measured on the host: one probe pass over the tree around it.
struct HostCache {
  std::vector<double> t, dot;
  std::vector<std::uint64_t> cnt;
};

int host_probe(pagani_ws* ws, const double* err, std::size_t n, double t, double low,
               double high, HostCache& cache)
{
  double node[kTreeNodes];
  int perm[kTreeNodes];
  ph::ThresholdSet S;
  search_tree(t, low, high, node, S, perm);
  HIP_TRY(ws, ph::multi_threshold(ws->stream, n, err, S, ws->d_mpart, ws->d_mout));
  HIP_TRY(ws, hipMemcpyAsync(ws->h_mout, ws->d_mout, 2 * S.k * sizeof(double),
                             hipMemcpyDeviceToHost, ws->stream));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  ws->stats.classify_passes++;
  for (int i = 0; i < S.k; ++i) {
    cache.t.push_back(node[perm[i]]);
    cache.cnt.push_back(static_cast<std::uint64_t>(ws->h_mout[i]));
    cache.dot.push_back(ws->h_mout[S.k + i]);
  }
  return PAGANI_OK;
}

// classify, :362-451.  The flags of the accepted threshold (*last_thr) are
// written by the caller.  have_minmax: d_scal already holds err's minimum
// and maximum (the loop's refinement computed them).
int classify(pagani_ws* ws, Classifier& c, const double* err, std::size_t n, double iter_err,
             double iter_fin_err, double tot_fin_err, pagani_class_res& res, double* last_thr,
             bool have_minmax)
{
  std::memset(&res, 0, sizeof(res));
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  SearchInputs in{};
  in.n = n;
  in.iter_err = iter_err;
  in.iter_fin_err = iter_fin_err;
  in.tot_fin_err = tot_fin_err;
  in.target_error = std::fabs(c.hist[2]) * c.epsrel;
  in.max_budget = c.max_budget;
  in.max_active = c.max_active;
  // min and max of the errors: from the refinement's records when nothing
  // has finalized since (an iteration that took the error-budget fix has)
  if (!have_minmax) {
    HIP_TRY(ws, ph::minmax(ws->stream, n, err, ws->part.p));
    HIP_TRY(ws, launch_finalize(ws, n, false, true));
  }
  SearchStep step{};
  HostCache cache;
  if (ws->device_search) {
    HIP_TRY(ws, ph::search_start(ws->stream, ws->d_search, ws->d_scal, in));
    for (;;) {
      for (int r = 0; r < kSearchRounds; ++r)
        HIP_TRY(ws, ph::search_round(ws->stream, n, err, ws->d_search, ws->d_mpart, ws->d_mout));
      HIP_TRY(ws, hipMemcpyAsync(ws->h_search, ws->d_search, sizeof(SearchState),
                                 hipMemcpyDeviceToHost, ws->stream));
      HIP_TRY(ws, hipStreamSynchronize(ws->stream));
      step = ws->h_search->step;
      if (step.done || step.overflow) break;
    }
    ws->stats.classify_passes += ws->h_search->passes;
    if (!step.done) {
      // finish on the host, from the start, with the device's measurements
      in = ws->h_search->in;
      const SearchState& hs = *ws->h_search;
      for (int i = 0; i < hs.len; ++i) {
        cache.t.push_back(hs.ct[i]);
        cache.cnt.push_back(hs.ccnt[i]);
        cache.dot.push_back(hs.cdot[i]);
      }
    }
  } else {
    HIP_TRY(ws, hipMemcpyAsync(ws->h_scal, ws->d_scal, sizeof(ph::Scalars),
                               hipMemcpyDeviceToHost, ws->stream));
    HIP_TRY(ws, hipStreamSynchronize(ws->stream));
    in.mn = ws->h_scal->mn;
    in.mx = ws->h_scal->mx;
  }
  if (!step.done) {
    for (;;) {
      auto find = [&](double x, std::uint64_t& cn, double& d) {
        for (std::size_t i = 0; i < cache.t.size(); ++i)
          if (cache.t[i] == x) {
            cn = cache.cnt[i];
            d = cache.dot[i];
            return true;
          }
        return false;
      };
      search_replay(in, find, step);
      if (step.done) break;
      if (step.overflow) return fail(ws, PAGANI_EHIP, "classifier search did not end");
      if ((rc = host_probe(ws, err, n, step.t, step.low, step.high, cache))) return rc;
    }
  }
  res = step.res;
  *last_thr = step.last_thr;
  c.max_active = .5;
  c.max_budget = .25;
  return PAGANI_OK;
}

// ---------------------------------------------------------------------------
// the loop

int check_volume(pagani_ws* ws, const double* lows, const double* highs)
{
  for (int j = 0; j < ws->ndim; ++j) {
    const double lo = lows ? lows[j] : 0.0, hi = highs ? highs[j] : 1.0;
    if (!std::isfinite(lo) || !std::isfinite(hi))
      return fail(ws, PAGANI_EBADARG, "non-finite integration bounds");
  }
  return PAGANI_OK;
}

// Why a run stopped early at an iteration boundary: the per-call iteration
// limit (pagani_ws_set_iteration_limit).
void suspend(pagani_ws* ws, const LoopState& st)
{
  ws->susp = st;
  ws->susp.held = true;
}

// Grow a buffer to at least `need` elements keeping its first `keep`.
template <class T>
int grow_keep(pagani_ws* ws, DevBuf<T>& b, std::size_t need, std::size_t keep, const char* what)
{
  if (need <= b.cap) return PAGANI_OK;
  DevBuf<T> nb;
  int rc = ensure(ws, nb, need, what);
  if (rc) return rc;
  if (keep && b.p)
    HIP_TRY(ws, hipMemcpyAsync(nb.p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, ws->stream));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  release(b);
  b = nb;
  return PAGANI_OK;
}

// The largest list size up to cap whose launches keep n's slice count:
// with a tuned table, the end of the run of size classes from n's own that
// share its entry (the table need not be monotone); with the rule, which is
// non-increasing in n, a binary search.
std::uint64_t round_cap(const pagani_ws* ws, std::uint64_t n, std::uint64_t cap, int slices)
{
  if (ws->slices_override <= 0)
    if (const std::uint8_t* t = tuned_table(ws)) {
      int b = slice_bucket(n);
      if (t[b] > 0) {
        while (b + 1 < kSliceBuckets && t[b + 1] == t[b] &&
               (std::uint64_t(1) << (b + 1)) <= cap)
          ++b;
        const std::uint64_t end = b + 1 < 64 ? (std::uint64_t(2) << b) - 1 : cap;
        return std::max<std::uint64_t>(n, std::min(end, cap));
      }
    }
  std::uint64_t lo = n, hi = cap;
  while (lo < hi) {
    const std::uint64_t mid = lo + (hi - lo + 1) / 2;
    if (choose_slices(ws, mid) == slices)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Iterations queued per round at most: a round that stops early leaves the
// rest as no-op launches, so the length trades round trips for those.
constexpr int kDevQueueMax = 16;

// Full-length rounds of the device-resident loop captured as HIP graphs:
// every launch of a round is fixed by the round's length, its grids (the
// largest list of its slice count) and its slice count, so a later round
// with the same three replays the graph (one launch instead of 48).  Kept
// for one segment (the buffers it points at are fixed for the segment).
struct RoundGraphs {
  struct Entry {
    std::uint64_t rcap;
    int slices;
    hipGraphExec_t exec;
  };
  std::vector<Entry> entries;
  hipGraphExec_t find(std::uint64_t rcap, int slices) const
  {
    for (const Entry& e : entries)
      if (e.rcap == rcap && e.slices == slices) return e.exec;
    return nullptr;
  }
  ~RoundGraphs()
  {
    for (Entry& e : entries) (void)hipGraphExecDestroy(e.exec);
  }
};

// What a device-resident segment ended with (see run_device_segment).
enum DevOutcome { kSegHost, kSegDone, kSegEmpty, kSegLimit };

// The loop variables the host loop and a device segment hand to each other.
struct LoopVars {
  std::uint64_t it;
  std::size_t n;
  bool has_parents;
  pagani_result* cum;
  Classifier* clf;
};

// Run iterations on the device, in rounds of 2, 4, 8, 16 queued
// iterations with one host round trip each, until the device loop stops.
// Each iteration is three launches: the evaluation, k_step_dev (refinement,
// the reductions and the decisions) and the filter + split, with fixed
// launch shapes for the round, reading the counts from the DevLoop
// descriptors.  On return the loop variables
// describe where the host continues; with kSegHost the iteration `v.it`
// has been evaluated and reduced (its sums are in d_scal) and the host
// finishes it.
int run_device_segment(pagani_ws* ws, pagani_eval_fn eval, void* ctx, double epsrel,
                       double epsabs, int relerr, bool user_regions, const double* lows,
                       const double* highs, LoopVars& v, DevOutcome& outcome)
{
  const int d = ws->ndim;
  const std::size_t cap = ph::kDevLoopMax;
  const int cur = ws->cur;
  int rc;
  // the live list (its own bounds, or its parents' and their halved axes)
  // and the parents' estimates are kept; everything else is sized for the
  // capacity
  const std::size_t rows = ws->split ? v.n / 2 : v.n;
  if ((rc = grow_keep(ws, ws->L[cur], cap * d, rows * d, "region bounds")) ||
      (rc = grow_keep(ws, ws->H[cur], cap * d, rows * d, "region bounds")) ||
      (rc = grow_keep(ws, ws->pdim[cur], cap / 2, ws->split ? rows : 0, "split axes")) ||
      (rc = grow_keep(ws, ws->pest, cap / 2, v.has_parents ? v.n / 2 : 0, "parent estimates")) ||
      (rc = ensure(ws, ws->L[1 - cur], cap * d, "region bounds")) ||
      (rc = ensure(ws, ws->H[1 - cur], cap * d, "region bounds")) ||
      (rc = ensure(ws, ws->pdim[1 - cur], cap / 2, "split axes")) ||
      (rc = ensure(ws, ws->est, cap, "estimates")) || (rc = ensure(ws, ws->err0, cap, "errors")) ||
      (rc = ensure(ws, ws->err1, cap, "errors")) || (rc = ensure(ws, ws->dim, cap, "split axes")) ||
      (rc = ensure(ws, ws->active, cap, "flags")) || (rc = ensure_chunks(ws, cap)))
    return rc;

  DevLoop& h = *ws->h_loop;
  std::memset(&h, 0, sizeof(h));
  h.it.n = v.n;
  h.it.cur = static_cast<std::uint32_t>(cur);
  h.it.run = 1;
  h.it.has_parents = v.has_parents ? 1u : 0u;
  h.it.slices = static_cast<std::uint32_t>(choose_slices(ws, v.n));
  h.it.split = ws->split ? 1u : 0u;
  h.num_cus = ws->num_cus;
  h.slices_override = ws->slices_override;
  if (const std::uint8_t* t = tuned_table(ws)) {
    h.has_table = 1;
    std::memcpy(h.slice_table, t, sizeof(h.slice_table));
  }
  h.generic = ws->rule.feval - ws->rule.set_begin[3];
  h.n_cap = cap;
  h.mem_total = v.clf->total;
  h.epsrel = epsrel;
  h.epsabs = epsabs;
  h.d = d;
  h.mem_model = v.clf->model;
  h.relerr = relerr;
  h.user_regions = user_regions ? 1 : 0;
  h.iter = v.it;
  h.cum_est = v.cum->estimate;
  h.cum_err = v.cum->errorest;
  h.cum_nregions = v.cum->nregions;
  h.cum_iters = v.cum->iters;
  h.cum_status = v.cum->status;
  h.status = kDevRun;
  for (int i = 0; i < 3; ++i) h.hist[i] = v.clf->hist[i];
  h.iters_collected = v.clf->iters_collected;

  pagani_eval_args args{};
  args.L = ws->L[0].p;
  args.H = ws->H[0].p;
  args.L2 = ws->L[1].p;
  args.H2 = ws->H[1].p;
  args.pdim = ws->pdim[0].p;
  args.pdim2 = ws->pdim[1].p;
  args.est = ws->est.p;
  args.err = ws->err0.p;
  args.dim = ws->dim.p;
  args.stream = ws->stream;
  args.dev = &ws->d_loop->it;
  make_consts(ws, lows, highs, cap, args.consts, &ws->lv);
  RoundGraphs graphs;

  for (int queued = 2;; queued = std::min(queued * 2, kDevQueueMax)) {
    // A round's launches have as many waves per workgroup as its first
    // iteration's slice count, and grids for the largest list with that
    // slice count (workgroups past the actual list return at once; queued
    // iterations after the loop stops return at once too, so the grids are
    // kept no larger than they can be used).  decide_step pauses the round
    // (it.run = 0, status kDevRun) before an iteration with another slice
    // count.
    h.it.run = 1;
    h.launch_slices = static_cast<std::int32_t>(h.it.slices);
    h.round_cap = round_cap(ws, h.it.n, cap, h.launch_slices);
    set_slices(args.consts, h.launch_slices);
    args.n = h.round_cap;
    const std::size_t rcap = h.round_cap;
    HIP_TRY(ws, hipMemcpyAsync(ws->d_loop, ws->h_loop, sizeof(DevLoop), hipMemcpyHostToDevice,
                               ws->stream));
    if (ws->timing) record(ws, EV_DEV0);
    // the round's launches: `queued` iterations of evaluation, refinement
    // with reductions and decisions, filter + split
    auto enqueue = [&]() -> int {
      for (int q = 0; q < queued; ++q) {
        const int erc = eval(ctx, &args);
        if (erc) return fail(ws, erc < 0 ? erc : PAGANI_EHIP, "evaluation callback failed");
        HIP_TRY(ws, ph::step_dev(ws->stream, rcap, ws->d_loop, ws->est.p, ws->err0.p,
                                 ws->err1.p, ws->pest.p, epsrel, relerr, ws->active.p, ws->part.p,
                                 ws->cnt.p, ws->off.p, ws->fpart.p, ws->fcnt.p, ws->foff.p,
                                 ws->d_scal, &ws->d_loop->ticket));
        HIP_TRY(ws, ph::compact_dev(ws->stream, rcap, &ws->d_loop->cp, d, ws->L[0].p,
                                    ws->H[0].p, ws->L[1].p, ws->H[1].p, ws->pdim[0].p,
                                    ws->pdim[1].p, ws->active.p, ws->est.p, ws->err0.p,
                                    ws->err1.p, ws->dim.p, ws->off.p, ws->foff.p, ws->pest.p,
                                    nullptr, ws->lv));
      }
      return PAGANI_OK;
    };
    if (ws->graphs && queued == kDevQueueMax) {
      hipGraphExec_t exec = graphs.find(rcap, h.launch_slices);
      if (!exec) {
        HIP_TRY(ws, hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ws->stream, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) return hip_fail(ws, e, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return hip_fail(ws, ie, "hipGraphInstantiate");
        graphs.entries.push_back({rcap, h.launch_slices, exec});
      }
      HIP_TRY(ws, hipGraphLaunch(exec, ws->stream));
      ws->stats.graph_launches++;
    } else if ((rc = enqueue())) {
      return rc;
    }
    if (ws->timing) record(ws, EV_DEV1);
    HIP_TRY(ws, hipMemcpyAsync(ws->h_loop, ws->d_loop, sizeof(DevLoop), hipMemcpyDeviceToHost,
                               ws->stream));
    HIP_TRY(ws, hipStreamSynchronize(ws->stream));
    if (ws->timing) ws->stats.device_ms += elapsed(ws, EV_DEV0, EV_DEV1);
    ws->stats.device_rounds++;
    ws->stats.device_iterations += h.iterations;
    ws->stats.device_regions += h.evaluated;
    ws->stats.regions_evaluated += h.evaluated;
    ws->stats.eval_launches += h.iterations;
    ws->stats.iterations += h.iterations;
    ws->stats.max_regions = std::max<std::uint64_t>(ws->stats.max_regions, h.max_n);
    for (std::uint32_t i = 0; i < h.ntrace; ++i) ws->trace.push_back(h.trace[i]);
    h.ntrace = 0;
    h.evaluated = 0;
    h.iterations = 0;
    if (h.status != kDevRun) break;
  }

  v.it = h.iter;
  v.n = h.it.n;
  v.has_parents = h.it.has_parents != 0;
  ws->cur = static_cast<int>(h.it.cur);
  ws->split = h.it.split != 0;
  v.cum->estimate = h.cum_est;
  v.cum->errorest = h.cum_err;
  v.cum->nregions = h.cum_nregions;
  v.cum->iters = h.cum_iters;
  v.cum->status = h.cum_status;
  for (int i = 0; i < 3; ++i) v.clf->hist[i] = h.hist[i];
  v.clf->iters_collected = h.iters_collected;
  switch (h.status) {
    case kDevDone: outcome = kSegDone; break;
    case kDevDoneEmpty: outcome = kSegEmpty; break;
    case kDevLimit: outcome = kSegLimit; break;
    default: outcome = kSegHost; break;
  }
  return PAGANI_OK;
}

// The iteration loop, Workspace.cuh:148-358 / :362-494.  `resume` continues
// a suspended or loaded run instead of starting one (then epsrel, epsabs,
// lows, highs, relerr, init and init_k are taken from it).
int integrate_impl(pagani_ws* ws, pagani_eval_fn eval, void* ctx, double epsrel, double epsabs,
                   const double* lows_in, const double* highs_in, int relerr,
                   const pagani_region_list* init, std::size_t init_k, const LoopState* resume,
                   const pagani_genz* genz, int precision, pagani_result* out)
{
  if (!ws || !eval || !out) return fail(ws, PAGANI_EBADARG, "null argument");
  const int d = ws->ndim;
  LoopState st;
  if (resume) {
    st = *resume;
  } else {
    if (!(epsrel >= 0.0) || !(epsabs >= 0.0) || std::isnan(epsrel))
      return fail(ws, PAGANI_EBADARG, "epsrel and epsabs must be >= 0");
    int rc0 = check_volume(ws, lows_in, highs_in);
    if (rc0) return rc0;
    st.epsrel = epsrel;
    st.epsabs = epsabs;
    // kRelerrPredictSplit: the flag of each iteration follows the rule of
    // integrate<IntegT, predict_split=true> (Workspace.cuh:206-210)
    st.relerr = ws->predict_split ? kRelerrPredictSplit : (relerr ? 1 : 0);
    for (int j = 0; j < d; ++j) {
      st.lows[j] = lows_in ? lows_in[j] : 0.0;
      st.highs[j] = highs_in ? highs_in[j] : 1.0;
    }
    st.user_regions = init != nullptr;
    st.mem_model = ws->mem_model;
    st.mem_total = ws->mem_model_total ? ws->mem_model_total : ws->hbm_total;
    if (genz) st.genz = *genz;
    st.precision = precision;
    std::memset(&st.cum, 0, sizeof st.cum);
    st.cum.status = 1;
    st.cum.lastPhase = -1;
  }
  epsrel = st.epsrel;
  epsabs = st.epsabs;
  relerr = st.relerr;
  ws->tune_family = st.genz.family;
  ws->tune_precision = st.precision == PAGANI_FP32 ? PAGANI_FP32 : PAGANI_FP64;
  const double* lows = st.lows;
  const double* highs = st.highs;
  const bool user_regions = st.user_regions != 0;
  int rc = PAGANI_OK;
  const auto wall0 = std::chrono::steady_clock::now();
  std::memset(out, 0, sizeof(*out));
  out->status = 1;
  out->lastPhase = -1;
  if (!resume) ws->trace.clear();
  ws->susp.held = false;
  ws->live_evaluated = false;
  harvest_compact(ws);

  // initial region list
  std::size_t n = 0;
  if (resume) {
    n = st.n;
    ws->split = st.split != 0;
    ws->lv = st.lv;
  } else if (init) {
    ws->split = false;
    ws->lv = ph::Levels{};
    n = init->n;
    if (n == 0 || !init->L || !init->H) return fail(ws, PAGANI_EBADARG, "empty region list");
    if ((rc = ensure(ws, ws->L[ws->cur], n * d, "region bounds"))) return rc;
    if ((rc = ensure(ws, ws->H[ws->cur], n * d, "region bounds"))) return rc;
    HIP_TRY(ws, hipMemcpyAsync(ws->L[ws->cur].p, init->L, n * d * sizeof(double),
                               hipMemcpyDeviceToDevice, ws->stream));
    HIP_TRY(ws, hipMemcpyAsync(ws->H[ws->cur].p, init->H, n * d * sizeof(double),
                               hipMemcpyDeviceToDevice, ws->stream));
  } else {
    // a fresh uniform start always begins in pair 0, so repeated
    // integrations alternate the ping-pong pairs identically and a pool
    // grown by the first finds every buffer at its size in the next
    ws->cur = 0;
    ws->split = false;
    std::size_t k = init_k ? init_k : (d < 5 ? 4 : (d <= 10 ? 2 : 1));
    // every width of the run is 1/k halved some number of times: H holds
    // that number, one uint16_t per axis and region (pagani_eval_args)
    ws->lv = ph::Levels{};
    ws->lv.on = ws->level_form ? 1 : 0;
    for (int j = 0; j < d; ++j) ws->lv.h0[j] = 1. / static_cast<double>(k);
    n = 1;
    for (int j = 0; j < d; ++j) {
      if (n > (std::size_t(1) << 40) / k) return fail(ws, PAGANI_EBADARG, "initial split too large");
      n *= k;
    }
    if ((rc = ensure(ws, ws->L[ws->cur], n * d, "region bounds"))) return rc;
    if ((rc = ensure(ws, ws->H[ws->cur], n * d, "region bounds"))) return rc;
    HIP_TRY(ws, ph::uniform_split(ws->stream, d, k, n, ws->L[ws->cur].p, ws->H[ws->cur].p,
                                  ws->lv.on != 0));
  }

  const std::size_t total = st.mem_total;
  Classifier clf(epsrel, epsabs, total, d, st.mem_model);
  clf.hist[0] = st.hist[0];
  clf.hist[1] = st.hist[1];
  clf.hist[2] = st.hist[2];
  clf.iters_collected = st.iters_collected;
  pagani_result cum = st.cum;
  bool has_parents = st.has_parents != 0;
  int status_code = PAGANI_OK;
  const int mem_model = st.mem_model;
  std::uint64_t iters_this_call = 0;
  double* errs = ws->err0.p;

  // the reference's debug recorders (pagani_ws_set_debug); they hold files
  // open for this call only
  DebugRecorder rec;
  if (ws->dbg.level > 0) {
    std::string why;
    if (!rec.open(ws->dbg, ws->stream, d, epsrel, ws->rule, lows, highs, &why))
      return fail(ws, PAGANI_EBADARG, why);
  }

  std::uint64_t it = st.it;
  for (; it < 700 && n > 0; it++) {
    ws->scal_minmax = false;  // d_scal's min / max belong to an earlier list
    if (ws->iter_limit && iters_this_call == ws->iter_limit) {
      st.it = it;
      st.n = n;
      st.has_parents = has_parents ? 1 : 0;
      st.split = ws->split ? 1 : 0;
      st.lv = ws->lv;
      st.cum = cum;
      st.hist[0] = clf.hist[0];
      st.hist[1] = clf.hist[1];
      st.hist[2] = clf.hist[2];
      st.iters_collected = clf.iters_collected;
      suspend(ws, st);
      ws->live_n = n;
      HIP_TRY(ws, hipStreamSynchronize(ws->stream));
      harvest_compact(ws);
      ws->stats.wall_ms += std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - wall0)
                               .count();
      *out = cum;
      return PAGANI_SUSPENDED;
    }
    // Small lists far from the memory budget: iterate on the device until
    // it needs the host (or finishes).
    bool reduced = false;  // this iteration's sums are already in d_scal
    if (ws->device_loop && !ws->iter_limit && ws->snapshot_iter == ~std::uint64_t(0) &&
        relerr != kRelerrPredictSplit && rec.level() == 0 &&
        n <= ph::kDevLoopMax && mem_ratio(n, d, total, mem_model) <= .1) {
      LoopVars v{it, n, has_parents, &cum, &clf};
      DevOutcome outcome;
      if ((rc = run_device_segment(ws, eval, ctx, epsrel, epsabs, relerr, user_regions, lows,
                                   highs, v, outcome)))
        return rc;
      it = v.it;
      n = v.n;
      has_parents = v.has_parents;
      if (outcome == kSegDone) {
        errs = has_parents ? ws->err1.p : ws->err0.p;
        goto done_evaluated;
      }
      if (outcome == kSegEmpty || outcome == kSegLimit) break;
      reduced = true;
    }
    if (!reduced) {
      iters_this_call++;
      if ((rc = ensure(ws, ws->est, n, "estimates")) || (rc = ensure(ws, ws->err0, n, "errors")) ||
          (rc = ensure(ws, ws->err1, n, "errors")) || (rc = ensure(ws, ws->dim, n, "split axes")) ||
          (rc = ensure(ws, ws->active, n, "flags")) || (rc = ensure_chunks(ws, n)))
        return rc;

      // evaluate (Cubature_rules::apply_cubature_integration_rules)
      pagani_eval_args args{};  // dev = NULL: an ordinary launch
      args.n = n;
      args.L = ws->L[ws->cur].p;
      args.H = ws->H[ws->cur].p;
      args.pdim = ws->split ? ws->pdim[ws->cur].p : nullptr;
      args.est = ws->est.p;
      args.err = ws->err0.p;
      args.dim = ws->dim.p;
      args.stream = ws->stream;
      make_consts(ws, lows, highs, n, args.consts, &ws->lv);
      if (rec.points(n)) {
        if ((rc = ensure(ws, ws->dfe, n * static_cast<std::size_t>(ws->rule.feval),
                         "point values")))
          return rc;
        args.fevals = ws->dfe.p;
      }
      rec.mark();
      record(ws, EV_EVAL0);
      rc = eval(ctx, &args);
      if (rc) return fail(ws, rc < 0 ? rc : PAGANI_EHIP, "evaluation callback failed");
      record(ws, EV_EVAL1);
      if (rec.level() > 0) {
        rec.phase(it, "apply_cubature_rules");
        HIP_TRY(ws, rec.regions(it, n, ws->est.p, ws->err0.p));
        HIP_TRY(ws, rec.points(n, args.L, args.H, args.pdim, ws->lv, ws->est.p, ws->err0.p,
                               args.fevals));
        rec.mark();
      }
      ws->stats.regions_evaluated += n;
      ws->stats.eval_launches++;
      ws->stats.max_regions = std::max<std::uint64_t>(ws->stats.max_regions, n);

      // two-level error + relerr flags + the four region sums + counts
      errs = has_parents ? ws->err1.p : ws->err0.p;
      // predict_split: relerr classification off while the list has at most
      // 15M regions, before iteration 15 and before any region finished,
      // on after that (Workspace.cuh:206-210, 415-420)
      const int relerr_it =
          relerr == kRelerrPredictSplit
              ? ((n <= 15000000 && it < 15 && cum.nregions == 0) ? 0 : 1)
              : relerr;
      record(ws, EV_REF0);
      HIP_TRY(ws, ph::refine(ws->stream, n, ws->est.p, ws->err0.p, errs,
                             has_parents ? ws->pest.p : nullptr, epsrel, relerr_it, ws->active.p,
                             ws->part.p, ws->cnt.p));
      HIP_TRY(ws, launch_finalize(ws, n, true, true, has_parents));
      ws->scal_minmax = true;
      record(ws, EV_REF1);
    }
    errs = has_parents ? ws->err1.p : ws->err0.p;
    HIP_TRY(ws, hipMemcpyAsync(ws->h_scal, ws->d_scal, sizeof(ph::Scalars),
                               hipMemcpyDeviceToHost, ws->stream));
    HIP_TRY(ws, hipStreamSynchronize(ws->stream));
    if (!reduced) {
      if (ws->timing) {
        ws->stats.eval_ms += elapsed(ws, EV_EVAL0, EV_EVAL1);
        ws->stats.refine_ms += elapsed(ws, EV_REF0, EV_REF1);
      }
      ws->stats.iterations++;
    }
    harvest_compact(ws);

    const double iter_est = ws->h_scal->sum[0];
    const double iter_err = ws->h_scal->sum[1];
    double act_est = ws->h_scal->sum[2];
    double act_err = ws->h_scal->sum[3];
    std::size_t m = ws->h_scal->count;
    if (user_regions) cum.iters++;
    rec.phase(it, "two_level_errorest");
    rec.iteration(it, cum.estimate + iter_est, cum.errorest + iter_err, cum.estimate,
                  cum.errorest, n, cum.nregions);

    // predict_split's snapshot of the list just evaluated, Workspace.cuh:244-248
    if (it == ws->snapshot_iter && cum.nregions == 0) {
      if ((rc = ensure(ws, ws->snapL, n * d, "snapshot")) ||
          (rc = ensure(ws, ws->snapH, n * d, "snapshot")))
        return rc;
      if (ws->split || ws->lv.on) {
        HIP_TRY(ws, ph::expand(ws->stream, d, n, ws->L[ws->cur].p, ws->H[ws->cur].p,
                               ws->split ? ws->pdim[ws->cur].p : nullptr, ws->lv, ws->snapL.p,
                               ws->snapH.p));
      } else {
        HIP_TRY(ws, hipMemcpyAsync(ws->snapL.p, ws->L[ws->cur].p, n * d * sizeof(double),
                                   hipMemcpyDeviceToDevice, ws->stream));
        HIP_TRY(ws, hipMemcpyAsync(ws->snapH.p, ws->H[ws->cur].p, n * d * sizeof(double),
                                   hipMemcpyDeviceToDevice, ws->stream));
      }
      ws->snap_n = n;
    }

    // accuracy_reached, PaganiUtils.cuh:387-394
    const double tot_est = std::fabs(cum.estimate + iter_est);
    const double tot_err = cum.errorest + iter_err;
    if (tot_err / tot_est <= epsrel || tot_err <= epsabs) {
      ws->trace.push_back({n, 0, cum.estimate + iter_est, tot_err});
      cum.estimate += iter_est;
      cum.errorest += iter_err;
      cum.status = 0;
      cum.nregions += n;
      goto done_evaluated;
    }
    clf.store(cum.estimate + iter_est);
    {
      // compute_finished_estimates, PaganiUtils.cuh:368-385
      double fin_est = iter_est - act_est;
      double fin_err = iter_err - act_err;
      bool recount = false;
      rec.phase(it, "compute_finished_estimates");
      // fix_error_budget_overflow, Workspace.cuh:121-146
      if (cum.errorest + fin_err > std::fabs(cum.estimate + iter_est) * epsrel) {
        ws->stats.budget_fixes++;
        HIP_TRY(ws, ph::fill_u8(ws->stream, n, ws->active.p, 1));
        fin_err = 0.;
        fin_est = 0.;
        recount = true;
      }
      rec.phase(it, "fix_error_budget_overflow");
      // heuristic_classify, Workspace.cuh:74-118
      const bool necessary = mem_ratio(n, d, total, mem_model) > 1.;
      bool terminate;
      if (!clf.criteria_met(n)) {
        terminate = necessary;
      } else {
        if ((rc = ensure(ws, ws->flags, n, "flags"))) return rc;
        pagani_class_res res;
        double thr = 0.0;
        record(ws, EV_CLS0);
        rc = classify(ws, clf, errs, n, iter_err, fin_err, cum.errorest, res, &thr,
                      ws->scal_minmax);
        if (rc) return rc;
        const bool ok = res.pass_mem && res.pass_errorest_budget;
        if (ok) {
          // the accepted threshold's flags replace the active flags
          // (Workspace.cuh:104-111), with their counts for the compaction
          HIP_TRY(ws, ph::threshold_flags(ws->stream, n, ws->est.p, errs, thr, ws->flags.p,
                                          ws->part.p, ws->cnt.p, has_parents));
          if ((rc = finalize_read(ws, n, true, false, has_parents))) return rc;
          if (ws->h_scal->count != res.num_active)
            return fail(ws, PAGANI_EHIP, "classifier flag count disagrees with its pass");
          std::swap(ws->active, ws->flags);
          act_est = ws->h_scal->sum[2];
          m = ws->h_scal->count;
          fin_est = iter_est - act_est;
          fin_err = res.finished_errorest;
          recount = false;
        }
        record(ws, EV_CLS1);
        terminate = (!ok && necessary) || res.num_active == 0;
        if (ws->timing) {
          HIP_TRY(ws, hipStreamSynchronize(ws->stream));
          ws->stats.classify_ms += elapsed(ws, EV_CLS0, EV_CLS1);
        }
      }
      if (recount) {
        HIP_TRY(ws, ph::count_flags(ws->stream, n, ws->active.p, ws->est.p, errs, ws->part.p,
                                    ws->cnt.p, has_parents));
        if ((rc = finalize_read(ws, n, true, false, has_parents))) return rc;
        m = ws->h_scal->count;
      }
      rec.phase(it, "heuristic_classify");
      if (terminate) {
        ws->trace.push_back({n, 0, cum.estimate + iter_est, tot_err});
        cum.estimate += iter_est;
        cum.errorest += iter_err;
        cum.nregions += n;
        goto done_evaluated;
      }
      if (m > 0) {
        const int nxt = 1 - ws->cur;
        if ((rc = ensure(ws, ws->L[nxt], m * d, "region bounds")) ||
            (rc = ensure(ws, ws->H[nxt], m * d, "region bounds")) ||
            (rc = ensure(ws, ws->pdim[nxt], m, "split axes")) ||
            (rc = ensure(ws, ws->pest, m, "parent estimates"))) {
          // the device cannot hold the next split: stop like the reference's
          // memory termination (status 1, all leaves summed)
          ws->trace.push_back({n, 0, cum.estimate + iter_est, tot_err});
          cum.estimate += iter_est;
          cum.errorest += iter_err;
          cum.nregions += n;
          status_code = rc;
          goto done_evaluated;
        }
      }
      ws->trace.push_back({n, m, cum.estimate + iter_est, tot_err});
      cum.estimate += fin_est;
      cum.errorest += fin_err;
    }
    // Sub_regions_filter::filter + Sub_region_splitter::split, fused
    cum.nregions += n - m;
    if (m == 0) {
      n = 0;
    } else {
      const int nxt = 1 - ws->cur;
      record(ws, EV_CMP0);
      // the kept regions become the next list's parents (split form)
      HIP_TRY(ws, ph::compact_parents(ws->stream, d, n, m, ws->L[ws->cur].p, ws->H[ws->cur].p,
                                      ws->split ? ws->pdim[ws->cur].p : nullptr, ws->active.p,
                                      ws->est.p, errs, ws->dim.p, ws->off.p, ws->foff.p,
                                      ws->L[nxt].p, ws->H[nxt].p, ws->pdim[nxt].p, ws->pest.p,
                                      nullptr, has_parents, ws->lv));
      record(ws, EV_CMP1);
      ws->compact_pending = ws->timing;
      ws->cur = nxt;
      ws->split = true;
      n = 2 * m;
      has_parents = true;
    }
    if (rec.level() > 0) {
      // Sub_regions_filter::filter and Sub_region_splitter::split are one
      // kernel here
      rec.phase(it, "region_filtering");
      rec.folded(it, "region_splitting");
    }
    if (!user_regions) cum.iters++;
  }
  cum.nregions += n;
  goto done;
done_evaluated:
  // the live list is the one evaluated last: its est / err / dim are valid
  ws->live_evaluated = true;
  ws->live_err1 = errs == ws->err1.p;
done:
  ws->live_n = n;
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  harvest_compact(ws);
  ws->stats.integrals++;
  ws->stats.wall_ms +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  *out = cum;
  return status_code;
}

struct GenzCtx {
  pagani_genz g;
  int precision;
};

int genz_trampoline(void* ctx, const pagani_eval_args* a)
{
  const GenzCtx* c = static_cast<const GenzCtx*>(ctx);
  return pagani_genz_eval_launch(&c->g, c->precision, a);
}

int check_genz(pagani_ws* ws, const pagani_genz* f, int precision)
{
  if (!f) return fail(ws, PAGANI_EBADARG, "null integrand");
  if (f->ndim != ws->ndim) return fail(ws, PAGANI_EBADARG, "integrand ndim != workspace ndim");
  if (f->ndim > PAGANI_GENZ_MAX_DIM)
    return fail(ws, PAGANI_EUNSUPPORTED, "Genz descriptors cover ndim 2..10");
  if (f->family < PAGANI_GENZ_OSCILLATORY || f->family > PAGANI_GENZ_DISCONTINUOUS)
    return fail(ws, PAGANI_EBADARG, "unknown integrand family");
  if (precision != PAGANI_FP64 && precision != PAGANI_FP32)
    return fail(ws, PAGANI_EBADARG, "precision must be 64 or 32");
  return PAGANI_OK;
}

}  // namespace internal
}  // namespace pagani

using namespace pagani::internal;

// ===========================================================================
// C ABI

extern "C" {

int pagani_rule_tables(int ndim, double* w, double* scale, double* norm, double* lambda,
                       int* set_begin, double* gtab)
{
  pagani::RuleTables R;
  if (!pagani::build_rule(ndim, R)) return PAGANI_EBADARG;
  for (int s = 0; s < PAGANI_NSETS; ++s)
    for (int r = 0; r < PAGANI_NRULES; ++r) {
      if (w) w[s * PAGANI_NRULES + r] = R.w[s][r];
      if (scale) scale[s * PAGANI_NRULES + r] = R.scale[s][r];
      if (norm) norm[s * PAGANI_NRULES + r] = R.norm[s][r];
    }
  if (lambda)
    for (int i = 0; i < 5; ++i) lambda[i] = R.lambda[i];
  if (set_begin)
    for (int s = 0; s <= PAGANI_NSETS; ++s) set_begin[s] = R.set_begin[s];
  if (gtab) std::memcpy(gtab, R.gtab.data(), R.gtab.size() * sizeof(double));
  return R.feval;
}

int pagani_rule_point(int ndim, int p, int axis, double* g)
{
  static std::mutex mu;
  static pagani::RuleTables tables[PAGANI_MAX_DIM + 1];
  if (!g || ndim < PAGANI_MIN_DIM || ndim > PAGANI_MAX_DIM || axis < 0 || axis >= ndim)
    return PAGANI_EBADARG;
  std::lock_guard<std::mutex> lock(mu);
  pagani::RuleTables& R = tables[ndim];
  if (R.feval == 0 && !pagani::build_rule(ndim, R)) return PAGANI_EBADARG;
  if (p < 0 || p >= R.feval) return PAGANI_EBADARG;
  *g = point_generator(R, p, axis);
  return PAGANI_OK;
}

static thread_local std::string g_create_error;

const char* pagani_create_error(void) { return g_create_error.c_str(); }

pagani_ws* pagani_ws_create(int ndim, int device)
{
  g_create_error.clear();
  auto bail = [](pagani_ws* ws, const std::string& why, hipError_t e) -> pagani_ws* {
    g_create_error = why + (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : "");
    (void)hipGetLastError();
    if (ws) pagani_ws_destroy(ws);
    return nullptr;
  };
  if (ndim < PAGANI_MIN_DIM || ndim > PAGANI_MAX_DIM)
    return bail(nullptr, "ndim outside [2, 16]", hipSuccess);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return bail(nullptr, "hipGetDeviceCount", e);
  if (device < 0 && (e = hipGetDevice(&device)) != hipSuccess)
    return bail(nullptr, "hipGetDevice", e);
  if (device >= ndev) return bail(nullptr, "device index out of range", hipSuccess);
  pagani_ws* ws = new pagani_ws();
  ws->ndim = ndim;
  ws->device = device;
  if ((e = hipSetDevice(device)) != hipSuccess) return bail(ws, "hipSetDevice", e);
  if ((e = hipStreamCreate(&ws->stream)) != hipSuccess) return bail(ws, "hipStreamCreate", e);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ws->hbm_total = prop.totalGlobalMem;
    ws->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }
  (void)hipGetLastError();
  if (!pagani::build_rule(ndim, ws->rule)) return bail(ws, "rule tables", hipSuccess);
  const std::size_t gbytes = ws->rule.gtab.size() * sizeof(double);
  if ((e = hipMalloc(&ws->d_gtab, gbytes)) != hipSuccess) return bail(ws, "hipMalloc(gtab)", e);
  if ((e = hipMemcpy(ws->d_gtab, ws->rule.gtab.data(), gbytes, hipMemcpyHostToDevice)) !=
      hipSuccess)
    return bail(ws, "hipMemcpy(gtab)", e);
  if ((e = hipMalloc(&ws->d_scal, sizeof(ph::Scalars))) != hipSuccess)
    return bail(ws, "hipMalloc(scalars)", e);
  if ((e = hipHostMalloc(&ws->h_scal, sizeof(ph::Scalars), hipHostMallocDefault)) != hipSuccess)
    return bail(ws, "hipHostMalloc(scalars)", e);
  const std::size_t mpart = std::size_t(ph::kMultiBlocks) * 3 * ph::kMaxThresholds;
  if ((e = hipMalloc(&ws->d_mpart, mpart * sizeof(double))) != hipSuccess)
    return bail(ws, "hipMalloc(threshold partials)", e);
  if ((e = hipMalloc(&ws->d_mout, 2 * ph::kMaxThresholds * sizeof(double))) != hipSuccess)
    return bail(ws, "hipMalloc(threshold results)", e);
  if ((e = hipHostMalloc(&ws->h_mout, 2 * ph::kMaxThresholds * sizeof(double),
                         hipHostMallocDefault)) != hipSuccess)
    return bail(ws, "hipHostMalloc(threshold results)", e);
  if ((e = hipMalloc(&ws->d_search, sizeof(SearchState))) != hipSuccess)
    return bail(ws, "hipMalloc(classifier search)", e);
  if ((e = hipHostMalloc(&ws->h_search, sizeof(SearchState), hipHostMallocDefault)) != hipSuccess)
    return bail(ws, "hipHostMalloc(classifier search)", e);
  if ((e = hipMalloc(&ws->d_loop, sizeof(DevLoop))) != hipSuccess)
    return bail(ws, "hipMalloc(device loop state)", e);
  if ((e = hipHostMalloc(&ws->h_loop, sizeof(DevLoop), hipHostMallocDefault)) != hipSuccess)
    return bail(ws, "hipHostMalloc(device loop state)", e);
  for (auto& ev : ws->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(ws, "hipEventCreate", e);
  return ws;
}

void pagani_ws_destroy(pagani_ws* ws)
{
  if (!ws) return;
  if (ws->stream) (void)hipStreamSynchronize(ws->stream);
  for_each_pool(ws, [](auto& b) { release(b); });
  if (ws->d_mpart) (void)hipFree(ws->d_mpart);
  if (ws->d_mout) (void)hipFree(ws->d_mout);
  if (ws->h_mout) (void)hipHostFree(ws->h_mout);
  if (ws->d_search) (void)hipFree(ws->d_search);
  if (ws->h_search) (void)hipHostFree(ws->h_search);
  if (ws->d_loop) (void)hipFree(ws->d_loop);
  if (ws->h_loop) (void)hipHostFree(ws->h_loop);
  if (ws->d_gtab) (void)hipFree(ws->d_gtab);
  if (ws->d_scal) (void)hipFree(ws->d_scal);
  if (ws->h_scal) (void)hipHostFree(ws->h_scal);
  for (auto& e : ws->ev)
    if (e) (void)hipEventDestroy(e);
  if (ws->stream) (void)hipStreamDestroy(ws->stream);
  delete ws;
}

const char* pagani_ws_last_error(const pagani_ws* ws)
{
  return ws ? ws->error.c_str() : "null workspace";
}

int pagani_ws_ndim(const pagani_ws* ws) { return ws ? ws->ndim : PAGANI_EBADARG; }

int pagani_ws_set_device_mem(pagani_ws* ws, size_t total_bytes)
{
  if (!ws) return PAGANI_EBADARG;
  ws->mem_model_total = total_bytes;
  return PAGANI_OK;
}

int pagani_ws_set_mem_model(pagani_ws* ws, int model)
{
  if (!ws || (model != PAGANI_MEM_REFERENCE && model != PAGANI_MEM_ACTUAL))
    return fail(ws, PAGANI_EBADARG, "unknown memory model");
  ws->mem_model = model;
  return PAGANI_OK;
}

size_t pagani_ws_device_mem(const pagani_ws* ws)
{
  if (!ws) return 0;
  return ws->mem_model_total ? ws->mem_model_total : ws->hbm_total;
}

int pagani_ws_set_device_loop(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->device_loop = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_set_slices(pagani_ws* ws, int slices)
{
  if (!ws || slices < 0 || slices > PAGANI_MAX_SLICES || (slices & (slices - 1)))
    return fail(ws, PAGANI_EBADARG, "slices must be 0 or a power of two <= 8");
  ws->slices_override = slices;
  return PAGANI_OK;
}

int pagani_ws_set_timing(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->timing = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_get_stats(const pagani_ws* ws, pagani_stats* out)
{
  if (!ws || !out) return PAGANI_EBADARG;
  *out = ws->stats;
  return PAGANI_OK;
}

int pagani_ws_reset_stats(pagani_ws* ws)
{
  if (!ws) return PAGANI_EBADARG;
  std::memset(&ws->stats, 0, sizeof(ws->stats));
  return PAGANI_OK;
}

int pagani_ws_eval_consts(pagani_ws* ws, const double* lows, const double* highs, size_t n,
                          pagani_eval_consts* out)
{
  if (!ws || !out) return PAGANI_EBADARG;
  make_consts(ws, lows, highs, n, *out);
  return PAGANI_OK;
}

void* pagani_ws_stream(pagani_ws* ws) { return ws ? ws->stream : nullptr; }

size_t pagani_ws_trace(const pagani_ws* ws, size_t cap, uint64_t* n, uint64_t* m, double* est,
                       double* err)
{
  if (!ws) return 0;
  const std::size_t k = std::min(cap, ws->trace.size());
  for (std::size_t i = 0; i < k; ++i) {
    if (n) n[i] = ws->trace[i].n;
    if (m) m[i] = ws->trace[i].m;
    if (est) est[i] = ws->trace[i].estimate;
    if (err) err[i] = ws->trace[i].errorest;
  }
  return ws->trace.size();
}

int pagani_ws_live_regions(pagani_ws* ws, size_t* n, const double** L, const double** H)
{
  if (!ws || !n || !L || !H) return PAGANI_EBADARG;
  if ((ws->split || ws->lv.on) && ws->live_n > 0) {
    // write the list out flat, widths as doubles, into the other pair, which
    // then holds it
    if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
    const int d = ws->ndim, nxt = 1 - ws->cur;
    int rc;
    if ((rc = ensure(ws, ws->L[nxt], ws->live_n * d, "region bounds")) ||
        (rc = ensure(ws, ws->H[nxt], ws->live_n * d, "region bounds")))
      return rc;
    HIP_TRY(ws, ph::expand(ws->stream, d, ws->live_n, ws->L[ws->cur].p, ws->H[ws->cur].p,
                           ws->split ? ws->pdim[ws->cur].p : nullptr, ws->lv, ws->L[nxt].p,
                           ws->H[nxt].p));
    HIP_TRY(ws, hipStreamSynchronize(ws->stream));
    ws->cur = nxt;
    ws->split = false;
    ws->lv = ph::Levels{};
    if (ws->susp.held) {
      ws->susp.split = 0;
      ws->susp.lv = ph::Levels{};
    }
  }
  *n = ws->live_n;
  *L = ws->L[ws->cur].p;
  *H = ws->H[ws->cur].p;
  return PAGANI_OK;
}

int pagani_integrate_cb(pagani_ws* ws, pagani_eval_fn eval, void* ctx, double epsrel,
                        double epsabs, const double* lows, const double* highs,
                        int relerr_classification, const pagani_region_list* init, size_t init_k,
                        pagani_result* out)
{
  if (!ws) return PAGANI_EBADARG;
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  return integrate_impl(ws, eval, ctx, epsrel, epsabs, lows, highs, relerr_classification, init,
                        init_k, nullptr, nullptr, 0, out);
}

int pagani_integrate_genz(pagani_ws* ws, const pagani_genz* f, int precision, double epsrel,
                          double epsabs, const double* lows, const double* highs,
                          int relerr_classification, size_t init_k, pagani_result* out)
{
  if (!ws) return PAGANI_EBADARG;
  int rc = check_genz(ws, f, precision);
  if (rc) return rc;
  GenzCtx ctx{*f, precision};
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  return integrate_impl(ws, genz_trampoline, &ctx, epsrel, epsabs, lows, highs,
                        relerr_classification, nullptr, init_k, nullptr, f, precision, out);
}

size_t pagani_ws_footprint(const pagani_ws* ws)
{
  if (!ws) return 0;
  std::size_t bytes = 0;
  for_each_pool(const_cast<pagani_ws*>(ws), [&](auto& b) { bytes += b.cap * sizeof(*b.p); });
  return bytes;
}

int pagani_ws_trim(pagani_ws* ws)
{
  if (!ws) return PAGANI_EBADARG;
  if (ws->susp.held) return fail(ws, PAGANI_EBADARG, "a suspended run holds the region pool");
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  harvest_compact(ws);
  for_each_pool(ws, [](auto& b) { release(b); });
  ws->live_n = 0;
  ws->live_evaluated = false;
  ws->snap_n = 0;
  return PAGANI_OK;
}

int pagani_ws_set_iteration_limit(pagani_ws* ws, uint64_t iterations)
{
  if (!ws) return PAGANI_EBADARG;
  ws->iter_limit = iterations;
  return PAGANI_OK;
}

int pagani_ws_suspended(const pagani_ws* ws) { return ws ? (ws->susp.held ? 1 : 0) : 0; }

int pagani_integrate_cb_resume(pagani_ws* ws, pagani_eval_fn eval, void* ctx, pagani_result* out)
{
  if (!ws) return PAGANI_EBADARG;
  if (!ws->susp.held) return fail(ws, PAGANI_EBADARG, "no suspended run to resume");
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  return integrate_impl(ws, eval, ctx, 0, 0, nullptr, nullptr, 0, nullptr, 0, &ws->susp,
                        nullptr, 0, out);
}

int pagani_integrate_genz_resume(pagani_ws* ws, const pagani_genz* f, int precision,
                                 pagani_result* out)
{
  if (!ws) return PAGANI_EBADARG;
  int rc = check_genz(ws, f, precision);
  if (rc) return rc;
  if (!ws->susp.held) return fail(ws, PAGANI_EBADARG, "no suspended run to resume");
  const pagani_genz& g = ws->susp.genz;
  if (g.family != 0) {
    bool same = g.family == f->family && g.ndim == f->ndim && ws->susp.precision == precision;
    for (int j = 0; same && j < f->ndim; ++j)
      same = g.a[j] == f->a[j] && g.u[j] == f->u[j];
    if (!same) return fail(ws, PAGANI_EBADARG, "integrand differs from the suspended run's");
  }
  GenzCtx ctx{*f, precision};
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  return integrate_impl(ws, genz_trampoline, &ctx, 0, 0, nullptr, nullptr, 0, nullptr, 0,
                        &ws->susp, nullptr, 0, out);
}

int pagani_ws_set_graphs(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->graphs = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_set_tuned_slices(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->tuned = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_set_predict_split(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->predict_split = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_set_level_form(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->level_form = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_set_device_search(pagani_ws* ws, int enable)
{
  if (!ws) return PAGANI_EBADARG;
  ws->device_search = enable != 0;
  return PAGANI_OK;
}

int pagani_ws_set_debug(pagani_ws* ws, int level, const char* dir, const char* id, int append)
{
  if (!ws || level < 0 || level > 3) return PAGANI_EBADARG;
  ws->dbg.level = level;
  ws->dbg.dir = dir && *dir ? dir : ".";
  ws->dbg.plain = id == nullptr;
  ws->dbg.id = id ? id : "";
  ws->dbg.append = append != 0;
  return PAGANI_OK;
}

int pagani_ws_set_snapshot_iter(pagani_ws* ws, uint64_t iteration)
{
  if (!ws) return PAGANI_EBADARG;
  ws->snapshot_iter = iteration;
  ws->snap_n = 0;
  return PAGANI_OK;
}

int pagani_ws_snapshot_regions(const pagani_ws* ws, size_t* n, const double** L,
                               const double** H)
{
  if (!ws || !n || !L || !H) return PAGANI_EBADARG;
  *n = ws->snap_n;
  *L = ws->snap_n ? ws->snapL.p : nullptr;
  *H = ws->snap_n ? ws->snapH.p : nullptr;
  return PAGANI_OK;
}

int pagani_uniform_split_dev(pagani_ws* ws, size_t k, double* L, double* H)
{
  if (!ws || !L || !H || k == 0) return fail(ws, PAGANI_EBADARG, "bad uniform split arguments");
  std::size_t n = 1;
  for (int j = 0; j < ws->ndim; ++j) n *= k;
  HIP_TRY(ws, ph::uniform_split(ws->stream, ws->ndim, k, n, L, H));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

int pagani_eval_genz_dev(pagani_ws* ws, const pagani_genz* f, int precision, size_t n,
                         const double* L, const double* H, const double* lows,
                         const double* highs, double* est, double* err, uint8_t* dim)
{
  if (!ws) return PAGANI_EBADARG;
  int rc = check_genz(ws, f, precision);
  if (rc) return rc;
  if (n == 0) return PAGANI_OK;
  if (!L || !H || !est || !err || !dim) return fail(ws, PAGANI_EBADARG, "null device pointer");
  if ((rc = check_volume(ws, lows, highs))) return rc;
  pagani_eval_args args{};  // dev = NULL: an ordinary launch
  args.n = n;
  args.L = L;
  args.H = H;
  args.est = est;
  args.err = err;
  args.dim = dim;
  args.stream = ws->stream;
  ws->tune_family = f->family;
  ws->tune_precision = precision;
  make_consts(ws, lows, highs, n, args.consts);
  record(ws, EV_EVAL0);
  rc = pagani_genz_eval_launch(f, precision, &args);
  if (rc) return fail(ws, rc, "evaluation launch failed");
  record(ws, EV_EVAL1);
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  if (ws->timing) ws->stats.eval_ms += elapsed(ws, EV_EVAL0, EV_EVAL1);
  ws->stats.regions_evaluated += n;
  ws->stats.eval_launches++;
  return PAGANI_OK;
}

int pagani_refine_dev(pagani_ws* ws, size_t n, const double* est, const double* err_in,
                      double* err_out, const double* parent_est, double epsrel,
                      int relerr_classification, uint8_t* active, double* sums,
                      size_t* num_active)
{
  if (!ws || !est || !err_in || !err_out || !active)
    return fail(ws, PAGANI_EBADARG, "null argument");
  if (parent_est && (n % 2)) return fail(ws, PAGANI_EBADARG, "refinement needs an even count");
  if (parent_est && err_in == err_out)
    return fail(ws, PAGANI_EBADARG, "refinement cannot run in place");
  if (n == 0) return PAGANI_OK;
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::refine(ws->stream, n, est, err_in, err_out, parent_est, epsrel,
                         relerr_classification, active, ws->part.p, ws->cnt.p));
  if ((rc = finalize_read(ws, n, true, false, parent_est != nullptr))) return rc;
  if (sums)
    for (int i = 0; i < 4; ++i) sums[i] = ws->h_scal->sum[i];
  if (num_active) *num_active = ws->h_scal->count;
  return PAGANI_OK;
}

int pagani_filter_dev(pagani_ws* ws, size_t n, const double* L, const double* H,
                      const uint8_t* active, const double* est, const double* err,
                      const uint8_t* dim, double* L2, double* H2, double* parent_est,
                      double* parent_err, uint8_t* dim2, size_t* m_out)
{
  if (!ws || !L || !H || !active || !est || !err || !dim || !L2 || !H2 || !m_out)
    return fail(ws, PAGANI_EBADARG, "null argument");
  *m_out = 0;
  if (n == 0) return PAGANI_OK;
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::count_flags(ws->stream, n, active, est, err, ws->part.p, ws->cnt.p, false));
  if ((rc = finalize_read(ws, n, true, false))) return rc;
  const std::size_t m = ws->h_scal->count;
  *m_out = m;
  if (m == 0) return PAGANI_OK;
  HIP_TRY(ws, ph::compact(ws->stream, 0, ws->ndim, n, m, L, H, active, est, err, dim, ws->off.p,
                          ws->foff.p, L2, H2, parent_est, parent_err, dim2));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

int pagani_split_dev(pagani_ws* ws, size_t m, const double* L, const double* H,
                     const uint8_t* dim, double* L2, double* H2)
{
  if (!ws || !L || !H || !dim || !L2 || !H2) return fail(ws, PAGANI_EBADARG, "null argument");
  HIP_TRY(ws, ph::split_all(ws->stream, ws->ndim, m, L, H, dim, L2, H2));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

int pagani_filter_split_dev(pagani_ws* ws, size_t n, const double* L, const double* H,
                            const uint8_t* active, const double* est, const double* err,
                            const uint8_t* dim, double* L2, double* H2, double* parent_est,
                            double* parent_err, size_t* m_out)
{
  if (!ws || !L || !H || !active || !est || !err || !dim || !L2 || !H2 || !m_out)
    return fail(ws, PAGANI_EBADARG, "null argument");
  *m_out = 0;
  if (n == 0) return PAGANI_OK;
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::count_flags(ws->stream, n, active, est, err, ws->part.p, ws->cnt.p, false));
  if ((rc = finalize_read(ws, n, true, false))) return rc;
  const std::size_t m = ws->h_scal->count;
  *m_out = m;
  if (m == 0) return PAGANI_OK;
  HIP_TRY(ws, ph::compact(ws->stream, 1, ws->ndim, n, m, L, H, active, est, err, dim, ws->off.p,
                          ws->foff.p, L2, H2, parent_est, parent_err, nullptr));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

int pagani_sum_dev(pagani_ws* ws, const double* x, size_t n, double* out)
{
  if (!ws || !out || (n && !x)) return fail(ws, PAGANI_EBADARG, "null argument");
  *out = 0.0;
  if (n == 0) return PAGANI_OK;
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::sum_dot(ws->stream, n, x, nullptr, ws->part.p, ws->cnt.p));
  if ((rc = finalize_read(ws, n, false, false))) return rc;
  *out = ws->h_scal->sum[0];
  return PAGANI_OK;
}

int pagani_dot_dev(pagani_ws* ws, const uint8_t* flags, const double* x, size_t n, double* out)
{
  if (!ws || !out || (n && (!x || !flags))) return fail(ws, PAGANI_EBADARG, "null argument");
  *out = 0.0;
  if (n == 0) return PAGANI_OK;
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::sum_dot(ws->stream, n, x, flags, ws->part.p, ws->cnt.p));
  if ((rc = finalize_read(ws, n, false, false))) return rc;
  *out = ws->h_scal->sum[2];
  return PAGANI_OK;
}

int pagani_minmax_dev(pagani_ws* ws, const double* x, size_t n, double* mn, double* mx)
{
  if (!ws || !x || !mn || !mx || n == 0) return fail(ws, PAGANI_EBADARG, "bad minmax arguments");
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::minmax(ws->stream, n, x, ws->part.p));
  if ((rc = finalize_read(ws, n, false, true))) return rc;
  *mn = ws->h_scal->mn;
  *mx = ws->h_scal->mx;
  return PAGANI_OK;
}

int pagani_scan_flags_dev(pagani_ws* ws, const uint8_t* flags, size_t n, uint64_t* offsets,
                          size_t* total)
{
  if (!ws || !flags || !offsets || !total) return fail(ws, PAGANI_EBADARG, "null argument");
  *total = 0;
  if (n == 0) return PAGANI_OK;
  int rc = ensure_chunks(ws, n);
  if (rc) return rc;
  HIP_TRY(ws, ph::sum_dot(ws->stream, n, nullptr, flags, ws->part.p, ws->cnt.p));
  if ((rc = finalize_read(ws, n, true, false))) return rc;
  *total = ws->h_scal->count;
  HIP_TRY(ws, ph::scan_flags(ws->stream, n, flags, ws->off.p, ws->foff.p, offsets));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

int pagani_classify_dev(pagani_ws* ws, const double* err, size_t n, double epsrel,
                        const double* est_hist, double iter_errorest,
                        double iter_finished_errorest, double total_finished_errorest,
                        uint8_t* flags, pagani_class_res* out)
{
  if (!ws || !err || !flags || !out || n == 0 || !est_hist)
    return fail(ws, PAGANI_EBADARG, "bad classify arguments");
  const std::size_t total = ws->mem_model_total ? ws->mem_model_total : ws->hbm_total;
  Classifier c(epsrel, 0.0, total, ws->ndim, ws->mem_model);
  c.hist[0] = est_hist[0];
  c.hist[1] = est_hist[1];
  c.hist[2] = est_hist[2];
  c.iters_collected = 3;
  double thr = 0.0;
  int rc = classify(ws, c, err, n, iter_errorest, iter_finished_errorest, total_finished_errorest,
                    *out, &thr);
  if (rc) return rc;
  // the flags of the last applied threshold (Classification_res::active_flags)
  HIP_TRY(ws, ph::threshold(ws->stream, n, err, thr, flags, ws->part.p, ws->cnt.p));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

}  // extern "C"
