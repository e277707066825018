// ws_internal.hpp -- the workspace object and the helpers shared by the
// single-integral driver (driver.hip) and the batched driver (batch.hip).
// Library-internal; not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "classify_search.hpp"
#include "pagani.h"
#include "phases.hpp"
#include "recorder.hpp"
#include "rule.hpp"

namespace pagani {
namespace internal {

namespace ph = pagani::phases;

// Tuned launch shapes: slices per power-of-two size class of the list
// (bucket b holds 2^b <= n < 2^(b+1)), measured by gpuintegration_amd/
// tune.py per integrand family, dimension and precision; 0 = no
// measurement (choose_slices_n's rule decides).
constexpr int kSliceBuckets = 32;

template <class T>
struct DevBuf {
  T* p = nullptr;
  std::size_t cap = 0;
};

struct TraceRec {
  std::uint64_t n, m;
  double estimate, errorest;
};

// State of the device-resident loop (device memory, mirrored in pinned host
// memory between rounds).  decide_step (devloop_decide.hpp, run by the last
// workgroup of k_step_dev) makes the host loop's per-iteration decisions on
// it; see integrate_impl for when the loop runs on the device.
constexpr int kDevTrace = 64;  // trace records per round (a round is at most 64 iterations)
enum : std::int32_t {
  kDevRun = 0,        // every queued iteration committed; more may follow
  kDevDone = 1,       // converged: the last list was evaluated
  kDevDoneEmpty = 2,  // no region stayed active: the list is empty
  kDevHost = 3,       // the host finishes iteration `iter` from its scalars
  kDevLimit = 4,      // the iteration cap was reached after a split
};

struct DevLoop {
  pagani_dev_iter it;          // the next iteration
  pagani::phases::DevCompact cp;  // this iteration's filter + split
  // fixed for the run
  std::uint64_t n_cap, mem_total;
  double epsrel, epsabs;
  std::int32_t d, mem_model, relerr, user_regions;
  // the loop's running state
  std::uint64_t iter;
  double cum_est, cum_err;
  std::uint64_t cum_nregions, cum_iters;
  std::int32_t cum_status, status;
  double hist[3];
  std::uint64_t iters_collected;
  // the evaluation's slice choice (choose_slices, with the integrand's
  // tuned table when has_table) and this round's launch: its waves per
  // workgroup and the largest list its grids cover
  std::int32_t num_cus, slices_override, generic, launch_slices;
  std::int32_t has_table, pad_table;
  std::uint8_t slice_table[kSliceBuckets];
  std::uint64_t round_cap;
  // per round
  std::uint64_t evaluated, iterations, max_n;
  std::uint32_t ntrace;
  // k_step_dev's workgroup counter (zero between launches)
  std::uint32_t ticket;
  TraceRec trace[kDevTrace];
};

// LoopState::relerr value of a run whose relerr flag follows the
// predict_split rule, iteration by iteration (0 and 1 are the caller's flag).
constexpr int kRelerrPredictSplit = 2;

// Everything the iteration loop carries from one iteration to the next,
// besides the region list itself (L[cur], H[cur]: n regions) and the
// parents' estimates (pest: n/2 when has_parents).  A run suspended at an
// iteration boundary keeps it in the workspace; a checkpoint file holds it
// with the arrays.
struct LoopState {
  bool held = false;
  std::uint64_t it = 0;  // index of the next iteration
  std::uint64_t n = 0;   // regions to evaluate next
  int has_parents = 0;
  int split = 0;         // the list is held in split form
  phases::Levels lv{};   // the list's widths are levels (the library's uniform start)
  int user_regions = 0;  // Sub_regions& overload bookkeeping
  int relerr = 1;
  double epsrel = 0, epsabs = 0;
  double lows[PAGANI_MAX_DIM] = {}, highs[PAGANI_MAX_DIM] = {};
  pagani_result cum{};
  double hist[3] = {0, 0, 0};  // classifier's last three estimates
  std::uint64_t iters_collected = 0;
  int mem_model = PAGANI_MEM_REFERENCE;
  std::uint64_t mem_total = 0;
  // the Genz descriptor when the run integrates one (family 0: a callback)
  pagani_genz genz{};
  int precision = 0;
};

}  // namespace internal
}  // namespace pagani

struct pagani_ws {
  int ndim = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  std::size_t hbm_total = 0;
  std::size_t mem_model_total = 0;  // 0 = hbm_total
  int mem_model = PAGANI_MEM_REFERENCE;
  int num_cus = 256;
  int slices_override = 0;
  bool timing = false;
  pagani_stats stats{};
  pagani::RuleTables rule;
  double* d_gtab = nullptr;

  pagani::internal::DevBuf<double> L[2], H[2];
  pagani::internal::DevBuf<std::uint8_t> pdim[2];  // halved axes of a list in split form
  int cur = 0;
  bool split = false;  // the list in L[cur], H[cur] is in split form (pagani_eval_args)
  // H[cur] holds levels, not widths (pagani_eval_args); level_form: the
  // uniform start uses them (pagani_ws_set_level_form)
  pagani::phases::Levels lv{};
  bool level_form = true;
  pagani::internal::DevBuf<double> est, err0, err1, pest, perr;
  pagani::internal::DevBuf<std::uint8_t> dim, active, flags;
  pagani::internal::DevBuf<double> part, fpart;
  pagani::internal::DevBuf<std::uint32_t> cnt, fcnt;
  pagani::internal::DevBuf<std::uint64_t> off, foff;
  pagani::phases::Scalars* d_scal = nullptr;
  pagani::phases::Scalars* h_scal = nullptr;
  // multi-threshold pass: per-block partials, results (device / pinned host)
  double* d_mpart = nullptr;
  double* d_mout = nullptr;
  double* h_mout = nullptr;
  // the classifier's device search (classify_search.hpp) and its host copy
  pagani::internal::SearchState* d_search = nullptr;
  pagani::internal::SearchState* h_search = nullptr;
  bool device_search = true;
  // d_scal holds the min and max of the iteration's refined errors (set
  // after the refinement's finalize, cleared by every other finalize)
  bool scal_minmax = false;
  hipEvent_t ev[10] = {};
  // device-resident loop, and its full-length rounds as HIP graphs
  bool device_loop = true;
  bool graphs = true;
  pagani::internal::DevLoop* d_loop = nullptr;
  pagani::internal::DevLoop* h_loop = nullptr;
  bool compact_pending = false;
  std::vector<pagani::internal::TraceRec> trace;
  std::size_t live_n = 0;  // regions in L[cur], H[cur] after the last call
  // est / err0-or-err1 / dim hold the live list's values (a finished call
  // whose last iteration evaluated that list)
  bool live_evaluated = false;
  bool live_err1 = false;
  // suspension and checkpoints
  std::uint64_t iter_limit = 0;  // iterations per call before suspending; 0 = none
  pagani::internal::LoopState susp;
  // snapshot of the region list evaluated in iteration snapshot_iter when no
  // region had finished by then (Workspace.cuh:244-248, predict_split)
  std::uint64_t snapshot_iter = ~std::uint64_t(0);
  // relerr classification by the predict_split rule (Workspace.cuh:206-210)
  bool predict_split = false;
  // launch shapes from the tuned table (pagani_ws_set_tuned_slices), for
  // the integrand of the running call: its family (0: a callback) and
  // precision
  bool tuned = true;
  int tune_family = 0, tune_precision = PAGANI_FP64;
  pagani::internal::DevBuf<double> snapL, snapH;
  std::size_t snap_n = 0;
  // debug recording of the following integrate calls (pagani_ws_set_debug)
  // and the device buffer of its point values
  pagani::internal::DebugConfig dbg;
  pagani::internal::DevBuf<double> dfe;
};

namespace pagani {
namespace internal {

inline int fail(pagani_ws* ws, int code, const std::string& msg)
{
  if (ws) ws->error = msg;
  return code;
}

inline int hip_fail(pagani_ws* ws, hipError_t e, const char* where)
{
  return fail(ws, PAGANI_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ws, expr)                                  \
  do {                                                     \
    const hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(ws, e_, #expr);  \
  } while (0)

// Grow a buffer to at least n elements (contents are not kept: every
// buffer is rewritten before it is read).  Region lists roughly double per
// iteration, so while memory is plentiful a buffer grows to twice the
// request: it then fits the next iteration as well and is reallocated every
// other iteration at most.  Once twice the request would take more than an
// eighth of the free device memory, buffers are sized exactly, so headroom
// kept by one buffer never starves the next.
template <class T>
int ensure(pagani_ws* ws, DevBuf<T>& b, std::size_t n, const char* what)
{
  if (n <= b.cap) return PAGANI_OK;
  const auto t0 = std::chrono::steady_clock::now();
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
  }
  std::size_t want = n;
  std::size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    if (2 * n * sizeof(T) <= free_b / 8) want = 2 * n;
  } else {
    (void)hipGetLastError();
  }
  void* p = nullptr;
  if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    want = n;
    if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ws, PAGANI_ENOMEM,
                  std::string("out of device memory allocating ") + what + " (" +
                      std::to_string(n * sizeof(T)) + " bytes)");
    }
  }
  b.p = static_cast<T*>(p);
  b.cap = want;
  ws->stats.alloc_ms +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  ws->stats.allocations++;
  return PAGANI_OK;
}

template <class T>
inline void release(DevBuf<T>& b)
{
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

// Every growable buffer of two workspaces, pairwise (same field, same type).
template <class F>
inline void for_each_pool_pair(pagani_ws* a, pagani_ws* b, F&& f)
{
  for (int i = 0; i < 2; ++i) {
    f(a->L[i], b->L[i]);
    f(a->H[i], b->H[i]);
    f(a->pdim[i], b->pdim[i]);
  }
  f(a->est, b->est);
  f(a->err0, b->err0);
  f(a->err1, b->err1);
  f(a->pest, b->pest);
  f(a->perr, b->perr);
  f(a->dim, b->dim);
  f(a->active, b->active);
  f(a->flags, b->flags);
  f(a->part, b->part);
  f(a->fpart, b->fpart);
  f(a->cnt, b->cnt);
  f(a->fcnt, b->fcnt);
  f(a->off, b->off);
  f(a->foff, b->foff);
  f(a->snapL, b->snapL);
  f(a->snapH, b->snapH);
  f(a->dfe, b->dfe);
}

// Move the larger of each pair of pool buffers from `from` to `to` (both on
// one device, neither suspended); both workspaces' region lists are
// dropped.  The batch passes one worker's pool on to the workspace of the
// next dimension instead of freeing it and growing a new one.
int pool_handover(pagani_ws* from, pagani_ws* to);

// Every growable buffer of the workspace, for trimming and accounting.
template <class F>
inline void for_each_pool(pagani_ws* ws, F&& f)
{
  f(ws->L[0]);
  f(ws->L[1]);
  f(ws->H[0]);
  f(ws->H[1]);
  f(ws->pdim[0]);
  f(ws->pdim[1]);
  f(ws->est);
  f(ws->err0);
  f(ws->err1);
  f(ws->pest);
  f(ws->perr);
  f(ws->dim);
  f(ws->active);
  f(ws->flags);
  f(ws->part);
  f(ws->fpart);
  f(ws->cnt);
  f(ws->fcnt);
  f(ws->off);
  f(ws->foff);
  f(ws->snapL);
  f(ws->snapH);
  f(ws->dfe);
}

inline int ensure_chunks(pagani_ws* ws, std::size_t n)
{
  // a paired list (ph::num_chunks(n, true)) may have one chunk more
  const std::size_t nch = ph::num_chunks(n) + 1;
  const std::size_t nfold = std::max<std::size_t>(ph::num_folds(nch), 1);
  int rc;
  if ((rc = ensure(ws, ws->part, nch * ph::kPartDoubles, "chunk sums")) ||
      (rc = ensure(ws, ws->cnt, nch, "chunk counts")) ||
      (rc = ensure(ws, ws->off, nch, "chunk offsets")) ||
      (rc = ensure(ws, ws->fpart, nfold * ph::kPartDoubles, "fold sums")) ||
      (rc = ensure(ws, ws->fcnt, nfold, "fold counts")))
    return rc;
  return ensure(ws, ws->foff, nfold, "fold offsets");
}

inline hipError_t launch_finalize(pagani_ws* ws, std::size_t n, bool counts, bool minmax,
                                  bool paired = false)
{
  ws->scal_minmax = false;
  return ph::finalize(ws->stream, ph::num_chunks(n, paired), ws->part.p,
                      counts ? ws->cnt.p : nullptr,
                      counts ? ws->off.p : nullptr, ws->fpart.p, ws->fcnt.p, ws->foff.p,
                      ws->d_scal, minmax ? 1 : 0);
}

// Finalize the chunk records of an n-region pass (in the chunks of
// ph::num_chunks(n, paired)) and read the scalars back.
inline int finalize_read(pagani_ws* ws, std::size_t n, bool counts, bool minmax,
                         bool paired = false)
{
  HIP_TRY(ws, launch_finalize(ws, n, counts, minmax, paired));
  HIP_TRY(ws, hipMemcpyAsync(ws->h_scal, ws->d_scal, sizeof(ph::Scalars),
                             hipMemcpyDeviceToHost, ws->stream));
  HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  return PAGANI_OK;
}

enum { EV_EVAL0, EV_EVAL1, EV_REF0, EV_REF1, EV_CMP0, EV_CMP1, EV_CLS0, EV_CLS1, EV_DEV0, EV_DEV1 };

inline void record(pagani_ws* ws, int i)
{
  if (ws->timing) (void)hipEventRecord(ws->ev[i], ws->stream);
}

inline double elapsed(pagani_ws* ws, int a, int b)
{
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ws->ev[a], ws->ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0.0;
  }
  return ms;
}

__host__ __device__ inline int slice_bucket(std::size_t n)
{
  int b = 0;
  while (b + 1 < kSliceBuckets && (n >> (b + 1)) != 0) ++b;
  return b;
}

// Waves per batch of 64 regions for an n-region evaluation: the tuned
// table's entry for n's size class when there is one; otherwise more waves
// share a batch while there are too few batches to give every CU four
// waves per SIMD and each wave would still get at least 16 generic points.
// The rule is non-increasing in n; a table need not be (round_cap in
// driver.hip finds how far a launch shape holds either way).
__host__ __device__ inline int choose_slices_n(std::size_t n, int num_cus, int override_,
                                               int generic,
                                               const std::uint8_t* table = nullptr)
{
  if (override_ > 0) return override_ < PAGANI_MAX_SLICES ? override_ : PAGANI_MAX_SLICES;
  if (table) {
    const int t = table[slice_bucket(n)];
    if (t > 0) return t < PAGANI_MAX_SLICES ? t : PAGANI_MAX_SLICES;
  }
  const std::size_t batches = (n + 63) / 64;
  const std::size_t target = static_cast<std::size_t>(num_cus) * 4 * 4;
  int s = 1;
  while (s < PAGANI_MAX_SLICES && batches * s < target && generic / (2 * s) >= 16) s *= 2;
  return s;
}

__host__ __device__ inline std::size_t mem_required(std::size_t n, int d, int model)
{
  const std::size_t nd = n * static_cast<std::size_t>(d);
  if (model == PAGANI_MEM_ACTUAL) {
    // lists are held in split form (parents' bounds + halved axis): the
    // list's n/2 parents (8d per region) + the next list's n parents (16d)
    // + the next list's est/err0/err1 (48) and dim/active/flags (6) + both
    // lists' halved axes (1.5) + the next parents' estimates (8)
    return 8 * nd + 16 * nd + 48 * n + 6 * n + 2 * n + 8 * n;
  }
  // num_doubles_needed / num_ints_needed / device_mem_required_for_full_split
  const std::size_t dbl = 8 * n + 10 * nd;
  const std::size_t ints = 4 * n;
  return 8 * dbl + 4 * ints;
}

__host__ __device__ inline std::size_t free_mem(std::size_t n, int d, std::size_t total,
                                                int model)
{
  const std::size_t occ = mem_required(n, d, model);
  return total > occ ? total - occ : 1;
}

__host__ __device__ inline double mem_ratio(std::size_t n, int d, std::size_t total, int model)
{
  return static_cast<double>(mem_required(n, d, model)) /
         static_cast<double>(free_mem(n, d, total, model));
}

struct Classifier {
  double epsrel = 0, epsabs = 0;
  int required_digits = 0;
  double hist[3] = {0, 0, 0};  // uninitialised in the reference; zeroed here
  std::size_t iters_collected = 0;
  double max_budget = .25, max_active = .5;
  std::size_t total = 0;
  int d = 0;
  int model = PAGANI_MEM_REFERENCE;

  Classifier(double rel, double abs_, std::size_t tot, int ndim, int mem_model)
      : epsrel(rel), epsabs(abs_), total(tot), d(ndim), model(mem_model)
  {
    required_digits = static_cast<int>(std::ceil(std::log10(1 / epsrel)));
  }

  void store(double e)
  {
    hist[0] = hist[1];
    hist[1] = hist[2];
    hist[2] = e;
    iters_collected++;
  }

  // sigDigitsSame, :170-203
  bool sig_digits_same() const
  {
    double third = std::fabs(hist[0]), second = std::fabs(hist[1]), first = std::fabs(hist[2]);
    while (first != 0. && first < 1.) first *= 10;
    while (second != 0. && second < 1.) second *= 10;
    while (third != 0. && third < 1.) third *= 10;
    char a[512] = {0}, b[512] = {0}, c[512] = {0};
    std::snprintf(a, sizeof a - 1, "%.15f", third);
    std::snprintf(b, sizeof b - 1, "%.15f", second);
    std::snprintf(c, sizeof c - 1, "%.15f", first);
    bool verdict = true;
    int sig = 0;
    for (int i = 0; i < required_digits + 1 && sig < required_digits && verdict && i < 500; ++i) {
      verdict = c[i] == b[i] && b[i] == a[i];
      sig += (verdict && c[i] != '.') ? 1 : 0;
    }
    return verdict;
  }

  bool estimate_converged() const
  {
    return !(iters_collected - 1 < 1 || !sig_digits_same());
  }

  bool criteria_met(std::size_t n) const
  {
    const double ratio = mem_ratio(n, d, total, model);
    if (ratio > 1.) return true;
    return ratio > .1 && estimate_converged();
  }
};

int classify(pagani_ws* ws, Classifier& c, const double* err, std::size_t n, double iter_err,
             double iter_fin_err, double tot_fin_err, pagani_class_res& res, double* last_thr,
             bool have_minmax = false);

}  // namespace internal
}  // namespace pagani
