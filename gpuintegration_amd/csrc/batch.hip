// batch.hip -- independent integrals from one call (pagani_integrate_batch).
//
// The reference integrates a set of integrands with a host loop of
// Workspace::integrate calls (new_time_and_call.cuh:156-167, driven by
// randomized_integrands.cu:80-105).  Here a batch owns `workers` host
// threads; each has one workspace per dimension (created on first use, kept
// for later calls) and so its own stream.  Items are taken largest-first
// (higher dimension, then smaller epsrel) from a shared counter.  While one
// worker waits on its integral's per-iteration host round trip, the others'
// kernels keep the device busy.
//
// Every item runs the single-integral loop unchanged, on a workspace whose
// memory-model total is the batch's, so an item's result depends on nothing
// but the item and that total: not on the worker, the order or the overlap.
// The default total is the device's HBM divided by the number of workers,
// so that the pools the memory model admits fit on the device together.
//
// With pagani_batch_set_segments, the short FP64 Genz items first run whole
// in the persistent kernel (segments.hip); the ones it hands over continue
// on the workers from the state it reached.  Their results then agree with
// the single-integral loop's to rounding rather than to the bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "pagani.h"
#include "segments.hpp"
#include "ws_internal.hpp"

// concurrent launches of the persistent kernel
constexpr int kSegStreams = 4;

struct pagani_batch {
  int device = 0;
  int workers = 1;
  std::size_t hbm = 0;
  std::size_t mem_total = 0;  // 0 = hbm / workers
  int mem_model = PAGANI_MEM_REFERENCE;
  bool timing = false;
  bool tuned = true;
  // ws[w][d]: worker w's workspace for dimension d
  std::vector<std::array<pagani_ws*, PAGANI_MAX_DIM + 1>> ws;
  std::mutex mu;
  std::string error;
  // the persistent kernel (pagani_batch_set_segments): regions per segment
  // (0: off), the pool and the item arrays, kept between calls
  std::size_t seg_size = 0;
  int num_cus = 256;
  char* seg_pool = nullptr;
  std::size_t seg_pool_bytes = 0;
  void* seg_dev = nullptr;  // items, outputs and the counter
  std::size_t seg_dev_bytes = 0;
  std::uint64_t seg_integrals = 0;  // finished in segments (statistics)
  std::uint64_t seg_evaluated = 0;  // regions the kernel evaluated (statistics)
  std::uint64_t seg_iterations = 0;
  double seg_ms = 0;                // host time of the segment launches
  hipStream_t seg_stream[kSegStreams] = {};
  hipEvent_t seg_event[kSegStreams] = {};
};

namespace {

thread_local std::string g_batch_create_error;

std::size_t item_total(const pagani_batch* b)
{
  return b->mem_total ? b->mem_total : b->hbm / static_cast<std::size_t>(b->workers);
}

void set_error(pagani_batch* b, const std::string& msg)
{
  std::lock_guard<std::mutex> g(b->mu);
  if (b->error.empty()) b->error = msg;
}

// One item on worker w; fills *out and returns the integrate call's code.
int run_item(pagani_batch* b, int w, const pagani_batch_item& it, pagani_result* out)
{
  std::memset(out, 0, sizeof(*out));
  out->status = 1;
  out->lastPhase = -1;
  const int d = it.f.ndim;
  if (d < PAGANI_MIN_DIM || d > PAGANI_GENZ_MAX_DIM) {
    set_error(b, "batch item ndim outside [2, 10]");
    return PAGANI_EBADARG;
  }
  pagani_ws*& ws = b->ws[w][d];
  if (!ws) {
    ws = pagani_ws_create(d, b->device);
    if (!ws) {
      set_error(b, std::string("workspace creation failed: ") + pagani_create_error());
      return PAGANI_EHIP;
    }
  }
  pagani_ws_set_device_mem(ws, item_total(b));
  pagani_ws_set_mem_model(ws, b->mem_model);
  pagani_ws_set_timing(ws, b->timing ? 1 : 0);
  pagani_ws_set_tuned_slices(ws, b->tuned ? 1 : 0);
  // a graph capture on one worker's stream is broken by another worker's
  // device-wide synchronisation (a buffer freed while it grows): graphs
  // only when one worker runs
  pagani_ws_set_graphs(ws, b->workers == 1 ? 1 : 0);
  const int rc = pagani_integrate_genz(ws, &it.f, it.precision, it.epsrel, it.epsabs, nullptr,
                                       nullptr, it.relerr_classification, it.init_k, out);
  if (rc < 0 && rc != PAGANI_ENOMEM) set_error(b, pagani_ws_last_error(ws));
  return rc;
}

// An item the persistent kernel handed over (its state before iteration
// o.it, the list in the segment at `base`) continues on worker w: the list
// and the parents' estimates are copied into the workspace's pool, the
// loop state is held as a suspended run's, and the run resumes.
int run_item_resume(pagani_batch* b, int w, const pagani_batch_item& it,
                    const pagani::internal::SegOut& o, char* base, pagani_result* out)
{
  namespace I = pagani::internal;
  std::memset(out, 0, sizeof(*out));
  out->status = 1;
  out->lastPhase = -1;
  const int d = it.f.ndim;
  pagani_ws*& ws = b->ws[w][d];
  if (!ws) {
    ws = pagani_ws_create(d, b->device);
    if (!ws) {
      set_error(b, std::string("workspace creation failed: ") + pagani_create_error());
      return PAGANI_EHIP;
    }
  }
  pagani_ws_set_device_mem(ws, item_total(b));
  pagani_ws_set_mem_model(ws, b->mem_model);
  pagani_ws_set_timing(ws, b->timing ? 1 : 0);
  pagani_ws_set_tuned_slices(ws, b->tuned ? 1 : 0);
  pagani_ws_set_graphs(ws, b->workers == 1 ? 1 : 0);
  const std::size_t n = o.n, np = o.has_parents ? n / 2 : 0;
  const I::SegView v = I::segment_view(base, d, b->seg_size);
  ws->susp.held = false;
  ws->live_n = 0;
  ws->live_evaluated = false;
  ws->snap_n = 0;
  ws->trace.clear();
  ws->cur = 0;
  int rc;
  if ((rc = I::ensure(ws, ws->L[0], d * n, "region bounds")) ||
      (rc = I::ensure(ws, ws->H[0], d * n, "region bounds")) ||
      (np && (rc = I::ensure(ws, ws->pest, np, "parent estimates")))) {
    if (rc != PAGANI_ENOMEM) set_error(b, pagani_ws_last_error(ws));
    return rc;
  }
  if (hipMemcpyAsync(ws->L[0].p, v.L[o.cur], d * n * sizeof(double), hipMemcpyDeviceToDevice,
                     ws->stream) != hipSuccess ||
      hipMemcpyAsync(ws->H[0].p, v.H[o.cur], d * n * sizeof(double), hipMemcpyDeviceToDevice,
                     ws->stream) != hipSuccess ||
      (np && hipMemcpyAsync(ws->pest.p, v.pest, np * sizeof(double), hipMemcpyDeviceToDevice,
                            ws->stream) != hipSuccess)) {
    (void)hipGetLastError();
    set_error(b, "copy of a segment's list failed");
    return PAGANI_EHIP;
  }
  I::LoopState st;
  st.held = true;
  st.it = o.it;
  st.n = n;
  st.has_parents = o.has_parents;
  st.split = 0;
  st.lv = pagani::phases::Levels{};
  st.user_regions = 0;
  st.relerr = it.relerr_classification ? 1 : 0;
  st.epsrel = it.epsrel;
  st.epsabs = it.epsabs;
  for (int j = 0; j < d; ++j) {
    st.lows[j] = 0.0;
    st.highs[j] = 1.0;
  }
  st.cum = o.r;
  st.cum.status = 1;
  st.cum.lastPhase = -1;
  for (int k = 0; k < 3; ++k) st.hist[k] = o.hist[k];
  st.iters_collected = o.iters_collected;
  st.mem_model = b->mem_model;
  st.mem_total = item_total(b);
  st.genz = it.f;
  st.precision = it.precision;
  ws->susp = st;
  ws->live_n = n;
  ws->split = false;
  ws->lv = st.lv;
  rc = pagani_integrate_genz_resume(ws, &it.f, it.precision, out);
  if (rc < 0 && rc != PAGANI_ENOMEM) set_error(b, pagani_ws_last_error(ws));
  return rc;
}

// Grow a device buffer (contents not kept).
int grow(pagani_batch* b, void** p, std::size_t* cap, std::size_t need)
{
  if (need <= *cap) return PAGANI_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, need) != hipSuccess) {
    (void)hipGetLastError();
    set_error(b, "hipMalloc of the segment pool failed");
    return PAGANI_ENOMEM;
  }
  *cap = need;
  return PAGANI_OK;
}

// The items of one persistent-kernel launch: sel[first, first + count).
struct SegGroup {
  int d, fam;
  std::size_t first, count;
  unsigned grid;
  std::size_t bytes, pool_off;
  pagani_eval_consts C;
};

// Streams and events of the persistent kernel's launches, made on first use.
int seg_streams(pagani_batch* b)
{
  if (b->seg_stream[0]) return PAGANI_OK;
  for (int i = 0; i < kSegStreams; ++i)
    if (hipStreamCreateWithFlags(&b->seg_stream[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->seg_event[i], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      set_error(b, "stream creation for the segment kernel failed");
      return PAGANI_EHIP;
    }
  return PAGANI_OK;
}

// The persistent kernel over the eligible items (FP64 Genz descriptors whose
// initial split fits a segment); done[i] is set for the items it finished.
// One launch per (dimension, family), all at once on kSegStreams streams,
// highest dimension first, so that while one launch waits for its slowest
// integral the others' workgroups take the freed compute units; within a
// launch the tightest epsrel goes first.
int run_segments(pagani_batch* b, const pagani_batch_item* items, std::size_t count,
                 pagani_result* results, int* codes, std::vector<char>& done,
                 std::vector<pagani::internal::SegOut>& handoff, std::vector<char*>& hbase)
{
  namespace I = pagani::internal;
  const auto t0 = std::chrono::steady_clock::now();
  const std::size_t S = b->seg_size;
  std::vector<I::SegItem> sel;
  std::vector<std::size_t> idx;
  std::vector<SegGroup> groups;
  std::size_t pool = 0;
  for (int d = PAGANI_GENZ_MAX_DIM; d >= PAGANI_MIN_DIM; --d)
    for (int fam = PAGANI_GENZ_OSCILLATORY; fam <= PAGANI_GENZ_DISCONTINUOUS; ++fam) {
      std::vector<std::size_t> mine;
      for (std::size_t i = 0; i < count; ++i) {
        const pagani_batch_item& it = items[i];
        if (it.f.ndim != d || it.f.family != fam || it.precision != PAGANI_FP64) continue;
        if (!(it.epsrel >= 0.0) || !(it.epsabs >= 0.0)) continue;
        // the corner peak is evaluated point by point (1105 points per 8D
        // region, against per-axis factors for the other families): from 8D
        // one compute unit per integral takes longer than the regular loop
        // on the whole device (measured: 165 ms for the 8D items of the
        // 512-integral batch, 4 ms each on the regular loop)
        if (fam == PAGANI_GENZ_CORNER_PEAK && d >= 8) continue;
        const std::size_t k = it.init_k ? it.init_k : (d < 5 ? 4 : (d <= 10 ? 2 : 1));
        std::size_t n = 1;
        bool fits = true;
        for (int j = 0; j < d && fits; ++j) {
          n *= k;
          fits = n <= S;
        }
        if (fits) mine.push_back(i);
      }
      if (mine.empty()) continue;
      std::stable_sort(mine.begin(), mine.end(), [&](std::size_t x, std::size_t y) {
        return items[x].epsrel < items[y].epsrel;
      });
      // the rule's launch constants at dimension d
      pagani_ws*& ws = b->ws[0][d];
      if (!ws) ws = pagani_ws_create(d, b->device);
      if (!ws) {
        set_error(b, std::string("workspace creation failed: ") + pagani_create_error());
        return PAGANI_EHIP;
      }
      SegGroup g{};
      g.d = d;
      g.fam = fam;
      g.first = sel.size();
      g.count = mine.size();
      if (pagani_ws_eval_consts(ws, nullptr, nullptr, 64, &g.C) != PAGANI_OK) {
        set_error(b, "launch constants failed");
        return PAGANI_EHIP;
      }
      g.grid = I::segment_grid(d, fam,
                               static_cast<unsigned>(std::min<std::size_t>(
                                   mine.size(), 4 * static_cast<std::size_t>(b->num_cus))));
      g.bytes = I::segment_bytes(d, S);
      g.pool_off = pool;
      pool += g.grid * g.bytes;
      groups.push_back(g);
      for (std::size_t i : mine) {
        const pagani_batch_item& it = items[i];
        I::SegItem s{};
        s.f = it.f;
        s.epsrel = it.epsrel;
        s.epsabs = it.epsabs;
        s.init_k = it.init_k ? it.init_k : (d < 5 ? 4 : (d <= 10 ? 2 : 1));
        s.relerr = it.relerr_classification ? 1 : 0;
        sel.push_back(s);
        idx.push_back(i);
      }
    }
  if (groups.empty()) return PAGANI_OK;
  // the pool takes at most a quarter of the device: fewer workgroups per
  // launch when the segments would need more (they then take more items
  // each), and none at all -- every item on the regular loop -- when even
  // that cannot be allocated
  const std::size_t cap = b->hbm / 4;
  if (pool > cap) {
    const std::size_t share = cap / groups.size();
    pool = 0;
    for (SegGroup& g : groups) {
      const std::size_t fit = share / g.bytes;
      g.grid = static_cast<unsigned>(std::max<std::size_t>(1, std::min<std::size_t>(g.grid, fit)));
      g.pool_off = pool;
      pool += g.grid * g.bytes;
    }
  }
  int rc;
  if ((rc = seg_streams(b))) return rc;
  if (grow(b, reinterpret_cast<void**>(&b->seg_pool), &b->seg_pool_bytes, pool) != PAGANI_OK) {
    std::lock_guard<std::mutex> g(b->mu);
    b->error.clear();
    return PAGANI_OK;
  }
  // device: the items and the launches' counters (one copy in), then the
  // outputs (one copy out)
  const std::size_t ib = sel.size() * sizeof(I::SegItem);
  const std::size_t cb = (groups.size() * sizeof(std::uint32_t) + 255) / 256 * 256;
  const std::size_t ob = sel.size() * sizeof(I::SegOut);
  if ((rc = grow(b, &b->seg_dev, &b->seg_dev_bytes, ib + cb + ob))) return rc;
  std::vector<char> in(ib + cb, 0);
  std::memcpy(in.data(), sel.data(), ib);
  char* dev = static_cast<char*>(b->seg_dev);
  auto* ditems = reinterpret_cast<I::SegItem*>(dev);
  auto* dnext = reinterpret_cast<std::uint32_t*>(dev + ib);
  auto* dout = reinterpret_cast<I::SegOut*>(dev + ib + cb);
  std::vector<I::SegOut> out(sel.size());
  hipStream_t* st = b->seg_stream;
  bool ok = hipMemcpyAsync(dev, in.data(), ib + cb, hipMemcpyHostToDevice, st[0]) == hipSuccess &&
            hipEventRecord(b->seg_event[0], st[0]) == hipSuccess;
  for (int k = 1; ok && k < kSegStreams; ++k)
    ok = hipStreamWaitEvent(st[k], b->seg_event[0], 0) == hipSuccess;
  for (std::size_t q = 0; ok && q < groups.size(); ++q) {
    const SegGroup& g = groups[q];
    ok = I::launch_segments(st[q % kSegStreams], g.d, g.fam, ditems + g.first,
                            static_cast<std::uint32_t>(g.count), dnext + q,
                            b->seg_pool + g.pool_off, g.bytes, S, g.grid, g.C, item_total(b),
                            b->mem_model, dout + g.first) == hipSuccess;
  }
  for (int k = 1; ok && k < kSegStreams; ++k)
    ok = hipEventRecord(b->seg_event[k], st[k]) == hipSuccess &&
         hipStreamWaitEvent(st[0], b->seg_event[k], 0) == hipSuccess;
  ok = ok &&
       hipMemcpyAsync(out.data(), dout, ob, hipMemcpyDeviceToHost, st[0]) == hipSuccess &&
       hipStreamSynchronize(st[0]) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    for (int k = 0; k < kSegStreams; ++k) (void)hipStreamSynchronize(st[k]);
    set_error(b, "segment kernel failed");
    return PAGANI_EHIP;
  }
  std::size_t gq = 0;
  for (std::size_t q = 0; q < sel.size(); ++q) {
    while (q >= groups[gq].first + groups[gq].count) ++gq;
    const I::SegOut& o = out[q];
    b->seg_evaluated += o.evaluated;
    b->seg_iterations += o.iterations;
    if (o.outcome != 0) {
      if (o.n > 0) {
        const SegGroup& g = groups[gq];
        handoff[idx[q]] = o;
        hbase[idx[q]] = b->seg_pool + g.pool_off + static_cast<std::size_t>(o.block) * g.bytes;
      }
      continue;
    }
    results[idx[q]] = o.r;
    codes[idx[q]] = PAGANI_OK;
    done[idx[q]] = 1;
    b->seg_integrals++;
  }
  b->seg_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  return PAGANI_OK;
}

}  // namespace

extern "C" {

pagani_batch* pagani_batch_create(int device, int workers)
{
  g_batch_create_error.clear();
  if (workers < 1 || workers > 16) {
    g_batch_create_error = "workers outside [1, 16]";
    return nullptr;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    g_batch_create_error = std::string("hipGetDeviceCount: ") + hipGetErrorString(e);
    return nullptr;
  }
  if (device < 0 && hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    g_batch_create_error = "hipGetDevice failed";
    return nullptr;
  }
  if (device >= ndev) {
    g_batch_create_error = "device index out of range";
    return nullptr;
  }
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) {
    (void)hipGetLastError();
    g_batch_create_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
    return nullptr;
  }
  pagani_batch* b = new pagani_batch();
  b->device = device;
  b->workers = workers;
  b->hbm = prop.totalGlobalMem;
  b->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  std::array<pagani_ws*, PAGANI_MAX_DIM + 1> none;
  none.fill(nullptr);
  b->ws.assign(static_cast<std::size_t>(workers), none);
  return b;
}

void pagani_batch_destroy(pagani_batch* b)
{
  if (!b) return;
  for (auto& row : b->ws)
    for (pagani_ws* w : row)
      if (w) pagani_ws_destroy(w);
  if (b->seg_pool) (void)hipFree(b->seg_pool);
  if (b->seg_dev) (void)hipFree(b->seg_dev);
  for (int k = 0; k < kSegStreams; ++k) {
    if (b->seg_event[k]) (void)hipEventDestroy(b->seg_event[k]);
    if (b->seg_stream[k]) (void)hipStreamDestroy(b->seg_stream[k]);
  }
  delete b;
}

const char* pagani_batch_last_error(const pagani_batch* b)
{
  return b ? b->error.c_str() : (g_batch_create_error.empty() ? "null batch"
                                                              : g_batch_create_error.c_str());
}

int pagani_batch_set_device_mem(pagani_batch* b, size_t total_bytes)
{
  if (!b) return PAGANI_EBADARG;
  b->mem_total = total_bytes;
  return PAGANI_OK;
}

int pagani_batch_set_mem_model(pagani_batch* b, int model)
{
  if (!b) return PAGANI_EBADARG;
  if (model != PAGANI_MEM_REFERENCE && model != PAGANI_MEM_ACTUAL) {
    b->error = "unknown memory model";
    return PAGANI_EBADARG;
  }
  b->mem_model = model;
  return PAGANI_OK;
}

int pagani_batch_set_timing(pagani_batch* b, int enable)
{
  if (!b) return PAGANI_EBADARG;
  b->timing = enable != 0;
  return PAGANI_OK;
}

int pagani_batch_set_segments(pagani_batch* b, size_t regions)
{
  if (!b) return PAGANI_EBADARG;
  if (regions != 0 && (regions < 256 || regions > (std::size_t(1) << 22))) {
    b->error = "segment size outside [256, 2^22] regions";
    return PAGANI_EBADARG;
  }
  b->seg_size = regions;
  return PAGANI_OK;
}

int pagani_batch_set_tuned_slices(pagani_batch* b, int enable)
{
  if (!b) return PAGANI_EBADARG;
  b->tuned = enable != 0;
  return PAGANI_OK;
}

int pagani_batch_get_stats(const pagani_batch* b, pagani_stats* out)
{
  if (!b || !out) return PAGANI_EBADARG;
  std::memset(out, 0, sizeof(*out));
  for (const auto& row : b->ws)
    for (const pagani_ws* w : row) {
      if (!w) continue;
      pagani_stats s;
      if (pagani_ws_get_stats(w, &s) != PAGANI_OK) continue;
      out->eval_ms += s.eval_ms;
      out->refine_ms += s.refine_ms;
      out->finalize_ms += s.finalize_ms;
      out->compact_ms += s.compact_ms;
      out->classify_ms += s.classify_ms;
      out->wall_ms += s.wall_ms;
      out->regions_evaluated += s.regions_evaluated;
      out->eval_launches += s.eval_launches;
      out->iterations += s.iterations;
      out->classify_passes += s.classify_passes;
      out->max_regions = std::max(out->max_regions, s.max_regions);
      out->integrals += s.integrals;
      out->alloc_ms += s.alloc_ms;
      out->allocations += s.allocations;
      out->device_ms += s.device_ms;
      out->device_iterations += s.device_iterations;
      out->device_regions += s.device_regions;
      out->device_rounds += s.device_rounds;
      out->budget_fixes += s.budget_fixes;
      out->graph_launches += s.graph_launches;
    }
  out->segment_integrals = b->seg_integrals;
  out->segment_ms = b->seg_ms;
  out->integrals += b->seg_integrals;
  out->regions_evaluated += b->seg_evaluated;
  out->iterations += b->seg_iterations;
  return PAGANI_OK;
}

int pagani_batch_reset_stats(pagani_batch* b)
{
  if (!b) return PAGANI_EBADARG;
  for (auto& row : b->ws)
    for (pagani_ws* w : row)
      if (w) pagani_ws_reset_stats(w);
  b->seg_integrals = 0;
  b->seg_evaluated = 0;
  b->seg_iterations = 0;
  b->seg_ms = 0;
  return PAGANI_OK;
}

int pagani_integrate_batch(pagani_batch* b, const pagani_batch_item* items, size_t count,
                           pagani_result* results, int* codes)
{
  if (!b) return PAGANI_EBADARG;
  if (count == 0) return PAGANI_OK;
  if (!items || !results || !codes) {
    b->error = "null argument";
    return PAGANI_EBADARG;
  }
  b->error.clear();
  std::vector<char> done(count, 0);
  // items the persistent kernel handed over (hbase[i] != NULL)
  std::vector<pagani::internal::SegOut> handoff(b->seg_size ? count : 0);
  std::vector<char*> hbase(b->seg_size ? count : 0, nullptr);
  if (b->seg_size) {
    if (hipSetDevice(b->device) != hipSuccess) {
      (void)hipGetLastError();
      b->error = "hipSetDevice failed";
      return PAGANI_EHIP;
    }
    const int rc = run_segments(b, items, count, results, codes, done, handoff, hbase);
    if (rc) return rc;
  }
  // the rest, largest first: higher dimension, then tighter epsrel, then
  // input order
  std::vector<std::size_t> order;
  for (std::size_t i = 0; i < count; ++i)
    if (!done[i]) order.push_back(i);
  const std::size_t nrest = order.size();
  std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) {
    if (items[x].f.ndim != items[y].f.ndim) return items[x].f.ndim > items[y].f.ndim;
    return items[x].epsrel < items[y].epsrel;
  });
  std::atomic<std::size_t> next{0};
  auto work = [&](int w) {
    if (hipSetDevice(b->device) != hipSuccess) {
      (void)hipGetLastError();
      set_error(b, "hipSetDevice failed in a batch worker");
      for (std::size_t k; (k = next.fetch_add(1)) < nrest;) {
        const std::size_t i = order[k];
        std::memset(&results[i], 0, sizeof(results[i]));
        results[i].status = 1;
        codes[i] = PAGANI_EHIP;
      }
      return;
    }
    // One pool per worker: when a worker moves to another dimension (items
    // come in decreasing dimension), its other workspaces hand their
    // buffers to that dimension's workspace and release the rest, so idle
    // pools never crowd out the running one and the pool grows once per
    // worker instead of once per dimension.
    int last_d = 0;
    for (std::size_t k; (k = next.fetch_add(1)) < nrest;) {
      const std::size_t i = order[k];
      const int d = items[i].f.ndim;
      if (d != last_d && d >= PAGANI_MIN_DIM && d <= PAGANI_MAX_DIM) {
        pagani_ws*& to = b->ws[w][d];
        if (!to) to = pagani_ws_create(d, b->device);  // failure: run_item reports it
        for (int x = PAGANI_MIN_DIM; x <= PAGANI_MAX_DIM; ++x) {
          if (x == d || !b->ws[w][x]) continue;
          if (to) (void)pagani::internal::pool_handover(b->ws[w][x], to);
          (void)pagani_ws_trim(b->ws[w][x]);
        }
      }
      last_d = d;
      codes[i] = !hbase.empty() && hbase[i]
                     ? run_item_resume(b, w, items[i], handoff[i], hbase[i], &results[i])
                     : run_item(b, w, items[i], &results[i]);
    }
  };
  const int nthreads = static_cast<int>(std::min<std::size_t>(b->workers, nrest));
  if (nthreads == 0) {
  } else if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    pool.reserve(nthreads);
    for (int w = 0; w < nthreads; ++w) pool.emplace_back(work, w);
    for (auto& t : pool) t.join();
  }
  for (std::size_t i = 0; i < count; ++i)
    if (codes[i] < 0) return codes[i];
  return PAGANI_OK;
}

}  // extern "C"
