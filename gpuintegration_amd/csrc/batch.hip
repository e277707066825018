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

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "pagani.h"
#include "ws_internal.hpp"

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
  return PAGANI_OK;
}

int pagani_batch_reset_stats(pagani_batch* b)
{
  if (!b) return PAGANI_EBADARG;
  for (auto& row : b->ws)
    for (pagani_ws* w : row)
      if (w) pagani_ws_reset_stats(w);
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
  // largest first: higher dimension, then tighter epsrel, then input order
  std::vector<std::size_t> order(count);
  std::iota(order.begin(), order.end(), std::size_t(0));
  std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) {
    if (items[x].f.ndim != items[y].f.ndim) return items[x].f.ndim > items[y].f.ndim;
    return items[x].epsrel < items[y].epsrel;
  });
  std::atomic<std::size_t> next{0};
  auto work = [&](int w) {
    if (hipSetDevice(b->device) != hipSuccess) {
      (void)hipGetLastError();
      set_error(b, "hipSetDevice failed in a batch worker");
      for (std::size_t k; (k = next.fetch_add(1)) < count;) {
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
    for (std::size_t k; (k = next.fetch_add(1)) < count;) {
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
      codes[i] = run_item(b, w, items[i], &results[i]);
    }
  };
  const int nthreads = static_cast<int>(std::min<std::size_t>(b->workers, count));
  if (nthreads == 1) {
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
