// workspace.hpp -- the reference's C++ integrator surface on the MI355X
// library (drop-in for cuda/pagani/quad/GPUquad/Workspace.cuh:23-71).
//
//   Workspace<double, ndim> pagani;
//   numint::integration_result r = pagani.integrate(f, epsrel, epsabs, vol);
//
// IntegT is any copyable class with `__device__ double operator()(double...)`
// taking ndim arguments (const or not), as in the reference.  The region
// evaluation kernel is instantiated for IntegT in the including translation
// unit (compile it with hipcc); everything else runs in libpagani.so behind
// the C ABI of pagani.h.  The integrand is passed to the kernel by value on
// every launch (the reference copies it once into managed memory,
// cudaMemoryUtil.h:181-196, 345-353); state it points to must be device
// memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "pagani.h"
#include "pagani/eval_kernel.hpp"
#include "pagani/integration_result.hh"
#include "pagani/volume.hh"

namespace pagani {

inline void throw_on(int rc, pagani_ws* ws, const char* what)
{
  if (rc < 0)
    throw std::runtime_error(std::string(what) + ": " +
                             (ws ? pagani_ws_last_error(ws) : "no workspace"));
}

inline numint::integration_result to_result(const pagani_result& r)
{
  numint::integration_result o;
  o.estimate = r.estimate;
  o.errorest = r.errorest;
  o.neval = r.neval;
  o.nregions = r.nregions;
  o.nFinishedRegions = r.nFinishedRegions;
  o.status = r.status;
  o.lastPhase = r.lastPhase;
  o.chi_sq = r.chi_sq;
  o.iters = r.iters;
  return o;
}

template <int D, class IntegT>
int eval_callback(void* ctx, const pagani_eval_args* args)
{
  return eval_regions<D, IntegT, double>(*static_cast<const IntegT*>(ctx), args);
}

}  // namespace pagani

// Sub_regions<T, ndim> (Sub_regions.cuh:38-258): a device SoA region list,
// dLeftCoord[j*size + r] and dLength[j*size + r] in unit-cube coordinates.
template <typename T, std::size_t ndim>
struct Sub_regions {
  static_assert(std::is_same<T, double>::value, "regions are FP64");

  Sub_regions() = default;
  explicit Sub_regions(std::size_t partitions_per_axis) { uniform_split(partitions_per_axis); }
  Sub_regions(const Sub_regions&) = delete;
  Sub_regions& operator=(const Sub_regions&) = delete;
  ~Sub_regions()
  {
    free_device();
    free_snapshot();
  }

  // k boxes per axis (create_uniform_split, Sub_regions.cuh:11-36)
  void uniform_split(std::size_t k)
  {
    std::size_t n = 1;
    for (std::size_t j = 0; j < ndim; ++j) n *= k;
    std::vector<T> L(n * ndim), H(n * ndim);
    const T len = 1. / static_cast<T>(k);
    for (std::size_t t = 0; t < n; ++t) {
      std::size_t q = t;
      for (std::size_t j = 0; j < ndim; ++j) {
        L[j * n + t] = static_cast<T>(q % k) * len;
        H[j * n + t] = len;
        q /= k;
      }
    }
    assign_host(L.data(), H.data(), n);
  }

  // replace the list with n regions given on the host
  void assign_host(const T* L, const T* H, std::size_t n)
  {
    device_init(n);
    if (n) {
      copy(dLeftCoord, L, n * ndim, hipMemcpyHostToDevice);
      copy(dLength, H, n * ndim, hipMemcpyHostToDevice);
    }
  }

  // replace the list with n regions given on the device
  void assign_device(const T* L, const T* H, std::size_t n)
  {
    device_init(n);
    if (n) {
      copy(dLeftCoord, L, n * ndim, hipMemcpyDeviceToDevice);
      copy(dLength, H, n * ndim, hipMemcpyDeviceToDevice);
    }
  }

  void device_init(std::size_t n)
  {
    free_device();
    size = n;
    if (n == 0) return;
    if (hipMalloc(&dLeftCoord, n * ndim * sizeof(T)) != hipSuccess ||
        hipMalloc(&dLength, n * ndim * sizeof(T)) != hipSuccess)
      throw std::bad_alloc();
  }

  // host copies of the bounds (refresh_host_device, Sub_regions.cuh:82-87)
  void to_host(std::vector<T>& L, std::vector<T>& H) const
  {
    L.resize(size * ndim);
    H.resize(size * ndim);
    if (size) {
      copy(L.data(), dLeftCoord, size * ndim, hipMemcpyDeviceToHost);
      copy(H.data(), dLength, size * ndim, hipMemcpyDeviceToHost);
    }
  }

  // take_snapshot / load_snapshot (Sub_regions.cuh:220-239): one device
  // copy of the list, swapped back in by load_snapshot
  void take_snapshot() { set_snapshot(dLeftCoord, dLength, size); }

  void set_snapshot(const T* L, const T* H, std::size_t n)
  {
    free_snapshot();
    snapshot_size = n;
    if (n == 0) return;
    if (hipMalloc(&snapshot_dLeftCoord, n * ndim * sizeof(T)) != hipSuccess ||
        hipMalloc(&snapshot_dLength, n * ndim * sizeof(T)) != hipSuccess) {
      free_snapshot();
      throw std::bad_alloc();
    }
    copy(snapshot_dLeftCoord, L, n * ndim, hipMemcpyDeviceToDevice);
    copy(snapshot_dLength, H, n * ndim, hipMemcpyDeviceToDevice);
  }

  void load_snapshot()
  {
    free_device();
    dLeftCoord = snapshot_dLeftCoord;
    dLength = snapshot_dLength;
    size = snapshot_size;
    snapshot_dLeftCoord = snapshot_dLength = nullptr;
    snapshot_size = 0;
  }

  T* dLeftCoord = nullptr;
  T* dLength = nullptr;
  std::size_t size = 0;
  T* snapshot_dLeftCoord = nullptr;
  T* snapshot_dLength = nullptr;
  std::size_t snapshot_size = 0;

 private:
  void free_snapshot()
  {
    if (snapshot_dLeftCoord) (void)hipFree(snapshot_dLeftCoord);
    if (snapshot_dLength) (void)hipFree(snapshot_dLength);
    snapshot_dLeftCoord = snapshot_dLength = nullptr;
    snapshot_size = 0;
  }

  static void copy(T* dst, const T* src, std::size_t count, hipMemcpyKind kind)
  {
    if (hipMemcpy(dst, src, count * sizeof(T), kind) != hipSuccess)
      throw std::runtime_error("Sub_regions: hipMemcpy failed");
  }
  void free_device()
  {
    if (dLeftCoord) (void)hipFree(dLeftCoord);
    if (dLength) (void)hipFree(dLength);
    dLeftCoord = dLength = nullptr;
    size = 0;
  }
};

// Workspace (Workspace.cuh:23-71).  debug > 0 turns on the reference's
// debug recorders for this workspace's integrate calls (CSV files in the
// working directory, pagani_ws_set_debug; collect_mult_runs appends to
// them); use_custom is accepted for source compatibility -- the library
// always uses its own deterministic reductions.
template <typename T, std::size_t ndim, int debug = 0, bool use_custom = false,
          bool collect_mult_runs = false>
class Workspace {
  static_assert(std::is_same<T, double>::value, "the integrator surface is FP64");
  static_assert(ndim >= PAGANI_MIN_DIM && ndim <= PAGANI_MAX_DIM, "unsupported ndim");

 public:
  Workspace() : ws_(pagani_ws_create(static_cast<int>(ndim), -1))
  {
    if (!ws_) throw std::runtime_error("pagani_ws_create failed (no HIP device?)");
  }
  ~Workspace() { pagani_ws_destroy(ws_); }
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;

  template <typename IntegT, bool predict_split = false, bool collect_iters = false>
  numint::integration_result integrate(const IntegT& integrand, T epsrel, T epsabs,
                                       quad::Volume<T, static_cast<int>(ndim)> const& vol,
                                       bool relerr_classification = true)
  {
    IntegT f = integrand;
    pagani_result r;
    if constexpr (debug > 0)
      pagani_ws_set_debug(ws_, debug < 3 ? debug : 3, ".", nullptr, collect_mult_runs ? 1 : 0);
    // this overload has no caller-owned list to keep a predict_split snapshot in
    pagani_ws_set_snapshot_iter(ws_, ~std::uint64_t(0));
    pagani_ws_set_predict_split(ws_, predict_split ? 1 : 0);
    const int rc = pagani_integrate_cb(ws_, &pagani::eval_callback<static_cast<int>(ndim), IntegT>,
                                       &f, epsrel, epsabs, vol.lows, vol.highs,
                                       relerr_classification ? 1 : 0, nullptr, 0, &r);
    pagani::throw_on(rc == PAGANI_ENOMEM ? 0 : rc, ws_, "Workspace::integrate");
    return pagani::to_result(r);
  }

  // Sub_regions& overload (Workspace.cuh:148-358): starts from the caller's
  // list and leaves the final list in it.
  template <typename IntegT, bool predict_split = false, bool collect_iters = false>
  numint::integration_result integrate(const IntegT& integrand, Sub_regions<T, ndim>& subregions,
                                       T epsrel, T epsabs,
                                       quad::Volume<T, static_cast<int>(ndim)> const& vol,
                                       bool relerr_classification = true,
                                       const std::string& optional = "default")
  {
    IntegT f = integrand;
    pagani_result r;
    if constexpr (debug > 0)
      pagani_ws_set_debug(ws_, debug < 3 ? debug : 3, ".", optional.c_str(),
                          collect_mult_runs ? 1 : 0);
    else
      (void)optional;
    pagani_region_list init{subregions.size, subregions.dLeftCoord, subregions.dLength};
    // predict_split: the list of iteration 15 if no region finished before
    // (Workspace.cuh:244-248)
    pagani_ws_set_snapshot_iter(ws_, predict_split ? 15 : ~std::uint64_t(0));
    pagani_ws_set_predict_split(ws_, predict_split ? 1 : 0);
    const int rc = pagani_integrate_cb(ws_, &pagani::eval_callback<static_cast<int>(ndim), IntegT>,
                                       &f, epsrel, epsabs, vol.lows, vol.highs,
                                       relerr_classification ? 1 : 0, &init, 0, &r);
    pagani::throw_on(rc == PAGANI_ENOMEM ? 0 : rc, ws_, "Workspace::integrate");
    std::size_t n = 0;
    const double *L = nullptr, *H = nullptr;
    pagani::throw_on(pagani_ws_live_regions(ws_, &n, &L, &H), ws_, "live regions");
    subregions.assign_device(L, H, n);
    if (predict_split) {
      pagani::throw_on(pagani_ws_snapshot_regions(ws_, &n, &L, &H), ws_, "snapshot");
      if (n) subregions.set_snapshot(L, H, n);
    }
    return pagani::to_result(r);
  }

  pagani_ws* handle() { return ws_; }

 private:
  pagani_ws* ws_ = nullptr;
};
