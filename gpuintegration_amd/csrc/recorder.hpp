// recorder.hpp -- the reference's debug recorders for one integrate call
// (pagani_ws_set_debug).
//
// Workspace<T, ndim, debug, use_custom, collect_mult_runs> writes, when
// debug > 0, a per-iteration CSV and a per-phase time breakdown
// (Workspace.cuh:171-179, 199-352), and its Cubature_rules writes every
// region's estimate and error (debug >= 2) and the generator table and every
// rule point's value (debug > 2, lists under 1024 regions) (PaganiUtils.cuh:
// 50-241, Func_Eval.cuh).  This class keeps the same files, columns and
// phase names.  Phase times are host milliseconds between stream
// synchronisations at the phase boundaries, as the reference's host timers
// measure around its synchronous phases.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "pagani.h"
#include "phases.hpp"
#include "rule.hpp"

namespace pagani {
namespace internal {

struct DebugConfig {
  int level = 0;
  bool append = false;
  // the recorders of integrate(f, epsrel, epsabs, vol) (Workspace.cuh:
  // 377-431: cuda_iters.csv only) instead of those of the Sub_regions&
  // overload (:171-352, with the finished totals and the time breakdown)
  bool plain = false;
  std::string dir = ".";
  std::string id = "default";
};

class DebugRecorder {
 public:
  DebugRecorder() = default;
  DebugRecorder(const DebugRecorder&) = delete;
  DebugRecorder& operator=(const DebugRecorder&) = delete;
  ~DebugRecorder() { close(); }

  // Opens the files of `cfg.level` (nothing at level 0).  Returns false with
  // *why set when a file cannot be opened.
  bool open(const DebugConfig& cfg, hipStream_t stream, int ndim, double epsrel,
            const RuleTables& rule, const double* lows, const double* highs, std::string* why)
  {
    close();
    level_ = cfg.level;
    if (level_ <= 0) return true;
    stream_ = stream;
    ndim_ = ndim;
    epsrel_ = epsrel;
    id_ = cfg.id;
    rule_ = &rule;
    lows_.assign(lows, lows + ndim);
    highs_.assign(highs, highs + ndim);
    const char* mode = cfg.append ? "a" : "w";
    plain_ = cfg.plain;
    auto path = [&](const char* name) { return cfg.dir + "/" + name; };
    auto fopen_ = [&](const char* name, std::FILE*& f) {
      f = std::fopen(path(name).c_str(), mode);
      if (!f) *why = "cannot open " + path(name);
      return f != nullptr;
    };
    if (plain_) {
      if (!fopen_("cuda_iters.csv", iters_)) return false;
      if (!cfg.append) std::fprintf(iters_, "it, estimate, errorest, nregions\n");
    } else {
      if (!fopen_("cuda_pagani_iters.csv", iters_) ||
          !fopen_("cuda_pagani_time_breakdown.csv", times_))
        return false;
      if (!cfg.append) {
        std::fprintf(iters_,
                     "it, estimate, errorest, festimate, ferrorest, nregions, nfregions\n");
        std::fprintf(times_, "id, ndim, epsrel, it, name, time\n");
      }
    }
    if (level_ >= 2 && !fopen_("cuda_regions.csv", regions_)) return false;
    if (level_ >= 3) {
      if (!fopen_("cuda_fevals.csv", fevals_) || !fopen_("cuda_generators.csv", gens_))
        return false;
      if (!cfg.append) {
        std::fprintf(fevals_, "reg, fid,");
        for (int j = 0; j < ndim; ++j) std::fprintf(fevals_, "dim%dlow,dim%dhigh,", j, j);
        for (int j = 0; j < ndim; ++j) std::fprintf(fevals_, "gdim%dlow,gdim%dhigh,", j, j);
        for (int j = 0; j < ndim; ++j) std::fprintf(fevals_, "dim%d,", j);
        std::fprintf(fevals_, "feval, estimate, errorest\n");
      }
      // the generator table, i = feval * axis + point (PaganiUtils.cuh:187-200)
      std::fprintf(gens_, "i, gen\n");
      for (int j = 0; j < ndim; ++j)
        for (int p = 0; p < rule.feval; ++p)
          std::fprintf(gens_, "%d,%.17g\n", rule.feval * j + p, point_generator(rule, p, j));
    }
    mark();
    return true;
  }

  void close()
  {
    for (std::FILE** f : {&iters_, &times_, &regions_, &fevals_, &gens_})
      if (*f) {
        std::fclose(*f);
        *f = nullptr;
      }
    level_ = 0;
  }

  int level() const { return level_; }
  bool points(std::size_t n) const { return level_ >= 3 && n < 1024; }

  // Start timing a phase.
  void mark()
  {
    (void)hipStreamSynchronize(stream_);
    t0_ = std::chrono::steady_clock::now();
  }

  // The phase that started at the last mark ends here.
  void phase(std::uint64_t it, const char* name)
  {
    if (level_ <= 0 || !times_) return;
    (void)hipStreamSynchronize(stream_);
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(times_, "%s,%d,%.17g,%llu,%s,%.6f\n", id_.c_str(), ndim_, epsrel_,
                 static_cast<unsigned long long>(it), name,
                 std::chrono::duration<double, std::milli>(t1 - t0_).count());
    t0_ = t1;
  }

  // A phase this library does inside another one's kernel.
  void folded(std::uint64_t it, const char* name)
  {
    if (level_ <= 0 || !times_) return;
    std::fprintf(times_, "%s,%d,%.17g,%llu,%s,%.6f\n", id_.c_str(), ndim_, epsrel_,
                 static_cast<unsigned long long>(it), name, 0.0);
  }

  // Workspace.cuh:225-231: the totals with and without this iteration's
  // regions, the list size and the regions finished before (:428-431 for the
  // plain overload: the totals and the list size).
  void iteration(std::uint64_t it, double est, double err, double fest, double ferr,
                 std::size_t n, std::uint64_t nf)
  {
    if (level_ <= 0) return;
    if (plain_)
      std::fprintf(iters_, "%llu,%.17g,%.17g,%zu\n", static_cast<unsigned long long>(it), est,
                   err, n);
    else
      std::fprintf(iters_, "%llu,%.17g,%.17g,%.17g,%.17g,%zu,%llu\n",
                   static_cast<unsigned long long>(it), est, err, fest, ferr, n,
                   static_cast<unsigned long long>(nf));
  }

  // Print_region_evals (PaganiUtils.cuh:175-184): the rule's estimate and
  // error of every region, before the two-level refinement.
  hipError_t regions(std::uint64_t it, std::size_t n, const double* est, const double* err)
  {
    if (level_ < 2) return hipSuccess;
    hipError_t e = copy(n, est, err);
    if (e != hipSuccess) return e;
    for (std::size_t t = 0; t < n; ++t)
      std::fprintf(regions_, "%llu,%zu,%.17g,%.17g\n", static_cast<unsigned long long>(it), t,
                   h_est_[t], h_err_[t]);
    return hipSuccess;
  }

  // Print_func_evals (PaganiUtils.cuh:130-173): every rule point of every
  // region -- the region's bounds in the unit cube, the integration box, the
  // point, its value, the region's estimate and error.  L, H, pdim, lv: the
  // list (split form when pdim is not NULL, widths as levels when lv.on),
  // fev: the point values (pagani_eval_args::fevals).
  hipError_t points(std::size_t n, const double* L, const double* H, const std::uint8_t* pdim,
                    const phases::Levels& lv, const double* est, const double* err,
                    const double* fev)
  {
    if (!points(n)) return hipSuccess;
    const int F = rule_->feval;
    const std::size_t rows = pdim ? n / 2 : n;
    std::vector<double> hL(rows * ndim_), hH(rows * ndim_), hF(n * F);
    std::vector<std::uint8_t> hP(pdim ? rows : 0);
    hipError_t e = copy(n, est, err);
    if (e == hipSuccess)
      e = hipMemcpy(hL.data(), L, hL.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && lv.on) {
      std::vector<std::uint16_t> lev(hH.size());
      e = hipMemcpy(lev.data(), H, lev.size() * sizeof(std::uint16_t), hipMemcpyDeviceToHost);
      for (std::size_t i = 0; i < lev.size(); ++i)
        hH[i] = std::ldexp(lv.h0[i / rows], -static_cast<int>(lev[i]));
    } else if (e == hipSuccess) {
      e = hipMemcpy(hH.data(), H, hH.size() * sizeof(double), hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && pdim) e = hipMemcpy(hP.data(), pdim, rows, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(hF.data(), fev, hF.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    std::vector<double> lo(ndim_), len(ndim_);
    for (std::size_t t = 0; t < n; ++t) {
      const std::size_t row = pdim ? t % rows : t;
      for (int j = 0; j < ndim_; ++j) {
        lo[j] = hL[j * rows + row];
        len[j] = hH[j * rows + row];
        if (pdim && hP[row] == j) {  // the child's half along the parent's axis
          const double half = len[j] / 2;
          if (t >= rows) lo[j] = lo[j] + half;
          len[j] = half;
        }
      }
      for (int p = 0; p < F; ++p) {
        std::fprintf(fevals_, "%zu,%d,", t, p);
        for (int j = 0; j < ndim_; ++j)
          std::fprintf(fevals_, "%.17g,%.17g,", lo[j], lo[j] + len[j]);
        for (int j = 0; j < ndim_; ++j) std::fprintf(fevals_, "%.17g,%.17g,", lows_[j], highs_[j]);
        for (int j = 0; j < ndim_; ++j) {
          // the point in user coordinates, formed as the kernel forms it
          const double up = lo[j] + len[j];
          const double range = highs_[j] - lows_[j];
          const double c = std::fma(0.5 * (lo[j] + up), range, lows_[j]);
          const double w = (up - lo[j]) * range;
          std::fprintf(fevals_, "%.17g,", std::fma(-point_generator(*rule_, p, j), w, c));
        }
        std::fprintf(fevals_, "%.17g,%.17g,%.17g\n", hF[t * F + p], h_est_[t], h_err_[t]);
      }
    }
    return hipSuccess;
  }

 private:
  hipError_t copy(std::size_t n, const double* est, const double* err)
  {
    hipError_t e0 = hipStreamSynchronize(stream_);
    if (e0 != hipSuccess) return e0;
    h_est_.resize(n);
    h_err_.resize(n);
    hipError_t e = hipMemcpy(h_est_.data(), est, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(h_err_.data(), err, n * sizeof(double), hipMemcpyDeviceToHost);
    return e;
  }

  int level_ = 0;
  bool plain_ = false;
  hipStream_t stream_ = nullptr;
  int ndim_ = 0;
  double epsrel_ = 0;
  std::string id_;
  const RuleTables* rule_ = nullptr;
  std::vector<double> lows_, highs_, h_est_, h_err_;
  std::FILE* iters_ = nullptr;
  std::FILE* times_ = nullptr;
  std::FILE* regions_ = nullptr;
  std::FILE* fevals_ = nullptr;
  std::FILE* gens_ = nullptr;
  std::chrono::steady_clock::time_point t0_;
};

}  // namespace internal
}  // namespace pagani
