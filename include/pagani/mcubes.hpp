// mcubes.hpp -- the reference's m-Cubes C++ surface on the MI355X library
// (drop-in for cuda/mcubes/vegasT.cuh:1003-1112, vegasT1D.cuh:735-803 and
// cuda/mcubes/mcubes.cuh:9-41).
//
//   quad::Volume<double, 5> vol;
//   auto r = cuda_mcubes::integrate<F, 5, false, Custom_generator>(f, epsrel, epsabs,
//                                                                  ncall, &vol, 15, 10, 5);
//
// IntegT is any copyable class with `__device__ double operator()(double...)`
// taking NDIM arguments (const or not), as in the reference.  The sampling
// kernel is instantiated for IntegT in the including translation unit (compile
// it with hipcc); the driver runs in libpagani.so behind pagani_mc_* of
// pagani.h.  The integrand is passed to the kernel by value on every launch.
//
// GeneratorType: Custom_generator is the reference's per-cube LCG (same
// numbers in the same order); Curand_generator, the reference's default,
// selects this library's default streams (xoroshiro128++ per iteration and
// cube, seeded from a fixed seed) in place of curand's clock-seeded XORWOW,
// so runs are reproducible.  DEBUG_MCUBES is accepted and ignored (the
// reference writes sample dumps to files with it).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "pagani.h"
#include "pagani/integration_result.hh"
#include "pagani/mcubes_kernel.hpp"
#include "pagani/volume.hh"

// Generator tags with the reference's names (vegas_utils.cuh:24-70).
class Custom_generator {
 public:
  using stream = pagani::mc::LcgStream;
  static constexpr int kind = PAGANI_MC_RNG_LCG;
};

class Curand_generator {
 public:
  using stream = pagani::mc::XoroStream;
  static constexpr int kind = PAGANI_MC_RNG_DEFAULT;
};

namespace pagani {
namespace mc {

// RAII owner of a pagani_mc handle on the current device.
class Handle {
 public:
  explicit Handle(int ndim) : h_(pagani_mc_create(ndim, -1))
  {
    if (!h_)
      throw std::runtime_error(std::string("pagani_mc_create: ") + pagani_mc_last_error(nullptr));
  }
  ~Handle() { pagani_mc_destroy(h_); }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  pagani_mc* get() const { return h_; }

 private:
  pagani_mc* h_;
};

template <class IntegT, int NDIM, class GeneratorType>
numint::integration_result run(IntegT const& integrand, double epsrel, double epsabs,
                               double ncall, quad::Volume<double, NDIM> const* volume,
                               int total_iters, int adjust_iters, int skip_iters, bool one_axis,
                               bool adjust_params)
{
  static_assert(NDIM >= 1 && NDIM <= PAGANI_MC_MAX_DIM, "m-Cubes supports 1..16 dimensions");
  Handle h(NDIM);
  pagani_mc_params p{};
  p.ncall = ncall;
  p.total_iters = total_iters;
  p.adjust_iters = adjust_iters;
  p.skip_iters = skip_iters;
  p.generator = GeneratorType::kind;
  p.one_axis = one_axis ? 1 : 0;
  p.adjust_params = adjust_params ? 1 : 0;
  p.seed = 0;
  quad::Volume<double, NDIM> unit;
  const quad::Volume<double, NDIM>* v = volume ? volume : &unit;
  pagani_mc_result r{};
  IntegT f = integrand;
  const int rc = pagani_mc_integrate_cb(
      h.get(), &sample_callback<NDIM, IntegT, typename GeneratorType::stream>, &f, epsrel,
      epsabs, v->lows, v->highs, &p, &r);
  if (rc < 0) throw std::runtime_error(std::string("m-Cubes: ") + pagani_mc_last_error(h.get()));
  numint::integration_result o;
  o.estimate = r.estimate;
  o.errorest = r.errorest;
  o.chi_sq = r.chi_sq;
  o.status = r.status;
  o.iters = r.iters;
  o.neval = r.neval;
  return o;
}

// Number of arguments of F's (non-template) call operator, the role of
// cubacpp::arity in the reference (cubacpp/arity.hh:36-41); read from the
// member function's type, so __device__-only operators count too.
template <class T>
struct call_arity;
template <class C, class R, class... A>
struct call_arity<R (C::*)(A...)> {
  static constexpr int value = sizeof...(A);
};
template <class C, class R, class... A>
struct call_arity<R (C::*)(A...) const> {
  static constexpr int value = sizeof...(A);
};

template <class F>
constexpr int arity()
{
  return call_arity<decltype(&F::operator())>::value;
}

}  // namespace mc
}  // namespace pagani

namespace cuda_mcubes {

// vegas() once (vegasT.cuh:1003-1026).
template <typename IntegT, int NDIM, bool DEBUG_MCUBES = false,
          typename GeneratorType = Curand_generator>
numint::integration_result integrate(IntegT const& ig, double epsrel, double epsabs,
                                     double ncall, quad::Volume<double, NDIM> const* volume,
                                     int totalIters = 15, int adjustIters = 15,
                                     int skipIters = 5)
{
  return pagani::mc::run<IntegT, NDIM, GeneratorType>(ig, epsrel, epsabs, ncall, volume,
                                                      totalIters, adjustIters, skipIters,
                                                      false, false);
}

// vegas() rerun with AdjustParams until converged or out of budget
// (vegasT.cuh:1082-1112).
template <typename IntegT, int NDIM, bool DEBUG_MCUBES = false,
          typename GeneratorType = Curand_generator>
numint::integration_result simple_integrate(IntegT const& integrand, double epsrel,
                                            double epsabs, double ncall,
                                            quad::Volume<double, NDIM> const* volume,
                                            int totalIters = 15, int adjustIters = 15,
                                            int skipIters = 5)
{
  return pagani::mc::run<IntegT, NDIM, GeneratorType>(integrand, epsrel, epsabs, ncall, volume,
                                                      totalIters, adjustIters, skipIters,
                                                      false, true);
}

// vegasT1D.cuh: one grid, adapted on axis 0 and shared by every axis
// (integrands symmetric in their arguments).
template <typename IntegT, int NDIM, bool DEBUG_MCUBES = false,
          typename GeneratorType = Curand_generator>
numint::integration_result integrate1D(IntegT const& integrand, double epsrel, double epsabs,
                                       double ncall, quad::Volume<double, NDIM> const* volume,
                                       int titer, int itmax, int skip)
{
  return pagani::mc::run<IntegT, NDIM, GeneratorType>(integrand, epsrel, epsabs, ncall, volume,
                                                      titer, itmax, skip, true, false);
}

template <typename IntegT, int NDIM, bool DEBUG_MCUBES = false,
          typename GeneratorType = Curand_generator>
numint::integration_result simple_integrate1D(IntegT const& integrand, double epsrel,
                                              double epsabs, double ncall,
                                              quad::Volume<double, NDIM> const* volume,
                                              int titer, int itmax, int skip)
{
  return pagani::mc::run<IntegT, NDIM, GeneratorType>(integrand, epsrel, epsabs, ncall, volume,
                                                      titer, itmax, skip, true, true);
}

// vegasT.cuh:1028-1080
class mcubes {
 public:
  mcubes(double epsrel, double epsabs, double ncall, int totalIters = 15, int adjustIters = 15,
         int skipIters = 5)
      : epsrel_(epsrel),
        epsabs_(epsabs),
        ncall_(ncall),
        totalIters_(totalIters),
        adjustIters_(adjustIters),
        skipIters_(skipIters)
  {
  }

  template <typename F, int NDIM>
  numint::integration_result integrate(F& integrand, quad::Volume<double, NDIM> const* volume)
  {
    return cuda_mcubes::integrate<F, NDIM>(integrand, epsrel_, epsabs_, ncall_, volume,
                                           totalIters_, adjustIters_, skipIters_);
  }

 private:
  double epsrel_;
  double epsabs_;
  double ncall_;
  int totalIters_;
  int adjustIters_;
  int skipIters_;
};

}  // namespace cuda_mcubes

namespace quad {

// cuda/mcubes/mcubes.cuh:9-41: simple_integrate with the arity of F.
struct mcubes {
  long long int maxcalls = 1000 * 1000ULL;
  int total_iters = 70;
  int adjust_iters = 20;
  int skip_iters = 0;

  template <typename F>
  numint::integration_result integrate(
      F const& f, double epsabs, double epsrel,
      quad::Volume<double, pagani::mc::arity<F>()> const* pvol) const
  {
    constexpr int ndim = pagani::mc::arity<F>();
    static_assert(ndim >= 1, "integrand must take 1..16 doubles");
    return cuda_mcubes::simple_integrate<F, ndim>(f, epsrel, epsabs,
                                                  static_cast<double>(maxcalls), pvol,
                                                  total_iters, adjust_iters, skip_iters);
  }
};

}  // namespace quad
