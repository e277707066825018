// genz_visit.hpp -- builds the Genz functor a descriptor names and hands it
// to a generic callable (the PAGANI evaluation and the m-Cubes sampling
// dispatch both start here).  Parameters are converted as the reference's
// integrands.cuh functors hold them: 1/a^2 for the product peak, a^2 for the
// Gaussian, 2 pi u_0 for the oscillatory phase.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "pagani.h"
#include "pagani/genz.hpp"

namespace pagani {
namespace internal {

// Host and device (the batch's segment kernel builds the functors on the
// device from the descriptors in its item list).
template <class Real, int D, class Fn>
__host__ __device__ int visit_genz(const pagani_genz* g, Fn&& fn)
{
  switch (g->family) {
    case PAGANI_GENZ_OSCILLATORY: {
      pagani::Oscillatory<Real, D> f;
      for (int i = 0; i < D; ++i) f.a[i] = static_cast<Real>(g->a[i]);
      f.phase = static_cast<Real>(2.0 * M_PI * g->u[0]);
      return fn(f);
    }
    case PAGANI_GENZ_PRODUCT_PEAK: {
      pagani::ProductPeak<Real, D> f;
      f.k2 = 1.0;
      for (int i = 0; i < D; ++i) {
        f.ia2[i] = static_cast<Real>(1.0 / (g->a[i] * g->a[i]));
        f.u[i] = static_cast<Real>(g->u[i]);
        f.a[i] = static_cast<Real>(g->a[i]);
        f.k2 *= g->a[i] * g->a[i];
      }
      return fn(f);
    }
    case PAGANI_GENZ_CORNER_PEAK: {
      pagani::CornerPeak<Real, D> f;
      for (int i = 0; i < D; ++i) f.a[i] = static_cast<Real>(g->a[i]);
      return fn(f);
    }
    case PAGANI_GENZ_GAUSSIAN: {
      pagani::Gaussian<Real, D> f;
      for (int i = 0; i < D; ++i) {
        f.a[i] = static_cast<Real>(g->a[i]);
        f.a2[i] = static_cast<Real>(g->a[i] * g->a[i]);
        f.u[i] = static_cast<Real>(g->u[i]);
      }
      return fn(f);
    }
    case PAGANI_GENZ_C0: {
      pagani::C0<Real, D> f;
      for (int i = 0; i < D; ++i) {
        f.a[i] = static_cast<Real>(g->a[i]);
        f.u[i] = static_cast<Real>(g->u[i]);
      }
      return fn(f);
    }
    case PAGANI_GENZ_DISCONTINUOUS: {
      pagani::Discontinuous<Real, D> f;
      f.c = 0.0;
      for (int i = 0; i < D; ++i) {
        f.a[i] = static_cast<Real>(g->a[i]);
        f.u[i] = static_cast<Real>(g->u[i]);
        f.c += g->a[i] * g->u[i];
      }
      return fn(f);
    }
    default:
      return PAGANI_EBADARG;
  }
}

}  // namespace internal
}  // namespace pagani
