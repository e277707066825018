// genz_eval.hip -- precompiled evaluation kernels for the Genz descriptor
// path: 6 families x ndim 2..8 x {FP64, FP32 integrand}.
#include <hip/hip_runtime.h>

#include <cmath>

#include "pagani/eval_kernel.hpp"
#include "pagani/genz.hpp"
#include "pagani_ext.h"

namespace {

template <class Real, int D>
int launch(const pagani_genz* g, const pagani_eval_args* a)
{
  switch (g->family) {
  case PAGANI_GENZ_OSCILLATORY: {
    pagani::Oscillatory<Real, D> f;
    for (int i = 0; i < D; ++i) f.a[i] = static_cast<Real>(g->a[i]);
    f.phase = static_cast<Real>(2.0 * M_PI * g->u[0]);
    return pagani::eval_regions<D, decltype(f), Real>(f, a);
  }
  case PAGANI_GENZ_PRODUCT_PEAK: {
    pagani::ProductPeak<Real, D> f;
    for (int i = 0; i < D; ++i) {
      f.ia2[i] = static_cast<Real>(1.0 / (g->a[i] * g->a[i]));
      f.u[i] = static_cast<Real>(g->u[i]);
    }
    return pagani::eval_regions<D, decltype(f), Real>(f, a);
  }
  case PAGANI_GENZ_CORNER_PEAK: {
    pagani::CornerPeak<Real, D> f;
    for (int i = 0; i < D; ++i) f.a[i] = static_cast<Real>(g->a[i]);
    return pagani::eval_regions<D, decltype(f), Real>(f, a);
  }
  case PAGANI_GENZ_GAUSSIAN: {
    pagani::Gaussian<Real, D> f;
    for (int i = 0; i < D; ++i) {
      f.a2[i] = static_cast<Real>(g->a[i] * g->a[i]);
      f.u[i] = static_cast<Real>(g->u[i]);
    }
    return pagani::eval_regions<D, decltype(f), Real>(f, a);
  }
  case PAGANI_GENZ_C0: {
    pagani::C0<Real, D> f;
    for (int i = 0; i < D; ++i) {
      f.a[i] = static_cast<Real>(g->a[i]);
      f.u[i] = static_cast<Real>(g->u[i]);
    }
    return pagani::eval_regions<D, decltype(f), Real>(f, a);
  }
  case PAGANI_GENZ_DISCONTINUOUS: {
    pagani::Discontinuous<Real, D> f;
    for (int i = 0; i < D; ++i) {
      f.a[i] = static_cast<Real>(g->a[i]);
      f.u[i] = static_cast<Real>(g->u[i]);
    }
    return pagani::eval_regions<D, decltype(f), Real>(f, a);
  }
  default:
    return PAGANI_EBADARG;
  }
}

template <class Real>
int launch_dim(const pagani_genz* g, const pagani_eval_args* a)
{
  switch (g->ndim) {
  case 2: return launch<Real, 2>(g, a);
  case 3: return launch<Real, 3>(g, a);
  case 4: return launch<Real, 4>(g, a);
  case 5: return launch<Real, 5>(g, a);
  case 6: return launch<Real, 6>(g, a);
  case 7: return launch<Real, 7>(g, a);
  case 8: return launch<Real, 8>(g, a);
  default: return PAGANI_EUNSUPPORTED;
  }
}

}  // namespace

extern "C" int pagani_genz_eval_launch(const pagani_genz* g, int precision,
                                       const pagani_eval_args* a)
{
  if (!g || !a) return PAGANI_EBADARG;
  if (precision == PAGANI_FP64) return launch_dim<double>(g, a);
  if (precision == PAGANI_FP32) return launch_dim<float>(g, a);
  return PAGANI_EBADARG;
}
