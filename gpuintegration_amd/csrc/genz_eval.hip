// genz_eval.hip -- precompiled evaluation kernels for the Genz descriptor
// path: 6 families x ndim 2..10 x {FP64, FP32 integrand}.
#include <hip/hip_runtime.h>

#include <cmath>

#include "pagani/eval_kernel.hpp"
#include "pagani/genz.hpp"
#include "genz_visit.hpp"
#include "pagani_ext.h"

namespace {

template <class Real, int D>
int launch(const pagani_genz* g, const pagani_eval_args* a)
{
  return pagani::internal::visit_genz<Real, D>(g, [&](const auto& f) {
    return pagani::eval_regions<D, std::decay_t<decltype(f)>, Real>(f, a);
  });
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
  case 9: return launch<Real, 9>(g, a);
  case 10: return launch<Real, 10>(g, a);
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
