// interop.hip -- the reference's C interop shim
// (test/cuda/pagani/do_integration_from_c.cu:20-38) on the C++ header API:
// integrate x*y over the unit square to epsrel 1e-6 and hand the estimate
// to a C caller.
#include <exception>

#include "pagani/workspace.hpp"

namespace {
struct XY {
  __device__ double operator()(double x, double y) const { return x * y; }
};
}  // namespace

extern "C" int do_integration_from_c(double* res)
{
  if (!res) return 1;
  try {
    constexpr int ndim = 2;
    quad::Volume<double, ndim> vol;
    Workspace<double, ndim> pagani;
    const auto result = pagani.integrate<XY>(XY{}, 1.0e-6, 1.0e-12, vol);
    const int rc = result.status;
    if (rc == 0) *res = result.estimate;
    return rc;
  } catch (const std::exception&) {
    return 1;
  }
}
