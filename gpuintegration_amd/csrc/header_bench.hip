// header_bench.hip -- the drop-in C++ surface as a user compiles it, for the
// benchmark's generic leg: the 8D Genz Gaussian written the way the
// reference writes its integrands (common/cuda/integrands.cuh:1303-1934, a
// functor with only operator()), integrated through Workspace<double, 8>
// (include/pagani/workspace.hpp).  The evaluation kernel instantiated here
// is the plain path: every point mapped and the integrand called point by
// point, with none of the centred, scaled or factored forms the library's
// Genz descriptors use.  Built into libpagani_header_bench.so.
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>

#include "pagani/workspace.hpp"

namespace {

// f(x) = exp(-sum a_i^2 (x_i - u_i)^2), the reference's Gaussian formula
struct PlainGaussian8 {
  double a[8], u[8];
  __device__ double operator()(double x0, double x1, double x2, double x3, double x4, double x5,
                               double x6, double x7) const
  {
    const double x[8] = {x0, x1, x2, x3, x4, x5, x6, x7};
    double s = 0.0;
    for (int i = 0; i < 8; ++i) s += a[i] * a[i] * ((x[i] - u[i]) * (x[i] - u[i]));
    return exp(-s);
  }
};

}  // namespace

// `warmup` untimed integrations, then `steps` timed ones (host clock around
// the calls, each call blocking); kernel times and counts of the timed calls
// in *stats.  Returns 0, or a negative status.
extern "C" int pagani_header_bench_gaussian8(const double* a, const double* u, double epsrel,
                                             double epsabs, int steps, int warmup,
                                             pagani_result* out, pagani_stats* stats,
                                             double* seconds)
{
  if (!a || !u || !out || !stats || !seconds || steps < 1 || warmup < 0) return PAGANI_EBADARG;
  try {
    PlainGaussian8 f;
    for (int i = 0; i < 8; ++i) {
      f.a[i] = a[i];
      f.u[i] = u[i];
    }
    Workspace<double, 8> pagani;
    quad::Volume<double, 8> vol;
    // the initial split of the workload: 4 per axis (the Sub_regions& overload)
    numint::integration_result r{};
    for (int w = 0; w < warmup; ++w) {
      Sub_regions<double, 8> regions(4);
      r = pagani.integrate(f, regions, epsrel, epsabs, vol);
    }
    pagani_ws_set_timing(pagani.handle(), 1);
    pagani_ws_reset_stats(pagani.handle());
    double total = 0.0;
    for (int s = 0; s < steps; ++s) {
      Sub_regions<double, 8> regions(4);  // the list is built outside the timed call
      const auto t0 = std::chrono::steady_clock::now();
      r = pagani.integrate(f, regions, epsrel, epsabs, vol);
      total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    *seconds = total;
    std::memset(out, 0, sizeof(*out));
    out->estimate = r.estimate;
    out->errorest = r.errorest;
    out->nregions = r.nregions;
    out->status = r.status;
    out->iters = r.iters;
    return pagani_ws_get_stats(pagani.handle(), stats);
  } catch (const std::exception&) {
    return PAGANI_EHIP;
  }
}

// The plain path's evaluation kernel alone, for hardware-counter runs:
// `launches` evaluations of the n regions (L, H: SoA bounds on the device,
// unit-cube integration box) with the launch constants the workspace gives
// a callback, on the workspace's stream; blocks until they finish.
extern "C" int pagani_header_bench_eval_gaussian8(pagani_ws* ws, const double* a, const double* u,
                                                  size_t n, const double* L, const double* H,
                                                  double* est, double* err, uint8_t* dim,
                                                  int launches)
{
  if (!ws || !a || !u || !L || !H || !est || !err || !dim || launches < 1) return PAGANI_EBADARG;
  PlainGaussian8 f;
  for (int i = 0; i < 8; ++i) {
    f.a[i] = a[i];
    f.u[i] = u[i];
  }
  pagani_eval_args args{};
  int rc = pagani_ws_eval_consts(ws, nullptr, nullptr, n, &args.consts);
  if (rc) return rc;
  args.n = n;
  args.L = L;
  args.H = H;
  args.est = est;
  args.err = err;
  args.dim = dim;
  args.stream = pagani_ws_stream(ws);
  for (int l = 0; l < launches; ++l)
    if ((rc = pagani::eval_regions<8, PlainGaussian8>(f, &args))) return rc;
  return hipStreamSynchronize(static_cast<hipStream_t>(args.stream)) == hipSuccess ? PAGANI_OK
                                                                                : PAGANI_EHIP;
}
