// reftests.hip -- the reference's test integrands driven through the C++
// drop-in header (include/pagani/workspace.hpp), exported to the pytest
// suite through extern "C".  Built into tests/cxx/libreftests.so.
//
// Integrands (formulas as in the reference's tests):
//   PTest/NTest/ZTest            Easy_Integrals.cu:25-52, RegionSampling.cu
//   F_2_6D F_3_8D F_4_5D F_5_8D F_6_6D   common/cuda/integrands.cuh
//   BoxIntegral8_15/22/25        cuda/pagani/demos/function.cuh:113-172
//   Fun6                         accuracy_improves_with_epsrel.cu:13-20
//   NaN_Integral                 NanIntegral.cu:16-24
//   XY                           do_integration_from_c.cu:10-16
#include <cmath>
#include <cstring>
#include <exception>
#include <vector>

#include "pagani/workspace.hpp"

namespace {

// whether the drivers below let the workspace run its device-resident loop
int g_device_loop = 1;

class PTest {
 public:
  __device__ __host__ double operator()(double, double) { return 15.37; }
};
class NTest {
 public:
  __device__ __host__ double operator()(double, double) { return -15.37; }
};
class ZTest {
 public:
  __device__ __host__ double operator()(double, double) { return 0.; }
};
struct XY {
  __device__ double operator()(double x, double y) const { return x * y; }
};
class NaN_Integral {
 public:
  __device__ double operator()(double, double) { return NAN; }
};

class F_2_6D {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double w, double v, double u)
  {
    const double a = 50., b = .5;
    const double xs[6] = {x, y, z, w, v, u};
    double val = 1.;
    for (double t : xs) val *= 1. / ((1. / (a * a)) + (t - b) * (t - b));
    return val;
  }
};

class F_3_8D {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double w, double v, double u,
                                        double t, double s)
  {
    return pow(1. + 8. * s + 7. * t + 6. * u + 5. * v + 4. * w + 3. * x + 2. * y + z, -9.);
  }
};

class F_4_5D {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double w, double v)
  {
    const double beta = .5, a2 = 25. * 25.;
    const double xs[5] = {x, y, z, w, v};
    double s = 0.;
    for (double t : xs) s += a2 * ((t - beta) * (t - beta));
    return exp(-1.0 * s);
  }
};

class F_5_8D {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double k, double m, double n,
                                        double p, double q)
  {
    const double beta = .5;
    const double xs[8] = {x, y, z, k, m, n, p, q};
    double t1 = 0.;
    for (double t : xs) t1 -= 10. * fabs(t - beta);
    return exp(t1);
  }
};

class F_6_6D {
 public:
  __device__ __host__ double operator()(double y, double v, double x, double w, double z, double u)
  {
    if (y > .9 || v > .8 || x > .7 || w > .6 || z > .5 || u > .4) return 0.;
    return exp(10. * y + 9. * v + 8. * x + 7. * w + 6. * z + 5. * u);
  }
};

template <int S>
class BoxIntegral8 {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double k, double l, double m,
                                        double n, double o)
  {
    const double sum = x * x + y * y + z * z + k * k + l * l + m * m + n * n + o * o;
    return pow(sum, S / 2.);
  }
};

struct Fun6 {
  __device__ __host__ double operator()(double u, double v, double w, double x, double y, double z)
  {
    return (12.0 / (7.0 - 6 * log(2.0) * log(2.0) + log(64.0))) *
           (u * v + (pow(w, y) * x * y) / (1 + u) + z * z);
  }
};

template <int D, class F>
int run(double epsrel, double epsabs, pagani_result* out)
{
  Workspace<double, D> pagani;
  pagani_ws_set_device_loop(pagani.handle(), g_device_loop);
  quad::Volume<double, D> vol;
  const numint::integration_result r = pagani.integrate(F{}, epsrel, epsabs, vol);
  out->estimate = r.estimate;
  out->errorest = r.errorest;
  out->neval = r.neval;
  out->nregions = r.nregions;
  out->nFinishedRegions = r.nFinishedRegions;
  out->status = r.status;
  out->lastPhase = r.lastPhase;
  out->chi_sq = r.chi_sq;
  out->iters = r.iters;
  return 0;
}

template <int D, class F>
int run_split(int k, double epsrel, double epsabs, std::size_t mem_total, pagani_result* out,
              std::size_t* final_n)
{
  Workspace<double, D> pagani;
  pagani_ws_set_device_loop(pagani.handle(), g_device_loop);
  if (mem_total) pagani_ws_set_device_mem(pagani.handle(), mem_total);
  quad::Volume<double, D> vol;
  Sub_regions<double, D> regions(static_cast<std::size_t>(k));
  const numint::integration_result r = pagani.integrate(F{}, regions, epsrel, epsabs, vol);
  out->estimate = r.estimate;
  out->errorest = r.errorest;
  out->nregions = r.nregions;
  out->status = r.status;
  out->iters = r.iters;
  *final_n = regions.size;
  return 0;
}

// predict_split (Workspace.cuh:244-248): the list evaluated in iteration 15
// is kept in the caller's Sub_regions when no region finished before it;
// load_snapshot (Sub_regions.cuh:232-239) then restarts from that list.
template <int D, class F>
int run_predict_split(int k, double epsrel, double epsabs, int relerr, pagani_result* out,
                      std::size_t* snap_n, std::uint64_t* trace_n, std::uint64_t* trace_m,
                      std::size_t cap, std::size_t* ntrace, pagani_result* out2)
{
  Workspace<double, D> pagani;
  quad::Volume<double, D> vol;
  Sub_regions<double, D> regions(static_cast<std::size_t>(k));
  numint::integration_result r = pagani.template integrate<F, true>(
      F{}, regions, epsrel, epsabs, vol, relerr != 0);
  out->estimate = r.estimate;
  out->errorest = r.errorest;
  out->nregions = r.nregions;
  out->status = r.status;
  out->iters = r.iters;
  *snap_n = regions.snapshot_size;
  *ntrace = pagani_ws_trace(pagani.handle(), cap, trace_n, trace_m, nullptr, nullptr);
  std::memset(out2, 0, sizeof(*out2));
  if (regions.snapshot_size) {
    regions.load_snapshot();
    if (regions.size != *snap_n || regions.snapshot_size != 0) return -3;
    r = pagani.integrate(F{}, regions, epsrel, epsabs, vol, relerr != 0);
    out2->estimate = r.estimate;
    out2->errorest = r.errorest;
    out2->nregions = r.nregions;
    out2->status = r.status;
    out2->iters = r.iters;
  }
  return 0;
}

// RegionSampling.cu: evaluate every region of a k^D uniform split once.
template <int D, class F>
int sample(int k, double* est_host, double* err_host, std::size_t cap)
{
  Workspace<double, D> pagani;
  Sub_regions<double, D> regions(static_cast<std::size_t>(k));
  const std::size_t n = regions.size;
  if (n > cap) return -1;
  double *est = nullptr, *err = nullptr;
  std::uint8_t* dim = nullptr;
  if (hipMalloc(&est, n * sizeof(double)) != hipSuccess ||
      hipMalloc(&err, n * sizeof(double)) != hipSuccess || hipMalloc(&dim, n) != hipSuccess)
    return -3;
  pagani_eval_args args{};  // dev = NULL: an ordinary launch
  args.n = n;
  args.L = regions.dLeftCoord;
  args.H = regions.dLength;
  args.est = est;
  args.err = err;
  args.dim = dim;
  args.stream = pagani_ws_stream(pagani.handle());
  int rc = pagani_ws_eval_consts(pagani.handle(), nullptr, nullptr, n, &args.consts);
  if (rc == 0) rc = pagani::eval_regions<D, F, double>(F{}, &args);
  if (rc == 0 && hipStreamSynchronize(static_cast<hipStream_t>(args.stream)) != hipSuccess) rc = -2;
  if (rc == 0 && (hipMemcpy(est_host, est, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(err_host, err, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    rc = -2;
  (void)hipFree(est);
  (void)hipFree(err);
  (void)hipFree(dim);
  return rc == 0 ? static_cast<int>(n) : rc;
}

// cos(sum (i+1) x_i), the 8D sum-of-cosines workload at any dimension
// (integrands.cuh:1322 has the 8D instance)
template <int D>
struct CosSum {
  template <class... X>
  __device__ double operator()(X... xs) const
  {
    const double x[D] = {static_cast<double>(xs)...};
    double s = 0.0;
    for (int i = 0; i < D; ++i) s += (i + 1) * x[i];
    return cos(s);
  }
};

// exp(-10 (|x - .3| + |y - .5| + |z - .7|)): a 3D C0 integrand that runs
// past iteration 15 under predict_split at epsrel 1e-6
struct C0Three {
  __device__ double operator()(double x, double y, double z) const
  {
    return exp(-10. * fabs(x - .3) - 10. * fabs(y - .5) - 10. * fabs(z - .7));
  }
};

// cos(s + t + ... + b): the reference's 10D oscillatory demo integrand
// (integrands.cuh:1144-1167, Oscillatory_10D; driven by
// cuda/pagani/demos/oscillatory_integrands.cu over [pi, 1.1 pi]^10)
struct Oscillatory10 {
  __device__ double operator()(double s, double t, double u, double v, double w, double x,
                               double y, double z, double c, double b) const
  {
    return cos(s + t + u + v + w + x + y + z + c + b);
  }
};

// A whole integration through the header with its per-iteration trace;
// predict_split selects integrate<F, true>.
template <int D, class F>
int run_traced(double epsrel, double epsabs, std::size_t mem_total, int predict_split,
               int relerr, pagani_result* out, std::uint64_t* tn, std::uint64_t* tm,
               std::size_t cap, std::size_t* ntrace)
{
  Workspace<double, D> pagani;
  pagani_ws_set_device_loop(pagani.handle(), g_device_loop);
  if (mem_total) pagani_ws_set_device_mem(pagani.handle(), mem_total);
  quad::Volume<double, D> vol;
  const numint::integration_result r =
      predict_split ? pagani.template integrate<F, true>(F{}, epsrel, epsabs, vol, relerr != 0)
                    : pagani.integrate(F{}, epsrel, epsabs, vol, relerr != 0);
  std::memset(out, 0, sizeof(*out));
  out->estimate = r.estimate;
  out->errorest = r.errorest;
  out->nregions = r.nregions;
  out->status = r.status;
  out->iters = r.iters;
  *ntrace = pagani_ws_trace(pagani.handle(), cap, tn, tm, nullptr, nullptr);
  return 0;
}

// The six Genz families as a user would write them: the formulas of the
// reference's integrands (common/cuda/integrands.cuh) in a functor with only
// operator(), so the kernel takes its plain path (x = c - g h per point, no
// origin, no axis scale, no factored sums).
template <int D, int Fam>
struct PlainGenz {
  double a[D], u[D];
  template <class... X>
  __device__ double operator()(X... xs) const
  {
    const double x[D] = {static_cast<double>(xs)...};
    double s = 0.0;
    if constexpr (Fam == PAGANI_GENZ_OSCILLATORY) {
      s = 2.0 * M_PI * u[0];
      for (int i = 0; i < D; ++i) s += a[i] * x[i];
      return cos(s);
    } else if constexpr (Fam == PAGANI_GENZ_PRODUCT_PEAK) {
      double v = 1.0;
      for (int i = 0; i < D; ++i) v *= 1.0 / (1.0 / (a[i] * a[i]) + (x[i] - u[i]) * (x[i] - u[i]));
      return v;
    } else if constexpr (Fam == PAGANI_GENZ_CORNER_PEAK) {
      s = 1.0;
      for (int i = 0; i < D; ++i) s += a[i] * x[i];
      return pow(s, -(D + 1.0));
    } else if constexpr (Fam == PAGANI_GENZ_GAUSSIAN) {
      for (int i = 0; i < D; ++i) s += a[i] * a[i] * ((x[i] - u[i]) * (x[i] - u[i]));
      return exp(-s);
    } else if constexpr (Fam == PAGANI_GENZ_C0) {
      for (int i = 0; i < D; ++i) s += a[i] * fabs(x[i] - u[i]);
      return exp(-s);
    } else {
      for (int i = 0; i < D; ++i)
        if (x[i] > u[i]) return 0.0;
      for (int i = 0; i < D; ++i) s += a[i] * x[i];
      return exp(s);
    }
  }
};

// One evaluation of n host-given regions (SoA [j*n + t]) with a plain
// functor, at `slices` waves per batch (0: the library's choice).
template <int D, int Fam>
int sample_plain(const double* a, const double* u, std::size_t n, const double* Lh,
                 const double* Hh, int slices, double* est_h, double* err_h, int* dim_h)
{
  PlainGenz<D, Fam> f;
  for (int i = 0; i < D; ++i) {
    f.a[i] = a[i];
    f.u[i] = u[i];
  }
  Workspace<double, D> pagani;
  pagani_ws_set_slices(pagani.handle(), slices);
  double *L = nullptr, *H = nullptr, *est = nullptr, *err = nullptr;
  std::uint8_t* dim = nullptr;
  int rc = 0;
  if (hipMalloc(&L, n * D * sizeof(double)) != hipSuccess ||
      hipMalloc(&H, n * D * sizeof(double)) != hipSuccess ||
      hipMalloc(&est, n * sizeof(double)) != hipSuccess ||
      hipMalloc(&err, n * sizeof(double)) != hipSuccess || hipMalloc(&dim, n) != hipSuccess)
    rc = -3;
  if (rc == 0 && (hipMemcpy(L, Lh, n * D * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(H, Hh, n * D * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
    rc = -2;
  pagani_eval_args args{};
  args.n = n;
  args.L = L;
  args.H = H;
  args.est = est;
  args.err = err;
  args.dim = dim;
  args.stream = pagani_ws_stream(pagani.handle());
  if (rc == 0) rc = pagani_ws_eval_consts(pagani.handle(), nullptr, nullptr, n, &args.consts);
  if (rc == 0) rc = pagani::eval_regions<D, PlainGenz<D, Fam>, double>(f, &args);
  if (rc == 0 && hipStreamSynchronize(static_cast<hipStream_t>(args.stream)) != hipSuccess) rc = -2;
  std::vector<std::uint8_t> dh(n);
  if (rc == 0 && (hipMemcpy(est_h, est, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(err_h, err, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(dh.data(), dim, n, hipMemcpyDeviceToHost) != hipSuccess))
    rc = -2;
  for (std::size_t t = 0; rc == 0 && t < n; ++t) dim_h[t] = dh[t];
  (void)hipFree(L);
  (void)hipFree(H);
  (void)hipFree(est);
  (void)hipFree(err);
  (void)hipFree(dim);
  return rc;
}

template <int D>
int sample_plain_family(int fam, const double* a, const double* u, std::size_t n,
                        const double* L, const double* H, int slices, double* est, double* err,
                        int* dim)
{
  switch (fam) {
    case PAGANI_GENZ_OSCILLATORY:
      return sample_plain<D, PAGANI_GENZ_OSCILLATORY>(a, u, n, L, H, slices, est, err, dim);
    case PAGANI_GENZ_PRODUCT_PEAK:
      return sample_plain<D, PAGANI_GENZ_PRODUCT_PEAK>(a, u, n, L, H, slices, est, err, dim);
    case PAGANI_GENZ_CORNER_PEAK:
      return sample_plain<D, PAGANI_GENZ_CORNER_PEAK>(a, u, n, L, H, slices, est, err, dim);
    case PAGANI_GENZ_GAUSSIAN:
      return sample_plain<D, PAGANI_GENZ_GAUSSIAN>(a, u, n, L, H, slices, est, err, dim);
    case PAGANI_GENZ_C0:
      return sample_plain<D, PAGANI_GENZ_C0>(a, u, n, L, H, slices, est, err, dim);
    case PAGANI_GENZ_DISCONTINUOUS:
      return sample_plain<D, PAGANI_GENZ_DISCONTINUOUS>(a, u, n, L, H, slices, est, err, dim);
    default:
      return -1;
  }
}

}  // namespace

extern "C" int reftest_sample_plain_genz(int fam, int d, const double* a, const double* u,
                                         std::size_t n, const double* L, const double* H,
                                         int slices, double* est, double* err, int* dim)
{
  if (!a || !u || !L || !H || !est || !err || !dim || n == 0) return -1;
  try {
    switch (d) {
      case 2: return sample_plain_family<2>(fam, a, u, n, L, H, slices, est, err, dim);
      case 3: return sample_plain_family<3>(fam, a, u, n, L, H, slices, est, err, dim);
      case 4: return sample_plain_family<4>(fam, a, u, n, L, H, slices, est, err, dim);
      case 5: return sample_plain_family<5>(fam, a, u, n, L, H, slices, est, err, dim);
      case 6: return sample_plain_family<6>(fam, a, u, n, L, H, slices, est, err, dim);
      case 7: return sample_plain_family<7>(fam, a, u, n, L, H, slices, est, err, dim);
      case 8: return sample_plain_family<8>(fam, a, u, n, L, H, slices, est, err, dim);
      case 9: return sample_plain_family<9>(fam, a, u, n, L, H, slices, est, err, dim);
      case 10: return sample_plain_family<10>(fam, a, u, n, L, H, slices, est, err, dim);
      default: return -1;
    }
  } catch (const std::exception&) {
    return -2;
  }
}

extern "C" void reftest_set_device_loop(int on) { g_device_loop = on ? 1 : 0; }

// The demo's call: Sub_regions with 2 partitions per axis (get_partitions_
// per_axis, new_time_and_call.cuh:116-126), epsabs 1e-40, the box
// [low, high]^10, with the per-iteration trace.
extern "C" int reftest_oscillatory10(double low, double high, double epsrel, pagani_result* out,
                                     std::uint64_t* tn, std::uint64_t* tm, std::size_t cap,
                                     std::size_t* ntrace)
{
  if (!out || !tn || !tm || !ntrace) return -1;
  try {
    Workspace<double, 10> workspace;
    pagani_ws_set_device_mem(workspace.handle(), static_cast<std::size_t>(16e9));
    quad::Volume<double, 10> vol(low, high);
    Sub_regions<double, 10> sub_regions(2);
    const numint::integration_result r =
        workspace.template integrate<Oscillatory10, false, false>(Oscillatory10{}, sub_regions,
                                                                   epsrel, 1.0e-40, vol, true);
    std::memset(out, 0, sizeof(*out));
    out->estimate = r.estimate;
    out->errorest = r.errorest;
    out->nregions = r.nregions;
    out->status = r.status;
    out->iters = r.iters;
    *ntrace = pagani_ws_trace(workspace.handle(), cap, tn, tm, nullptr, nullptr);
    return 0;
  } catch (const std::exception&) {
    return -2;
  }
}

extern "C" int reftest_integrate_traced(const char* name, double epsrel, double epsabs,
                                        std::size_t mem_total, int predict_split, int relerr,
                                        pagani_result* out, std::uint64_t* tn, std::uint64_t* tm,
                                        std::size_t cap, std::size_t* ntrace)
{
  if (!name || !out || !tn || !tm || !ntrace) return -1;
  try {
#define PAGANI_TRACED(N, D, F)                                                              \
  if (!std::strcmp(name, N))                                                                \
    return run_traced<D, F>(epsrel, epsabs, mem_total, predict_split, relerr, out, tn, tm, \
                            cap, ntrace);
    PAGANI_TRACED("F_2_6D", 6, F_2_6D)
    PAGANI_TRACED("F_3_8D", 8, F_3_8D)
    PAGANI_TRACED("F_4_5D", 5, F_4_5D)
    PAGANI_TRACED("F_5_8D", 8, F_5_8D)
    PAGANI_TRACED("F_6_6D", 6, F_6_6D)
    PAGANI_TRACED("CosSum5", 5, CosSum<5>)
    PAGANI_TRACED("CosSum6", 6, CosSum<6>)
    PAGANI_TRACED("C0Three", 3, C0Three)
#undef PAGANI_TRACED
  } catch (const std::exception&) {
    return -2;
  }
  return -1;
}

extern "C" int reftest_integrate(const char* name, double epsrel, double epsabs,
                                 pagani_result* out)
{
  if (!name || !out) return -1;
  std::memset(out, 0, sizeof(*out));
  try {
    if (!std::strcmp(name, "PTest")) return run<2, PTest>(epsrel, epsabs, out);
    if (!std::strcmp(name, "NTest")) return run<2, NTest>(epsrel, epsabs, out);
    if (!std::strcmp(name, "ZTest")) return run<2, ZTest>(epsrel, epsabs, out);
    if (!std::strcmp(name, "XY")) return run<2, XY>(epsrel, epsabs, out);
    if (!std::strcmp(name, "NaN_Integral")) return run<2, NaN_Integral>(epsrel, epsabs, out);
    if (!std::strcmp(name, "F_2_6D")) return run<6, F_2_6D>(epsrel, epsabs, out);
    if (!std::strcmp(name, "F_3_8D")) return run<8, F_3_8D>(epsrel, epsabs, out);
    if (!std::strcmp(name, "F_4_5D")) return run<5, F_4_5D>(epsrel, epsabs, out);
    if (!std::strcmp(name, "F_5_8D")) return run<8, F_5_8D>(epsrel, epsabs, out);
    if (!std::strcmp(name, "F_6_6D")) return run<6, F_6_6D>(epsrel, epsabs, out);
    if (!std::strcmp(name, "BoxIntegral8_15")) return run<8, BoxIntegral8<15>>(epsrel, epsabs, out);
    if (!std::strcmp(name, "BoxIntegral8_22")) return run<8, BoxIntegral8<22>>(epsrel, epsabs, out);
    if (!std::strcmp(name, "BoxIntegral8_25")) return run<8, BoxIntegral8<25>>(epsrel, epsabs, out);
    if (!std::strcmp(name, "Fun6")) return run<6, Fun6>(epsrel, epsabs, out);
  } catch (const std::exception&) {
    return -2;
  }
  return -1;
}

extern "C" int reftest_integrate_split(const char* name, int k, double epsrel, double epsabs,
                                       std::size_t mem_total, pagani_result* out,
                                       std::size_t* final_n)
{
  if (!name || !out || !final_n || k < 1) return -1;
  std::memset(out, 0, sizeof(*out));
  try {
    if (!std::strcmp(name, "XY"))
      return run_split<2, XY>(k, epsrel, epsabs, mem_total, out, final_n);
    if (!std::strcmp(name, "F_4_5D"))
      return run_split<5, F_4_5D>(k, epsrel, epsabs, mem_total, out, final_n);
  } catch (const std::exception&) {
    return -2;
  }
  return -1;
}

// Workspace<double, D, debug, use_custom, collect_mult_runs> with the debug
// recorders on: the CSV files go to the working directory.  form 0: the
// Sub_regions& overload at debug 1 with optional "hdr" (4 per axis); form 1:
// the plain overload at debug 1 with collect_mult_runs (appends); form 2: the
// Sub_regions& overload at debug 3 (regions and rule points; 2 per axis).
template <int debug, bool collect>
int run_debug_header(int form, double epsrel, pagani_result* out)
{
  Workspace<double, 2, debug, false, collect> pagani;
  quad::Volume<double, 2> vol;
  numint::integration_result r{};
  if (form == 1) {
    r = pagani.integrate(XY{}, epsrel, 1e-12, vol);
  } else {
    Sub_regions<double, 2> regions(form == 0 ? 4 : 2);
    r = pagani.integrate(XY{}, regions, epsrel, 1e-12, vol, true, "hdr");
  }
  out->estimate = r.estimate;
  out->errorest = r.errorest;
  out->nregions = r.nregions;
  out->status = r.status;
  out->iters = r.iters;
  return 0;
}

extern "C" int reftest_debug_header(int form, double epsrel, pagani_result* out)
{
  if (!out) return -1;
  std::memset(out, 0, sizeof(*out));
  try {
    if (form == 0) return run_debug_header<1, false>(form, epsrel, out);
    if (form == 1) return run_debug_header<1, true>(form, epsrel, out);
    if (form == 2) return run_debug_header<3, false>(form, epsrel, out);
  } catch (const std::exception&) {
    return -2;
  }
  return -1;
}

extern "C" int reftest_predict_split(const char* name, int k, double epsrel, double epsabs,
                                     int relerr, pagani_result* out, std::size_t* snap_n,
                                     std::uint64_t* trace_n, std::uint64_t* trace_m,
                                     std::size_t cap, std::size_t* ntrace, pagani_result* out2)
{
  if (!name || !out || !snap_n || !trace_n || !trace_m || !ntrace || !out2 || k < 1) return -1;
  try {
    if (!std::strcmp(name, "F_4_5D"))
      return run_predict_split<5, F_4_5D>(k, epsrel, epsabs, relerr, out, snap_n, trace_n,
                                          trace_m, cap, ntrace, out2);
  } catch (const std::exception&) {
    return -2;
  }
  return -1;
}

extern "C" int reftest_sample(const char* name, int k, double* est, double* err, std::size_t cap)
{
  if (!name || !est || !err || k < 1) return -1;
  try {
    if (!std::strcmp(name, "PTest")) return sample<2, PTest>(k, est, err, cap);
    if (!std::strcmp(name, "NTest")) return sample<2, NTest>(k, est, err, cap);
    if (!std::strcmp(name, "ZTest")) return sample<2, ZTest>(k, est, err, cap);
  } catch (const std::exception&) {
    return -2;
  }
  return -1;
}

// ---------------------------------------------------------------------------
// Interpolation tables (include/pagani/interp.hpp), driven like the
// reference's Interpolation1D/2D/3D.cu: evaluate on the device at given
// points, optionally through clamp().

#include <fstream>

#include "pagani/interp.hpp"
#include "pagani/str_to_doubles.hh"

namespace {

template <class T>
struct DevVec {
  T* p = nullptr;
  explicit DevVec(std::size_t n)
  {
    if (n && hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::bad_alloc();
  }
  ~DevVec()
  {
    if (p) (void)hipFree(p);
  }
  void put(const T* h, std::size_t n)
  {
    if (n && hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("copy");
  }
  void get(T* h, std::size_t n) const
  {
    if (n && hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("copy");
  }
};

__global__ void k_interp1(quad::Interp1D f, const double* x, double* out, std::size_t n,
                          int clamp)
{
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = clamp ? f.clamp(x[i]) : f(x[i]);
}

__global__ void k_interp2(quad::Interp2D f, const double* x, const double* y, double* out,
                          std::size_t n, int clamp)
{
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = clamp ? f.clamp(x[i], y[i]) : f(x[i], y[i]);
}

__global__ void k_interp3(quad::Interp3D f, const double* x, const double* y, const double* z,
                          double* out, std::size_t n, int clamp)
{
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = clamp ? f.clamp(x[i], y[i], z[i]) : f(x[i], y[i], z[i]);
}

unsigned blocks_for(std::size_t n) { return static_cast<unsigned>((n + 255) / 256); }

// an integrand holding a table, as the reference's model integrands do
struct TableIntegrand {
  quad::Interp2D table;
  __device__ double operator()(double x, double y) const { return table(x, y); }
};

// memory-usage check (interpolator_mem_usage.cu): an integrand with both
// kinds of table, copied into every launch
struct TwoTables {
  quad::Interp1D a;
  quad::Interp2D b;
  __device__ double operator()(double x, double y) const { return a(x) * b(x, y); }
};

bool load_table(const char* path, quad::Interp2D& t)
{
  std::ifstream in(path);
  if (!in) return false;
  in >> t;
  return static_cast<bool>(in) || in.eof();
}

}  // namespace

extern "C" int reftest_parse_doubles(const char* line, double* out, std::size_t cap,
                                     std::size_t* n)
{
  try {
    const std::vector<double> v = str_to_doubles(line);
    *n = v.size();
    for (std::size_t i = 0; i < v.size() && i < cap; ++i) out[i] = v[i];
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

extern "C" int reftest_interp1d(const double* xs, const double* zs, std::size_t cols,
                                const double* px, std::size_t n, int clamp, double* out)
{
  try {
    quad::Interp1D f(xs, zs, cols);
    DevVec<double> dx(n), dout(n);
    dx.put(px, n);
    hipLaunchKernelGGL(k_interp1, dim3(blocks_for(n)), dim3(256), 0, 0, f, dx.p, dout.p, n,
                       clamp);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
    dout.get(out, n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

extern "C" int reftest_interp2d(const double* xs, const double* ys, const double* zs,
                                std::size_t cols, std::size_t rows, const double* px,
                                const double* py, std::size_t n, int clamp, double* out)
{
  try {
    quad::Interp2D f(xs, ys, zs, cols, rows);
    DevVec<double> dx(n), dy(n), dout(n);
    dx.put(px, n);
    dy.put(py, n);
    hipLaunchKernelGGL(k_interp2, dim3(blocks_for(n)), dim3(256), 0, 0, f, dx.p, dy.p, dout.p,
                       n, clamp);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
    dout.get(out, n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

extern "C" int reftest_interp3d(const double* xs, const double* ys, const double* zs,
                                const double* vs, std::size_t nx, std::size_t ny, std::size_t nz,
                                const double* px, const double* py, const double* pz,
                                std::size_t n, int clamp, double* out)
{
  try {
    quad::Interp3D f(xs, ys, zs, vs, nx, ny, nz);
    DevVec<double> dx(n), dy(n), dz(n), dout(n);
    dx.put(px, n);
    dy.put(py, n);
    dz.put(pz, n);
    hipLaunchKernelGGL(k_interp3, dim3(blocks_for(n)), dim3(256), 0, 0, f, dx.p, dy.p, dz.p,
                       dout.p, n, clamp);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
    dout.get(out, n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// A table read from a stream-format file, evaluated at n points.
extern "C" int reftest_interp2d_file(const char* path, const double* px, const double* py,
                                     std::size_t n, double* out)
{
  try {
    quad::Interp2D f;
    if (!load_table(path, f)) return -3;
    DevVec<double> dx(n), dy(n), dout(n);
    dx.put(px, n);
    dy.put(py, n);
    hipLaunchKernelGGL(k_interp2, dim3(blocks_for(n)), dim3(256), 0, 0, f, dx.p, dy.p, dout.p,
                       n, 0);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
    dout.get(out, n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// PAGANI over the table's own rectangle, with the table inside the integrand.
extern "C" int reftest_integrate_table_file(const char* path, const double* lows,
                                            const double* highs, double epsrel, double epsabs,
                                            pagani_result* out)
{
  try {
    TableIntegrand f;
    if (!load_table(path, f.table)) return -3;
    Workspace<double, 2> pagani;
    quad::Volume<double, 2> vol(lows, highs);
    const numint::integration_result r = pagani.integrate(f, epsrel, epsabs, vol);
    out->estimate = r.estimate;
    out->errorest = r.errorest;
    out->nregions = r.nregions;
    out->status = r.status;
    out->iters = r.iters;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Device memory in use before and after `runs` integrations of an integrand
// holding two tables (interpolator_mem_usage.cu): copies must not leak.
extern "C" int reftest_interp_mem_usage(int runs, std::size_t* free_before,
                                        std::size_t* free_after, double* last)
{
  try {
    const double xs1[5] = {0., .25, .5, .75, 1.};
    const double zs1[5] = {1., 1.5, 2., 2.5, 3.};
    const double xs2[3] = {0., .5, 1.};
    const double ys2[2] = {0., 1.};
    const double zs2[6] = {1., 2., 3., 4., 5., 6.};
    Workspace<double, 2> pagani;
    quad::Volume<double, 2> vol;
    TwoTables f{quad::Interp1D(xs1, zs1, 5), quad::Interp2D(xs2, ys2, zs2, 3, 2)};
    (void)pagani.integrate(f, 1e-6, 1e-12, vol);  // warm up the pools
    std::size_t total = 0;
    if (hipMemGetInfo(free_before, &total) != hipSuccess) return -2;
    numint::integration_result r{};
    for (int i = 0; i < runs; ++i) r = pagani.integrate(f, 1e-6, 1e-12, vol);
    if (hipMemGetInfo(free_after, &total) != hipSuccess) return -2;
    *last = r.estimate;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// ---------------------------------------------------------------------------
// The non-positive-argument exponential of the Gaussian and C0 families
// (pagani::math::exp_neg) beside the device library's exp.

#include "pagani/genz.hpp"

namespace {
__global__ void k_exp_neg(const double* x, double* mine, double* lib, std::size_t n)
{
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    mine[i] = pagani::math::exp_neg(x[i]);
    lib[i] = ::exp(x[i]);
  }
}
}  // namespace

extern "C" int reftest_exp_neg(const double* x, std::size_t n, double* mine, double* lib)
{
  try {
    DevVec<double> dx(n), dm(n), dl(n);
    dx.put(x, n);
    hipLaunchKernelGGL(k_exp_neg, dim3(blocks_for(n)), dim3(256), 0, 0, dx.p, dm.p, dl.p, n);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
    dm.get(mine, n);
    dl.get(lib, n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// ---------------------------------------------------------------------------
// The bounded-argument sine and cosine of the cosine integrands' factored
// sums (pagani::math::sincos_bounded) beside the device library's.

#include "pagani/trig.hpp"

namespace {
__global__ void k_sincos(const double* x, double* s, double* c, double* ls, double* lc,
                         std::size_t n)
{
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    double a, b;
    pagani::math::sincos_bounded(x[i], a, b);
    s[i] = a;
    c[i] = b;
    ::sincos(x[i], &a, &b);
    ls[i] = a;
    lc[i] = b;
  }
}
}  // namespace

extern "C" int reftest_sincos(const double* x, std::size_t n, double* s, double* c, double* ls,
                              double* lc)
{
  try {
    DevVec<double> dx(n), ds(n), dc(n), dls(n), dlc(n);
    dx.put(x, n);
    hipLaunchKernelGGL(k_sincos, dim3(blocks_for(n)), dim3(256), 0, 0, dx.p, ds.p, dc.p, dls.p,
                       dlc.p, n);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
    ds.get(s, n);
    dc.get(c, n);
    dls.get(ls, n);
    dlc.get(lc, n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// ---------------------------------------------------------------------------
// m-Cubes through the reference's C++ surface (include/pagani/mcubes.hpp).
// The integrands are the ones of cuda/mcubes/demos (Genz2_6D.cu, Genz3_3D.cu,
// Genz4_5D.cu, Genz5_8D.cu, Genz6_6D.cu, Genz1_8D.cu, SinSum6D.cu, Gauss9D.cu).
// ---------------------------------------------------------------------------
#include "pagani/mcubes.hpp"

namespace {

class MC_F_3_3D {
 public:
  __device__ __host__ double operator()(double x, double y, double z)
  {
    return pow(1 + 3 * x + 2 * y + z, -4);
  }
};

class MC_F_6_6D {
 public:
  __device__ __host__ double operator()(double u, double v, double w, double x, double y, double z)
  {
    if (z > .9 || y > .8 || x > .7 || w > .6 || v > .5 || u > .4) return 0.;
    return exp(10 * z + 9 * y + 8 * x + 7 * w + 6 * v + 5 * u);
  }
};

class MC_F_1_8D {
 public:
  __device__ __host__ double operator()(double s, double t, double u, double v, double w,
                                        double x, double y, double z)
  {
    return cos(s + 2. * t + 3. * u + 4. * v + 5. * w + 6. * x + 7. * y + 8. * z);
  }
};

class SinSum6D {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double k, double l,
                                        double m)
  {
    return sin(x + y + z + k + l + m);
  }
};

class Gauss9D {
 public:
  __device__ __host__ double operator()(double x, double y, double z, double k, double l,
                                        double m, double n, double o, double p)
  {
    const double sum = x * x + y * y + z * z + k * k + l * l + m * m + n * n + o * o + p * p;
    const double s = 0.01;
    return exp(-1 * sum / (2 * s * s)) * (1 / pow(sqrt(2 * M_PI) * s, 9));
  }
};

class Gauss1D {
 public:
  __device__ double operator()(double x) const { return exp(-50.0 * (x - 0.3) * (x - 0.3)); }
};

// 12D and 16D products of cos: exact prod_j sin(1)/1 = sin(1)^d on [0, 1]^d
template <int D>
struct CosProd;
template <>
struct CosProd<12> {
  __device__ double operator()(double a, double b, double c, double d, double e, double f,
                               double g, double h, double i, double j, double k, double l) const
  {
    return cos(a) * cos(b) * cos(c) * cos(d) * cos(e) * cos(f) * cos(g) * cos(h) * cos(i) *
           cos(j) * cos(k) * cos(l);
  }
};
template <>
struct CosProd<16> {
  __device__ double operator()(double a, double b, double c, double d, double e, double f,
                               double g, double h, double i, double j, double k, double l,
                               double m, double n, double o, double p) const
  {
    return cos(a) * cos(b) * cos(c) * cos(d) * cos(e) * cos(f) * cos(g) * cos(h) * cos(i) *
           cos(j) * cos(k) * cos(l) * cos(m) * cos(n) * cos(o) * cos(p);
  }
};

template <int D, class F, class G>
int mc_run(int mode, double epsrel, double ncall, int titer, int itmax, int skip,
           const double* lows, const double* highs, pagani_result* out)
{
  F f;
  quad::Volume<double, D> vol;
  if (lows && highs) vol = quad::Volume<double, D>(lows, highs);
  numint::integration_result r;
  const double epsabs = 1e-20;
  switch (mode) {
    case 0:
      r = cuda_mcubes::integrate<F, D, false, G>(f, epsrel, epsabs, ncall, &vol, titer, itmax,
                                                 skip);
      break;
    case 1:
      r = cuda_mcubes::simple_integrate<F, D, false, G>(f, epsrel, epsabs, ncall, &vol, titer,
                                                        itmax, skip);
      break;
    case 2:
      r = cuda_mcubes::integrate1D<F, D, false, G>(f, epsrel, epsabs, ncall, &vol, titer, itmax,
                                                   skip);
      break;
    case 3: {
      // quad::mcubes (mcubes.cuh): arity deduced from F, default generator
      quad::mcubes m;
      m.maxcalls = static_cast<long long>(ncall);
      m.total_iters = titer;
      m.adjust_iters = itmax;
      m.skip_iters = skip;
      r = m.integrate(f, epsabs, epsrel, &vol);
      break;
    }
    case 4: {
      cuda_mcubes::mcubes m(epsrel, epsabs, ncall, titer, itmax, skip);
      r = m.integrate<F, D>(f, &vol);
      break;
    }
    default:
      return -1;
  }
  std::memset(out, 0, sizeof(*out));
  out->estimate = r.estimate;
  out->errorest = r.errorest;
  out->chi_sq = r.chi_sq;
  out->status = r.status;
  out->iters = r.iters;
  out->neval = r.neval;
  return 0;
}

template <class G>
int mc_dispatch(const char* name, int mode, double epsrel, double ncall, int titer, int itmax,
                int skip, pagani_result* out)
{
  if (!std::strcmp(name, "XY")) return mc_run<2, XY, G>(mode, epsrel, ncall, titer, itmax, skip,
                                                         nullptr, nullptr, out);
  if (!std::strcmp(name, "F_2_6D"))
    return mc_run<6, F_2_6D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr, out);
  if (!std::strcmp(name, "F_3_3D"))
    return mc_run<3, MC_F_3_3D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr,
                                   out);
  if (!std::strcmp(name, "F_4_5D"))
    return mc_run<5, F_4_5D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr, out);
  if (!std::strcmp(name, "F_5_8D"))
    return mc_run<8, F_5_8D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr, out);
  if (!std::strcmp(name, "F_6_6D"))
    return mc_run<6, MC_F_6_6D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr,
                                   out);
  if (!std::strcmp(name, "F_1_8D"))
    return mc_run<8, MC_F_1_8D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr,
                                   out);
  if (!std::strcmp(name, "SinSum6D")) {
    const double lo[6] = {0, 0, 0, 0, 0, 0}, hi[6] = {10, 10, 10, 10, 10, 10};
    return mc_run<6, SinSum6D, G>(mode, epsrel, ncall, titer, itmax, skip, lo, hi, out);
  }
  if (!std::strcmp(name, "Gauss1D"))
    return mc_run<1, Gauss1D, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr, out);
  if (!std::strcmp(name, "CosProd12"))
    return mc_run<12, CosProd<12>, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr,
                                      out);
  if (!std::strcmp(name, "CosProd16"))
    return mc_run<16, CosProd<16>, G>(mode, epsrel, ncall, titer, itmax, skip, nullptr, nullptr,
                                      out);
  if (!std::strcmp(name, "Gauss9D")) {
    double lo[9], hi[9];
    for (int i = 0; i < 9; ++i) {
      lo[i] = -1;
      hi[i] = 1;
    }
    return mc_run<9, Gauss9D, G>(mode, epsrel, ncall, titer, itmax, skip, lo, hi, out);
  }
  return -1;
}

}  // namespace

// generator: 0 = Curand_generator (the library's default streams), 1 =
// Custom_generator; mode: 0 integrate, 1 simple_integrate, 2 integrate1D,
// 3 quad::mcubes, 4 cuda_mcubes::mcubes.
extern "C" int reftest_mcubes(const char* name, int generator, int mode, double epsrel,
                              double ncall, int titer, int itmax, int skip, pagani_result* out)
{
  if (!name || !out) return -1;
  try {
    if (generator == 1)
      return mc_dispatch<Custom_generator>(name, mode, epsrel, ncall, titer, itmax, skip, out);
    return mc_dispatch<Curand_generator>(name, mode, epsrel, ncall, titer, itmax, skip, out);
  } catch (const std::exception&) {
    return -2;
  }
}
