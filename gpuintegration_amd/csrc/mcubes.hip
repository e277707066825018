// mcubes.hip -- the m-Cubes driver (pagani_mc_* entry points).
//
// One vegas() run (cuda/mcubes/vegasT.cuh:640-1000): the run constants and
// the uniform grid on the host; per iteration one sampling launch over every
// cube (the integrand's kernel, behind a pagani_mc_sample_fn), one
// reduction launch that folds the workgroup partials in a fixed order, one
// copy of the sums and histograms to the host, then the host combines the
// iteration into the running estimate and, in adjusting iterations, refines
// the grid (mcubes_grid.cpp).  simple_integrate's retries with AdjustParams
// (vegasT.cuh:1082-1112) wrap the run.  All device memory is allocated at
// creation: the largest launch has kMaxGrid workgroups.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "genz_visit.hpp"
#include "mcubes_grid.hpp"
#include "pagani.h"
#include "pagani/mcubes_kernel.hpp"

namespace {

using pagani::mc::kBins;
using pagani::mc::kThreads;

constexpr std::uint32_t kMaxGrid = 1024;
constexpr int kOutDoubles = 2 + PAGANI_MC_MAX_DIM * kBins;

thread_local std::string g_mc_create_error;

// Sums over the workgroups: out[0] = sum fb, out[1] = sum f2b, out[2 + i] =
// histogram entry i.  Workgroup b folds 64 entries: wave w of its 16 adds
// the partials of workgroups w, w + 16, ... in that order, then the wave
// sums are added in wave order -- a fixed order for any run.
constexpr int kRedWaves = 16;

__global__ void __launch_bounds__(64 * kRedWaves)
    k_mc_reduce(const double* __restrict__ part, const double* __restrict__ hist,
                std::uint32_t grid, std::uint32_t nhist, double* __restrict__ out)
{
  __shared__ double ws[kRedWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const std::uint32_t t = blockIdx.x * 64 + lane;  // output entry
  double s = 0.0;
  if (t < nhist + 2) {
    if (t < 2) {
      for (std::uint32_t g = wave; g < grid; g += kRedWaves) s += part[2 * g + t];
    } else {
      const std::uint32_t i = t - 2;
      for (std::uint32_t g = wave; g < grid; g += kRedWaves)
        s += hist[static_cast<std::size_t>(g) * nhist + i];
    }
  }
  ws[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && t < nhist + 2) {
    double r = ws[0][lane];
#pragma unroll
    for (int w = 1; w < kRedWaves; ++w) r += ws[w][lane];
    out[t] = r;
  }
}

}  // namespace

struct pagani_mc {
  int ndim = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  double* d_xi = nullptr;
  double* d_part = nullptr;
  double* d_hist = nullptr;
  double* d_out = nullptr;
  double* h_out = nullptr;  // pinned
  double* h_xi = nullptr;   // pinned
  std::vector<double> ti, tsi;
  std::vector<double> grid;  // last run's final grid
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  pagani_mc_stats stats{};
};

namespace {

int fail(pagani_mc* mc, int code, const std::string& msg)
{
  if (mc) mc->error = msg;
  return code;
}

int hip_fail(pagani_mc* mc, hipError_t e, const char* where)
{
  (void)hipGetLastError();
  return fail(mc, PAGANI_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define MC_TRY(mc, expr)                                  \
  do {                                                    \
    const hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(mc, e_, #expr); \
  } while (0)

struct Launch {
  std::uint32_t grid = 0, cpt = 0;
};

// cubes_per_thread so that at most kMaxGrid workgroups cover every cube
Launch shape(std::uint64_t ncubes)
{
  const std::uint64_t per_wg_max = static_cast<std::uint64_t>(kThreads) * kMaxGrid;
  Launch l;
  l.cpt = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, (ncubes + per_wg_max - 1) / per_wg_max));
  const std::uint64_t per_wg = static_cast<std::uint64_t>(kThreads) * l.cpt;
  l.grid = static_cast<std::uint32_t>((ncubes + per_wg - 1) / per_wg);
  return l;
}

int check_setup(pagani_mc* mc, const pagani::mc::Setup& s)
{
  if (!(s.ng >= 1) || !(s.ncubes >= 1))
    return fail(mc, PAGANI_EBADARG, "ncall too small: no stratum per axis");
  if (s.ncubes >= 4294967296.0)
    return fail(mc, PAGANI_EUNSUPPORTED, "more than 2^32 cubes (the reference numbers cubes in 32 bits)");
  return PAGANI_OK;
}

void fill_args(pagani_mc* mc, const pagani::mc::Setup& s, const Launch& l, int generator,
               pagani_mc_args& a)
{
  std::memset(&a, 0, sizeof(a));
  a.xi = mc->d_xi;
  for (int j = 0; j < mc->ndim; ++j) {
    a.lo[j] = s.lo[j];
    a.dx[j] = s.dx[j];
  }
  a.xjac = s.xjac;
  a.dxg = s.dxg;
  a.xnd = s.xnd;
  a.ncubes = static_cast<std::uint32_t>(s.ncubes);
  a.ng = static_cast<std::uint32_t>(s.ng);
  a.npg = static_cast<std::uint32_t>(s.npg);
  a.cubes_per_thread = l.cpt;
  a.grid = l.grid;
  a.generator = generator;
  a.part = mc->d_part;
  a.hist = mc->d_hist;
  a.stream = mc->stream;
}

// One sampling launch and its reduction; on return h_out holds the sums.
int sample_pass(pagani_mc* mc, pagani_mc_sample_fn sample, void* ctx, pagani_mc_args& a,
                std::uint64_t samples)
{
  const std::uint32_t nhist = a.hist_dims * kBins;
  if (mc->timing) MC_TRY(mc, hipEventRecord(mc->ev0, mc->stream));
  const int rc = sample(ctx, &a);
  if (rc) return fail(mc, rc < 0 ? rc : PAGANI_EHIP, "sampling callback failed");
  if (mc->timing) MC_TRY(mc, hipEventRecord(mc->ev1, mc->stream));
  hipLaunchKernelGGL(k_mc_reduce, dim3((nhist + 2 + 63) / 64), dim3(64 * kRedWaves), 0,
                     mc->stream, mc->d_part, mc->d_hist, a.grid, nhist, mc->d_out);
  MC_TRY(mc, hipGetLastError());
  MC_TRY(mc, hipMemcpyAsync(mc->h_out, mc->d_out, (nhist + 2) * sizeof(double),
                            hipMemcpyDeviceToHost, mc->stream));
  MC_TRY(mc, hipStreamSynchronize(mc->stream));
  if (mc->timing) {
    float ms = 0;
    MC_TRY(mc, hipEventElapsedTime(&ms, mc->ev0, mc->ev1));
    mc->stats.sample_ms += ms;
  }
  mc->stats.samples += samples;
  mc->stats.launches++;
  return PAGANI_OK;
}

int upload_grid(pagani_mc* mc, const std::vector<double>& xi)
{
  std::memcpy(mc->h_xi, xi.data(), xi.size() * sizeof(double));
  MC_TRY(mc, hipMemcpyAsync(mc->d_xi, mc->h_xi, xi.size() * sizeof(double),
                            hipMemcpyHostToDevice, mc->stream));
  return PAGANI_OK;
}

// One vegas() run.
int run(pagani_mc* mc, pagani_mc_sample_fn sample, void* ctx, double epsrel, double epsabs,
        const double* lows, const double* highs, const pagani_mc_params& p, double ncall,
        int titer, pagani_mc_result* out)
{
  const int d = mc->ndim;
  pagani::mc::Setup s;
  pagani::mc::setup(s, d, ncall, lows, highs);
  int rc = check_setup(mc, s);
  if (rc) return rc;
  const int hd = p.one_axis ? 1 : d;
  std::vector<double> xi(static_cast<std::size_t>(d) * (kBins + 1));
  pagani::mc::grid_init(d, xi.data());
  const Launch l = shape(static_cast<std::uint64_t>(s.ncubes));
  pagani_mc_args a;
  fill_args(mc, s, l, p.generator, a);
  const std::uint64_t per_iter = static_cast<std::uint64_t>(s.calls);
  double si = 0.0, swgt = 0.0, schi = 0.0;
  std::vector<double> hist(static_cast<std::size_t>(hd) * kBins);
  int it;
  for (it = 1; it <= p.adjust_iters && out->status == 1; out->iters++, it++) {
    if ((rc = upload_grid(mc, xi))) return rc;
    a.hist_dims = static_cast<std::uint32_t>(hd);
    a.key = pagani::mc::iter_key(p.seed, it);
    if ((rc = sample_pass(mc, sample, ctx, a, per_iter))) return rc;
    out->neval += per_iter;
    const double ti = mc->h_out[0];
    const double tsi = mc->h_out[1] * s.dv2g;
    mc->ti.push_back(ti);
    mc->tsi.push_back(tsi);
    if (it > p.skip_iters) {
      pagani::mc::combine(ti, tsi, it, si, schi, swgt, out->estimate, out->chi_sq, out->errorest);
      out->status = pagani::mc::get_status(out->estimate, out->errorest, it, epsrel, epsabs);
    }
    std::memcpy(hist.data(), mc->h_out + 2, hist.size() * sizeof(double));
    pagani::mc::refine(d, hd, hist.data(), xi.data());
  }
  if ((rc = upload_grid(mc, xi))) return rc;
  for (it = p.adjust_iters + 1; it <= titer && out->status; out->iters++, it++) {
    a.hist_dims = 0;
    a.key = pagani::mc::iter_key(p.seed, it);
    if ((rc = sample_pass(mc, sample, ctx, a, per_iter))) return rc;
    out->neval += per_iter;
    const double ti = mc->h_out[0];
    const double tsi = mc->h_out[1] * s.dv2g;
    mc->ti.push_back(ti);
    mc->tsi.push_back(tsi);
    pagani::mc::combine(ti, tsi, it, si, schi, swgt, out->estimate, out->chi_sq, out->errorest);
    out->status = pagani::mc::get_status(out->estimate, out->errorest, it, epsrel, epsabs);
  }
  mc->grid = xi;
  return PAGANI_OK;
}

int validate(pagani_mc* mc, const pagani_mc_params* p, const double* lows, const double* highs)
{
  if (!p) return fail(mc, PAGANI_EBADARG, "null params");
  if (!(p->ncall >= 1) || p->total_iters < 0 || p->adjust_iters < 0 || p->skip_iters < 0)
    return fail(mc, PAGANI_EBADARG, "bad m-Cubes parameters");
  if (p->generator != PAGANI_MC_RNG_DEFAULT && p->generator != PAGANI_MC_RNG_LCG)
    return fail(mc, PAGANI_EBADARG, "unknown generator");
  for (int j = 0; j < mc->ndim; ++j) {
    const double lo = lows ? lows[j] : 0.0, hi = highs ? highs[j] : 1.0;
    if (!(hi > lo)) return fail(mc, PAGANI_EBADARG, "empty integration box");
  }
  return PAGANI_OK;
}

template <int D>
int genz_sample(void* ctx, const pagani_mc_args* a)
{
  const pagani_genz* g = static_cast<const pagani_genz*>(ctx);
  return pagani::internal::visit_genz<double, D>(
      g, [&](const auto& f) { return pagani::mc::launch_any<D>(f, a); });
}

pagani_mc_sample_fn genz_sampler(int ndim)
{
  switch (ndim) {
    case 2: return &genz_sample<2>;
    case 3: return &genz_sample<3>;
    case 4: return &genz_sample<4>;
    case 5: return &genz_sample<5>;
    case 6: return &genz_sample<6>;
    case 7: return &genz_sample<7>;
    case 8: return &genz_sample<8>;
    default: return nullptr;
  }
}

}  // namespace

extern "C" {

pagani_mc* pagani_mc_create(int ndim, int device)
{
  g_mc_create_error.clear();
  if (ndim < 1 || ndim > PAGANI_MC_MAX_DIM) {
    g_mc_create_error = "ndim outside [1, 16]";
    return nullptr;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    g_mc_create_error = std::string("hipGetDeviceCount: ") + hipGetErrorString(e);
    return nullptr;
  }
  if (device < 0 && hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    g_mc_create_error = "hipGetDevice failed";
    return nullptr;
  }
  if (device >= ndev) {
    g_mc_create_error = "device index out of range";
    return nullptr;
  }
  if ((e = hipSetDevice(device)) != hipSuccess) {
    (void)hipGetLastError();
    g_mc_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return nullptr;
  }
  pagani_mc* mc = new pagani_mc();
  mc->ndim = ndim;
  mc->device = device;
  const std::size_t nxi = static_cast<std::size_t>(ndim) * (kBins + 1);
  const std::size_t nhist = static_cast<std::size_t>(kMaxGrid) * ndim * kBins;
  bool ok = hipStreamCreateWithFlags(&mc->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&mc->d_xi, nxi * sizeof(double)) == hipSuccess &&
            hipMalloc(&mc->d_part, 2 * kMaxGrid * sizeof(double)) == hipSuccess &&
            hipMalloc(&mc->d_hist, nhist * sizeof(double)) == hipSuccess &&
            hipMalloc(&mc->d_out, kOutDoubles * sizeof(double)) == hipSuccess &&
            hipHostMalloc(&mc->h_out, kOutDoubles * sizeof(double), hipHostMallocDefault) ==
                hipSuccess &&
            hipHostMalloc(&mc->h_xi, nxi * sizeof(double), hipHostMallocDefault) == hipSuccess &&
            hipEventCreate(&mc->ev0) == hipSuccess && hipEventCreate(&mc->ev1) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    g_mc_create_error = "device allocation failed";
    pagani_mc_destroy(mc);
    return nullptr;
  }
  return mc;
}

void pagani_mc_destroy(pagani_mc* mc)
{
  if (!mc) return;
  (void)hipSetDevice(mc->device);
  if (mc->stream) (void)hipStreamSynchronize(mc->stream);
  if (mc->d_xi) (void)hipFree(mc->d_xi);
  if (mc->d_part) (void)hipFree(mc->d_part);
  if (mc->d_hist) (void)hipFree(mc->d_hist);
  if (mc->d_out) (void)hipFree(mc->d_out);
  if (mc->h_out) (void)hipHostFree(mc->h_out);
  if (mc->h_xi) (void)hipHostFree(mc->h_xi);
  if (mc->ev0) (void)hipEventDestroy(mc->ev0);
  if (mc->ev1) (void)hipEventDestroy(mc->ev1);
  if (mc->stream) (void)hipStreamDestroy(mc->stream);
  delete mc;
}

const char* pagani_mc_last_error(const pagani_mc* mc)
{
  return mc ? mc->error.c_str()
            : (g_mc_create_error.empty() ? "null m-Cubes handle" : g_mc_create_error.c_str());
}

int pagani_mc_set_timing(pagani_mc* mc, int enable)
{
  if (!mc) return PAGANI_EBADARG;
  mc->timing = enable != 0;
  return PAGANI_OK;
}

int pagani_mc_get_stats(const pagani_mc* mc, pagani_mc_stats* out)
{
  if (!mc || !out) return PAGANI_EBADARG;
  *out = mc->stats;
  return PAGANI_OK;
}

int pagani_mc_reset_stats(pagani_mc* mc)
{
  if (!mc) return PAGANI_EBADARG;
  mc->stats = pagani_mc_stats{};
  return PAGANI_OK;
}

int pagani_mc_integrate_cb(pagani_mc* mc, pagani_mc_sample_fn sample, void* ctx, double epsrel,
                           double epsabs, const double* lows, const double* highs,
                           const pagani_mc_params* params, pagani_mc_result* out)
{
  if (!mc) return PAGANI_EBADARG;
  if (!out || !sample) return fail(mc, PAGANI_EBADARG, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->status = 1;
  int rc = validate(mc, params, lows, highs);
  if (rc) return rc;
  MC_TRY(mc, hipSetDevice(mc->device));
  const auto t0 = std::chrono::steady_clock::now();
  mc->ti.clear();
  mc->tsi.clear();
  double ncall = params->ncall;
  int titer = params->total_iters;
  do {
    out->runs++;
    out->ncall = ncall;
    rc = run(mc, sample, ctx, epsrel, epsabs, lows, highs, *params, ncall, titer, out);
    if (rc) {
      out->status = 1;
      return rc;
    }
  } while (params->adjust_params && out->status == 1 && pagani::mc::adjust_params(ncall, titer));
  mc->stats.integrals++;
  mc->stats.wall_ms +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return PAGANI_OK;
}

int pagani_mc_integrate_genz(pagani_mc* mc, const pagani_genz* f, double epsrel,
                             double epsabs, const double* lows, const double* highs,
                             const pagani_mc_params* params, pagani_mc_result* out)
{
  if (!mc) return PAGANI_EBADARG;
  if (!f || !out) return fail(mc, PAGANI_EBADARG, "null argument");
  if (f->ndim != mc->ndim) return fail(mc, PAGANI_EBADARG, "descriptor ndim differs from the handle's");
  const pagani_mc_sample_fn fn = genz_sampler(f->ndim);
  if (!fn) return fail(mc, PAGANI_EUNSUPPORTED, "Genz descriptors cover ndim 2..8");
  if (f->family < PAGANI_GENZ_OSCILLATORY || f->family > PAGANI_GENZ_DISCONTINUOUS)
    return fail(mc, PAGANI_EBADARG, "unknown Genz family");
  pagani_genz g = *f;
  return pagani_mc_integrate_cb(mc, fn, &g, epsrel, epsabs, lows, highs, params, out);
}

int pagani_mc_trace(const pagani_mc* mc, double* ti, double* tsi, size_t cap, size_t* len)
{
  if (!mc) return PAGANI_EBADARG;
  const std::size_t n = mc->ti.size();
  for (std::size_t i = 0; i < n && i < cap; ++i) {
    if (ti) ti[i] = mc->ti[i];
    if (tsi) tsi[i] = mc->tsi[i];
  }
  if (len) *len = n;
  return PAGANI_OK;
}

int pagani_mc_grid(const pagani_mc* mc, double* xi)
{
  if (!mc || !xi) return PAGANI_EBADARG;
  if (mc->grid.empty()) return PAGANI_EBADARG;
  std::memcpy(xi, mc->grid.data(), mc->grid.size() * sizeof(double));
  return PAGANI_OK;
}

int pagani_mc_sample_genz(pagani_mc* mc, const pagani_genz* f, double ncall, const double* lows,
                          const double* highs, const double* xi, uint64_t key, int generator,
                          int hist_dims, double* ti, double* tsi, double* hist)
{
  if (!mc) return PAGANI_EBADARG;
  if (!f || !xi || !ti || !tsi || (hist_dims && !hist))
    return fail(mc, PAGANI_EBADARG, "null argument");
  if (f->ndim != mc->ndim) return fail(mc, PAGANI_EBADARG, "descriptor ndim differs from the handle's");
  if (hist_dims != 0 && hist_dims != 1 && hist_dims != mc->ndim)
    return fail(mc, PAGANI_EBADARG, "hist_dims must be 0, 1 or ndim");
  const pagani_mc_sample_fn fn = genz_sampler(f->ndim);
  if (!fn) return fail(mc, PAGANI_EUNSUPPORTED, "Genz descriptors cover ndim 2..8");
  pagani_mc_params p{};
  p.ncall = ncall;
  p.generator = generator;
  int rc = validate(mc, &p, lows, highs);
  if (rc) return rc;
  MC_TRY(mc, hipSetDevice(mc->device));
  pagani::mc::Setup s;
  pagani::mc::setup(s, mc->ndim, ncall, lows, highs);
  if ((rc = check_setup(mc, s))) return rc;
  std::vector<double> g(xi, xi + static_cast<std::size_t>(mc->ndim) * (kBins + 1));
  if ((rc = upload_grid(mc, g))) return rc;
  const Launch l = shape(static_cast<std::uint64_t>(s.ncubes));
  pagani_mc_args a;
  fill_args(mc, s, l, generator, a);
  a.hist_dims = static_cast<std::uint32_t>(hist_dims);
  a.key = key;
  pagani_genz desc = *f;
  if ((rc = sample_pass(mc, fn, &desc, a, static_cast<std::uint64_t>(s.calls)))) return rc;
  *ti = mc->h_out[0];
  *tsi = mc->h_out[1];
  if (hist_dims)
    std::memcpy(hist, mc->h_out + 2, static_cast<std::size_t>(hist_dims) * kBins * sizeof(double));
  return PAGANI_OK;
}

}  // extern "C"
