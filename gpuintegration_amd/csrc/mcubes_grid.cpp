// mcubes_grid.cpp -- see mcubes_grid.hpp.  Built with -ffp-contract=off.
#include "mcubes_grid.hpp"

#include <cmath>
#include <cstddef>

namespace pagani {
namespace mc {

namespace {

constexpr int kNd = PAGANI_MC_BINS;
constexpr double kAlpha = 1.5;  // Internal_Vegas_Params alph, vegas_utils.cuh:158

// rebin (vegasT.cuh:617-637) on one axis: xi[0..nd] with xi[0] = 0, weights
// r[1..nd]; the new right edges put rc of the weight in every bin, linear
// inside the old bins.
void rebin(double rc, int nd, const double* r, double* xin, double* xi)
{
  int k = 0;
  double dr = 0.0, xn = 0.0, xo = 0.0;
  for (int i = 1; i < nd; ++i) {
    while (rc > dr) dr += r[++k];
    if (k > 1) xo = xi[k - 1];
    xn = xi[k];
    dr -= rc;
    xin[i] = xn - (xn - xo) * dr / r[k];
  }
  for (int i = 1; i < nd; ++i) xi[i] = xin[i];
  xi[nd] = 1.0;
}

}  // namespace

void setup(Setup& s, int ndim, double ncall, const double* lows, const double* highs)
{
  const std::size_t ng = static_cast<std::size_t>(std::pow(ncall / 2.0 + 0.25, 1.0 / ndim));
  double k = 1;
  for (int i = 1; i < ndim; ++i) k *= ng;
  k *= ng;
  s.ng = ng;
  s.ncubes = k;
  const double q = ncall / k;  // IMAX(ncall / k, 2), stored into an int
  s.npg = static_cast<std::uint64_t>(static_cast<int>(q > 2 ? q : 2));
  s.calls = static_cast<double>(s.npg) * k;
  const double dxg = 1.0 / ng;
  double dv2g = 1;
  for (int i = 1; i <= ndim; ++i) dv2g *= dxg;
  const double npg = static_cast<double>(s.npg);
  s.dv2g = (s.calls * dv2g * s.calls * dv2g) / npg / npg / (npg - 1.0);
  s.xnd = kNd;
  s.dxg = dxg * s.xnd;
  s.xjac = 1.0 / s.calls;
  for (int j = 0; j < ndim; ++j) {
    s.lo[j] = lows ? lows[j] : 0.0;
    s.dx[j] = (highs ? highs[j] : 1.0) - s.lo[j];
    s.xjac *= s.dx[j];
  }
}

void grid_init(int ndim, double* xi)
{
  double r[kNd + 1], xin[kNd + 1];
  for (int i = 1; i <= kNd; ++i) r[i] = 1.0;
  for (int j = 0; j < ndim; ++j) {
    double* x = xi + static_cast<std::size_t>(j) * (kNd + 1);
    for (int i = 0; i <= kNd; ++i) x[i] = 0.0;
    x[1] = 1.0;
    rebin(1 / static_cast<double>(kNd), kNd, r, xin, x);
    x[0] = 0.0;
  }
}

void refine(int ndim, int hist_dims, double* hist, double* xi)
{
  const double xnd = kNd;
  double r[kNd + 1], xin[kNd + 1];
  for (int j = 0; j < hist_dims; ++j) {
    double* h = hist + static_cast<std::size_t>(j) * kNd - 1;  // bins h[1..kNd]
    double xo = h[1], xn = h[2], rc, dt;
    h[1] = (xo + xn) / 2.0;
    dt = h[1];
    for (int i = 2; i < kNd; ++i) {
      rc = xo + xn;
      xo = xn;
      xn = h[i + 1];
      h[i] = (rc + xn) / 3.0;
      dt += h[i];
    }
    h[kNd] = (xo + xn) / 2.0;
    dt += h[kNd];
    if (dt > 0.0) {
      rc = 0.0;
      for (int i = 1; i <= kNd; ++i) {
        r[i] = std::pow((1.0 - h[i] / dt) / (std::log(dt) - std::log(h[i])), kAlpha);
        rc += r[i];
      }
      rebin(rc / xnd, kNd, r, xin, xi + static_cast<std::size_t>(j) * (kNd + 1));
    }
  }
  if (hist_dims == 1)
    for (int j = 1; j < ndim; ++j)
      for (int i = 1; i <= kNd; ++i) xi[static_cast<std::size_t>(j) * (kNd + 1) + i] = xi[i];
}

void combine(double ti, double tsi, int iteration, double& si, double& schi, double& swgt,
             double& estimate, double& chi_sq, double& errorest)
{
  const double wgt = 1.0 / tsi;
  si += wgt * ti;
  schi += wgt * ti * ti;
  swgt += wgt;
  estimate = si / swgt;
  chi_sq = (schi - si * estimate) / (static_cast<double>(iteration) - 0.9999);
  if (chi_sq < 0.0) chi_sq = 0.0;
  errorest = std::sqrt(1.0 / swgt);
}

int get_status(double estimate, double errorest, int iteration, double epsrel, double epsabs)
{
  const bool ok = std::fabs(errorest / estimate) <= epsrel || errorest <= epsabs;
  return (ok && iteration >= 5) ? 0 : 1;
}

bool adjust_params(double& ncall, int& total_iters)
{
  if (ncall >= 8.e9 && total_iters >= 100) return false;
  if (ncall >= 8.e9) {
    total_iters += 10;
    return true;
  }
  if (ncall >= 1.e9) {
    ncall += 1.e9;
    return true;
  }
  ncall *= 10.;
  return true;
}

}  // namespace mc
}  // namespace pagani
