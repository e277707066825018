/*
 * mcubes_oracle.c -- CPU restatement of the m-Cubes VEGAS integrator.
 *
 * TEST INFRASTRUCTURE ONLY (see mcubes_oracle.h).  Plain C99, sequential:
 * cubes in index order, every sum a left-to-right loop.
 *
 * Reference files restated here (paths relative to the reference tree):
 *   cuda/mcubes/vegasT.cuh            get_indx, the sampling kernels, rebin,
 *                                     the vegas() driver, simple_integrate
 *   cuda/mcubes/vegasT1D.cuh          one-axis grid refinement
 *   cuda/mcubes/util/vegas_utils.cuh  Custom_generator, ComputeNcubes,
 *                                     GetStatus, AdjustParams
 *
 * Arithmetic notes.  The reference's device code is compiled by nvcc, which
 * contracts a*b + c into one fused multiply-add; the three point-mapping
 * expressions of Setup_Integrand_Eval (vegasT.cuh:219, :231, :238) are
 * written here as fma() accordingly.  Everything else is plain IEEE double
 * in the reference's operation order (this file is compiled without
 * contraction).
 */
#include "mcubes_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pagani_oracle.h"

#define NB ORC_MC_BINS

/* ======================================================================== */
/* Random numbers                                                            */
/* ======================================================================== */

/* Custom_generator::operator() (vegas_utils.cuh:34-40): the ANSI C LCG
 * s' = (1103515245 s + 12345) mod 2^31, computed in 32-bit unsigned
 * arithmetic and masked to 31 bits; the double is s' / 2^31. */
uint32_t orc_mc_lcg_step(uint32_t s)
{
  const uint32_t t = 1103515245u * s + 12345u;
  return t & 0x7fffffffu;
}

/* SplitMix64 (Steele, Lea and Flood 2014; Vigna's public-domain
 * splitmix64.c): the seeding generator of the default streams. */
uint64_t orc_mc_splitmix64(uint64_t* z)
{
  uint64_t x = (*z += 0x9e3779b97f4a7c15ull);
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

/* Key of iteration `it` (1-based) of a run seeded with `seed`. */
uint64_t orc_mc_iter_key(uint64_t seed, int it)
{
  uint64_t z = seed ^ ((uint64_t)(uint32_t)it << 32);
  return orc_mc_splitmix64(&z);
}

/* Stream of one cube.  LCG: the reference's Custom_generator, reseeded with
 * the cube index before each cube (vegasT.cuh:351-353) -- the same numbers
 * every iteration.  Default: xoroshiro128++ (Blackman and Vigna 2018)
 * seeded by two SplitMix64 outputs of key + cube * gamma; the reference's
 * default, curand's XORWOW seeded from the clock, is not reproducible and
 * is replaced by this counter-seeded stream (results depend on the seed and
 * the cube numbering only, not on the launch shape). */
void orc_mc_rng_init(orc_mc_rng* g, int kind, uint64_t key, uint64_t cube)
{
  g->kind = kind;
  g->lcg = (uint32_t)cube;
  uint64_t z = key + (uint64_t)(uint32_t)cube * 0x9e3779b97f4a7c15ull;
  g->s0 = orc_mc_splitmix64(&z);
  g->s1 = orc_mc_splitmix64(&z);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xoro_next(orc_mc_rng* g)
{
  const uint64_t s0 = g->s0;
  uint64_t s1 = g->s1;
  const uint64_t r = rotl64(s0 + s1, 17) + s0;
  s1 ^= s0;
  g->s0 = rotl64(s0, 49) ^ s1 ^ (s1 << 21);
  g->s1 = rotl64(s1, 28);
  return r;
}

/* One sample's n numbers in axis order.  Default stream: each 64-bit output
 * gives two, from its low then its high 32 bits x: (x | 1) 2^-32, the
 * midpoints of 2^31 equal cells of (0, 1); with n odd the last high half is
 * unused. */
void orc_mc_rng_fill(orc_mc_rng* g, int n, double* r)
{
  if (g->kind == ORC_MC_RNG_LCG) {
    for (int j = 0; j < n; ++j) {
      g->lcg = orc_mc_lcg_step(g->lcg);
      r[j] = (double)g->lcg / 2147483648.0;
    }
    return;
  }
  for (int j = 0; j < n; j += 2) {
    const uint64_t v = xoro_next(g);
    r[j] = (double)((uint32_t)v | 1u) * 0x1.0p-32;
    if (j + 1 < n) r[j + 1] = (double)((uint32_t)(v >> 32) | 1u) * 0x1.0p-32;
  }
}

/* ======================================================================== */
/* Grid                                                                      */
/* ======================================================================== */

/* get_indx (vegasT.cuh:137-156): mixed-radix digits of m in base ng, most
 * significant first, plus one. */
void orc_mc_get_indx(uint64_t m, uint32_t* kg, int d, uint32_t ng)
{
  uint64_t dp[ORC_MC_MAXDIM];
  dp[0] = 1;
  for (int j = 1; j < d; ++j) dp[j] = dp[j - 1] * ng;
  for (int j = 0; j < d; ++j) {
    const uint64_t t0 = dp[d - j - 1];
    const uint64_t t1 = m / t0;
    kg[j] = (uint32_t)(1 + t1);
    m -= t1 * t0;
  }
}

/* rebin (vegasT.cuh:617-637): new right edges so that each new bin holds
 * rc of the weights r[1..nd], interpolating linearly inside old bins. */
void orc_mc_rebin(double rc, int nd, const double* r, double* xin, double* xi)
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

/* Starting grid (vegasT.cuh:677-680, :718-723): xi[1] = 1 on every axis,
 * then one rebin of unit weights with rc = 1/nd. */
void orc_mc_grid_init(int d, double* xi)
{
  double r[NB + 1], xin[NB + 1];
  for (int i = 1; i <= NB; ++i) r[i] = 1.0;
  for (int j = 0; j < d; ++j) {
    double* x = xi + (size_t)j * (NB + 1);
    for (int i = 0; i <= NB; ++i) x[i] = 0.0;
    x[1] = 1.0;
    orc_mc_rebin(1 / (double)NB, NB, r, xin, x);
    x[0] = 0.0;
  }
}

/* Grid refinement after an adjusting iteration (vegasT.cuh:880-925): a
 * three-point smoothing of each axis' histogram, weights
 * r_i = ((1 - d_i/dt) / (ln dt - ln d_i))^alpha, then rebin.  hd = 1 is
 * vegasT1D.cuh:606-636: axis 0 only, its grid copied to the other axes.
 * hist is smoothed in place. */
void orc_mc_refine(int d, int hd, double* hist, double* xi)
{
  const double xnd = NB;
  double r[NB + 1], xin[NB + 1];
  for (int j = 0; j < hd; ++j) {
    double* h = hist + (size_t)j * NB - 1; /* h[1..NB] */
    double xo = h[1], xn = h[2], rc, dt;
    h[1] = (xo + xn) / 2.0;
    dt = h[1];
    for (int i = 2; i < NB; ++i) {
      rc = xo + xn;
      xo = xn;
      xn = h[i + 1];
      h[i] = (rc + xn) / 3.0;
      dt += h[i];
    }
    h[NB] = (xo + xn) / 2.0;
    dt += h[NB];
    if (dt > 0.0) {
      rc = 0.0;
      for (int i = 1; i <= NB; ++i) {
        r[i] = pow((1.0 - h[i] / dt) / (log(dt) - log(h[i])), ORC_MC_ALPH);
        rc += r[i];
      }
      orc_mc_rebin(rc / xnd, NB, r, xin, xi + (size_t)j * (NB + 1));
    }
  }
  if (hd == 1)
    for (int j = 1; j < d; ++j)
      for (int i = 1; i <= NB; ++i) xi[(size_t)j * (NB + 1) + i] = xi[i];
}

/* ======================================================================== */
/* Sampling                                                                  */
/* ======================================================================== */

/* Constants of vegas() (vegasT.cuh:686-716; ComputeNcubes and
 * Compute_samples_per_cube in vegas_utils.cuh repeat them). */
void orc_mc_setup_init(orc_mc_setup* s, int d, double ncall, const double* lows,
                       const double* highs)
{
  const size_t ng = (size_t)pow(ncall / 2.0 + 0.25, 1.0 / d);
  double k = 1;
  for (int i = 1; i < d; ++i) k *= ng;
  k *= ng;
  s->ng = (uint32_t)ng;
  s->ncubes = k;
  const double q = ncall / k;
  s->npg = (uint32_t)(int)(q > 2 ? q : 2); /* IMAX(ncall / k, 2) into an int */
  s->calls = (double)s->npg * k;
  double dxg = 1.0 / ng, dv2g = 1;
  for (int i = 1; i <= d; ++i) dv2g *= dxg;
  s->dv2g = (s->calls * dv2g * s->calls * dv2g) / s->npg / s->npg / (s->npg - 1.0);
  s->xnd = NB;
  s->dxg = dxg * s->xnd;
  s->xjac = 1.0 / s->calls;
  for (int j = 0; j < d; ++j) {
    s->lo[j] = lows ? lows[j] : 0.0;
    s->dx[j] = (highs ? highs[j] : 1.0) - s->lo[j];
    s->xjac *= s->dx[j];
  }
}

/* One pass of vegas_kernel (hd > 0) or vegas_kernelF (hd = 0) over every
 * cube, vegasT.cuh:197-398 / :490-612. */
void orc_mc_sample(int fam, int d, const double* prm, const orc_mc_setup* s,
                   const double* xi, uint64_t key, int generator, int hd, double* ti,
                   double* tsi, double* hist)
{
  const size_t nc = (size_t)s->ncubes;
  double fbg = 0.0, f2bg = 0.0;
  uint32_t kg[ORC_MC_MAXDIM];
  int ia[ORC_MC_MAXDIM];
  double x[ORC_MC_MAXDIM];
  orc_mc_get_indx(0, kg, d, s->ng);
  for (size_t m = 0; m < nc; ++m) {
    orc_mc_rng g;
    orc_mc_rng_init(&g, generator, key, m);
    double fb = 0.0, f2b = 0.0;
    for (uint32_t k = 1; k <= s->npg; ++k) {
      double wgt = s->xjac;
      double ran[ORC_MC_MAXDIM];
      orc_mc_rng_fill(&g, d, ran);
      for (int j = 0; j < d; ++j) {
        const double* xj = xi + (size_t)j * (NB + 1);
        const double ran00 = ran[j];
        const double xn = fma((double)kg[j] - ran00, s->dxg, 1.0);
        int a = (int)xn;
        if (a > NB) a = NB;
        if (a < 1) a = 1;
        ia[j] = a;
        double xo, rc;
        if (a > 1) {
          xo = xj[a] - xj[a - 1];
          rc = fma(xn - a, xo, xj[a - 1]);
        } else {
          xo = xj[a];
          rc = (xn - a) * xo;
        }
        x[j] = fma(rc, s->dx[j], s->lo[j]);
        wgt *= xo * s->xnd;
      }
      const double f = wgt * orc_integrand(fam, d, prm, x);
      const double f2 = f * f;
      fb += f;
      f2b += f2;
      for (int j = 0; j < hd; ++j) hist[(size_t)j * NB + ia[j] - 1] += f2;
    }
    f2b = sqrt(f2b * s->npg);
    f2b = (f2b - fb) * (f2b + fb);
    if (f2b <= 0.0) f2b = ORC_MC_TINY;
    fbg += fb;
    f2bg += f2b;
    /* odometer increment, vegasT.cuh:386-391 */
    for (int k = d - 1; k >= 0; --k) {
      kg[k] %= s->ng;
      if (++kg[k] != 1) break;
    }
  }
  *ti = fbg;
  *tsi = f2bg;
}

/* ======================================================================== */
/* Driver                                                                    */
/* ======================================================================== */

/* GetStatus / PrecisionAchieved, vegas_utils.cuh:233-254 */
static int get_status(double est, double err, int it, double epsrel, double epsabs)
{
  const int ok = fabs(err / est) <= epsrel || err <= epsabs;
  return (ok && it >= 5) ? 0 : 1;
}

/* AdjustParams, vegas_utils.cuh:283-298 */
static int adjust_params(double* ncall, int* titer)
{
  if (*ncall >= 8.e9 && *titer >= 100) return 0;
  if (*ncall >= 8.e9) {
    *titer += 10;
    return 1;
  }
  if (*ncall >= 1.e9) {
    *ncall += 1.e9;
    return 1;
  }
  *ncall *= 10.;
  return 1;
}

static void trace_push(orc_mc_trace* tr, double ti, double tsi)
{
  if (!tr || tr->len >= tr->cap) return;
  tr->ti[tr->len] = ti;
  tr->tsi[tr->len] = tsi;
  tr->len++;
}

/* One vegas() run, vegasT.cuh:640-1000. */
static int vegas_run(int fam, int d, const double* prm, double epsrel, double epsabs,
                     const double* lows, const double* highs, const orc_mc_params* p,
                     double ncall, int titer, orc_mc_result* out, orc_mc_trace* tr,
                     double* xi)
{
  orc_mc_setup s;
  orc_mc_setup_init(&s, d, ncall, lows, highs);
  const int hd = p->one_axis ? 1 : d;
  double* hist = (double*)malloc(sizeof(double) * (size_t)hd * NB);
  if (!hist) return -1;
  orc_mc_grid_init(d, xi);
  double si = 0.0, swgt = 0.0, schi = 0.0;
  int it;
  for (it = 1; it <= p->adjust_iters && out->status == 1; out->iters++, it++) {
    memset(hist, 0, sizeof(double) * (size_t)hd * NB);
    double ti, tsi;
    orc_mc_sample(fam, d, prm, &s, xi, orc_mc_iter_key(p->seed, it), p->generator, hd, &ti,
                  &tsi, hist);
    tsi *= s.dv2g;
    trace_push(tr, ti, tsi);
    if (it > p->skip_iters) {
      const double wgt = 1.0 / tsi;
      si += wgt * ti;
      schi += wgt * ti * ti;
      swgt += wgt;
      out->estimate = si / swgt;
      out->chi_sq = (schi - si * out->estimate) / ((double)it - 0.9999);
      if (out->chi_sq < 0.0) out->chi_sq = 0.0;
      out->errorest = sqrt(1.0 / swgt);
      out->status = get_status(out->estimate, out->errorest, it, epsrel, epsabs);
    }
    orc_mc_refine(d, hd, hist, xi);
  }
  for (it = p->adjust_iters + 1; it <= titer && out->status; out->iters++, it++) {
    double ti, tsi;
    orc_mc_sample(fam, d, prm, &s, xi, orc_mc_iter_key(p->seed, it), p->generator, 0, &ti,
                  &tsi, NULL);
    tsi *= s.dv2g;
    trace_push(tr, ti, tsi);
    const double wgt = 1.0 / tsi;
    si += wgt * ti;
    schi += wgt * ti * ti;
    swgt += wgt;
    out->estimate = si / swgt;
    out->chi_sq = (schi - si * out->estimate) / (it - 0.9999);
    if (out->chi_sq < 0.0) out->chi_sq = 0.0;
    out->errorest = sqrt(1.0 / swgt);
    out->status = get_status(out->estimate, out->errorest, it, epsrel, epsabs);
  }
  free(hist);
  return 0;
}

/* integrate / simple_integrate, vegasT.cuh:1003-1026 / :1082-1112 */
int orc_mc_integrate(int fam, int d, const double* prm, double epsrel, double epsabs,
                     const double* lows, const double* highs, const orc_mc_params* p,
                     orc_mc_result* out, orc_mc_trace* trace, double* xi_out)
{
  if (d < 1 || d > ORC_MC_MAXDIM) return -1;
  memset(out, 0, sizeof(*out));
  out->status = 1;
  double* xi = (double*)malloc(sizeof(double) * (size_t)d * (NB + 1));
  if (!xi) return -1;
  double ncall = p->ncall;
  int titer = p->total_iters;
  do {
    out->runs++;
    out->ncall = ncall;
    if (vegas_run(fam, d, prm, epsrel, epsabs, lows, highs, p, ncall, titer, out, trace, xi)) {
      free(xi);
      return -1;
    }
  } while (p->adjust_params && out->status == 1 && adjust_params(&ncall, &titer));
  if (xi_out) memcpy(xi_out, xi, sizeof(double) * (size_t)d * (NB + 1));
  free(xi);
  return 0;
}
