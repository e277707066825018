/*
 * mcubes_oracle.h -- CPU restatement of the m-Cubes VEGAS integrator.
 *
 * TEST INFRASTRUCTURE ONLY, like pagani_oracle.h: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it;
 * the product library never links it and never falls back to it.
 *
 * Restated from the reference's cuda/mcubes/vegasT.cuh (vegas(),
 * vegas_kernel, vegas_kernelF, rebin, get_indx), cuda/mcubes/vegasT1D.cuh
 * (the one-axis grid variant) and cuda/mcubes/util/vegas_utils.cuh
 * (Custom_generator, GetStatus, AdjustParams).  Each function cites its
 * lines in mcubes_oracle.c.
 *
 * Conventions (shared with the GPU library):
 *   - the importance grid of axis j is xi[j*(ORC_MC_BINS+1) + i],
 *     i = 0..ORC_MC_BINS: the right edge of bin i in [0, 1], with
 *     xi[j*(ORC_MC_BINS+1)] = 0 (the reference leaves index 0 unused and
 *     branches on bin 1 instead; a left edge of 0 gives the same values);
 *   - the bin histogram of axis j is hist[j*ORC_MC_BINS + (i-1)] for bin i;
 *   - cubes are numbered in get_indx order (last axis fastest).
 */
#ifndef MCUBES_ORACLE_H
#define MCUBES_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MC_BINS 500     /* Internal_Vegas_Params ndmx */
#define ORC_MC_MAXDIM 16
#define ORC_MC_TINY 1.0e-30 /* TINY */
#define ORC_MC_ALPH 1.5     /* Internal_Vegas_Params alph */

enum { ORC_MC_RNG_DEFAULT = 0, ORC_MC_RNG_LCG = 1 };

/* random streams: one per (iteration key, cube) */
typedef struct {
  int kind;
  uint32_t lcg;
  uint64_t s0, s1;
} orc_mc_rng;

uint32_t orc_mc_lcg_step(uint32_t s);
uint64_t orc_mc_splitmix64(uint64_t* z);
uint64_t orc_mc_iter_key(uint64_t seed, int it);
void orc_mc_rng_init(orc_mc_rng* g, int kind, uint64_t key, uint64_t cube);
/* the n numbers of one sample, in axis order */
void orc_mc_rng_fill(orc_mc_rng* g, int n, double* r);

/* kg[0..d-1] = 1-based stratum coordinates of cube m (last axis fastest) */
void orc_mc_get_indx(uint64_t m, uint32_t* kg, int d, uint32_t ng);

/* rebin one axis: xi[0..nd] (xi[0] = 0), r[1..nd], xin scratch [nd+1] */
void orc_mc_rebin(double rc, int nd, const double* r, double* xin, double* xi);

/* the uniform starting grid of every axis */
void orc_mc_grid_init(int d, double* xi);

/* grid refinement from the bin histograms (smoothing, weights, rebin);
 * hd = 1: axis 0's grid is refined and copied to every axis */
void orc_mc_refine(int d, int hd, double* hist, double* xi);

typedef struct {
  double ncall;
  int total_iters;  /* titer */
  int adjust_iters; /* itmax */
  int skip_iters;   /* skip */
  int generator;    /* ORC_MC_RNG_* */
  int one_axis;     /* vegasT1D.cuh grid */
  int adjust_params;/* simple_integrate's retries */
  uint64_t seed;
} orc_mc_params;

typedef struct {
  uint32_t ng, npg;
  double ncubes, calls, dv2g, dxg, xnd, xjac;
  double lo[ORC_MC_MAXDIM], dx[ORC_MC_MAXDIM];
} orc_mc_setup;

void orc_mc_setup_init(orc_mc_setup* s, int d, double ncall, const double* lows,
                       const double* highs);

/* one sampling pass over every cube: ti = sum fb, tsi = sum f2b (not yet
 * scaled by dv2g); hist (hd axes, may be NULL when hd = 0) accumulates
 * f^2 per bin. */
void orc_mc_sample(int fam, int d, const double* prm, const orc_mc_setup* s,
                   const double* xi, uint64_t key, int generator, int hd, double* ti,
                   double* tsi, double* hist);

typedef struct {
  double estimate, errorest, chi_sq;
  int status;
  size_t iters;
  size_t runs;     /* vegas() calls (simple_integrate retries + 1) */
  double ncall;    /* ncall of the last run */
} orc_mc_result;

/* per-iteration record: ti, tsi (scaled) of every iteration */
typedef struct {
  size_t cap, len;
  double* ti;
  double* tsi;
} orc_mc_trace;

/* vegas() (vegasT.cuh:640-1000) with simple_integrate's retries when
 * params->adjust_params; fam/prm select the integrand (orc_integrand).
 * xi_out (d*(ORC_MC_BINS+1), may be NULL) receives the final grid. */
int orc_mc_integrate(int fam, int d, const double* prm, double epsrel, double epsabs,
                     const double* lows, const double* highs, const orc_mc_params* p,
                     orc_mc_result* out, orc_mc_trace* trace, double* xi_out);

#ifdef __cplusplus
}
#endif

#endif
