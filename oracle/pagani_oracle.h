/*
 * pagani_oracle.h -- CPU restatement of the PAGANI adaptive cubature path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never
 * links it and never falls back to it.
 *
 * Every function restates the algorithm of a reference file (cited per
 * function in pagani_oracle.c).  Paths are relative to the reference tree
 * (marcpaterno/gpuintegration).
 *
 * Conventions
 *   - region bounds are SoA in unit-cube coordinates:
 *       L[j*n + t] = left coordinate, H[j*n + t] = length   (dim j, region t)
 *     (cuda/pagani/quad/GPUquad/Sub_regions.cuh:242-255)
 *   - active flags are uint8_t 0/1 (the reference stores them as double
 *     0./1., Region_characteristics.cuh:14; only the value matters)
 *   - split dimensions are int32.
 *   - integrands are selected by a family code + a parameter vector
 *     params = {a_0..a_{d-1}, u_0..u_{d-1}} (see ORC_FAM_* below).
 */
#ifndef PAGANI_ORACLE_H
#define PAGANI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NSETS 9
#define ORC_NRULES 5
#define ORC_MAXDIM 16

/* integrand families (same codes as the GPU library's pagani_family) */
enum {
  ORC_FAM_OSCILLATORY = 1,   /* cos(2*pi*u_0 + sum a_i x_i)                 */
  ORC_FAM_PRODUCT_PEAK = 2,  /* prod 1/(a_i^-2 + (x_i-u_i)^2)               */
  ORC_FAM_CORNER_PEAK = 3,   /* (1 + sum a_i x_i)^-(d+1)                    */
  ORC_FAM_GAUSSIAN = 4,      /* exp(-sum a_i^2 (x_i-u_i)^2)                 */
  ORC_FAM_C0 = 5,            /* exp(-sum a_i |x_i-u_i|)                     */
  ORC_FAM_DISCONTINUOUS = 6, /* 0 if any x_i > u_i else exp(sum a_i x_i)    */
  ORC_FAM_CONST = 10,        /* a_0                                          */
  ORC_FAM_PRODUCT = 11,      /* prod x_i   (the C-interop test's x*y)        */
  ORC_FAM_FUN6 = 12,         /* accuracy_improves_with_epsrel.cu Fun6 (6D)  */
  ORC_FAM_BOX = 13,          /* (sum x_i^2)^(a_0/2)                          */
  ORC_FAM_NAN = 14,          /* NaN everywhere                               */
  ORC_FAM_MONOMIAL = 15      /* prod x_i^{a_i} (a_i small integers)          */
};

/* ---- rule tables (Rule.cuh) -------------------------------------------- */
int orc_feval(int ndim);
/* w/scale/norm: [set*5 + rule]; lambda[5]; count[9] (points per set) */
int orc_rule_tables(int ndim, double* w, double* lambda, double* scale,
                    double* norm, int* count);
/* gen: feval*ndim generator coordinates (point-major), set: feval */
int orc_generators(int ndim, double* gen, int* set);

/* ---- integrands ---------------------------------------------------------- */
double orc_integrand(int family, int ndim, const double* params,
                     const double* x);
/* analytic value over [0,1]^d (families with a closed form); NaN otherwise */
double orc_exact(int family, int ndim, const double* params);

/* ---- per-region evaluation (Sample.cuh / Phases.cuh) --------------------- */
/* est/err/dim per region.  fdiff (n*ndim, may be NULL) receives the fourth
 * differences; rsum (n*ORC_RSUM, may be NULL) per region:
 *   [0..4]  the raw rule sums s_0..s_4 (Jacobian included, volume not)
 *   [5..9]  A_r = sum_p |w_r f_p|, the magnitudes that bound their rounding
 *   [10]    the region volume
 *   [11..13] the null-rule errors e_1, e_2, e_3 before the final branch
 *   [14]    the largest magnitude of the terms of a fourth difference,
 *           which bounds its rounding
 *   [15]    (orc_eval_regions_kform only; 0 otherwise) sum_p |w_0 J| c_p, where
 *           c_p is f's sensitivity to one rounding of its inner sum at point
 *           p (|f| times the magnitude of the exponent for the exponential
 *           families, |sin| times the angle's for the oscillatory one, ...):
 *           the floor of any FP64 evaluation of f at that point */
#define ORC_RSUM 16
int orc_eval_regions(int ndim, int family, const double* params,
                     const double* lows, const double* highs, size_t n,
                     const double* L, const double* H, double* est,
                     double* err, int* dim, double* fdiff, double* rsum);
/* The same evaluation at the library kernel's own points (pagani_oracle.c
 * explains the difference), with f and every sum in long double.  plain = 1
 * is an integrand without origin and axis scale: the reference's formula of
 * the family at x = fma(-g, h, c). */
int orc_eval_regions_kform(int ndim, int family, const double* params, int plain,
                           const double* lows, const double* highs, size_t n,
                           const double* L, const double* H, double* est,
                           double* err, int* dim, double* fdiff, double* rsum);

/* ---- region list phases -------------------------------------------------- */
void orc_uniform_split(int ndim, size_t k, double* L, double* H);
/* two-level error refinement + relerr classification (in place on err) */
void orc_refine(size_t n, const double* est, double* err,
                const double* parent_est, double epsrel, int relerr,
                uint8_t* active);
/* order-preserving compaction; returns m (number of active regions) */
size_t orc_filter(int ndim, size_t n, const double* L, const double* H,
                  const uint8_t* active, const double* est, const double* err,
                  const int* dim, double* L2, double* H2, double* pest,
                  double* perr, int* dim2);
/* bisection: m parents -> 2m children, [left m | right m] */
void orc_split(int ndim, size_t m, const double* L, const double* H,
               const int* dim, double* L2, double* H2);

/* ---- heuristic classifier (heuristic_classifier.cuh) --------------------- */
typedef struct {
  double epsrel, epsabs;
  int required_digits;
  double hist[3];
  size_t iters_collected;
  double max_budget, max_active;
  size_t total_mem;
  int ndim;
} orc_classifier;

typedef struct {
  int pass_mem, pass_errorest_budget;
  double threshold, errorest_budget_covered, percent_mem_active;
  double range_low, range_high;
  size_t num_active;
  double finished_errorest;
  double max_budget_perc_to_cover, max_active_perc;
  int threshold_passes;
} orc_class_res;

void orc_clf_init(orc_classifier* c, int ndim, double epsrel, double epsabs,
                  size_t total_mem);
void orc_clf_store(orc_classifier* c, double estimate);
int orc_clf_sig_digits_same(const orc_classifier* c);
int orc_clf_estimate_converged(const orc_classifier* c);
size_t orc_mem_required(size_t n, int ndim);
size_t orc_free_mem(size_t n, int ndim, size_t total_mem);
int orc_clf_criteria_met(const orc_classifier* c, size_t n);
void orc_clf_classify(orc_classifier* c, const double* err, size_t n,
                      double iter_errorest, double iter_finished_errorest,
                      double total_finished_errorest, uint8_t* flags,
                      orc_class_res* res);

/* ---- the integration loop (Workspace.cuh) -------------------------------- */
typedef struct {
  double estimate, errorest;
  size_t neval, nregions, nFinishedRegions;
  int status, lastPhase;
  double chi_sq;
  size_t iters;
} orc_result;

typedef struct {
  size_t cap;       /* capacity of the arrays below */
  size_t len;       /* iterations recorded */
  size_t* n;        /* regions evaluated per iteration */
  size_t* m;        /* regions kept active per iteration (0 at exit) */
  double* estimate; /* cumulative + iteration estimate */
  double* errorest; /* cumulative + iteration errorest */
} orc_trace;

/* init_k: regions per axis of the initial uniform split (0 = the
 * reference's default, 4 for d<5, 2 for d<=10, 1 otherwise).
 * user_regions: 1 selects the Sub_regions& overload's bookkeeping
 * (iters counted before the convergence check). */
int orc_integrate(int ndim, int family, const double* params, double epsrel,
                  double epsabs, const double* lows, const double* highs,
                  int relerr, size_t total_mem, size_t init_k,
                  int user_regions, orc_result* out, orc_trace* trace);

/* Options of orc_integrate_ex (all zero: orc_integrate).
 *   predict_split: the relerr flag of every iteration follows the rule of
 *     Workspace::integrate<IntegT, predict_split=true>
 *     (Workspace.cuh:206-210, 415-420) instead of `relerr`.
 *   max_iters > 0: stop before iteration max_iters runs (return value 2);
 *     the list it would evaluate is the final list.
 *   keep_list: hand the final list back (malloc'd, freed with orc_free):
 *     list_n regions, list_L/list_H SoA; when that list was evaluated (the
 *     run converged or terminated) list_est/list_err/list_dim too, else
 *     NULL. */
typedef struct {
  int predict_split;
  size_t max_iters;
  int keep_list;
  size_t list_n;
  double* list_L;
  double* list_H;
  double* list_est;
  double* list_err;
  int* list_dim;
  size_t budget_fixes; /* out: iterations that took fix_error_budget_overflow */
} orc_opts;

/* returns 0, 2 when stopped by max_iters, negative on a bad argument */
int orc_integrate_ex(int ndim, int family, const double* params, double epsrel,
                     double epsabs, const double* lows, const double* highs,
                     int relerr, size_t total_mem, size_t init_k,
                     int user_regions, orc_result* out, orc_trace* trace,
                     orc_opts* opts);
void orc_free(void* p);

/* sequential globally adaptive cubature (Cuhre-equivalent CPU baseline):
 * a heap of regions, always bisect the region with the largest error along
 * its fourth-difference axis, same degree-9 rule. */
int orc_cuhre_seq(int ndim, int family, const double* params, double epsrel,
                  double epsabs, size_t maxeval, orc_result* out);

/* number of OpenMP threads the oracle uses */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
