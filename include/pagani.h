/*
 * pagani.h -- C ABI of the MI355X-native PAGANI adaptive cubature library
 * (libpagani.so, HIP for gfx950).
 *
 * The boundary replaces the reference's C++ integrator surface
 *   Workspace<T,ndim>::integrate(...)        cuda/pagani/quad/GPUquad/Workspace.cuh:23-71
 *   numint::integration_result               common/integration_result.hh:11-21
 * and its phase classes (cited per entry point below).  Plain C types only:
 * pointers, sizes, PODs.  Every entry point returns an int status
 * (PAGANI_OK == 0, negative on error) and never throws or aborts; the text
 * of the last error of a workspace is available from pagani_ws_last_error.
 *
 * Device pointers ("_dev" entry points) are HIP device allocations on the
 * workspace's device; those entry points run on the workspace's stream and
 * return after it has drained.
 *
 * Data layout (HBM, structure of arrays, unit-cube coordinates):
 *   L[j*n + t]  left coordinate of region t along axis j   (FP64)
 *   H[j*n + t]  length of region t along axis j            (FP64)
 *   est[t], err[t]                                          (FP64)
 *   dim[t]      split axis, uint8                           (reference: int)
 *   active[t]   0/1 flag, uint8                             (reference: double)
 */
#ifndef PAGANI_H
#define PAGANI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAGANI_MIN_DIM 2
#define PAGANI_MAX_DIM 16      /* workspaces and the C++ header (Workspace<double, ndim>) */
#define PAGANI_GENZ_MAX_DIM 10 /* the precompiled Genz descriptor path */
#define PAGANI_NSETS 9
#define PAGANI_NRULES 5
#define PAGANI_MAX_SLICES 8

/* status codes */
#define PAGANI_OK 0
#define PAGANI_NOT_CONVERGED 1
#define PAGANI_SUSPENDED 2 /* iteration limit reached; the run can be resumed */
#define PAGANI_EBADARG (-1)
#define PAGANI_EHIP (-2)
#define PAGANI_ENOMEM (-3)
#define PAGANI_EUNSUPPORTED (-4)

/* evaluation precision of the integrand and the rule sums */
#define PAGANI_FP64 64
#define PAGANI_FP32 32 /* FP32 integrand, FP64 accumulation and error logic */

/* integrand families of the precompiled descriptor path */
#define PAGANI_GENZ_OSCILLATORY 1   /* cos(2 pi u_0 + sum a_i x_i)              */
#define PAGANI_GENZ_PRODUCT_PEAK 2  /* prod 1/(a_i^-2 + (x_i - u_i)^2)          */
#define PAGANI_GENZ_CORNER_PEAK 3   /* (1 + sum a_i x_i)^-(d+1)                 */
#define PAGANI_GENZ_GAUSSIAN 4      /* exp(-sum a_i^2 (x_i - u_i)^2)            */
#define PAGANI_GENZ_C0 5            /* exp(-sum a_i |x_i - u_i|)                */
#define PAGANI_GENZ_DISCONTINUOUS 6 /* 0 if some x_i > u_i else exp(sum a_i x_i) */

/* numint::integration_result, common/integration_result.hh:11-21.
 * PAGANI fills estimate, errorest, nregions, status, iters. */
typedef struct pagani_result {
  double estimate;
  double errorest;
  size_t neval;
  size_t nregions;
  size_t nFinishedRegions;
  int status; /* 0 converged, 1 not converged / memory termination */
  int lastPhase;
  double chi_sq;
  size_t iters;
} pagani_result;

/* Precompiled Genz integrand descriptor (POD).  The reference's fixed
 * functors F_k_dD (common/cuda/integrands.cuh:1303-1934) are instances.
 * ndim 2..PAGANI_GENZ_MAX_DIM; a and u hold ndim entries. */
typedef struct pagani_genz {
  int family;
  int ndim;
  double a[PAGANI_MAX_DIM];
  double u[PAGANI_MAX_DIM];
} pagani_genz;

/* Layout version of pagani_eval_consts / pagani_eval_args and of the region
 * lists they describe.  The library stamps it into every launch's constants
 * and pagani::eval_regions (compiled into the caller's code) refuses a
 * launch whose stamp differs from the version it was compiled with, so a
 * caller built against other headers fails with PAGANI_EUNSUPPORTED instead
 * of reading a list laid out differently. */
#define PAGANI_EVAL_ABI_VERSION 4

/* Uniform data of one evaluation launch: the degree-9 rule (Rule.cuh),
 * the integration volume and the launch shape.  Built by the library; an
 * evaluation callback passes it unchanged to pagani::eval_regions. */
typedef struct pagani_eval_consts {
  double w[PAGANI_NSETS][PAGANI_NRULES];     /* rule weights              */
  double scale[PAGANI_NSETS][PAGANI_NRULES]; /* null-rule scale           */
  double norm[PAGANI_NSETS][PAGANI_NRULES];  /* null-rule norm            */
  double lows[PAGANI_MAX_DIM];               /* integration volume        */
  double range[PAGANI_MAX_DIM];              /* highs - lows              */
  double jac;                                /* prod range                */
  double lambda1, lambda2, ratio;            /* 4th-difference constants  */
  const double* gtab; /* device: generator coordinates of points 1+4d..F-1,
                         point-major, ndim doubles per point               */
  int ndim, feval;
  int set_begin[PAGANI_NSETS + 1]; /* first point index of each set       */
  int slices;                      /* waves sharing one batch of 64 regions */
  int slice_begin[PAGANI_MAX_SLICES + 1]; /* generic-point slice bounds     */
  double lambda3, lambda4, lambda5; /* generator magnitudes of sets 3, 4, 8 */
  double h0[PAGANI_MAX_DIM]; /* level-form lists: every axis's unit-cube width at level 0 */
  uint32_t abi_version;      /* PAGANI_EVAL_ABI_VERSION of the library */
  uint32_t levels;           /* the list's widths are levels (pagani_eval_args) */
} pagani_eval_consts;

/* Iteration descriptor of the device-resident loop, in device memory: the
 * regions to evaluate, which SoA pair holds them, whether the iteration
 * runs at all (a loop that has stopped turns the queued iterations into
 * no-ops). */
typedef struct pagani_dev_iter {
  uint64_t n;
  uint32_t cur;         /* 0: (L, H), 1: (L2, H2) */
  uint32_t run;         /* 0: skip */
  uint32_t has_parents; /* the refinement has parents' estimates */
  uint32_t slices;      /* waves per batch of 64 regions (see consts.slices) */
  uint32_t split;       /* the list is in split form (see pagani_eval_args) */
  uint32_t pad;
} pagani_dev_iter;

/* One evaluation request: regions [0, n) of the SoA pool.  When dev is not
 * NULL the request belongs to the device-resident loop: n is an upper bound,
 * the count and the pair (L, H) or (L2, H2) are read from *dev on the
 * device, and the launch must not depend on them (pagani::eval_regions
 * handles both forms).
 *
 * A list is flat or in split form.  Flat: L, H hold the n regions' own
 * bounds, [j*n + t].  Split form (pdim not NULL; in the device-resident loop,
 * dev->split, with pdim2 for the pair (L2, H2)): the list is the two
 * children of each of n/2 parents, L, H hold the parents' bounds
 * [j*(n/2) + i] and pdim[i] the axis parent i is halved along; region t is
 * child t / (n/2) of parent t mod (n/2), the left half along that axis
 * for child 0 and the right half for child 1 (Sub_region_splitter.cuh:8-46
 * writes those children out; this library derives them where they are
 * read).
 *
 * Widths are doubles, or levels when consts.levels is set (the library's
 * own uniform start): H (H2) then points at one uint16_t per axis and row,
 * the number of times the row's width along that axis was halved from
 * consts.h0[j], so the width is h0[j] * 2^-level -- the same double that
 * halving h0[j] that many times gives (a run has at most 700 iterations,
 * each halving a region along one axis). */
typedef struct pagani_eval_args {
  size_t n;
  const double* L;
  const double* H;
  double* est;
  double* err;
  uint8_t* dim;
  void* stream; /* hipStream_t */
  pagani_eval_consts consts;
  const pagani_dev_iter* dev;
  const double* L2;
  const double* H2;
  const uint8_t* pdim;
  const uint8_t* pdim2;
  /* Optional (the debug recorder's point dump; host-driven launches only):
   * when not NULL, the value of every rule point, fevals[t * feval + p], with
   * p in the library's point order (pagani_rule_point).  pagani::eval_regions
   * writes them with a second kernel. */
  double* fevals;
} pagani_eval_args;

/* An evaluation callback launches the region-evaluation kernel for one
 * integrand on args->stream (pagani::eval_regions<ndim, IntegT> from
 * <pagani/eval_kernel.hpp> does it) and returns 0 or a negative status. */
typedef int (*pagani_eval_fn)(void* ctx, const pagani_eval_args* args);

/* Initial region list for the Sub_regions& overload (device SoA). */
typedef struct pagani_region_list {
  size_t n;
  const double* L;
  const double* H;
} pagani_region_list;

/* Per-phase kernel time (HIP events on the workspace stream, summed over
 * the calls since the last reset) and work counters. */
typedef struct pagani_stats {
  double eval_ms;
  double refine_ms;
  double finalize_ms;
  double compact_ms;
  double classify_ms;
  double wall_ms;
  uint64_t regions_evaluated;
  uint64_t eval_launches;
  uint64_t iterations;
  uint64_t classify_passes;
  uint64_t max_regions;
  uint64_t integrals;
  double alloc_ms;      /* host time spent growing device buffers */
  uint64_t allocations; /* device buffers (re)allocated */
  double device_ms;     /* time of the device-resident loop's rounds (timing on) */
  uint64_t device_iterations; /* iterations run by the device-resident loop */
  uint64_t device_regions;    /* regions it evaluated (included in regions_evaluated) */
  uint64_t device_rounds;     /* its host round trips */
  uint64_t budget_fixes;      /* iterations that took fix_error_budget_overflow */
  uint64_t graph_launches;    /* device-loop rounds replayed from a HIP graph */
  uint64_t segment_integrals; /* batch items finished by the persistent kernel
                                 (included in integrals; its regions and
                                 iterations are in regions_evaluated and
                                 iterations) */
  double segment_ms;          /* host time of the persistent kernel's launches */
} pagani_stats;

/* Result of Heuristic_classifier::classify, heuristic_classifier.cuh:22-58 */
typedef struct pagani_class_res {
  int pass_mem;
  int pass_errorest_budget;
  double threshold;
  double percent_mem_active;
  double range_low, range_high;
  size_t num_active;
  double finished_errorest;
  double max_budget_perc_to_cover;
  double max_active_perc;
  int threshold_passes;
} pagani_class_res;

typedef struct pagani_ws pagani_ws;

/* ---- rule tables (host only, no device needed) ------------------------------ */
/* quad::Rule<T>::Init / Rule9Generate, Rule.cuh:24-270, 370-510: weights,
 * null-rule scale/norm ([set*5 + rule], 45 each), lambda[5], the first point
 * index of each set (set_begin[10]) and, when gtab is not NULL, the generator
 * coordinates of points 1+4*ndim .. feval-1 (ndim doubles each).  Returns the
 * number of points per region, or a negative status. */
int pagani_rule_tables(int ndim, double* w, double* scale, double* norm, double* lambda,
                       int* set_begin, double* gtab);

/* ---- workspace ----------------------------------------------------------- */
/* Replaces the Workspace constructor (Workspace.cuh:50) and the rule setup of
 * Cubature_rules (PaganiUtils.cuh:52-96).  device < 0 selects the current
 * device.  Returns NULL on failure. */
pagani_ws* pagani_ws_create(int ndim, int device);
/* why the calling thread's last pagani_ws_create returned NULL */
const char* pagani_create_error(void);
void pagani_ws_destroy(pagani_ws* ws);
const char* pagani_ws_last_error(const pagani_ws* ws);
int pagani_ws_ndim(const pagani_ws* ws);
/* Total device memory assumed by the classifier's memory model
 * (total_device_mem(), heuristic_classifier.cuh:89-94, hard-coded 16e9 in
 * the reference).  0 selects the device's HBM size (the default). */
int pagani_ws_set_device_mem(pagani_ws* ws, size_t total_bytes);
size_t pagani_ws_device_mem(const pagani_ws* ws);
/* Bytes the classifier's memory model charges for splitting n regions:
 * PAGANI_MEM_REFERENCE (default) is device_mem_required_for_full_split
 * (heuristic_classifier.cuh:96-133, 8(10d+8)+16 bytes per region);
 * PAGANI_MEM_ACTUAL is this library's own footprint (lists are held as
 * parents plus the axis each is halved along: the list's parents, the next
 * list's parents, the next list's per-region arrays and the parents'
 * estimates: 24d+64 bytes per region), which lets a pool grow about 2.7x
 * larger at 8D on the same device before classification is forced. */
#define PAGANI_MEM_REFERENCE 0
#define PAGANI_MEM_ACTUAL 1
int pagani_ws_set_mem_model(pagani_ws* ws, int model);
/* Device-resident loop (on by default): while the region list is at most
 * 2^20 regions and the memory model is below a tenth of its budget,
 * iterations are decided on the device and enqueued in rounds of growing
 * length, with one host round trip per round instead of one per
 * iteration.  Results are identical either way; the host takes over for
 * the heuristic classifier, the error-budget fix and larger lists. */
int pagani_ws_set_device_loop(pagani_ws* ws, int enable);
/* Full-length rounds of the device-resident loop (16 iterations, 48
 * launches) are captured once as a HIP graph and replayed for later rounds
 * of the same shape (enable = 1, the default); results are identical
 * either way.  No reference counterpart. */
int pagani_ws_set_graphs(pagani_ws* ws, int enable);
/* The heuristic classifier's threshold search (heuristic_classifier.cuh:
 * 362-451) runs on the device by default: the host queues rounds of probe
 * passes and replays of the reference's loop and reads the outcome once.
 * enable = 0 runs the same replay on the host, one probe pass per round trip.
 * Both classify identically. */
int pagani_ws_set_device_search(pagani_ws* ws, int enable);
/* The library's uniform start keeps each region's widths as levels (two
 * bytes per axis: the number of halvings of 1/k, pagani_eval_args) instead of
 * doubles, so the filter + split and the evaluation move 6 bytes less per
 * axis and region; results are the same bits.  enable = 0 keeps doubles
 * (lists given by the caller always keep doubles). */
int pagani_ws_set_level_form(pagani_ws* ws, int enable);
/* waves per batch of 64 regions in the evaluation kernel; 0 = automatic */
int pagani_ws_set_slices(pagani_ws* ws, int slices);
/* No reference counterpart (the reference fixes n x 64 threads,
 * PaganiUtils.cuh:313-331).  Automatic launch shapes come from the table
 * the autotuner (gpuintegration_amd/tune.py) measured for the integrand's
 * family, dimension and precision, per size class of the list, where it
 * has an entry (enable = 1, the default), else from a fixed rule (enable =
 * 0, or an integrand without measurements, such as a callback). */
int pagani_ws_set_tuned_slices(pagani_ws* ws, int enable);
int pagani_ws_set_timing(pagani_ws* ws, int enable);
int pagani_ws_get_stats(const pagani_ws* ws, pagani_stats* out);
int pagani_ws_reset_stats(pagani_ws* ws);
/* the uniform launch data for n regions (for evaluation callbacks/tests) */
int pagani_ws_eval_consts(pagani_ws* ws, const double* lows,
                          const double* highs, size_t n,
                          pagani_eval_consts* out);
void* pagani_ws_stream(pagani_ws* ws); /* hipStream_t */
/* Per-iteration record of the last integrate call (the reference's debug
 * recorder, Workspace.cuh:171-242): regions evaluated, regions kept active
 * (0 on the final iteration), cumulative estimate and error.  Copies up to
 * cap records (any pointer may be NULL) and returns the number recorded. */
size_t pagani_ws_trace(const pagani_ws* ws, size_t cap, uint64_t* n, uint64_t* m,
                       double* estimate, double* errorest);
/* The region list the last integrate call ended with (the reference leaves
 * it in the caller's Sub_regions, Sub_region_filter.cuh:129-135,
 * Sub_region_splitter.cuh:81-84): *n regions, flat device SoA pointers
 * owned by the workspace and valid until its next call (a list held in
 * split form is written out flat first, on the workspace's stream). */
int pagani_ws_live_regions(pagani_ws* ws, size_t* n, const double** L, const double** H);

/* ---- integration ------------------------------------------------------ */
/* Workspace::integrate(integrand, epsrel, epsabs, vol, relerr_classification)
 * Workspace.cuh:361-494, for a precompiled Genz integrand.  lows/highs may
 * be NULL (unit cube).  init_k = regions per axis of the initial uniform
 * split, 0 = the reference's default (4 for d<5, 2 for d<=10). */
int pagani_integrate_genz(pagani_ws* ws, const pagani_genz* f, int precision,
                          double epsrel, double epsabs, const double* lows,
                          const double* highs, int relerr_classification,
                          size_t init_k, pagani_result* out);
/* Same loop for any integrand through an evaluation callback.  init = NULL
 * uses the default uniform split (Workspace.cuh:362-494); otherwise the
 * given region list, with the Sub_regions& overload's bookkeeping
 * (Workspace.cuh:148-358). */
int pagani_integrate_cb(pagani_ws* ws, pagani_eval_fn eval, void* ctx,
                        double epsrel, double epsabs, const double* lows,
                        const double* highs, int relerr_classification,
                        const pagani_region_list* init, size_t init_k,
                        pagani_result* out);

/* Device bytes held by the workspace's region pool, and their release (the
 * pool regrows on the next call).  The reference allocates and frees its
 * buffers inside every integrate call; a workspace keeps them between calls
 * unless trimmed.  Trimming fails while a suspended run holds the pool. */
size_t pagani_ws_footprint(const pagani_ws* ws);
int pagani_ws_trim(pagani_ws* ws);

/* ---- suspend, checkpoint, resume ----------------------------------------- */
/* No reference counterpart beyond the in-memory region-list copy
 * (Sub_regions::take_snapshot / load_snapshot, Sub_regions.cuh:220-239): a
 * long run can stop at an iteration boundary, be written to a file and be
 * continued later with exactly the result of the uninterrupted run.
 *
 * With a limit k > 0, an integrate or resume call returns PAGANI_SUSPENDED
 * after k iterations, with *out holding the running sums (status 1); the
 * workspace keeps the loop state until its next integrate call. */
int pagani_ws_set_iteration_limit(pagani_ws* ws, uint64_t iterations);
int pagani_ws_suspended(const pagani_ws* ws); /* 1 when a run is held */
int pagani_integrate_cb_resume(pagani_ws* ws, pagani_eval_fn eval, void* ctx,
                               pagani_result* out);
/* The integrand must be the suspended run's (checked for Genz descriptors). */
int pagani_integrate_genz_resume(pagani_ws* ws, const pagani_genz* f, int precision,
                                 pagani_result* out);
/* Checkpoint file of the held run (layout in checkpoint.hip's header
 * comment: a versioned header, the loop state, the trace, the SoA region
 * list and the parents' estimates, and a checksum).  Loading verifies the
 * size, the dimension and the checksum, then holds the run for a resume. */
int pagani_ws_save_state(pagani_ws* ws, const char* path);
int pagani_ws_load_state(pagani_ws* ws, const char* path);
/* The live region list as CSV, one row per region in the column order of
 * Snapshot<NDIM>::Save (quad.h:107-118): estimate, error, upper and lower
 * bound per axis, split axis, -1. */
int pagani_ws_write_regions_csv(pagani_ws* ws, const char* path);
/* predict_split (Workspace.cuh:244-248): copy the region list evaluated in
 * iteration `iteration` when no region has finished by then (the reference
 * uses 15); UINT64_MAX (the default) disables.  The copy stays in the
 * workspace (device SoA, *n regions) until the next snapshot. */
int pagani_ws_set_snapshot_iter(pagani_ws* ws, uint64_t iteration);
/* Workspace::integrate<IntegT, predict_split=true> also overrides the
 * relerr classification flag every iteration (Workspace.cuh:206-210,
 * 415-420): off while the list has at most 15M regions, the iteration is
 * before 15 and no region has finished, on otherwise.  enable = 1 applies
 * that rule to the following integrate calls instead of their
 * relerr_classification argument (a suspended run keeps its rule). */
int pagani_ws_set_predict_split(pagani_ws* ws, int enable);
/* The reference's debug recorders (Workspace<T, ndim, debug,
 * use_custom, collect_mult_runs>: Workspace.cuh:171-179, 199-352, 377-431;
 * PaganiUtils.cuh:50-241).  For the following integrate calls, level
 *   0  records nothing (the default);
 *   1  with id not NULL (the Sub_regions& overload, whose `optional`
 *      argument is the id), writes <dir>/cuda_pagani_iters.csv, one line
 *      per iteration after the error refinement ("it, estimate, errorest,
 *      festimate, ferrorest, nregions, nfregions": the running totals with
 *      and without this iteration's regions, the list size, the regions
 *      finished before it), and <dir>/cuda_pagani_time_breakdown.csv ("id,
 *      ndim, epsrel, it, name, time"): milliseconds per phase under the
 *      reference's names apply_cubature_rules, two_level_errorest,
 *      compute_finished_estimates, fix_error_budget_overflow,
 *      heuristic_classify, region_filtering and region_splitting (host time
 *      between stream synchronisations at the phase boundaries; the filter
 *      and the split are one kernel here, timed as region_filtering, with
 *      region_splitting 0); with id NULL (the plain overload) writes
 *      <dir>/cuda_iters.csv ("it, estimate, errorest, nregions") only;
 *   2  also <dir>/cuda_regions.csv: every region's estimate and rule error
 *      as evaluated ("it, reg, estimate, errorest"; reg in list order);
 *   3  also <dir>/cuda_generators.csv (the rule's generator table, "i, gen",
 *      i = feval * axis + point) and, for lists under 1024 regions,
 *      <dir>/cuda_fevals.csv: every rule point of every region (region and
 *      box bounds, the point, the value, the region's estimate and error).
 * append != 0 appends without header lines (collect_mult_runs).
 * Values carry 17 significant digits.  Recording runs the host loop only (no
 * device-resident rounds or graphs) with one stream sync per iteration; the
 * results are those of an unrecorded run.  dir NULL means ".". */
int pagani_ws_set_debug(pagani_ws* ws, int level, const char* dir, const char* id, int append);
/* Generator coordinate g_axis of rule point p (library order: the centre;
 * +-lambda1 and +-lambda2 per axis, axis-major, + before -; then the table of
 * sets 3-8), for ndim-D rules; the point of a region with centre c and width
 * w is c - g w.  Returns 0, or PAGANI_EBADARG. */
int pagani_rule_point(int ndim, int p, int axis, double* g);
int pagani_ws_snapshot_regions(const pagani_ws* ws, size_t* n, const double** L,
                               const double** H);

/* ---- batches of independent integrals ---------------------------------- */
/* The reference integrates a set of integrands with a host loop of
 * Workspace::integrate calls (new_time_and_call.cuh:156-167, driven by
 * randomized_integrands.cu:80-105).  A batch runs the same integrals from one
 * call: `workers` host threads, each with its own stream and one workspace
 * per dimension, take items largest-first (by dimension, then epsrel), so
 * one integral's host round trips overlap another's kernels.  Each item's
 * result is the one pagani_integrate_genz returns for it on a workspace whose
 * memory-model total is the batch's (set_device_mem; default: device HBM /
 * workers, so the concurrent pools fit together) -- independent of the
 * order and interleaving. */
typedef struct pagani_batch_item {
  pagani_genz f;
  double epsrel, epsabs;
  int precision;             /* PAGANI_FP64 or PAGANI_FP32 */
  int relerr_classification; /* 1 = the reference's default */
  size_t init_k;             /* 0 = default initial split */
} pagani_batch_item;

typedef struct pagani_batch pagani_batch;

/* NULL on failure (reason from pagani_batch_last_error(NULL)); workers in
 * [1, 16] */
pagani_batch* pagani_batch_create(int device, int workers);
void pagani_batch_destroy(pagani_batch* b);
const char* pagani_batch_last_error(const pagani_batch* b);
int pagani_batch_set_device_mem(pagani_batch* b, size_t total_bytes);
int pagani_batch_set_mem_model(pagani_batch* b, int model);
int pagani_batch_set_timing(pagani_batch* b, int enable);
/* pagani_ws_set_tuned_slices for every workspace of the batch */
int pagani_batch_set_tuned_slices(pagani_batch* b, int enable);
/* No reference counterpart (the reference runs every integral through its
 * host loop).  regions > 0 turns on the persistent kernel: FP64 Genz items
 * whose initial split has at most `regions` regions run first, whole, in one
 * launch per (dimension, family) -- one workgroup per integral, each with a
 * segment of `regions` regions of one pool, no host round trip per
 * iteration.  An item whose list would outgrow its segment, that needs the
 * error-budget fix (Workspace.cuh:121-146) or that comes past a tenth of its
 * memory budget (where the classifier may run) is handed over: its list and
 * loop state move to a worker's workspace and the regular loop continues
 * from that iteration.  The iterations are those of pagani_integrate_genz;
 * the results agree with it to rounding (the sums are taken in another
 * order), not to the bit.  0 (the default) = off, else regions in
 * [256, 2^22]. */
int pagani_batch_set_segments(pagani_batch* b, size_t regions);
/* statistics summed over all workspaces of the batch */
int pagani_batch_get_stats(const pagani_batch* b, pagani_stats* out);
int pagani_batch_reset_stats(pagani_batch* b);
/* Integrate count items; results[i] and codes[i] (the return code of the
 * item's integrate call) are filled for every item.  Returns PAGANI_OK, or
 * the first negative code of an item (the other items still run). */
int pagani_integrate_batch(pagani_batch* b, const pagani_batch_item* items,
                           size_t count, pagani_result* results, int* codes);

/* ---- phases on caller-owned device arrays --------------------------------- */
/* Sub_regions::uniform_split / create_uniform_split, Sub_regions.cuh:11-36,
 * 165-185: k^ndim boxes into L, H (capacity k^ndim regions). */
int pagani_uniform_split_dev(pagani_ws* ws, size_t k, double* L, double* H);
/* Cubature_rules::apply_cubature_integration_rules, PaganiUtils.cuh:291-343:
 * per-region estimate, error and split axis for a Genz integrand. */
int pagani_eval_genz_dev(pagani_ws* ws, const pagani_genz* f, int precision,
                         size_t n, const double* L, const double* H,
                         const double* lows, const double* highs, double* est,
                         double* err, uint8_t* dim);
/* two_level_errorest_and_relerr_classify / RefineError, hybrid.cuh:11-45,
 * Phases.cuh:219-268.  parent_est = NULL is iteration 0 (no refinement, all
 * regions active).  sums[4] (host, may be NULL) = {sum est, sum err_out,
 * sum active*est, sum active*err_out}; *num_active (may be NULL). */
int pagani_refine_dev(pagani_ws* ws, size_t n, const double* est,
                      const double* err_in, double* err_out,
                      const double* parent_est, double epsrel,
                      int relerr_classification, uint8_t* active,
                      double* sums, size_t* num_active);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, serving loops that forbid hipMalloc):
 * pagani_refine_dev's refinement, flags and sums with every intermediate
 * record in a fixed 1.5 MB device array of the library (one per GPU, held
 * until the process ends) instead of workspace buffers that grow on demand.
 * n is at most PAGANI_NOALLOC_MAX_REGIONS (else PAGANI_EBADARG).  The call
 * allocates nothing, only enqueues on the workspace's stream
 * (pagani_ws_stream) and returns: sums_dev[4] = {sum est, sum err_out, sum
 * active*est, sum active*err_out} and *num_active_dev (device pointers, may
 * be NULL) are valid once that stream has drained.  Every call on one GPU
 * shares its array: calls that overlap on one GPU race. */
#define PAGANI_NOALLOC_MAX_REGIONS ((size_t)1 << 26)
int pagani_refine_dev_noalloc(pagani_ws* ws, size_t n, const double* est,
                              const double* err_in, double* err_out,
                              const double* parent_est, double epsrel,
                              int relerr_classification, uint8_t* active,
                              double* sums_dev, uint64_t* num_active_dev);
/* Sub_regions_filter::filter, Sub_region_filter.cuh:86-139: order-preserving
 * compaction of the active regions (bounds at stride m, parent estimates,
 * split axes).  Returns m in *m_out. */
int pagani_filter_dev(pagani_ws* ws, size_t n, const double* L,
                      const double* H, const uint8_t* active,
                      const double* est, const double* err,
                      const uint8_t* dim, double* L2, double* H2,
                      double* parent_est, double* parent_err, uint8_t* dim2,
                      size_t* m_out);
/* Sub_region_splitter::split, Sub_region_splitter.cuh:48-87: m regions ->
 * 2m children [left m | right m] (L2/H2 at stride 2m). */
int pagani_split_dev(pagani_ws* ws, size_t m, const double* L,
                     const double* H, const uint8_t* dim, double* L2,
                     double* H2);
/* Filter and split fused (the library's own phase): the active regions'
 * children at stride 2m, their parents' estimates at [0, m). */
int pagani_filter_split_dev(pagani_ws* ws, size_t n, const double* L,
                            const double* H, const uint8_t* active,
                            const double* est, const double* err,
                            const uint8_t* dim, double* L2, double* H2,
                            double* parent_est, double* parent_err,
                            size_t* m_out);
/* thrust_utils.cuh:16-88 equivalents (deterministic, compensated sums) */
int pagani_sum_dev(pagani_ws* ws, const double* x, size_t n, double* out);
int pagani_dot_dev(pagani_ws* ws, const uint8_t* flags, const double* x,
                   size_t n, double* out);
int pagani_minmax_dev(pagani_ws* ws, const double* x, size_t n, double* mn,
                      double* mx);
/* exclusive scan of 0/1 flags (Sub_region_filter.cuh:62-82 scans them):
 * offsets[t] = number of set flags before t; *total = all set flags */
int pagani_scan_flags_dev(pagani_ws* ws, const uint8_t* flags, size_t n,
                          uint64_t* offsets, size_t* total);
/* Heuristic_classifier::classify, heuristic_classifier.cuh:362-451, with
 * the threshold passes on the device.  est_hist[3] = the classifier's last
 * three stored estimates (oldest first); flags receives the threshold
 * classification of the last pass. */
int pagani_classify_dev(pagani_ws* ws, const double* err, size_t n,
                        double epsrel, const double* est_hist,
                        double iter_errorest, double iter_finished_errorest,
                        double total_finished_errorest, uint8_t* flags,
                        pagani_class_res* out);

/* ---- m-Cubes: VEGAS with stratified sampling ------------------------------- */
/* The reference's second integrator: cuda/mcubes/vegasT.cuh (vegas(),
 * vegas_kernel, vegas_kernelF, integrate, simple_integrate, class mcubes),
 * cuda/mcubes/vegasT1D.cuh (one grid shared by all axes) and
 * cuda/mcubes/mcubes.cuh (quad::mcubes).  The C++ surface is
 * include/pagani/mcubes.hpp; these are the entry points under it.
 *
 * Importance grid (host and device): axis j has PAGANI_MC_BINS bins with
 * right edges xi[j*(PAGANI_MC_BINS+1) + i], i = 1..PAGANI_MC_BINS, in
 * [0, 1], and xi[j*(PAGANI_MC_BINS+1)] = 0.  Bin histograms:
 * hist[j*PAGANI_MC_BINS + i - 1].  Cubes (ng strata per axis) are numbered
 * with the last axis fastest (get_indx, vegasT.cuh:137-156). */
#define PAGANI_MC_BINS 500 /* Internal_Vegas_Params ndmx, vegas_utils.cuh:157 */
#define PAGANI_MC_MAX_DIM 16
/* random streams, one per (iteration, cube) */
#define PAGANI_MC_RNG_DEFAULT 0 /* xoroshiro128++ seeded by SplitMix64 of (seed, iteration, cube);
                                   stands in for the reference's clock-seeded curand */
#define PAGANI_MC_RNG_LCG 1     /* Custom_generator, vegas_utils.cuh:24-48: the C LCG
                                   reseeded with the cube index (same numbers every iteration) */

typedef struct pagani_mc_params {
  double ncall;      /* calls per iteration (ncall) */
  int total_iters;   /* titer */
  int adjust_iters;  /* itmax: the leading iterations that refine the grid */
  int skip_iters;    /* skip: leading iterations left out of the estimate */
  int generator;     /* PAGANI_MC_RNG_* */
  int one_axis;      /* 1: vegasT1D.cuh -- axis 0's grid is adapted and copied to all axes */
  int adjust_params; /* 1: simple_integrate -- rerun with AdjustParams while status is 1 */
  uint64_t seed;     /* PAGANI_MC_RNG_DEFAULT streams */
} pagani_mc_params;

/* numint::integration_result as m-Cubes fills it (estimate, errorest,
 * chi_sq, status, iters), plus the run's size. */
typedef struct pagani_mc_result {
  double estimate;
  double errorest;
  double chi_sq;
  int status;     /* 0: |errorest/estimate| <= epsrel or errorest <= epsabs after >= 5
                     iterations (GetStatus, vegas_utils.cuh:233-254); 1 otherwise */
  int pad;
  uint64_t iters; /* iterations over all runs */
  uint64_t runs;  /* vegas() runs (simple_integrate retries + 1) */
  uint64_t neval; /* integrand evaluations over all runs */
  double ncall;   /* ncall of the last run */
} pagani_mc_result;

/* One sampling launch: every cube of the stratification, npg points each
 * (vegas_kernel when hist_dims > 0, vegas_kernelF when 0).  Workgroup g
 * handles cubes [g*kThreads*cubes_per_thread, ...) and writes its sums of
 * fb and f2b to part[2g], part[2g+1] and its bin histograms to
 * hist[g*hist_dims*PAGANI_MC_BINS ...]. */
typedef struct pagani_mc_args {
  const double* xi; /* device grid, ndim x (PAGANI_MC_BINS+1) */
  double lo[PAGANI_MC_MAX_DIM];
  double dx[PAGANI_MC_MAX_DIM];
  double xjac, dxg, xnd;
  uint64_t key;     /* the iteration's stream key (PAGANI_MC_RNG_DEFAULT) */
  uint32_t ncubes;
  uint32_t ng, npg;
  uint32_t cubes_per_thread;
  uint32_t hist_dims; /* 0, 1 (one_axis) or ndim */
  uint32_t grid;      /* workgroups */
  int32_t generator;  /* PAGANI_MC_RNG_* */
  int32_t pad;
  double* part;
  double* hist;
  void* stream;       /* hipStream_t */
} pagani_mc_args;

/* A sampling callback launches the m-Cubes kernel for one integrand on
 * args->stream (pagani::mc::sample_callback<ndim, IntegT, Generator> from
 * include/pagani/mcubes_kernel.hpp); 0 on success. */
typedef int (*pagani_mc_sample_fn)(void* ctx, const pagani_mc_args* args);

typedef struct pagani_mc pagani_mc;

typedef struct pagani_mc_stats {
  double sample_ms; /* sampling kernels (with timing enabled) */
  double wall_ms;   /* integrate calls */
  uint64_t samples; /* integrand evaluations */
  uint64_t launches;
  uint64_t integrals;
} pagani_mc_stats;

/* NULL on failure (reason from pagani_mc_last_error(NULL)); ndim in
 * [1, PAGANI_MC_MAX_DIM]; device < 0 is the current device. */
pagani_mc* pagani_mc_create(int ndim, int device);
void pagani_mc_destroy(pagani_mc* mc);
const char* pagani_mc_last_error(const pagani_mc* mc);
int pagani_mc_set_timing(pagani_mc* mc, int enable);
int pagani_mc_get_stats(const pagani_mc* mc, pagani_mc_stats* out);
int pagani_mc_reset_stats(pagani_mc* mc);

/* vegas() / integrate / simple_integrate (vegasT.cuh:640-1112) for any
 * integrand behind a sampling callback; lows/highs may be NULL (unit
 * cube).  Returns PAGANI_OK, or a negative code (then *out holds status 1). */
int pagani_mc_integrate_cb(pagani_mc* mc, pagani_mc_sample_fn sample, void* ctx,
                           double epsrel, double epsabs, const double* lows,
                           const double* highs, const pagani_mc_params* params,
                           pagani_mc_result* out);
/* The same for a precompiled Genz integrand (ndim 2..8, FP64; m-Cubes'
 * precompiled samplers stop at 8 dimensions). */
int pagani_mc_integrate_genz(pagani_mc* mc, const pagani_genz* f, double epsrel,
                             double epsabs, const double* lows, const double* highs,
                             const pagani_mc_params* params, pagani_mc_result* out);
/* Per-iteration sums of the last integrate call (ti, and tsi scaled by
 * dv2g, vegasT.cuh:842-845), up to cap entries; *len = iterations run. */
int pagani_mc_trace(const pagani_mc* mc, double* ti, double* tsi, size_t cap, size_t* len);
/* The grid at the end of the last integrate call (host array of
 * ndim*(PAGANI_MC_BINS+1)). */
int pagani_mc_grid(const pagani_mc* mc, double* xi);
/* One sampling pass of a Genz integrand on a given host grid: the phase
 * under vegas_kernel / vegas_kernelF.  ti = sum fb, tsi = sum f2b (not
 * scaled by dv2g); hist (hist_dims*PAGANI_MC_BINS, may be NULL when
 * hist_dims == 0) receives the bin sums of f^2. */
int pagani_mc_sample_genz(pagani_mc* mc, const pagani_genz* f, double ncall,
                          const double* lows, const double* highs, const double* xi,
                          uint64_t key, int generator, int hist_dims, double* ti,
                          double* tsi, double* hist);

/* ---- the reference's C interop shim ---------------------------------------- */
/* test/cuda/pagani/do_integration_from_c.cu:20-38: integrates x*y over the
 * unit square to epsrel 1e-6, epsabs 1e-12; writes *res only when the
 * status is 0; returns the status (1 on a NULL pointer). */
int do_integration_from_c(double* res);

#ifdef __cplusplus
}
#endif
#endif
