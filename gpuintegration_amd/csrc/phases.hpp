// phases.hpp -- launchers of the region-list kernels (everything except the
// rule evaluation).  All launches are asynchronous on the given stream.
//
// Work is cut into fixed chunks of kChunk consecutive regions, one workgroup
// per chunk.  A chunk kernel writes per-chunk partial results (four
// compensated double-double sums and a 32-bit count); finalize() reduces them
// in a fixed order and turns the counts into exclusive chunk offsets, so
// every sum is deterministic run to run and the compaction needs no atomics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "pagani.h"

namespace pagani {
namespace internal {
struct DevLoop;
struct SearchState;
struct SearchInputs;
}
namespace phases {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr std::size_t kChunk = static_cast<std::size_t>(kThreads) * kPerThread;
// finalize stage A folds this many chunk records per workgroup
constexpr std::size_t kFold = 256;

__host__ __device__ inline std::size_t num_chunks(std::size_t n)
{
  return (n + kChunk - 1) / kChunk;
}

// Chunks of a list whose regions have parents ("paired": n = 2h children,
// t and t + h the two children of parent t mod h) are cut at h as well:
// ceil(h / kChunk) chunks of each half, so chunk c and chunk L + c hold
// exactly each other's siblings, and a kernel that reads a region's sibling
// or parent finds them in the paired chunk, which runs beside it on the same
// XCD (paired_chunk).  The refinement, the flag counts of the classifier and
// the error-budget fix, their finalize and the compaction of one iteration
// all cut the list the same way.  At most one chunk more than num_chunks(n).
__host__ __device__ inline std::size_t num_chunks(std::size_t n, bool paired)
{
  return paired ? 2 * ((n / 2 + kChunk - 1) / kChunk) : num_chunks(n);
}
__host__ __device__ inline void chunk_range(std::size_t c, std::size_t n, bool paired,
                                            std::size_t& begin, std::size_t& end)
{
  const std::size_t h = paired ? n / 2 : n;
  const std::size_t L = (h + kChunk - 1) / kChunk;
  begin = c < L ? c * kChunk : h + (c - L) * kChunk;
  const std::size_t lim = c < L ? h : n;
  end = begin + kChunk < lim ? begin + kChunk : lim;
}
__host__ __device__ inline std::size_t num_folds(std::size_t nchunks)
{
  return (nchunks + kFold - 1) / kFold;
}

// The device-resident loop handles lists of at most one fold of chunks, so
// its finalize is a single fold plus the final workgroup.
constexpr std::size_t kDevLoopMax = kChunk * kFold;

// Filter + split descriptor of the device-resident loop (device memory):
// the list's count n, pair and form (split: held as parents + halved axes),
// the count m of regions kept, which error array holds the refined errors.
struct DevCompact {
  std::uint64_t n, m;
  std::uint32_t cur, run, has_parents, split;
};

// Candidate thresholds measured by one pass of the classifier probe (the
// host measures the states a few moves of the search can reach): sorted
// ascending, at most kMaxThresholds - 1 of them.
constexpr int kMaxThresholds = 16;
struct ThresholdSet {
  double t[kMaxThresholds];
  int k;
};
// fixed grid of the multi-threshold pass (deterministic partial order):
// one workgroup per CU slot -- its 45 KB of LDS buckets leave room for
// three per CU on the 256 CUs, so the whole grid runs in one wave
constexpr int kMultiBlocks = 768;

// Results of finalize(), read back by the host driver.
struct Scalars {
  double sum[4];   // the four chunk sums, compensated
  double mn, mx;   // min / max (minmax pass only)
  std::uint64_t count;
  std::uint64_t pad;
};

// per-chunk partial record: 4 (hi, lo) sums + min + max
constexpr int kPartDoubles = 10;

// Widths of a list in level form (pagani_eval_args): H holds one uint16_t
// per axis and row, the number of halvings from h0[j]; on = 0: H holds
// doubles.
struct Levels {
  double h0[PAGANI_MAX_DIM];
  int on;
};

// k^d regions of the unit cube, widths as doubles, or, with levels, as
// level 0 (the width 1/k is the caller's h0).
hipError_t uniform_split(hipStream_t s, int d, std::size_t k, std::size_t n, double* L,
                         double* H, bool levels = false);

// Device-resident loop variants (lists of at most kDevLoopMax regions):
// fixed launch shapes for n_cap regions, the actual count and selections
// read from the descriptors; a descriptor with run == 0 makes them no-ops.
hipError_t step_dev(hipStream_t s, std::size_t n_cap, internal::DevLoop* loop, const double* est,
                    double* err0, double* err1, const double* pest, double epsrel, int relerr,
                    std::uint8_t* active, double* part, std::uint32_t* cnt, std::uint64_t* off,
                    double* fpart, std::uint32_t* fcnt, std::uint64_t* foff, Scalars* sc,
                    std::uint32_t* ticket);
hipError_t compact_dev(hipStream_t s, std::size_t n_cap, const DevCompact* cp, int d,
                       double* L0, double* H0, double* L1, double* H1, std::uint8_t* P0,
                       std::uint8_t* P1, const std::uint8_t* active, const double* est,
                       const double* err0, const double* err1, const std::uint8_t* dim,
                       const std::uint64_t* off, const std::uint64_t* foff, double* pest,
                       double* perr, const Levels& lv);

// RefineError + relerr classification (parent_est == nullptr: iteration 0,
// err_out may equal err_in).  Partial sums: {est, err_out, act*est, act*err_out},
// one record per chunk of num_chunks(n, parent_est != nullptr).
hipError_t refine(hipStream_t s, std::size_t n, const double* est, const double* err_in,
                  double* err_out, const double* parent_est, double epsrel, int relerr,
                  std::uint8_t* active, double* part, std::uint32_t* cnt);

// Same partial sums for given flags (after the budget fix or a classification),
// in the chunks of num_chunks(n, paired).
hipError_t count_flags(hipStream_t s, std::size_t n, const std::uint8_t* flags,
                       const double* est, const double* err, double* part,
                       std::uint32_t* cnt, bool paired);

// flags = err > thr; partial sums {0, err, 0, flag*err}, count.
hipError_t threshold(hipStream_t s, std::size_t n, const double* err, double thr,
                     std::uint8_t* flags, double* part, std::uint32_t* cnt);

// min / max of x into the partial records.
hipError_t minmax(hipStream_t s, std::size_t n, const double* x, double* part);

// Sum of x (partial sum 0) and dot(flags, x) (partial sum 2) for the
// thrust-equivalent entry points.
hipError_t sum_dot(hipStream_t s, std::size_t n, const double* x,
                   const std::uint8_t* flags, double* part, std::uint32_t* cnt);

// flags = err > thr and the refine-style record {est, err, flag*est,
// flag*err} with the flag count (the classifier's accepted threshold).
hipError_t threshold_flags(hipStream_t s, std::size_t n, const double* est, const double* err,
                           double thr, std::uint8_t* flags, double* part, std::uint32_t* cnt,
                           bool paired);

// For every threshold t_k of T: count of err > t_k and sum of those err.
// Writes T.k counts then T.k sums (doubles) to out (device).  bpart holds
// kMultiBlocks * 3 * kMaxThresholds doubles.
// The heuristic classifier's threshold search on the device
// (classify_search.hpp): search_start begins a search from `in` with the
// errors' min and max from the last minmax finalize (sc) and queues its first
// probe pass; each search_round runs the queued pass over err (bpart,
// mout: the multi-threshold scratch) and the replay that queues the next
// one.  Rounds after the search has finished do nothing.
hipError_t search_start(hipStream_t s, internal::SearchState* st, const Scalars* sc,
                        const internal::SearchInputs& in);
hipError_t search_round(hipStream_t s, std::size_t n, const double* err,
                        internal::SearchState* st, double* bpart, double* mout);

hipError_t multi_threshold(hipStream_t s, std::size_t n, const double* err,
                           const ThresholdSet& T, double* bpart, double* out);

// Reduce nchunks partial records in two stages: workgroups of kFold chunk
// records (chunk offsets relative to their fold into off, fold records into
// fpart/fcnt), then one workgroup over the folds (fold offsets into foff,
// results into *out).  Counts / offsets are skipped when cnt is nullptr.
// The global exclusive offset of chunk c is off[c] + foff[c / kFold].
hipError_t finalize(hipStream_t s, std::size_t nchunks, const double* part,
                    const std::uint32_t* cnt, std::uint64_t* off, double* fpart,
                    std::uint32_t* fcnt, std::uint64_t* foff, Scalars* out, int with_minmax);

// Order-preserving compaction of the active regions (chunk offsets from
// finalize).  split == 0: filter only, bounds at stride m, split axes
// copied to dim2.  split == 1: the two children of every active region at
// stride 2m ([left m | right m]), the chosen axis halved.  Parent estimates
// and errors go to pest/perr[0, m).
hipError_t compact(hipStream_t s, int split, int d, std::size_t n, std::size_t m,
                   const double* L, const double* H, const std::uint8_t* active,
                   const double* est, const double* err, const std::uint8_t* dim,
                   const std::uint64_t* off, const std::uint64_t* foff, double* L2,
                   double* H2, double* pest, double* perr, std::uint8_t* dim2);

// The iteration loop's filter + split, Sub_region_filter.cuh:8-139 with
// Sub_region_splitter.cuh:8-87: the m active regions of an n-region list
// (flat, or in split form when pdim is not NULL) become the parents of the
// next list, written once -- their bounds at stride m into L2/H2, the axis
// each is to be halved along (its own fourth-difference choice, dim) into
// pdim2, their estimates and errors into pest/perr -- and the next list is
// their 2m children in split form.  The reference writes both children's
// full bounds (32d bytes per kept region) where this writes 16d + 1.  off /
// foff: the offsets of the chunks of num_chunks(n, paired).
hipError_t compact_parents(hipStream_t s, int d, std::size_t n, std::size_t m, const double* L,
                           const double* H, const std::uint8_t* pdim,
                           const std::uint8_t* active, const double* est, const double* err,
                           const std::uint8_t* dim, const std::uint64_t* off,
                           const std::uint64_t* foff, double* L2, double* H2,
                           std::uint8_t* pdim2, double* pest, double* perr, bool paired,
                           const Levels& lv);

// Any of the loop's lists (flat or split form, widths as doubles or levels)
// written out flat with double widths: n regions at [j*n + t] of L2, H2.
hipError_t expand(hipStream_t s, int d, std::size_t n, const double* L, const double* H,
                  const std::uint8_t* pdim, const Levels& lv, double* L2, double* H2);

// Children of all m regions (Sub_region_splitter::split without a filter);
// also writes a list held in split form out flat.
hipError_t split_all(hipStream_t s, int d, std::size_t m, const double* L, const double* H,
                     const std::uint8_t* dim, double* L2, double* H2);

// Per-element exclusive scan of 0/1 flags using chunk offsets.
hipError_t scan_flags(hipStream_t s, std::size_t n, const std::uint8_t* flags,
                      const std::uint64_t* off, const std::uint64_t* foff, std::uint64_t* out);

hipError_t fill_u8(hipStream_t s, std::size_t n, std::uint8_t* p, std::uint8_t v);

}  // namespace phases
}  // namespace pagani
