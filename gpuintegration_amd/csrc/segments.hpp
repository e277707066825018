// segments.hpp -- the batch's persistent kernel (segments.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "pagani.h"

namespace pagani {
namespace internal {

// One integral for the persistent kernel: a Genz descriptor, its tolerances,
// the initial split (k per axis, resolved) and its relerr flag.
struct SegItem {
  pagani_genz f;
  double epsrel, epsabs;
  std::uint64_t init_k;
  int relerr;
  int pad;
};

// Its result: outcome 0 when the kernel finished it (r), 1 when the
// regular loop takes it over.  Handed over with n > 0, the loop state
// before iteration `it`: r holds the running totals, the list of n regions
// (parents' estimates when has_parents) is in buffer pair `cur` of the
// segment of workgroup `block`, and hist / iters_collected are the
// classifier's stored estimates.  n == 0: the regular loop starts afresh.
struct SegOut {
  pagani_result r;
  int outcome;
  int has_parents;
  int cur;
  unsigned block;
  std::uint64_t it, n;
  std::uint64_t evaluated;  // regions evaluated in the kernel's finished iterations
  std::uint64_t iterations;
  double hist[3];
  std::uint64_t iters_collected;
};

// The parts of one segment (device pointers), for the host's hand-over.
struct SegView {
  double* L[2];
  double* H[2];
  double* pest;
};
SegView segment_view(char* base, int d, std::size_t regions);

// Bytes of one segment of `regions` regions at dimension d.
std::size_t segment_bytes(int d, std::size_t regions);

// `want`, cut to the workgroups of the kernel for (d, family) that the
// device holds at once (more would only wait for the first to finish).
unsigned segment_grid(int d, int family, unsigned want);

// One launch for the items of dimension d (d in [2, PAGANI_GENZ_MAX_DIM])
// and one Genz family: `grid` workgroups take items from *next (zero before
// the launch), each with the segment at blockIdx * bytes of pool; C: the launch constants of
// the rule at dimension d; mem_total / mem_model: the memory model the
// regular loop would use for the items.
hipError_t launch_segments(hipStream_t s, int d, int family, const SegItem* items,
                           std::uint32_t count, std::uint32_t* next, char* pool,
                           std::size_t bytes, std::size_t S, unsigned grid,
                           const pagani_eval_consts& C, std::size_t mem_total, int mem_model,
                           SegOut* out);

}  // namespace internal
}  // namespace pagani
