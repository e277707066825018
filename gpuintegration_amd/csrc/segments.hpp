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

// Its result: outcome 0 when the kernel finished it, 1 when the regular
// loop must integrate it (r is then meaningless).
struct SegOut {
  pagani_result r;
  int outcome;
  int pad;
};

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
