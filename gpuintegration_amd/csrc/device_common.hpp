// device_common.hpp -- deterministic reductions and the per-region
// refinement, shared by the phase kernels (phases.hip) and the batch's
// segment kernel (segments.hip).  Compile with -ffp-contract=off: the
// refinement is plain IEEE double in the reference's operation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "phases.hpp"

namespace pagani {
namespace phases {

struct DD {
  double hi, lo;
};

// Compensated addition (two-sum of the leading parts, then the tails).
// Non-finite sums fall back to the plain sum so inf/NaN propagate as in an
// uncompensated reduction.
__device__ __forceinline__ DD dd_add(DD a, DD b)
{
  const double s = a.hi + b.hi;
  if (!__builtin_isfinite(s)) return DD{s, 0.0};
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  const double hi = s + e;
  return DD{hi, e - (hi - s)};
}

__device__ __forceinline__ DD dd_add(DD a, double b) { return dd_add(a, DD{b, 0.0}); }

__device__ __forceinline__ DD wave_reduce(DD v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    DD o;
    o.hi = __shfl_xor(v.hi, off, 64);
    o.lo = __shfl_xor(v.lo, off, 64);
    v = dd_add(v, o);
  }
  return v;
}

// Reduce a DD over a 256-thread block; every thread gets the result.
__device__ __forceinline__ DD block_reduce(DD v, DD* sh)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_reduce(v);
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  DD r = sh[0];
#pragma unroll
  for (int w = 1; w < kThreads / 64; ++w) r = dd_add(r, sh[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ std::uint32_t block_count(std::uint32_t c, std::uint32_t* sh)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (lane == 0) sh[wave] = c;
  __syncthreads();
  std::uint32_t r = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) r += sh[w];
  __syncthreads();
  return r;
}

// RefineError (Phases.cuh:219-268) of one region (estimate e, error r)
// with its sibling (es, rs) and their parent's estimate pe, and the relerr
// classification: r_out the refined error, active unless relerr finishes it.
__device__ __forceinline__ void refine_region(double e, double r, double es, double rs, double pe,
                                              double epsrel, int relerr, double& r_out,
                                              bool& active)
{
#pragma clang fp contract(off)
  const double diff = fabs(.25 * ((es + e) - pe));
  const double E = r + rs;
  if (E > 0.0) {
    const double cc = 1 + 2 * diff / E;
    r *= cc;
  }
  r += diff;
  const bool pass = relerr && r < fmax(epsrel * fabs(e), 1e-200);
  r_out = r;
  active = !pass;
}

}  // namespace phases
}  // namespace pagani
