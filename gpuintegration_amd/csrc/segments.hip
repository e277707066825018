// segments.hip -- the batch's persistent kernel: whole integrals, one
// workgroup each, every integral in its own segment of one region pool
// (pagani_batch_set_segments).
//
// A batch of many small integrals spends its time on launches and host
// round trips, not on arithmetic: each iteration of each integral is a few
// launches over a few thousand regions.  Here one launch per (dimension,
// Genz family) runs a fixed grid of workgroups that take items from a
// shared counter; a workgroup runs an item's whole loop (Workspace.cuh:
// 362-494) in its segment -- uniform split, evaluation (its four waves take
// 64-region batches, one lane per region, the kernel of
// include/pagani/eval_kernel.hpp), RefineError with the relerr flags, the
// iteration's sums (compensated, in a fixed order), the loop's decisions,
// filter + split -- with no host involvement, and writes the result.  An
// item whose list would outgrow the segment, or that reaches a branch the
// host takes (the error-budget fix, Workspace.cuh:121-146, or a memory
// model past a tenth of its budget, where the heuristic classifier may
// run), stops before that iteration changes its totals and writes its loop
// state (SegOut); the batch hands the list and the state to the regular
// loop, which continues from there.  The decisions are those of the
// device-resident loop (devloop_decide.hpp) for the iterations it keeps.
//
// Results agree with a lone run of the item to rounding: the per-region
// arithmetic is the same, but the evaluation always takes one wave per
// 64-region batch (a lone run shares small batches between waves) and the
// sums are reduced in another order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "device_common.hpp"
#include "genz_visit.hpp"
#include "pagani/eval_kernel.hpp"
#include "segments.hpp"
#include "ws_internal.hpp"

namespace pagani {
namespace internal {
namespace {

namespace ph = pagani::phases;
using ph::DD;

constexpr int kSegThreads = ph::kThreads;  // the block reductions' width
constexpr int kSegWaves = kSegThreads / 64;

std::size_t seg_bytes(int d, std::size_t S)
{
  // L[2], H[2] (d*S doubles each), est, err0 (S), pest (S/2), dim, active
  // (S bytes); every part 256-byte aligned
  auto al = [](std::size_t b) { return (b + 255) / 256 * 256; };
  return 4 * al(static_cast<std::size_t>(d) * S * 8) + 2 * al(S * 8) + al(S / 2 * 8) + 2 * al(S);
}

struct SegPtrs {
  double* L[2];
  double* H[2];
  double *est, *err0, *pest;
  std::uint8_t *dim, *active;
};

__host__ __device__ SegPtrs seg_ptrs(char* base, int d, std::size_t S)
{
  auto al = [](std::size_t b) { return (b + 255) / 256 * 256; };
  SegPtrs p;
  const std::size_t lb = al(static_cast<std::size_t>(d) * S * 8);
  p.L[0] = reinterpret_cast<double*>(base);
  p.L[1] = reinterpret_cast<double*>(base + lb);
  p.H[0] = reinterpret_cast<double*>(base + 2 * lb);
  p.H[1] = reinterpret_cast<double*>(base + 3 * lb);
  char* q = base + 4 * lb;
  p.est = reinterpret_cast<double*>(q);
  p.err0 = reinterpret_cast<double*>(q + al(S * 8));
  q += 2 * al(S * 8);
  p.pest = reinterpret_cast<double*>(q);
  q += al(S / 2 * 8);
  p.dim = reinterpret_cast<std::uint8_t*>(q);
  p.active = reinterpret_cast<std::uint8_t*>(q + al(S));
  return p;
}

// The loop state thread 0 keeps and the workgroup reads after each barrier.
struct SegLoop {
  double cum_est, cum_err;
  std::uint64_t nregions, iters;
  std::uint64_t n, m;
  int status;   // the result's status
  int stop;     // 0 go on, 1 finished here, 2 give the item to the regular loop
  double hist[3];  // Heuristic_classifier::store_estimate
  std::uint64_t iters_collected;
  std::uint64_t evaluated, iterations;
};

template <int D, class F>
__device__ void run_item(const F& f0, const SegItem& it, const SegPtrs& p, std::size_t S,
                         const pagani_eval_consts& C, std::size_t mem_total, int mem_model,
                         SegOut* out, SegLoop& s, DD* shd, std::uint32_t* shc,
                         std::uint32_t* wsum)
{
  F f = f0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // uniform split (Sub_regions.cuh:11-36)
  const std::size_t k = it.init_k;
  std::size_t n = 1;
  for (int j = 0; j < D; ++j) n *= k;
  {
    const double len = 1. / static_cast<double>(k);
    for (std::size_t t = tid; t < n; t += kSegThreads) {
      std::size_t q = t;
      for (int j = 0; j < D; ++j) {
        const std::size_t id = q % k;
        q /= k;
        p.L[0][static_cast<std::size_t>(j) * n + t] = static_cast<double>(id) * len;
        p.H[0][static_cast<std::size_t>(j) * n + t] = len;
      }
    }
  }
  if (tid == 0) {
    s.cum_est = 0;
    s.cum_err = 0;
    s.nregions = 0;
    s.iters = 0;
    s.status = 1;
    s.stop = 0;
    s.hist[0] = s.hist[1] = s.hist[2] = 0.0;
    s.iters_collected = 0;
    s.evaluated = 0;
    s.iterations = 0;
  }
  int cur = 0;
  bool parents = false;
  int iter = 0;
  __syncthreads();
  for (; iter < 700; ++iter) {
    // evaluation: 64-region batches, one lane per region
    const std::size_t nb = (n + 63) / 64;
    for (std::size_t b = wave; b < nb; b += kSegWaves) {
      const std::size_t t0 = b * 64 + lane;
      const bool live = t0 < n;
      const std::size_t t = live ? t0 : n - 1;
      detail::eval_regions_wave<D, F, double>(f, p.L[cur], p.H[cur], nullptr, n, t, live, C.lows,
                                              C.range, C.jac, C, 1, 0, lane, nullptr, p.est,
                                              p.err0, p.dim);
    }
    __syncthreads();
    // RefineError with the relerr flags, and the iteration's sums
    DD se{0, 0}, sr{0, 0}, sae{0, 0}, sar{0, 0};
    std::uint32_t c = 0;
    const std::size_t h = n / 2;
    for (std::size_t t = tid; t < n; t += kSegThreads) {
      const double e = p.est[t];
      double r = p.err0[t];
      bool a = true;
      if (parents) {
        const bool right = t >= h;
        const std::size_t sib = right ? t - h : t + h;
        const std::size_t par = right ? t - h : t;
        ph::refine_region(e, r, p.est[sib], p.err0[sib], p.pest[par], it.epsrel, it.relerr, r, a);
      }
      p.active[t] = a ? 1 : 0;
      se = ph::dd_add(se, e);
      sr = ph::dd_add(sr, r);
      if (a) {
        sae = ph::dd_add(sae, e);
        sar = ph::dd_add(sar, r);
        ++c;
      }
    }
    se = ph::block_reduce(se, shd);
    sr = ph::block_reduce(sr, shd);
    sae = ph::block_reduce(sae, shd);
    sar = ph::block_reduce(sar, shd);
    c = ph::block_count(c, shc);
    if (tid == 0) {
#pragma clang fp contract(off)
      const double iter_est = se.hi + se.lo, iter_err = sr.hi + sr.lo;
      const double act_est = sae.hi + sae.lo, act_err = sar.hi + sar.lo;
      const std::uint64_t m = c;
      // accuracy_reached, PaganiUtils.cuh:387-394
      const double tot_est = fabs(s.cum_est + iter_est);
      const double tot_err = s.cum_err + iter_err;
      if (tot_err / tot_est <= it.epsrel || tot_err <= it.epsabs) {
        s.cum_est += iter_est;
        s.cum_err += iter_err;
        s.status = 0;
        s.nregions += n;
        s.stop = 1;
        s.evaluated += n;
        s.iterations++;
      } else {
        // compute_finished_estimates, PaganiUtils.cuh:368-385
        const double fin_est = iter_est - act_est;
        const double fin_err = iter_err - act_err;
        const bool overflow = s.cum_err + fin_err > fabs(s.cum_est + iter_est) * it.epsrel;
        if (overflow || mem_ratio(n, D, mem_total, mem_model) > .1 || 2 * m > S) {
          // the regular loop redoes this iteration from the state before it
          s.stop = 2;
        } else {
          s.evaluated += n;
          s.iterations++;
          s.hist[0] = s.hist[1];
          s.hist[1] = s.hist[2];
          s.hist[2] = s.cum_est + iter_est;
          s.iters_collected++;
          s.cum_est += fin_est;
          s.cum_err += fin_err;
          s.nregions += n - m;
          s.iters++;
          s.m = m;
          if (m == 0) {
            s.stop = 1;
          } else if (iter + 1 >= 700) {
            // the loop's iteration limit: the next list is summed unevaluated
            s.nregions += 2 * m;
            s.stop = 1;
          }
        }
      }
    }
    __syncthreads();
    if (s.stop) break;
    // filter + split (Sub_region_filter.cuh:8-47, Sub_region_splitter.cuh:
    // 8-46): the kept region at compacted index idx has children idx and
    // m + idx, halved along its own axis
    const std::size_t m = s.m;
    const int nxt = cur ^ 1;
    std::size_t run = 0;
    for (std::size_t tile = 0; tile < n; tile += kSegThreads) {
      const std::size_t t = tile + tid;
      const bool a = t < n && p.active[t] == 1;
      const unsigned long long ballot = __ballot(a);
      const unsigned long long below = lane ? (ballot & ((1ull << lane) - 1ull)) : 0ull;
      if (lane == 0) wsum[wave] = static_cast<std::uint32_t>(__popcll(ballot));
      __syncthreads();
      std::uint32_t woff = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < kSegWaves; ++w) {
        if (w < wave) woff += wsum[w];
        tot += wsum[w];
      }
      __syncthreads();
      if (a) {
        const std::size_t idx = run + woff + static_cast<std::uint32_t>(__popcll(below));
        const int b = p.dim[t];
        const std::size_t n2 = 2 * m;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double l = p.L[cur][static_cast<std::size_t>(j) * n + t];
          const double w = p.H[cur][static_cast<std::size_t>(j) * n + t];
          const std::size_t o = static_cast<std::size_t>(j) * n2 + idx;
          if (j == b) {
            const double half = w / 2;
            p.L[nxt][o] = l;
            p.L[nxt][o + m] = l + half;
            p.H[nxt][o] = half;
            p.H[nxt][o + m] = half;
          } else {
            p.L[nxt][o] = l;
            p.L[nxt][o + m] = l;
            p.H[nxt][o] = w;
            p.H[nxt][o + m] = w;
          }
        }
        p.pest[idx] = p.est[t];
      }
      run += tot;
    }
    __syncthreads();
    n = 2 * m;
    cur = nxt;
    parents = true;
  }
  if (tid == 0) {
    SegOut o{};
    o.r.status = s.status;
    o.r.lastPhase = -1;
    o.r.estimate = s.cum_est;
    o.r.errorest = s.cum_err;
    o.r.nregions = s.nregions;
    o.r.iters = s.iters;
    o.outcome = s.stop == 2 ? 1 : 0;
    o.evaluated = s.evaluated;
    o.iterations = s.iterations;
    if (s.stop == 2) {
      o.it = static_cast<std::uint64_t>(iter);
      o.n = n;
      o.has_parents = parents ? 1 : 0;
      o.cur = cur;
      o.block = blockIdx.x;
      for (int k = 0; k < 3; ++k) o.hist[k] = s.hist[k];
      o.iters_collected = s.iters_collected;
    }
    *out = o;
  }
  __syncthreads();
}

// The family a Genz functor type implements.
template <class Fn, int D>
constexpr int family_of()
{
  using T = std::decay_t<Fn>;
  if constexpr (std::is_same_v<T, pagani::Oscillatory<double, D>>) return PAGANI_GENZ_OSCILLATORY;
  if constexpr (std::is_same_v<T, pagani::ProductPeak<double, D>>) return PAGANI_GENZ_PRODUCT_PEAK;
  if constexpr (std::is_same_v<T, pagani::CornerPeak<double, D>>) return PAGANI_GENZ_CORNER_PEAK;
  if constexpr (std::is_same_v<T, pagani::Gaussian<double, D>>) return PAGANI_GENZ_GAUSSIAN;
  if constexpr (std::is_same_v<T, pagani::C0<double, D>>) return PAGANI_GENZ_C0;
  if constexpr (std::is_same_v<T, pagani::Discontinuous<double, D>>)
    return PAGANI_GENZ_DISCONTINUOUS;
  return -1;
}

// One kernel per (dimension, family): the loop around one integrand only
// keeps the registers that integrand needs.
template <int D, int Fam>
__global__ void __launch_bounds__(kSegThreads)
    k_segments(const SegItem* __restrict__ items, std::uint32_t count,
               std::uint32_t* __restrict__ next, char* __restrict__ pool, std::size_t bytes,
               std::size_t S, const pagani_eval_consts C, std::size_t mem_total, int mem_model,
               SegOut* __restrict__ out)
{
  __shared__ SegLoop s;
  __shared__ DD shd[kSegWaves];
  __shared__ std::uint32_t shc[kSegWaves];
  __shared__ std::uint32_t wsum[kSegWaves];
  __shared__ std::uint32_t item;
  const SegPtrs p = seg_ptrs(pool + static_cast<std::size_t>(blockIdx.x) * bytes, D, S);
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(next, 1u);
    __syncthreads();
    const std::uint32_t i = item;
    __syncthreads();
    if (i >= count) return;
    if (items[i].f.ndim != D) {
      if (threadIdx.x == 0) out[i].outcome = 1;
      continue;
    }
    const SegItem it = items[i];
    bool ran = false;
    visit_genz<double, D>(&it.f, [&](const auto& f) {
      if constexpr (family_of<decltype(f), D>() == Fam) {
        run_item<D>(f, it, p, S, C, mem_total, mem_model, &out[i], s, shd, shc, wsum);
        ran = true;
      }
      return 0;
    });
    // an item of another family or dimension: the regular loop takes it
    if (!ran && threadIdx.x == 0) out[i].outcome = 1;
  }
}

template <int D = PAGANI_MIN_DIM, class Fn>
hipError_t with_genz_dim(int d, Fn&& fn)
{
  if constexpr (D > PAGANI_GENZ_MAX_DIM) {
    return hipErrorInvalidValue;
  } else {
    if (d == D) return fn(std::integral_constant<int, D>{});
    return with_genz_dim<D + 1>(d, fn);
  }
}

template <int Fam = PAGANI_GENZ_OSCILLATORY, class Fn>
hipError_t with_genz_family(int fam, Fn&& fn)
{
  if constexpr (Fam > PAGANI_GENZ_DISCONTINUOUS) {
    return hipErrorInvalidValue;
  } else {
    if (fam == Fam) return fn(std::integral_constant<int, Fam>{});
    return with_genz_family<Fam + 1>(fam, fn);
  }
}

// grid, cut to the workgroups of k_segments<D, Fam> the device holds at once
// (more would only wait for the first to finish)
template <int D, int Fam>
unsigned resident(unsigned grid)
{
  static const unsigned cap = [] {
    int b = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_segments<D, Fam>, kSegThreads, 0) !=
            hipSuccess ||
        hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
      (void)hipGetLastError();
      return 0u;
    }
    return static_cast<unsigned>(std::max(b, 1) * std::max(prop.multiProcessorCount, 1));
  }();
  return cap ? std::min(grid, cap) : grid;
}

}  // namespace

std::size_t segment_bytes(int d, std::size_t regions) { return seg_bytes(d, regions); }

SegView segment_view(char* base, int d, std::size_t regions)
{
  const SegPtrs p = seg_ptrs(base, d, regions);
  SegView v;
  for (int k = 0; k < 2; ++k) {
    v.L[k] = p.L[k];
    v.H[k] = p.H[k];
  }
  v.pest = p.pest;
  return v;
}

hipError_t launch_segments(hipStream_t s, int d, int family, const SegItem* items,
                           std::uint32_t count, std::uint32_t* next, char* pool,
                           std::size_t bytes, std::size_t S, unsigned grid,
                           const pagani_eval_consts& C, std::size_t mem_total, int mem_model,
                           SegOut* out)
{
  return with_genz_dim(d, [&](auto D) {
    return with_genz_family(family, [&](auto F) {
      constexpr int Dv = decltype(D)::value, Fv = decltype(F)::value;
      hipLaunchKernelGGL((k_segments<Dv, Fv>), dim3(grid), dim3(kSegThreads), 0, s, items,
                         count, next, pool, bytes, S, C, mem_total, mem_model, out);
      return hipGetLastError();
    });
  });
}

unsigned segment_grid(int d, int family, unsigned want)
{
  unsigned g = want;
  (void)with_genz_dim(d, [&](auto D) {
    return with_genz_family(family, [&](auto F) {
      g = resident<decltype(D)::value, decltype(F)::value>(want);
      return hipSuccess;
    });
  });
  return g;
}

}  // namespace internal
}  // namespace pagani
