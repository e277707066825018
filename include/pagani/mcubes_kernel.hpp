// mcubes_kernel.hpp -- the m-Cubes sampling kernel for gfx950, instantiated
// per integrand in the including translation unit (compile with hipcc).
//
// One launch draws npg points in every cube of the ng^ndim stratification
// (vegas_kernel / vegas_kernelF, cuda/mcubes/vegasT.cuh:400-612).  Each
// thread takes a chunk of cubes_per_thread consecutive cubes, stepping their
// stratum coordinates like an odometer (vegasT.cuh:386-391); each cube draws
// from its own random stream, so the points do not depend on the launch
// shape (only the order of the sums does).
// The importance grid is staged in LDS (one table of right edges per axis)
// and the adjusting iterations' bin histograms of f^2 are accumulated in
// LDS (fp64 LDS atomics) and written once per workgroup; the reference adds
// every sample into global memory with atomicAdd.  Workgroup sums of fb and
// f2b are formed in a fixed order, so ti and tsi of a launch are
// reproducible; the histograms are reproducible up to the order of the LDS
// additions within a workgroup.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "pagani.h"

namespace pagani {
namespace mc {

constexpr int kBins = PAGANI_MC_BINS;
constexpr int kThreads = 256;
constexpr double kTiny = 1.0e-30;  // TINY, vegas_utils.cuh / seqCodesDefs.hh

// Random streams.  A stream yields one sample's D numbers as integers x
// whose value is x * kScale exactly; the kernel forms kg - ran as one fused
// multiply-add, which rounds exactly as the subtraction of the value would.
//
// LcgStream: the reference's Custom_generator (vegas_utils.cuh:24-48),
// s' = (1103515245 s + 12345) mod 2^31 in 32-bit unsigned arithmetic, the
// value s' / 2^31, reseeded with the cube index (vegasT.cuh:351-353).
struct LcgStream {
  static constexpr double kScale = 0x1.0p-31;
  std::uint32_t s;
  __device__ LcgStream(std::uint64_t, std::uint32_t cube) : s(cube) {}
  template <int D>
  __device__ void fill(std::uint32_t (&x)[D])
  {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      s = (1103515245u * s + 12345u) & 0x7fffffffu;
      x[j] = s;
    }
  }
};

__host__ __device__ inline std::uint64_t splitmix64(std::uint64_t& z)
{
  // SplitMix64 (Steele, Lea and Flood 2014)
  std::uint64_t x = (z += 0x9e3779b97f4a7c15ull);
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Key of iteration `it` (1-based) of a run seeded with `seed`.
__host__ __device__ inline std::uint64_t iter_key(std::uint64_t seed, int it)
{
  std::uint64_t z = seed ^ (static_cast<std::uint64_t>(static_cast<std::uint32_t>(it)) << 32);
  return splitmix64(z);
}

// Default stream: xoroshiro128++ (Blackman and Vigna 2018), state from two
// SplitMix64 outputs of key + cube * 2^64/phi.  Each 64-bit output gives
// two numbers, from its low then its high 32 bits x: (x | 1) 2^-32, the
// midpoints of 2^31 equal cells of (0, 1) (32 bits place a point within its
// stratum to 2^-31 of the stratum width).  A sample draws ceil(D/2)
// outputs; with D odd the last high half is unused.
struct XoroStream {
  static constexpr double kScale = 0x1.0p-32;
  std::uint64_t s0, s1;
  __device__ XoroStream(std::uint64_t key, std::uint32_t cube)
  {
    std::uint64_t z = key + static_cast<std::uint64_t>(cube) * 0x9e3779b97f4a7c15ull;
    s0 = splitmix64(z);
    s1 = splitmix64(z);
  }
  __device__ std::uint64_t next()
  {
    const std::uint64_t a = s0;
    const std::uint64_t t = s1 ^ a;
    const std::uint64_t sum = a + s1;
    const std::uint64_t r = ((sum << 17) | (sum >> 47)) + a;
    s0 = ((a << 49) | (a >> 15)) ^ t ^ (t << 21);
    s1 = (t << 28) | (t >> 36);
    return r;
  }
  template <int D>
  __device__ void fill(std::uint32_t (&x)[D])
  {
#pragma unroll
    for (int j = 0; j < D; j += 2) {
      const std::uint64_t v = next();
      x[j] = static_cast<std::uint32_t>(v) | 1u;
      if (j + 1 < D) x[j + 1] = static_cast<std::uint32_t>(v >> 32) | 1u;
    }
  }
};

namespace detail {

template <class F, std::size_t... I>
__device__ __forceinline__ double apply(F& f, const double* x, std::index_sequence<I...>)
{
  return static_cast<double>(f(x[I]...));
}

__device__ __forceinline__ double wave_sum(double v)
{
  // fixed butterfly: every lane ends with the same, order-determined sum
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace detail

// LDS bytes of one workgroup.
__host__ __device__ constexpr std::size_t lds_bytes(int ndim, int hist_dims)
{
  return (static_cast<std::size_t>(ndim) * (kBins + 1) +
          static_cast<std::size_t>(hist_dims) * kBins) *
         sizeof(double);
}

template <int D, class F, class Stream>
__global__ void __launch_bounds__(kThreads) sample_kernel(F f, pagani_mc_args a)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  double* sxi = lds;
  double* shist = lds + D * (kBins + 1);
  __shared__ double wsum[2][kThreads / 64];
  const int hd = static_cast<int>(a.hist_dims);
  const int tid = threadIdx.x;
  for (int i = tid; i < D * (kBins + 1); i += kThreads) sxi[i] = a.xi[i];
  for (int i = tid; i < hd * kBins; i += kThreads) shist[i] = 0.0;
  __syncthreads();

  // Thread g of the grid (wave W = g / 64, lane l) takes chunk
  // l * (T / 64) + W of the T chunks: the lanes of a wave work in cubes far
  // apart, in different strata of every axis, so their bins (LDS reads and
  // histogram additions) rarely collide.
  const std::uint32_t T = gridDim.x * kThreads;
  const std::uint32_t g = blockIdx.x * kThreads + tid;
  const std::uint32_t chunk = (g & 63u) * (T >> 6) + (g >> 6);
  const std::uint64_t first = static_cast<std::uint64_t>(chunk) * a.cubes_per_thread;
  double fbg = 0.0, f2bg = 0.0;
  if (first < a.ncubes) {
    const std::uint64_t last64 = first + a.cubes_per_thread;
    const std::uint32_t end = last64 < a.ncubes ? static_cast<std::uint32_t>(last64) : a.ncubes;
    const std::uint32_t ng = a.ng;
    // get_indx (vegasT.cuh:137-156): digits of the first cube, most
    // significant first, plus one
    std::uint32_t kg[D];
    double kgd[D];  // kg as doubles, kept in step
    {
      std::uint32_t m = static_cast<std::uint32_t>(first);
#pragma unroll
      for (int j = D - 1; j >= 0; --j) {
        const std::uint32_t q = m / ng;
        kg[j] = m - q * ng + 1;
        kgd[j] = static_cast<double>(kg[j]);
        m = q;
      }
    }
    const double dxg = a.dxg, xnd = a.xnd, xjac = a.xjac;
    const double dnpg = static_cast<double>(a.npg);
    for (std::uint32_t c = static_cast<std::uint32_t>(first); c < end; ++c) {
      Stream rng(a.key, c);
      double fb = 0.0, f2b = 0.0;
      for (std::uint32_t k = 0; k < a.npg; ++k) {
        double wgt = xjac;
        double x[D];
        std::uint32_t ran[D];
        int ia[D];
        rng.fill(ran);
        // Setup_Integrand_Eval, vegasT.cuh:213-241, in two passes so that
        // every axis' pair of bin edges is requested from LDS before the
        // first is used.  kg - ran00 is one FMA on the stream's integer; bin
        // 1 takes its left edge 0 from the table instead of a branch.
        double xn[D], lo[D], hi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double kr = __builtin_fma(-static_cast<double>(ran[j]), Stream::kScale, kgd[j]);
          xn[j] = __builtin_fma(kr, dxg, 1.0);
          int b = static_cast<int>(xn[j]);
          b = b > kBins ? kBins : b;
          b = b < 1 ? 1 : b;
          ia[j] = b;
          const double* t = sxi + j * (kBins + 1) + b;
          hi[j] = t[0];
          lo[j] = t[-1];
        }
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const double xo = hi[j] - lo[j];
          const double rc = __builtin_fma(xn[j] - static_cast<double>(ia[j]), xo, lo[j]);
          x[j] = __builtin_fma(rc, a.dx[j], a.lo[j]);
          wgt *= xo * xnd;
        }
        const double fx = wgt * detail::apply(f, x, std::make_index_sequence<D>());
        const double f2 = fx * fx;
        fb += fx;
        f2b += f2;
#pragma unroll
        for (int j = 0; j < D; ++j)
          if (j < hd)  // uniform: a scalar branch, ia[j] stays a register
            __hip_atomic_fetch_add(&shist[j * kBins + ia[j] - 1], f2, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      // vegasT.cuh:373-384
      f2b = __builtin_sqrt(f2b * dnpg);
      f2b = (f2b - fb) * (f2b + fb);
      if (f2b <= 0.0) f2b = kTiny;
      fbg += fb;
      f2bg += f2b;
#pragma unroll
      for (int j = D - 1; j >= 0; --j) {
        if (kg[j] != ng) {
          ++kg[j];
          kgd[j] = static_cast<double>(kg[j]);
          break;
        }
        kg[j] = 1;
        kgd[j] = 1.0;
      }
    }
  }
  // workgroup sums in a fixed order
  fbg = detail::wave_sum(fbg);
  f2bg = detail::wave_sum(f2bg);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) {
    wsum[0][wave] = fbg;
    wsum[1][wave] = f2bg;
  }
  __syncthreads();
  if (tid == 0) {
    double s0 = wsum[0][0], s1 = wsum[1][0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) {
      s0 += wsum[0][w];
      s1 += wsum[1][w];
    }
    a.part[2 * blockIdx.x] = s0;
    a.part[2 * blockIdx.x + 1] = s1;
  }
  if (hd) {
    double* out = a.hist + static_cast<std::size_t>(blockIdx.x) * hd * kBins;
    for (int i = tid; i < hd * kBins; i += kThreads) out[i] = shist[i];
  }
}

// Launch for one integrand; the grid and the per-thread cube count come
// from the args (the library sizes them).
template <int D, class F, class Stream>
int launch(const F& f, const pagani_mc_args* a)
{
  static_assert(D >= 1 && D <= PAGANI_MC_MAX_DIM, "m-Cubes supports 1..16 dimensions");
  if (!a || a->grid == 0 || a->hist_dims > static_cast<std::uint32_t>(D) ||
      !a->xi || !a->part || (a->hist_dims && !a->hist))
    return PAGANI_EBADARG;
  const std::size_t lds = lds_bytes(D, static_cast<int>(a->hist_dims));
  auto kern = sample_kernel<D, F, Stream>;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             static_cast<int>(lds));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return PAGANI_EUNSUPPORTED;
    }
  }
  hipLaunchKernelGGL(kern, dim3(a->grid), dim3(kThreads), lds,
                     static_cast<hipStream_t>(a->stream), f, *a);
  return hipGetLastError() == hipSuccess ? PAGANI_OK : PAGANI_EHIP;
}

// Dispatch on args->generator.
template <int D, class F>
int launch_any(const F& f, const pagani_mc_args* a)
{
  if (!a) return PAGANI_EBADARG;
  if (a->generator == PAGANI_MC_RNG_LCG) return launch<D, F, LcgStream>(f, a);
  if (a->generator == PAGANI_MC_RNG_DEFAULT) return launch<D, F, XoroStream>(f, a);
  return PAGANI_EBADARG;
}

// pagani_mc_sample_fn for an integrand passed as ctx (a host copy of F).
template <int D, class F, class Stream>
int sample_callback(void* ctx, const pagani_mc_args* a)
{
  return launch<D, F, Stream>(*static_cast<const F*>(ctx), a);
}

}  // namespace mc
}  // namespace pagani
