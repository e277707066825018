// phases.hip -- region-list kernels for gfx950 (see phases.hpp).
//
// Compiled with -ffp-contract=off: the refinement and classification
// arithmetic is plain IEEE double in the reference's operation order, so
// these phases are bit-reproducible against a CPU restatement.
#include "phases.hpp"

#include <cmath>
#include <type_traits>

#include "classify_search.hpp"
#include "device_common.hpp"
#include "devloop_decide.hpp"

namespace pagani {
namespace phases {

namespace {

// Write the chunk record: four DD sums and a count.
__device__ __forceinline__ void write_record(DD s0, DD s1, DD s2, DD s3, std::uint32_t c,
                                             double* part, std::uint32_t* cnt,
                                             std::size_t chunk = blockIdx.x)
{
  __shared__ DD sh[kThreads / 64];
  __shared__ std::uint32_t shc[kThreads / 64];
  s0 = block_reduce(s0, sh);
  s1 = block_reduce(s1, sh);
  s2 = block_reduce(s2, sh);
  s3 = block_reduce(s3, sh);
  c = block_count(c, shc);
  if (threadIdx.x == 0) {
    double* p = part + chunk * kPartDoubles;
    p[0] = s0.hi;
    p[1] = s0.lo;
    p[2] = s1.hi;
    p[3] = s1.lo;
    p[4] = s2.hi;
    p[5] = s2.lo;
    p[6] = s3.hi;
    p[7] = s3.lo;
    if (cnt) cnt[chunk] = c;
  }
}

__global__ void __launch_bounds__(256) k_uniform_split(int d, std::size_t k, std::size_t n,
                                                       double len, double* L, double* H,
                                                       int levels)
{
  const std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  std::size_t q = t;
  for (int j = 0; j < d; ++j) {
    const std::size_t id = q % k;
    q /= k;
    L[static_cast<std::size_t>(j) * n + t] = static_cast<double>(id) * len;
    if (levels)
      reinterpret_cast<std::uint16_t*>(H)[static_cast<std::size_t>(j) * n + t] = 0;
    else
      H[static_cast<std::size_t>(j) * n + t] = len;
  }
}

// Row `row` of a list's bounds along axis j at stride `stride`: lower bound
// and width (from the level when lv.on), and the width's level.
__device__ __forceinline__ void read_row(const double* __restrict__ L,
                                         const double* __restrict__ H, const Levels& lv,
                                         std::size_t o, int j, double& l, double& w, int& lev)
{
  l = L[o];
  if (lv.on) {
    lev = reinterpret_cast<const std::uint16_t*>(H)[o];
    w = __builtin_ldexp(lv.h0[j], -lev);
  } else {
    lev = 0;
    w = H[o];
  }
}

// RefineError (Phases.cuh:219-268) and the relerr flag; children t and t+n/2
// share parent t mod n/2.
__device__ __forceinline__ void refine_chunk(std::size_t n, const double* __restrict__ est,
                                             const double* err_in, double* err_out,
                                             const double* __restrict__ pest, double epsrel,
                                             int relerr, std::uint8_t* __restrict__ active,
                                             double* __restrict__ part,
                                             std::uint32_t* __restrict__ cnt,
                                             std::size_t chunk = blockIdx.x)
{
  std::size_t base, end;
  chunk_range(chunk, n, pest != nullptr, base, end);
  const std::size_t h = n / 2;
  DD se{0, 0}, sr{0, 0}, sae{0, 0}, sar{0, 0};
  std::uint32_t c = 0;
  double mn = HUGE_VAL, mx = -HUGE_VAL;  // of the refined errors
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    if (t < end) {
      const double e = est[t];
      double r = err_in[t];
      bool a = true;
      if (pest) {
        const bool right = t >= h;
        const std::size_t sib = right ? t - h : t + h;
        const std::size_t par = right ? t - h : t;
        refine_region(e, r, est[sib], err_in[sib], pest[par], epsrel, relerr, r, a);
      }
      err_out[t] = r;
      active[t] = a ? 1 : 0;
      se = dd_add(se, e);
      sr = dd_add(sr, r);
      mn = fmin(mn, r);
      mx = fmax(mx, r);
      if (a) {
        sae = dd_add(sae, e);
        sar = dd_add(sar, r);
        ++c;
      }
    }
  }
  write_record(se, sr, sae, sar, c, part, cnt, chunk);
  // the errors' minimum and maximum in the record's last two slots, as
  // k_minmax writes them: the classifier's search range without a pass of
  // its own when it runs on these errors
  __shared__ double smn[kThreads / 64], smx[kThreads / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, off, 64));
    mx = fmax(mx, __shfl_xor(mx, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smn[wave] = mn;
    smx[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      mn = fmin(mn, smn[w]);
      mx = fmax(mx, smx[w]);
    }
    double* p = part + chunk * kPartDoubles;
    p[8] = mn;
    p[9] = mx;
  }
}

// A region's sibling lies n/2 away, in chunk c + L of a paired list's
// chunks (num_chunks(n, true) = 2L): workgroups are paired so that a chunk
// and its sibling chunk are read by workgroups running at the same time on
// the same XCD (workgroups are dealt to the 8 XCDs round robin, so b and
// b + 8 share one), and the second read of the siblings and their parent
// comes from that XCD's L2.  In groups of 16 workgroups, the first 8 take 8
// chunks of the first half and the next 8 their sibling chunks; the chunks
// left over alternate between the halves.
__device__ __forceinline__ std::size_t paired_chunk(std::size_t b, std::size_t nch)
{
  const std::size_t half = (nch + 1) / 2, groups = (nch - half) / 8;
  if (b < 16 * groups) {
    const std::size_t g = b / 16, r = b % 16;
    return r < 8 ? g * 8 + r : half + g * 8 + (r - 8);
  }
  const std::size_t t = b - 16 * groups;
  return (t & 1) ? half + 8 * groups + (t >> 1) : 8 * groups + (t >> 1);
}

__global__ void __launch_bounds__(kThreads)
    k_refine(std::size_t n, const double* __restrict__ est, const double* err_in,
             double* err_out, const double* __restrict__ pest, double epsrel, int relerr,
             std::uint8_t* __restrict__ active, double* __restrict__ part,
             std::uint32_t* __restrict__ cnt)
{
  refine_chunk(n, est, err_in, err_out, pest, epsrel, relerr, active, part, cnt,
               paired_chunk(blockIdx.x, gridDim.x));
}

__global__ void __launch_bounds__(kThreads)
    k_count_flags(std::size_t n, const std::uint8_t* __restrict__ flags,
                  const double* __restrict__ est, const double* __restrict__ err,
                  double* __restrict__ part, std::uint32_t* __restrict__ cnt, int paired)
{
  std::size_t base, end;
  chunk_range(blockIdx.x, n, paired != 0, base, end);
  DD se{0, 0}, sr{0, 0}, sae{0, 0}, sar{0, 0};
  std::uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    if (t < end) {
      const double e = est[t], r = err[t];
      se = dd_add(se, e);
      sr = dd_add(sr, r);
      if (flags[t] == 1) {
        sae = dd_add(sae, e);
        sar = dd_add(sar, r);
        ++c;
      }
    }
  }
  write_record(se, sr, sae, sar, c, part, cnt);
}

// flags of the classifier's accepted threshold, with the refine-style
// record (device_set_true_for_larger_than, heuristic_classifier.cuh:60-87,
// and the dot products of Workspace.cuh:107-110).
__global__ void __launch_bounds__(kThreads)
    k_threshold_flags(std::size_t n, const double* __restrict__ est,
                      const double* __restrict__ err, double thr,
                      std::uint8_t* __restrict__ flags, double* __restrict__ part,
                      std::uint32_t* __restrict__ cnt, int paired)
{
  std::size_t base, end;
  chunk_range(blockIdx.x, n, paired != 0, base, end);
  DD se{0, 0}, sr{0, 0}, sae{0, 0}, sar{0, 0};
  std::uint32_t c = 0;
  // every load of the chunk first, so they are all in flight together
  double ev[kPerThread], rv[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    ev[k] = t < end ? est[t] : 0.0;
    rv[k] = t < end ? err[t] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    if (t < end) {
      const double e = ev[k], r = rv[k];
      const bool a = r > thr;
      flags[t] = a ? 1 : 0;
      se = dd_add(se, e);
      sr = dd_add(sr, r);
      if (a) {
        sae = dd_add(sae, e);
        sar = dd_add(sar, r);
        ++c;
      }
    }
  }
  write_record(se, sr, sae, sar, c, part, cnt);
}

// The classifier's candidate thresholds in a single read of the errors: per
// candidate t_k, the count of err > t_k (apply_threshold,
// heuristic_classifier.cuh:282-303) and the sum of those errors (the dot
// product of evaluate_error_budget, :305-325).  The candidates arrive sorted
// ascending and padded with +inf to K, at most K - 1 of them real.  Each
// thread keeps K - 1 buckets in LDS: bucket k holds its errors with
// t_k < err <= t_(k+1), found by a binary search over the candidates, so an
// error costs a few compares instead of one compare and masked add per
// candidate; the count and sum of err > t_k are then the thread's buckets
// k..K-2.  Fixed grid, grid-stride order, per-thread sums in element order,
// fixed reduction trees: deterministic.
template <int K>
__device__ __forceinline__ void multi_threshold_block(std::size_t n, const double* __restrict__ err,
                                                      const ThresholdSet& T,
                                                      double* __restrict__ bpart)
{
  static_assert((K & (K - 1)) == 0, "binary search over a power of two");
  __shared__ double ts[K];
  __shared__ double bsum[K - 1][kThreads];
  __shared__ std::uint32_t bcnt[K - 1][kThreads];
  __shared__ DD sh[kThreads / 64];
  __shared__ std::uint32_t shc[kThreads / 64];
  const int tid = threadIdx.x;
  if (tid < K) ts[tid] = T.t[tid];
#pragma unroll
  for (int k = 0; k < K - 1; ++k) {
    bsum[k][tid] = 0.0;
    bcnt[k][tid] = 0;
  }
  __syncthreads();
  auto add = [&](double r) {
    // b = #{k : r > t_k} over t_0..t_(K-2) (t_(K-1) is +inf)
    int b = 0;
#pragma unroll
    for (int step = K / 2; step > 0; step >>= 1)
      if (r > ts[b + step - 1]) b += step;
    if (b > 0) {
      bsum[b - 1][tid] += r;
      bcnt[b - 1][tid] += 1u;
    }
  };
  // Grid-stride order; sixteen loads are issued before their values are
  // consumed (in the same order): the LDS buckets leave room for three
  // workgroups per CU, so each lane keeps more bytes in flight.
  constexpr int kAhead = 16;
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * kThreads;
  std::size_t t = static_cast<std::size_t>(blockIdx.x) * kThreads + tid;
  for (; t + (kAhead - 1) * stride < n; t += kAhead * stride) {
    double r[kAhead];
#pragma unroll
    for (int i = 0; i < kAhead; ++i) r[i] = err[t + i * stride];
#pragma unroll
    for (int i = 0; i < kAhead; ++i) add(r[i]);
  }
  for (; t < n; t += stride) add(err[t]);
  // suffix sums: candidate k counts the buckets k..K-2
  double* out = bpart + static_cast<std::size_t>(blockIdx.x) * 3 * kMaxThresholds;
  double sum = 0.0;
  std::uint32_t cnt = 0;
  for (int k = K - 2; k >= 0; --k) {
    sum += bsum[k][tid];
    cnt += bcnt[k][tid];
    if (k < T.k) {  // uniform
      const DD s = block_reduce(DD{sum, 0.0}, sh);
      const std::uint32_t c = block_count(cnt, shc);
      if (tid == 0) {
        out[3 * k] = s.hi;
        out[3 * k + 1] = s.lo;
        out[3 * k + 2] = static_cast<double>(c);
      }
    }
  }
}

template <int K>
__global__ void __launch_bounds__(kThreads)
    k_multi_threshold(std::size_t n, const double* __restrict__ err, const ThresholdSet T,
                      double* __restrict__ bpart)
{
  multi_threshold_block<K>(n, err, T, bpart);
}

// The pass the device search queued (nothing when it queued none).
template <int K>
__global__ void __launch_bounds__(kThreads)
    k_multi_threshold_dev(std::size_t n, const double* __restrict__ err,
                          const internal::SearchState* __restrict__ st, double* __restrict__ bpart)
{
  if (!st->pending) return;
  // the candidates are read where they lie (a local copy of the set would
  // go to scratch memory, thread by thread)
  multi_threshold_block<K>(n, err, st->S, bpart);
}

// Fold the kMultiBlocks partials of k_multi_threshold: one wave per
// threshold (workgroup k), lane j folds blocks j, j+64, j+128, ... in order,
// then a fixed butterfly over the wave.
__device__ __forceinline__ void multi_reduce_wave(int nblocks, int nthr,
                                                  const double* __restrict__ bpart,
                                                  double* __restrict__ out)
{
  const int k = blockIdx.x, j = threadIdx.x;
  DD s{0.0, 0.0};
  double c = 0.0;  // counts are integers below 2^53: exact
  for (int b = j; b < nblocks; b += 64) {
    const double* p = bpart + static_cast<std::size_t>(b) * 3 * kMaxThresholds + 3 * k;
    s = dd_add(s, DD{p[0], p[1]});
    c += p[2];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    DD o;
    o.hi = __shfl_xor(s.hi, off, 64);
    o.lo = __shfl_xor(s.lo, off, 64);
    s = dd_add(s, o);
    c += __shfl_xor(c, off, 64);
  }
  if (j == 0) {
    out[k] = c;
    out[nthr + k] = s.hi + s.lo;
  }
}

__global__ void __launch_bounds__(64)
    k_multi_reduce(int nblocks, int nthr, const double* __restrict__ bpart,
                   double* __restrict__ out)
{
  multi_reduce_wave(nblocks, nthr, bpart, out);
}

__global__ void __launch_bounds__(64)
    k_multi_reduce_dev(int nblocks, const internal::SearchState* __restrict__ st,
                       const double* __restrict__ bpart, double* __restrict__ out)
{
  if (!st->pending || static_cast<int>(blockIdx.x) >= st->S.k) return;
  multi_reduce_wave(nblocks, st->S.k, bpart, out);
}

// The heuristic classifier's search on the device, one wave: first != 0
// starts a search from `in` with the errors' min and max from the minmax
// finalize (sc); otherwise the last queued pass's results (mout) join the
// measured states.  Then search_replay runs the reference's loop over them
// (every lane the same scalar work, the lookups spread over the lanes) and
// either records the result or queues the next pass.  A search that would
// outgrow the cache or the step limit stops with step.overflow set, for the
// host to finish.
__global__ void __launch_bounds__(64)
    k_search(internal::SearchState* __restrict__ st, const Scalars* __restrict__ sc,
             const double* __restrict__ mout, int first, internal::SearchInputs in)
{
  namespace I = internal;
  const int lane = threadIdx.x;
  int len = 0;
  bool fresh = false;  // the last pass's results, read from mout
  if (first) {
    in.mn = sc->mn;
    in.mx = sc->mx;
  } else {
    if (!st->pending) return;
    in = st->in;
    len = st->len;
    fresh = true;
  }
  // measured states: the cache of earlier launches, then the last pass
  // (slot i of the pass holds tree node perm[i]).  Each lane holds its share
  // of the states in registers (entry q*64 + lane, NaN where there is none,
  // which equals nothing, as a NaN state equals nothing), so a lookup is a
  // few compares and ballots; only a hit reads memory.
  static_assert(I::kSearchCache <= 4 * 64, "four cache entries per lane");
  double cv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = q * 64 + lane;
    cv[q] = i < len ? st->ct[i] : __builtin_nan("");
  }
  const double fv =
      fresh && lane < I::kTreeNodes ? st->node[st->perm[lane]] : __builtin_nan("");
  auto find = [&](double x, std::uint64_t& c, double& d) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const unsigned long long hit = __ballot(cv[q] == x);
      if (hit) {
        const int j = q * 64 + __ffsll(static_cast<long long>(hit)) - 1;
        c = st->ccnt[j];
        d = st->cdot[j];
        return true;
      }
    }
    const unsigned long long hit = __ballot(fv == x);
    if (hit) {
      const int slot = __ffsll(static_cast<long long>(hit)) - 1;
      c = static_cast<std::uint64_t>(mout[slot]);
      d = mout[st->S.k + slot];
      return true;
    }
    return false;
  };
  I::SearchStep step;
  I::search_replay(in, find, step);
  // every lane has read what it needs: the state may change now
  __syncthreads();
  if (lane == 0) {
    if (first) {
      st->in = in;
      st->passes = 0;
    }
    if (fresh) {
      for (int i = 0; i < I::kTreeNodes; ++i) {
        st->ct[len + i] = st->node[st->perm[i]];
        st->ccnt[len + i] = static_cast<std::uint64_t>(mout[i]);
        st->cdot[len + i] = mout[st->S.k + i];
      }
      len += I::kTreeNodes;
    }
    st->len = len;
    st->pending = 0;
    if (!step.done && !step.overflow && len + I::kTreeNodes > I::kSearchCache) step.overflow = 1;
    st->step = step;
    if (!step.done && !step.overflow) {
      I::search_tree(step.t, step.low, step.high, st->node, st->S, st->perm);
      st->pending = 1;
      st->passes += 1;
    }
  }
}

// Stage A of finalize: one workgroup folds kFold chunk records (fixed
// tree), writes the chunks' offsets relative to the fold and one fold
// record.
__device__ __forceinline__ void fold_block(std::size_t fold, std::size_t nchunks,
                                           const double* __restrict__ part,
                                           const std::uint32_t* __restrict__ cnt,
                                           std::uint64_t* __restrict__ off,
                                           double* __restrict__ fpart,
                                           std::uint32_t* __restrict__ fcnt, int with_minmax,
                                           DD* sums_out = nullptr,
                                           std::uint32_t* total_out = nullptr)
{
  __shared__ DD sh[kThreads / 64];
  __shared__ std::uint32_t wtot[kThreads / 64];
  __shared__ double smn[kThreads / 64], smx[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const std::size_t c = fold * kFold + threadIdx.x;
  const bool live = c < nchunks;
  DD s[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  double mn = HUGE_VAL, mx = -HUGE_VAL;
  std::uint32_t v = 0;
  if (live) {
    const double* p = part + c * kPartDoubles;
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = DD{p[2 * i], p[2 * i + 1]};
    if (with_minmax) {
      mn = p[8];
      mx = p[9];
    }
    if (cnt) v = cnt[c];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) s[i] = block_reduce(s[i], sh);
  if (cnt) {
    std::uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const std::uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    std::uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      if (w < wave) wbase += wtot[w];
      total += wtot[w];
    }
    if (live && off) off[c] = wbase + inc - v;
    if (threadIdx.x == 0) fcnt[fold] = total;
    if (total_out) *total_out = total;
  }
  if (with_minmax) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = fmin(mn, __shfl_xor(mn, o, 64));
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) {
      smn[wave] = mn;
      smx[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 1; w < kThreads / 64; ++w) {
        mn = fmin(mn, smn[w]);
        mx = fmax(mx, smx[w]);
      }
  }
  if (threadIdx.x == 0) {
    double* q = fpart + fold * kPartDoubles;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      q[2 * i] = s[i].hi;
      q[2 * i + 1] = s[i].lo;
    }
    q[8] = mn;
    q[9] = mx;
  }
  if (sums_out) {
#pragma unroll
    for (int i = 0; i < 4; ++i) sums_out[i] = s[i];
  }
}

__global__ void __launch_bounds__(kThreads)
    k_fold(std::size_t nchunks, const double* __restrict__ part,
           const std::uint32_t* __restrict__ cnt, std::uint64_t* __restrict__ off,
           double* __restrict__ fpart, std::uint32_t* __restrict__ fcnt, int with_minmax)
{
  fold_block(blockIdx.x, nchunks, part, cnt, off, fpart, fcnt, with_minmax);
}

// device_set_true_for_larger_than + the reductions of apply_threshold and
// evaluate_error_budget (heuristic_classifier.cuh:60-87, 282-325).
__global__ void __launch_bounds__(kThreads)
    k_threshold(std::size_t n, const double* __restrict__ err, double thr,
                std::uint8_t* __restrict__ flags, double* __restrict__ part,
                std::uint32_t* __restrict__ cnt)
{
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kChunk;
  DD z{0, 0}, sr{0, 0}, sar{0, 0};
  std::uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    if (t < n) {
      const double r = err[t];
      const bool a = r > thr;
      flags[t] = a ? 1 : 0;
      sr = dd_add(sr, r);
      if (a) {
        sar = dd_add(sar, r);
        ++c;
      }
    }
  }
  write_record(z, sr, z, sar, c, part, cnt);
}

__global__ void __launch_bounds__(kThreads)
    k_sum_dot(std::size_t n, const double* __restrict__ x,
              const std::uint8_t* __restrict__ flags, double* __restrict__ part,
              std::uint32_t* __restrict__ cnt)
{
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kChunk;
  DD z{0, 0}, sx{0, 0}, sfx{0, 0};
  std::uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    if (t < n) {
      const double v = x ? x[t] : 0.0;
      sx = dd_add(sx, v);
      if (flags) {
        sfx = dd_add(sfx, static_cast<double>(flags[t]) * v);
        c += flags[t] != 0;
      }
    }
  }
  write_record(sx, z, sfx, z, c, part, cnt);
}

__global__ void __launch_bounds__(kThreads)
    k_minmax(std::size_t n, const double* __restrict__ x, double* __restrict__ part)
{
  __shared__ double smn[kThreads / 64], smx[kThreads / 64];
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kChunk;
  double mn = HUGE_VAL, mx = -HUGE_VAL;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    if (t < n) {
      mn = fmin(mn, x[t]);
      mx = fmax(mx, x[t]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, off, 64));
    mx = fmax(mx, __shfl_xor(mx, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smn[wave] = mn;
    smx[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      mn = fmin(mn, smn[w]);
      mx = fmax(mx, smx[w]);
    }
    double* p = part + static_cast<std::size_t>(blockIdx.x) * kPartDoubles;
    p[8] = mn;
    p[9] = mx;
  }
}

// One workgroup reduces all chunk records in rounds of 1024 chunks: the
// counts are scanned round by round (chunk order) into exclusive offsets;
// each thread folds its chunks' sums (fixed strided order), then a fixed
// tree combines the threads.
constexpr int kFinThreads = 1024;

__device__ __forceinline__ void finalize_block(std::size_t nchunks, const double* __restrict__ part,
                                               const std::uint32_t* __restrict__ cnt,
                                               std::uint64_t* __restrict__ off,
                                               Scalars* __restrict__ out, int with_minmax)
{
  __shared__ std::uint64_t wtot[kFinThreads / 64];
  __shared__ DD tree[kFinThreads];
  __shared__ double tmn[kFinThreads], tmx[kFinThreads];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  DD s[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  double mn = HUGE_VAL, mx = -HUGE_VAL;
  std::uint64_t running = 0;
  for (std::size_t r0 = 0; r0 < nchunks; r0 += kFinThreads) {
    const std::size_t c = r0 + tid;
    std::uint64_t v = 0;
    if (c < nchunks) {
      const double* p = part + c * kPartDoubles;
#pragma unroll
      for (int i = 0; i < 4; ++i) s[i] = dd_add(s[i], DD{p[2 * i], p[2 * i + 1]});
      if (with_minmax) {
        mn = fmin(mn, p[8]);
        mx = fmax(mx, p[9]);
      }
      if (cnt) v = cnt[c];
    }
    if (cnt) {
      // inclusive scan within the wave
      std::uint64_t inc = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const std::uint64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
      }
      if (lane == 63) wtot[wave] = inc;
      __syncthreads();
      std::uint64_t wbase = 0, total = 0;
      for (int w = 0; w < kFinThreads / 64; ++w) {
        if (w < wave) wbase += wtot[w];
        total += wtot[w];
      }
      if (c < nchunks && off) off[c] = running + wbase + inc - v;
      running += total;
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    tree[tid] = s[i];
    __syncthreads();
    for (int stride = kFinThreads / 2; stride > 0; stride >>= 1) {
      if (tid < stride) tree[tid] = dd_add(tree[tid], tree[tid + stride]);
      __syncthreads();
    }
    if (tid == 0) out->sum[i] = tree[0].hi + tree[0].lo;
    __syncthreads();
  }
  if (with_minmax) {
    tmn[tid] = mn;
    tmx[tid] = mx;
    __syncthreads();
    for (int stride = kFinThreads / 2; stride > 0; stride >>= 1) {
      if (tid < stride) {
        tmn[tid] = fmin(tmn[tid], tmn[tid + stride]);
        tmx[tid] = fmax(tmx[tid], tmx[tid + stride]);
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    out->count = running;
    out->mn = with_minmax ? tmn[0] : 0.0;
    out->mx = with_minmax ? tmx[0] : 0.0;
    out->pad = 0;
  }
}

__global__ void __launch_bounds__(kFinThreads)
    k_finalize(std::size_t nchunks, const double* __restrict__ part,
               const std::uint32_t* __restrict__ cnt, std::uint64_t* __restrict__ off,
               Scalars* __restrict__ out, int with_minmax)
{
  finalize_block(nchunks, part, cnt, off, out, with_minmax);
}

// One step of the device-resident loop after the evaluation: the
// refinement of every chunk (k_refine's work), then -- in whichever
// workgroup finishes last -- the fold of the chunk records (k_fold's work
// for the single fold a list of at most kDevLoopMax regions has), the final
// sums exactly as k_finalize forms them from one fold record, and the
// iteration's decisions (decide_step).  All workgroups of the grid take a
// ticket, those past the list too, so the descriptor the decision rewrites
// has been read by every workgroup before it changes.
__global__ void __launch_bounds__(kThreads)
    k_step_dev(internal::DevLoop* __restrict__ loop, const double* __restrict__ est,
               double* err0, double* err1, const double* __restrict__ pest, double epsrel,
               int relerr, std::uint8_t* __restrict__ active, double* __restrict__ part,
               std::uint32_t* __restrict__ cnt, std::uint64_t* __restrict__ off,
               double* __restrict__ fpart, std::uint32_t* __restrict__ fcnt,
               std::uint64_t* __restrict__ foff, Scalars* __restrict__ sc,
               std::uint32_t* __restrict__ ticket)
{
  __shared__ int last;
  const pagani_dev_iter it = loop->it;
  if (!it.run) {
    if (blockIdx.x == 0 && threadIdx.x == 0) loop->cp.run = 0;
    return;
  }
  const std::size_t n = it.n;
  const bool parents = it.has_parents != 0;
  const std::size_t nch = num_chunks(n, parents);
  if (blockIdx.x < nch)
    refine_chunk(n, est, err0, parents ? err1 : err0, parents ? pest : nullptr, epsrel, relerr,
                 active, part, cnt, paired_chunk(blockIdx.x, nch));
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1 ? 1 : 0;
  __syncthreads();
  if (!last) return;
  __threadfence();
  DD fs[4];
  std::uint32_t total = 0;
  fold_block(0, nch, part, cnt, off, fpart, fcnt, 0, fs, &total);
  if (threadIdx.x == 0) {
    double sums[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // k_finalize over one fold record: its thread 0 adds the record to
      // zero, then meets only zeros on the way up its tree
      DD x = dd_add(DD{0.0, 0.0}, DD{fs[i].hi, fs[i].lo});
      for (int stride = kFinThreads / 2; stride > 0; stride >>= 1) x = dd_add(x, DD{0.0, 0.0});
      sums[i] = x.hi + x.lo;
      sc->sum[i] = sums[i];
    }
    foff[0] = 0;
    sc->count = total;
    sc->mn = 0.0;
    sc->mx = 0.0;
    sc->pad = 0;
    internal::decide_step(loop, sums[0], sums[1], sums[2], sums[3], total);
    *ticket = 0;
  }
}

// Chunk-local exclusive position of the active element t (ballot within the
// wave, wave totals through LDS) for a workgroup of T threads.  Returns the
// workgroup total.
template <int T = kThreads>
__device__ __forceinline__ std::uint32_t chunk_rank(bool a, std::uint32_t& rank,
                                                    std::uint32_t* wsum)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long ballot = __ballot(a);
  const unsigned long long below = lane ? (ballot & ((1ull << lane) - 1ull)) : 0ull;
  const std::uint32_t pre = static_cast<std::uint32_t>(__popcll(below));
  if (lane == 0) wsum[wave] = static_cast<std::uint32_t>(__popcll(ballot));
  __syncthreads();
  std::uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < T / 64; ++w) {
    if (w < wave) woff += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  rank = woff + pre;
  return tot;
}

// The compaction moves data only, so it can use wider workgroups than the
// reductions (whose per-thread order fixes their rounding): a chunk of
// kChunk regions is covered in kChunk / kCompactThreads steps.
constexpr int kCompactThreads = 1024;

// alignRegions (Sub_region_filter.cuh:8-47) and, with split, the bisection
// of divideIntervalsGPU (Sub_region_splitter.cuh:8-46) fused: the active
// region t at compacted index idx produces children idx and m+idx.
template <int D>
__device__ __forceinline__ void compact_chunk(
    int split, std::size_t n, std::size_t m, const double* __restrict__ L,
    const double* __restrict__ H, const std::uint8_t* __restrict__ active,
    const double* __restrict__ est, const double* __restrict__ err,
    const std::uint8_t* __restrict__ dim, const std::uint64_t* __restrict__ off,
    const std::uint64_t* __restrict__ foff, double* __restrict__ L2, double* __restrict__ H2,
    double* __restrict__ pest, double* __restrict__ perr, std::uint8_t* __restrict__ dim2)
{
  __shared__ std::uint32_t wsum[kCompactThreads / 64];
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kChunk;
  std::uint64_t run = off[blockIdx.x] + foff[blockIdx.x / kFold];
  for (int k = 0; k < static_cast<int>(kChunk / kCompactThreads); ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kCompactThreads + threadIdx.x;
    const bool a = t < n && active[t] == 1;
    // the split's reads do not wait for the rank: a region's bounds are
    // requested before the workgroup's scan (whole cache lines are read for
    // a partly active wave in any case)
    double lv[D], hv[D];
    double ev = 0.0, rv = 0.0;
    int b = 0;
    if (a) {
      if (pest) ev = est[t];
      if (perr) rv = err[t];
    }
    if (split && a) {
      b = dim[t];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        lv[j] = L[static_cast<std::size_t>(j) * n + t];
        hv[j] = H[static_cast<std::size_t>(j) * n + t];
      }
    }
    std::uint32_t rank;
    const std::uint32_t tot = chunk_rank<kCompactThreads>(a, rank, wsum);
    if (a) {
      const std::size_t idx = run + rank;
      if (idx < m) {
        if (split) {
          const std::size_t n2 = 2 * m;
#pragma unroll
          for (int j = 0; j < D; ++j) {
            const double l = lv[j];
            const double h = hv[j];
            const std::size_t o = static_cast<std::size_t>(j) * n2 + idx;
            if (j == b) {
              const double half = h / 2;
              L2[o] = l;
              L2[o + m] = l + half;
              H2[o] = half;
              H2[o + m] = half;
            } else {
              L2[o] = l;
              L2[o + m] = l;
              H2[o] = h;
              H2[o + m] = h;
            }
          }
        } else {
          for (int j = 0; j < D; ++j) {
            L2[static_cast<std::size_t>(j) * m + idx] = L[static_cast<std::size_t>(j) * n + t];
            H2[static_cast<std::size_t>(j) * m + idx] = H[static_cast<std::size_t>(j) * n + t];
          }
          if (dim2) dim2[idx] = dim[t];
        }
        if (pest) pest[idx] = ev;
        if (perr) perr[idx] = rv;
      }
    }
    run += tot;
  }
}

template <int D>
__global__ void __launch_bounds__(kCompactThreads)
    k_compact(int split, std::size_t n, std::size_t m, const double* __restrict__ L,
              const double* __restrict__ H, const std::uint8_t* __restrict__ active,
              const double* __restrict__ est, const double* __restrict__ err,
              const std::uint8_t* __restrict__ dim, const std::uint64_t* __restrict__ off,
              const std::uint64_t* __restrict__ foff, double* __restrict__ L2,
              double* __restrict__ H2, double* __restrict__ pest, double* __restrict__ perr,
              std::uint8_t* __restrict__ dim2)
{
  compact_chunk<D>(split, n, m, L, H, active, est, err, dim, off, foff, L2, H2, pest, perr, dim2);
}

// The active regions of an n-region list (flat, or in split form when pdim
// is given) become the parents of the next list: region t at compacted
// index idx writes its own bounds (derived from its parent's in split form,
// as the evaluation derives them) at [j*m + idx], the axis it is to be
// halved along (dim[t]) at pdim2[idx], and its estimate (and, when perr is
// given, its error) at pest/perr[idx].  The chunk's offsets come from finalize, so chunks may
// run in any order: sibling chunks (n/2 apart, one parent's two children)
// are processed together, and the second read of a parent hits the L2.
template <int D>
__device__ __forceinline__ void compact_parents_chunk(
    std::size_t n, std::size_t m, const double* __restrict__ L,
    const double* __restrict__ H, const std::uint8_t* __restrict__ pdim,
    const std::uint8_t* __restrict__ active, const double* __restrict__ est,
    const double* __restrict__ err, const std::uint8_t* __restrict__ dim,
    const std::uint64_t* __restrict__ off, const std::uint64_t* __restrict__ foff,
    double* __restrict__ L2, double* __restrict__ H2, std::uint8_t* __restrict__ pdim2,
    double* __restrict__ pest, double* __restrict__ perr, std::size_t chunk, bool paired,
    const Levels& lv)
{
  // The positions first (ballots per step and wave, one barrier): each kept
  // region's offset in the chunk, in step order, so the list is the one a
  // step-by-step scan gives.  Each kept region leaves its index in LDS at its
  // position, and after a second barrier every thread moves the regions at
  // positions tid, tid + 1024, ...: all lanes busy on kept regions and the
  // writes dense, whatever share of the chunk is kept.
  constexpr int kSteps = static_cast<int>(kChunk / kCompactThreads);
  constexpr int kWaves = kCompactThreads / 64;
  __shared__ std::uint32_t wsum[kSteps][kWaves];
  static_assert(kChunk <= 65536, "chunk offsets are 16-bit");
  __shared__ std::uint16_t src[kChunk];
  std::size_t base, end;
  chunk_range(chunk, n, paired, base, end);
  const std::size_t h = n / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bool a[kSteps];
  std::uint32_t pre[kSteps];
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kCompactThreads + threadIdx.x;
    a[k] = t < end && active[t] == 1;
    const unsigned long long ballot = __ballot(a[k]);
    const unsigned long long below = lane ? (ballot & ((1ull << lane) - 1ull)) : 0ull;
    pre[k] = static_cast<std::uint32_t>(__popcll(below));
    if (lane == 0) wsum[k][wave] = static_cast<std::uint32_t>(__popcll(ballot));
  }
  __syncthreads();
  std::uint32_t kept = 0;  // the chunk's kept regions
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    std::uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const std::uint32_t c = wsum[k][w];
      if (w < wave) woff += c;
      tot += c;
    }
    if (a[k])
      src[kept + woff + pre[k]] =
          static_cast<std::uint16_t>(k * kCompactThreads + static_cast<int>(threadIdx.x));
    kept += tot;
  }
  __syncthreads();
  const std::uint64_t first = off[chunk] + foff[chunk / kFold];
  // one region's move at a time: few VGPRs, two workgroups per CU
#pragma unroll 1
  for (std::uint32_t q = threadIdx.x; q < kept; q += kCompactThreads) {
    const std::size_t i = first + q;
    if (i >= m) break;
    const std::size_t t = base + src[q];
    const double ev = est[t];
    const double rv = perr ? err[t] : 0.0;
    const int b = dim[t];
    // the row this region's bounds come from, and the axis its parent
    // halves (none for a flat list)
    std::size_t row = t, stride = n;
    int axis = -1;
    bool right = false;
    if (pdim) {
      right = t >= h;
      row = right ? t - h : t;
      stride = h;
      axis = pdim[row];
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double l, w;
      int g;
      read_row(L, H, lv, static_cast<std::size_t>(j) * stride + row, j, l, w, g);
      if (j == axis) {
        const double half = w / 2;
        if (right) l = l + half;
        w = half;
        ++g;
      }
      const std::size_t o = static_cast<std::size_t>(j) * m + i;
      L2[o] = l;
      if (lv.on)
        reinterpret_cast<std::uint16_t*>(H2)[o] = static_cast<std::uint16_t>(g);
      else
        H2[o] = w;
    }
    pdim2[i] = static_cast<std::uint8_t>(b);
    pest[i] = ev;
    if (perr) perr[i] = rv;
  }
}

template <int D>
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_parents(std::size_t n, std::size_t m, const double* __restrict__ L,
                      const double* __restrict__ H, const std::uint8_t* __restrict__ pdim,
                      const std::uint8_t* __restrict__ active, const double* __restrict__ est,
                      const double* __restrict__ err, const std::uint8_t* __restrict__ dim,
                      const std::uint64_t* __restrict__ off,
                      const std::uint64_t* __restrict__ foff, double* __restrict__ L2,
                      double* __restrict__ H2, std::uint8_t* __restrict__ pdim2,
                      double* __restrict__ pest, double* __restrict__ perr, int paired,
                      const Levels lv)
{
  compact_parents_chunk<D>(n, m, L, H, pdim, active, est, err, dim, off, foff, L2, H2, pdim2,
                           pest, perr, paired_chunk(blockIdx.x, gridDim.x), paired != 0, lv);
}

// The device-resident loop's filter + split: the descriptor says whether it
// runs, the counts, which pair holds the list and in which form (the next
// parents go to the other pair), and which error array the refinement
// wrote.
template <int D>
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_dev(const DevCompact* __restrict__ cp, double* L0, double* H0, double* L1,
                  double* H1, std::uint8_t* P0, std::uint8_t* P1,
                  const std::uint8_t* __restrict__ active, const double* __restrict__ est,
                  const double* err0, const double* err1, const std::uint8_t* __restrict__ dim,
                  const std::uint64_t* __restrict__ off, const std::uint64_t* __restrict__ foff,
                  double* __restrict__ pest, double* __restrict__ perr, const Levels lv)
{
  if (!cp->run) return;
  const std::size_t n = cp->n, m = cp->m;
  const bool paired = cp->has_parents != 0;
  const std::size_t nch = num_chunks(n, paired);
  if (blockIdx.x >= nch) return;
  const bool c1 = cp->cur != 0;
  compact_parents_chunk<D>(n, m, c1 ? L1 : L0, c1 ? H1 : H0,
                           cp->split ? (c1 ? P1 : P0) : nullptr, active, est,
                           paired ? err1 : err0, dim, off, foff, c1 ? L0 : L1, c1 ? H0 : H1,
                           c1 ? P0 : P1, pest, perr, paired_chunk(blockIdx.x, nch), paired, lv);
}

__global__ void __launch_bounds__(256)
    k_split_all(int d, std::size_t m, const double* __restrict__ L, const double* __restrict__ H,
                const std::uint8_t* __restrict__ dim, double* __restrict__ L2,
                double* __restrict__ H2)
{
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const std::size_t n2 = 2 * m;
  const int b = dim[i];
  for (int j = 0; j < d; ++j) {
    const double l = L[static_cast<std::size_t>(j) * m + i];
    const double h = H[static_cast<std::size_t>(j) * m + i];
    const std::size_t o = static_cast<std::size_t>(j) * n2 + i;
    if (j == b) {
      const double half = h / 2;
      L2[o] = l;
      L2[o + m] = l + half;
      H2[o] = half;
      H2[o + m] = half;
    } else {
      L2[o] = l;
      L2[o + m] = l;
      H2[o] = h;
      H2[o + m] = h;
    }
  }
}

__global__ void __launch_bounds__(256)
    k_expand(int d, std::size_t n, const double* __restrict__ L, const double* __restrict__ H,
             const std::uint8_t* __restrict__ pdim, const Levels lv, double* __restrict__ L2,
             double* __restrict__ H2)
{
  const std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const std::size_t h = n / 2;
  const bool right = pdim && t >= h;
  const std::size_t row = pdim ? (right ? t - h : t) : t;
  const std::size_t stride = pdim ? h : n;
  const int axis = pdim ? pdim[row] : -1;
  for (int j = 0; j < d; ++j) {
    double l, w;
    int g;
    read_row(L, H, lv, static_cast<std::size_t>(j) * stride + row, j, l, w, g);
    if (j == axis) {
      const double half = w / 2;
      if (right) l = l + half;
      w = half;
    }
    L2[static_cast<std::size_t>(j) * n + t] = l;
    H2[static_cast<std::size_t>(j) * n + t] = w;
  }
}

__global__ void __launch_bounds__(kThreads)
    k_scan_flags(std::size_t n, const std::uint8_t* __restrict__ flags,
                 const std::uint64_t* __restrict__ off, const std::uint64_t* __restrict__ foff,
                 std::uint64_t* __restrict__ out)
{
  __shared__ std::uint32_t wsum[kThreads / 64];
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kChunk;
  std::uint64_t run = off[blockIdx.x] + foff[blockIdx.x / kFold];
  for (int k = 0; k < kPerThread; ++k) {
    const std::size_t t = base + static_cast<std::size_t>(k) * kThreads + threadIdx.x;
    const bool a = t < n && flags[t] != 0;
    std::uint32_t rank;
    const std::uint32_t tot = chunk_rank(a, rank, wsum);
    if (t < n) out[t] = run + rank;
    run += tot;
  }
}

__global__ void __launch_bounds__(256) k_fill_u8(std::size_t n, std::uint8_t* p, std::uint8_t v)
{
  const std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

inline dim3 grid_for(std::size_t n, unsigned threads)
{
  return dim3(static_cast<unsigned>((n + threads - 1) / threads));
}

// fn(std::integral_constant<int, d>) for d in [PAGANI_MIN_DIM, PAGANI_MAX_DIM]
template <int D = PAGANI_MIN_DIM, class Fn>
hipError_t with_dim(int d, Fn&& fn)
{
  if constexpr (D > PAGANI_MAX_DIM) {
    return hipErrorInvalidValue;
  } else {
    if (d == D) return fn(std::integral_constant<int, D>{});
    return with_dim<D + 1>(d, fn);
  }
}

}  // namespace

hipError_t uniform_split(hipStream_t s, int d, std::size_t k, std::size_t n, double* L,
                         double* H, bool levels)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_uniform_split, grid_for(n, 256), dim3(256), 0, s, d, k, n,
                     1. / static_cast<double>(k), L, H, levels ? 1 : 0);
  return hipGetLastError();
}

hipError_t refine(hipStream_t s, std::size_t n, const double* est, const double* err_in,
                  double* err_out, const double* parent_est, double epsrel, int relerr,
                  std::uint8_t* active, double* part, std::uint32_t* cnt)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_refine, dim3(static_cast<unsigned>(num_chunks(n, parent_est != nullptr))),
                     dim3(kThreads), 0, s, n, est, err_in, err_out, parent_est, epsrel, relerr,
                     active, part, cnt);
  return hipGetLastError();
}

hipError_t count_flags(hipStream_t s, std::size_t n, const std::uint8_t* flags,
                       const double* est, const double* err, double* part, std::uint32_t* cnt,
                       bool paired)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_flags, dim3(static_cast<unsigned>(num_chunks(n, paired))),
                     dim3(kThreads), 0, s, n, flags, est, err, part, cnt, paired ? 1 : 0);
  return hipGetLastError();
}

hipError_t threshold(hipStream_t s, std::size_t n, const double* err, double thr,
                     std::uint8_t* flags, double* part, std::uint32_t* cnt)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_threshold, dim3(static_cast<unsigned>(num_chunks(n))), dim3(kThreads), 0,
                     s, n, err, thr, flags, part, cnt);
  return hipGetLastError();
}

hipError_t minmax(hipStream_t s, std::size_t n, const double* x, double* part)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_minmax, dim3(static_cast<unsigned>(num_chunks(n))), dim3(kThreads), 0, s,
                     n, x, part);
  return hipGetLastError();
}

hipError_t sum_dot(hipStream_t s, std::size_t n, const double* x, const std::uint8_t* flags,
                   double* part, std::uint32_t* cnt)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sum_dot, dim3(static_cast<unsigned>(num_chunks(n))), dim3(kThreads), 0, s,
                     n, x, flags, part, cnt);
  return hipGetLastError();
}

hipError_t threshold_flags(hipStream_t s, std::size_t n, const double* est, const double* err,
                           double thr, std::uint8_t* flags, double* part, std::uint32_t* cnt,
                           bool paired)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_threshold_flags, dim3(static_cast<unsigned>(num_chunks(n, paired))),
                     dim3(kThreads), 0, s, n, est, err, thr, flags, part, cnt, paired ? 1 : 0);
  return hipGetLastError();
}

hipError_t search_start(hipStream_t s, internal::SearchState* st, const Scalars* sc,
                        const internal::SearchInputs& in)
{
  hipLaunchKernelGGL(k_search, dim3(1), dim3(64), 0, s, st, sc, nullptr, 1, in);
  return hipGetLastError();
}

hipError_t search_round(hipStream_t s, std::size_t n, const double* err,
                        internal::SearchState* st, double* bpart, double* mout)
{
  hipLaunchKernelGGL(k_multi_threshold_dev<kMaxThresholds>, dim3(kMultiBlocks), dim3(kThreads), 0,
                     s, n, err, st, bpart);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_multi_reduce_dev, dim3(kMaxThresholds - 1), dim3(64), 0, s, kMultiBlocks,
                     st, bpart, mout);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_search, dim3(1), dim3(64), 0, s, st, nullptr, mout, 0,
                     internal::SearchInputs{});
  return hipGetLastError();
}

hipError_t multi_threshold(hipStream_t s, std::size_t n, const double* err, const ThresholdSet& T,
                           double* bpart, double* out)
{
  if (T.k < 1 || T.k >= kMaxThresholds) return hipErrorInvalidValue;
  for (int k = 1; k < T.k; ++k)
    if (!(T.t[k - 1] <= T.t[k])) return hipErrorInvalidValue;  // sorted, no NaN
  ThresholdSet U = T;
  for (int k = U.k; k < kMaxThresholds; ++k) U.t[k] = __builtin_inf();
  hipLaunchKernelGGL(k_multi_threshold<kMaxThresholds>, dim3(kMultiBlocks), dim3(kThreads), 0, s,
                     n, err, U, bpart);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_multi_reduce, dim3(T.k), dim3(64), 0, s, kMultiBlocks, T.k, bpart, out);
  return hipGetLastError();
}

hipError_t finalize(hipStream_t s, std::size_t nchunks, const double* part,
                    const std::uint32_t* cnt, std::uint64_t* off, double* fpart,
                    std::uint32_t* fcnt, std::uint64_t* foff, Scalars* out, int with_minmax)
{
  const std::size_t nfold = num_folds(nchunks);
  if (nfold > 0) {
    hipLaunchKernelGGL(k_fold, dim3(static_cast<unsigned>(nfold)), dim3(kThreads), 0, s, nchunks,
                       part, cnt, off, fpart, cnt ? fcnt : nullptr, with_minmax);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, s, nfold, fpart,
                     cnt ? fcnt : nullptr, cnt ? foff : nullptr, out, with_minmax);
  return hipGetLastError();
}

hipError_t compact(hipStream_t s, int split, int d, std::size_t n, std::size_t m,
                   const double* L, const double* H, const std::uint8_t* active,
                   const double* est, const double* err, const std::uint8_t* dim,
                   const std::uint64_t* off, const std::uint64_t* foff, double* L2, double* H2,
                   double* pest, double* perr, std::uint8_t* dim2)
{
  if (n == 0 || m == 0) return hipSuccess;
  return with_dim(d, [&](auto D) {
    hipLaunchKernelGGL((k_compact<decltype(D)::value>), dim3(static_cast<unsigned>(num_chunks(n))),
                       dim3(kCompactThreads), 0, s, split, n, m, L, H, active, est, err, dim, off,
                       foff, L2, H2, pest, perr, dim2);
    return hipGetLastError();
  });
}

hipError_t step_dev(hipStream_t s, std::size_t n_cap, internal::DevLoop* loop, const double* est,
                    double* err0, double* err1, const double* pest, double epsrel, int relerr,
                    std::uint8_t* active, double* part, std::uint32_t* cnt, std::uint64_t* off,
                    double* fpart, std::uint32_t* fcnt, std::uint64_t* foff, Scalars* sc,
                    std::uint32_t* ticket)
{
  // one workgroup more than num_chunks(n_cap): a paired list may have one
  // chunk more than a flat one
  hipLaunchKernelGGL(k_step_dev, dim3(static_cast<unsigned>(num_chunks(n_cap) + 1)), dim3(kThreads),
                     0, s, loop, est, err0, err1, pest, epsrel, relerr, active, part, cnt, off,
                     fpart, fcnt, foff, sc, ticket);
  return hipGetLastError();
}

hipError_t compact_dev(hipStream_t s, std::size_t n_cap, const DevCompact* cp, int d,
                       double* L0, double* H0, double* L1, double* H1, std::uint8_t* P0,
                       std::uint8_t* P1, const std::uint8_t* active, const double* est,
                       const double* err0, const double* err1, const std::uint8_t* dim,
                       const std::uint64_t* off, const std::uint64_t* foff, double* pest,
                       double* perr, const Levels& lv)
{
  return with_dim(d, [&](auto D) {
    hipLaunchKernelGGL((k_compact_dev<decltype(D)::value>),
                       dim3(static_cast<unsigned>(num_chunks(n_cap) + 1)), dim3(kCompactThreads), 0, s,
                       cp, L0, H0, L1, H1, P0, P1, active, est, err0, err1, dim, off, foff, pest,
                       perr, lv);
    return hipGetLastError();
  });
}

hipError_t compact_parents(hipStream_t s, int d, std::size_t n, std::size_t m, const double* L,
                           const double* H, const std::uint8_t* pdim,
                           const std::uint8_t* active, const double* est, const double* err,
                           const std::uint8_t* dim, const std::uint64_t* off,
                           const std::uint64_t* foff, double* L2, double* H2,
                           std::uint8_t* pdim2, double* pest, double* perr, bool paired,
                           const Levels& lv)
{
  if (n == 0 || m == 0) return hipSuccess;
  return with_dim(d, [&](auto D) {
    hipLaunchKernelGGL((k_compact_parents<decltype(D)::value>),
                       dim3(static_cast<unsigned>(num_chunks(n, paired))), dim3(kCompactThreads),
                       0, s, n, m, L, H, pdim, active, est, err, dim, off, foff, L2, H2, pdim2,
                       pest, perr, paired ? 1 : 0, lv);
    return hipGetLastError();
  });
}

hipError_t expand(hipStream_t s, int d, std::size_t n, const double* L, const double* H,
                  const std::uint8_t* pdim, const Levels& lv, double* L2, double* H2)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_expand, grid_for(n, 256), dim3(256), 0, s, d, n, L, H, pdim, lv, L2, H2);
  return hipGetLastError();
}

hipError_t split_all(hipStream_t s, int d, std::size_t m, const double* L, const double* H,
                     const std::uint8_t* dim, double* L2, double* H2)
{
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_all, grid_for(m, 256), dim3(256), 0, s, d, m, L, H, dim, L2, H2);
  return hipGetLastError();
}

hipError_t scan_flags(hipStream_t s, std::size_t n, const std::uint8_t* flags,
                      const std::uint64_t* off, const std::uint64_t* foff, std::uint64_t* out)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_flags, dim3(static_cast<unsigned>(num_chunks(n))), dim3(kThreads), 0,
                     s, n, flags, off, foff, out);
  return hipGetLastError();
}

hipError_t fill_u8(hipStream_t s, std::size_t n, std::uint8_t* p, std::uint8_t v)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u8, grid_for(n, 256), dim3(256), 0, s, n, p, v);
  return hipGetLastError();
}

}  // namespace phases
}  // namespace pagani
