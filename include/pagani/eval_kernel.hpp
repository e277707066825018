// eval_kernel.hpp -- region evaluation with the degree-9 rule on gfx950.
//
// Replaces INTEGRATE_GPU_PHASE1 / INIT_REGION_POOL / SampleRegionBlock /
// computePermutation (cuda/pagani/quad/GPUquad/Phases.cuh:301-404,
// Sample.cuh:103-290).  Included by the library (precompiled Genz families)
// and by user translation units compiled with hipcc (any functor with
// `double operator()(double, ..., double)`), so one kernel template serves
// both.
//
// Mapping: one LANE per region.  A wave owns 64 consecutive regions; every
// lane walks the same sequence of rule points, so the point's generator
// coordinates, the weights and the null-rule constants are wave-uniform and
// come from scalar loads; the five rule sums, the fourth differences and the
// error estimate are per-lane registers and need no cross-lane reduction.
// For small region lists a workgroup holds `slices` waves that share one
// batch of 64 regions and split its points; their partial sums are combined
// through LDS in a fixed order.
//
// Per point the lane does d FMAs for the point map x = c - g*h (c, h: region
// centre and width in user coordinates), one integrand call and one add into
// the running sum of the point's symmetric set; the five rule sums are
// formed per set as sum_s W[s][r] * S_s (all points of a set share their
// weights, Rule.cuh:36-147).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "pagani.h"
#include "pagani/trig.hpp"

namespace pagani {

// The generic points [set_begin[3], feval) split over `slices` waves of a
// batch: wave 0 also evaluates the centre and the axis points of sets 0-2,
// so it takes a smaller share; the rest is spread evenly.  Host and device
// use this one definition, so a slice count always means the same split.
__host__ __device__ inline void slice_range(int slices, int wave, int p0, int feval, int& pa,
                                            int& pb)
{
  if (slices <= 1) {
    pa = p0;
    pb = feval;
    return;
  }
  const int generic = feval - p0;
  const int share0 = feval / slices - p0 > 0 ? feval / slices - p0 : 0;
  const int rest = generic - share0;
  const int q = rest / (slices - 1), r = rest % (slices - 1);
  if (wave == 0) {
    pa = p0;
    pb = p0 + share0;
    return;
  }
  const int w = wave - 1;
  pa = p0 + share0 + w * q + (w < r ? w : r);
  pb = p0 + share0 + (w + 1) * q + (w + 1 < r ? w + 1 : r);
}

namespace detail {

// Integrands that depend on x only through x - u may say so with
// `double origin(int i) const` (= u_i) and `centered(t...)` (= f(u + t)).
// The kernel then subtracts u from each region's centre once and maps
// points straight to t = x - u.
template <class F, class = void>
struct has_origin : std::false_type {};
template <class F>
struct has_origin<F, std::void_t<decltype(std::declval<const F&>().origin(0))>>
    : std::true_type {};

// A centred integrand may also give `double axis_scale(int i) const` (a_i):
// its `centered` then takes v = a_i (x_i - u_i), and the kernel scales the
// region's centre offset and half-widths by a_i once per region.
template <class F, class = void>
struct has_axis_scale : std::false_type {};
template <class F>
struct has_axis_scale<F, std::void_t<decltype(std::declval<const F&>().axis_scale(0))>>
    : std::true_type {};

// A centred, scaled integrand that is separable in v -- f = outer(sum_j
// term(j, v_j)), or outer(prod_j term(j, v_j)) (is_multiplicative) -- may
// say so with `term` and `outer`.  The kernel then evaluates the points of
// sets 1-7, which move at most three coordinates off the centre, as outer
// of the centre's combined terms joined with the changed axes' term
// differences (ratios): the centre's terms and the changes at +-lambda1
// and +-lambda2 are computed once per region.
// `static constexpr bool separable = false` turns the path off where the
// generic loop is faster (an instantiation whose separable form would not
// fit the registers).
template <class F, class = void>
struct separable_enabled : std::true_type {};
template <class F>
struct separable_enabled<F, std::void_t<decltype(F::separable)>>
    : std::bool_constant<F::separable> {};

template <class F, class = void>
struct is_separable : std::false_type {};
template <class F>
struct is_separable<F, std::void_t<decltype(std::declval<const F&>().term(0, 0.0)),
                                   decltype(std::declval<const F&>().outer(0.0))>>
    : std::bool_constant<has_origin<F>::value && has_axis_scale<F>::value &&
                         separable_enabled<F>::value> {};

// Multiplicatively separable integrands (f = outer(prod_j term(j, v_j)))
// say so with `static constexpr bool multiplicative = true`; the kernel then
// combines the centre's terms and the changed axes' term ratios by products.
template <class F, class = void>
struct is_multiplicative : std::false_type {};
template <class F>
struct is_multiplicative<F, std::void_t<decltype(F::multiplicative)>>
    : std::bool_constant<F::multiplicative> {};

template <class F, class = void>
struct heavy_outer : std::false_type {};
template <class F>
struct heavy_outer<F, std::void_t<decltype(F::heavy_outer)>> : std::bool_constant<F::heavy_outer> {};

// Masked integrands (`static constexpr bool masked = true`, `bool valid(j,
// v)`): f is zero unless every axis' coordinate is valid; `term` ignores
// validity.  The kernel tracks the axes invalid at the centre and gives a
// changed coordinate that is invalid the term change -infinity.
template <class F, class = void>
struct is_masked : std::false_type {};
template <class F>
struct is_masked<F, std::void_t<decltype(F::masked)>> : std::bool_constant<F::masked> {};

// Exponential separable integrands (`static constexpr bool exponential =
// true`: outer(s + d) = outer(s) * factor(d), with `double factor(double d)`
// and `bool factor_safe(double s0)`, true when every factor and product met
// in a region whose centre sums to s0 stays in the normal range) are
// evaluated as products: the centre's value outer(S0) times one factor per
// changed axis, so a point of sets 5-8 costs a multiply instead of an outer
// evaluation.  Regions that fail factor_safe take the generic loop.
// An exponential integrand may give the centre's value by `double
// centre_value(double s0)` (at a wider range than `outer`); else outer(s0).
template <class F, class = void>
struct has_centre_value : std::false_type {};
template <class F>
struct has_centre_value<F, std::void_t<decltype(std::declval<const F&>().centre_value(0.0))>>
    : std::true_type {};

// Exponential integrands may give the factors of the two points c -+ delta
// (delta = g h >= 0) of one axis together: `void factor_pair(int j, double
// c, double delta, double aux, double& f_minus, double& f_plus)`, with an
// optional per-axis `double axis_aux(int j, double c)` computed once for
// the five pairs of the axis; where the two factors are an exponential and
// its reciprocal, the pair costs one exponential instead of two.
template <class F, class = void>
struct has_factor_pair : std::false_type {};
template <class F>
struct has_factor_pair<
    F, std::void_t<decltype(std::declval<const F&>().factor_pair(
           0, 0.0, 0.0, 0.0, std::declval<double&>(), std::declval<double&>()))>>
    : std::true_type {};

template <class F, class = void>
struct has_axis_aux : std::false_type {};
template <class F>
struct has_axis_aux<F, std::void_t<decltype(std::declval<const F&>().axis_aux(0, 0.0))>>
    : std::true_type {};

template <class F, class = void>
struct is_exponential : std::false_type {};
template <class F>
struct is_exponential<F, std::void_t<decltype(F::exponential)>>
    : std::bool_constant<F::exponential> {};

// Cosine integrands (`static constexpr bool cosine = true`, centred and
// scaled, `double phase_angle()`): f = cos(phase_angle() + sum_j v_j).
template <class F, class = void>
struct is_cosine : std::false_type {};
template <class F>
struct is_cosine<F, std::void_t<decltype(F::cosine)>> : std::bool_constant<F::cosine> {};

// A multiplicative integrand whose outer is linear (outer(p) = k p, up to
// the rounding of its result) says so with `static constexpr bool
// linear_outer = true`.
template <class F, class = void>
struct has_linear_outer : std::false_type {};
template <class F>
struct has_linear_outer<F, std::void_t<decltype(F::linear_outer)>>
    : std::bool_constant<F::linear_outer> {};

template <bool Mul>
struct SepOp {
  // combine the centre's value with a change; the change of term t0 -> t
  static __device__ __forceinline__ double join(double a, double b) { return Mul ? a * b : a + b; }
  static __device__ __forceinline__ double change(double t, double t0)
  {
    return Mul ? t / t0 : t - t0;
  }
};

template <class F, class Real, int D, std::size_t... I>
__device__ __forceinline__ double call_at(F& f, const Real (&x)[D],
                                          std::index_sequence<I...>)
{
  if constexpr (has_origin<F>::value)
    return static_cast<double>(f.centered(x[I]...));
  else
    return static_cast<double>(f(x[I]...));
}

template <class F, class Real, int D>
__device__ __forceinline__ double call(F& f, const Real (&x)[D])
{
  return call_at(f, x, std::make_index_sequence<D>());
}

// One set's contribution from generic (table-driven) points [a, b).
template <int D, class F, class Real>
__device__ __forceinline__ double set_sum(F& f, const double* __restrict__ g,
                                          int a, int b, const double (&cR)[D],
                                          const double (&hR)[D])
{
  double acc = 0.0;
#pragma unroll 1
  for (int p = a; p < b; ++p, g += D) {
    Real x[D];
#pragma unroll
    for (int j = 0; j < D; ++j)
      x[j] = static_cast<Real>(__builtin_fma(-g[j], hR[j], cR[j]));
    acc += call(f, x);
  }
  return acc;
}

// Rule sums and split axis of one region for a factored separable
// integrand (exponential, or multiplicative with a linear outer), slices ==
// 1.  A point's value is the centre's value times the changed axes'
// factors, so every set's sum over its sign combinations factors per axis:
// e.g. the 4 points (+-l1, +-l1) on axes j, k sum to S0 s_j s_k with s_j =
// r(+l1)_j + r(-l1)_j, and the 2^d corners to S0 prod_j s5_j.  With masks, a
// factor is 0 where the moved coordinate is invalid and an unchanged axis
// contributes u_j = [valid at the centre]: the sum over the sets of m
// changed axes is the x^m coefficient of prod_j (u_j + x s_j).
// One pass over the axes after the centre: axis j's ten factors (+-lambda1
// .. +-lambda5) are used as soon as they are formed -- the fourth difference,
// the axis sums of sets 1-4, the running coefficients of sets 5-7 and the
// corners' product -- so no per-axis array outlives its axis and the region
// needs few registers.  Every sum adds its terms in axis order.
template <int D, class F>
__device__ __forceinline__ bool factored_sums(F& f, const double (&cR)[D], const double (&hR)[D],
                                              double jac, const pagani_eval_consts& C,
                                              double (&rs)[PAGANI_NRULES], int& bisect)
{
  constexpr bool kExp = is_exponential<F>::value;
  constexpr bool kMul = is_multiplicative<F>::value || kExp;
  using Op = SepOp<kMul>;
  constexpr bool kMasked = is_masked<F>::value;
  constexpr bool kPair = kExp && has_factor_pair<F>::value;
  const int bisect0 = bisect;
  // the centre's joined terms (exponential: the centre's value) and the
  // axes invalid at the centre (a point is zero unless it changes them all)
  std::uint32_t vmask = 0;
  double S0 = f.term(0, cR[0]);
#pragma unroll
  for (int j = 0; j < D; ++j) {
    if (j > 0) {
      const double t = f.term(j, cR[j]);
      S0 = kExp ? S0 + t : Op::join(S0, t);
    }
    if constexpr (kMasked)
      if (!f.valid(j, cR[j])) vmask |= 1u << j;
  }
  if constexpr (kExp) {
    if (!f.factor_safe(S0)) return false;
    if constexpr (has_centre_value<F>::value)
      S0 = f.centre_value(S0);
    else
      S0 = f.outer(S0);
  }
  auto gated = [&](std::uint32_t changed, double s) {
    if constexpr (kMasked)
      if (vmask & ~changed) return 0.0;
    if constexpr (kExp) return s;
    return f.outer(s);
  };
  auto value = [&](double p) {
    if constexpr (kExp) return p;
    return f.outer(p);
  };
  const double f0 = gated(0u, S0);
  const double base = (f0 * jac) * 2 * (1 - C.ratio);
  double s1 = 0.0, s2 = 0.0, maxdiff = 0.0;
  double acc3 = 0.0, E3 = 1.0, acc4 = 0.0, E4 = 1.0;  // sets 3, 4: x^1, x^0 coefficients
  // sets 5-7: the x^2, xy and x^3 coefficients of prod_j (u_j + x s1_j + y
  // s2_j), with the lower ones they are built from
  double e0 = 1.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, g10 = 0.0, g01 = 0.0, g11 = 0.0;
  double prod = S0;  // set 8
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const double t0 = f.term(j, cR[j]);
    // the factor (exponential) or term ratio of axis j moved to v
    auto change = [&](double v) {
      if constexpr (kMasked)
        if (!f.valid(j, v)) return 0.0;
      if constexpr (kExp) return f.factor(f.term(j, v) - t0);
      return Op::change(f.term(j, v), t0);
    };
    const double uj = kMasked ? (((vmask >> j) & 1u) ? 0.0 : 1.0) : 1.0;
    double aux = 0.0;
    if constexpr (kPair && has_axis_aux<F>::value) aux = f.axis_aux(j, cR[j]);
    // the factors (term changes) of the points c - g h and c + g h
    auto pair = [&](double g, double& fm, double& fp) {
      const double vm = __builtin_fma(-g, hR[j], cR[j]);
      const double vp = __builtin_fma(g, hR[j], cR[j]);
      if constexpr (kPair) {
        f.factor_pair(j, cR[j], g * hR[j], aux, fm, fp);
        if constexpr (kMasked) {
          if (!f.valid(j, vm)) fm = 0.0;
          if (!f.valid(j, vp)) fp = 0.0;
        }
      } else {
        fm = change(vm);
        fp = change(vp);
      }
    };
    double d1p, d1n, d2p, d2n;
    pair(C.lambda1, d1p, d1n);
    pair(C.lambda2, d2p, d2n);
    // sets 0-2 and the fourth difference, Sample.cuh:176-218
    {
      const double fa = gated(1u << j, Op::join(S0, d1p));
      const double fb = gated(1u << j, Op::join(S0, d1n));
      const double fc = gated(1u << j, Op::join(S0, d2p));
      const double fd = gated(1u << j, Op::join(S0, d2n));
      s1 += fa + fb;
      s2 += fc + fd;
      const double diff =
          __builtin_fabs(base + C.ratio * (fa * jac + fb * jac) - (fc * jac + fd * jac));
      if (diff > maxdiff) {
        maxdiff = diff;
        bisect = j;
      }
    }
    // sets 3-4: +-lambda3, +-lambda4 on one axis
    {
      double m3, p3, m4, p4;
      pair(C.lambda3, m3, p3);
      const double s3 = m3 + p3;
      acc3 = __builtin_fma(E3, s3, acc3 * uj);
      E3 *= uj;
      pair(C.lambda4, m4, p4);
      const double s4 = m4 + p4;
      acc4 = __builtin_fma(E4, s4, acc4 * uj);
      E4 *= uj;
    }
    // sets 5-7 (the pairs (j, k) and (k, j) with k < j)
    {
      const double a1 = d1p + d1n;
      const double a2 = d2p + d2n;
      e3 = __builtin_fma(e2, a1, e3 * uj);
      e2 = __builtin_fma(e1, a1, e2 * uj);
      g11 = __builtin_fma(g01, a1, __builtin_fma(g10, a2, g11 * uj));
      g10 = __builtin_fma(e0, a1, g10 * uj);
      g01 = __builtin_fma(e0, a2, g01 * uj);
      e1 = g10;
      e0 *= uj;
    }
    // set 8: the 2^D corners, g_j = +-lambda5 on every axis
    {
      double m5, p5;
      pair(C.lambda5, m5, p5);
      prod *= m5 + p5;
    }
  }
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r)
    rs[r] = C.w[0][r] * f0 + C.w[1][r] * s1 + C.w[2][r] * s2;
  const double v3 = value(S0 * acc3), v4 = value(S0 * acc4);
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[3][r], v3, rs[r]);
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[4][r], v4, rs[r]);
  const double v5 = value(S0 * e2), v6 = value(S0 * g11), v7 = value(S0 * e3);
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r)
    rs[r] = __builtin_fma(C.w[7][r], v7, __builtin_fma(C.w[6][r], v6,
                                                       __builtin_fma(C.w[5][r], v5, rs[r])));
  const double v8 = value(prod);
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[8][r], v8, rs[r]);
  if constexpr (kMasked && kExp) {
    // a masked exponential's factors are bounded only by the box: a region
    // whose sums left the range takes the generic loop
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < PAGANI_NRULES; ++r) t += __builtin_fabs(rs[r]);
    if (!__builtin_isfinite(t)) {
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = 0.0;
      bisect = bisect0;
      return false;
    }
  }
  return true;
}

// Rule sums and split axis of one region for a separable integrand
// (slices == 1: this wave evaluates all points).  Points are visited in the
// generic order of every set, and each set's values are summed in that
// order, as set_sum does; a point x = c - g h changes v_j by -g h_j.
// Returns false, having done nothing, for an exponential integrand whose
// region fails factor_safe.  Factored integrands (exponential, or
// multiplicative with a linear outer) take factored_sums; the rest join the
// centre's terms with the changed axes' term changes point by point.
template <int D, class F>
__device__ __forceinline__ bool separable_sums(F& f, const double (&cR)[D], const double (&hR)[D],
                                               double jac, const pagani_eval_consts& C,
                                               double (&rs)[PAGANI_NRULES], int& bisect)
{
  constexpr bool kExp = is_exponential<F>::value;
  constexpr bool kMul = is_multiplicative<F>::value || kExp;
  constexpr bool kMasked = is_masked<F>::value;
  constexpr bool kFactor = kExp || (kMul && has_linear_outer<F>::value);
  static_assert(!(kMasked && kMul && !kFactor), "masks need additive terms or factored sums");
  if constexpr (kFactor) {
    return factored_sums<D, F>(f, cR, hR, jac, C, rs, bisect);
  } else {
    using Op = SepOp<kMul>;
    double T0[D];
#pragma unroll
    for (int j = 0; j < D; ++j) T0[j] = f.term(j, cR[j]);
    // axes invalid at the centre: a point is zero unless it changes them all
    std::uint32_t vmask = 0;
    if constexpr (kMasked) {
#pragma unroll
      for (int j = 0; j < D; ++j)
        if (!f.valid(j, cR[j])) vmask |= 1u << j;
    }
    // the change of axis j's term when its coordinate moves to v
    auto change = [&](int j, double v) {
      if constexpr (kMasked)
        if (!f.valid(j, v)) return kMul ? 0.0 : -__builtin_inf();
      return Op::change(f.term(j, v), T0[j]);
    };
    // a point's value from its joined terms; zero when an unchanged axis is
    // invalid at the centre
    auto gated = [&](std::uint32_t changed, double s) {
      if constexpr (kMasked)
        if (vmask & ~changed) return 0.0;
      return f.outer(s);
    };
    double S0 = T0[0];
#pragma unroll
    for (int j = 1; j < D; ++j) S0 = Op::join(S0, T0[j]);
    // term changes at g = +lambda (x = c - lambda h) and g = -lambda
    double d1p[D], d1n[D], d2p[D], d2n[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      d1p[j] = change(j, __builtin_fma(-C.lambda1, hR[j], cR[j]));
      d1n[j] = change(j, __builtin_fma(C.lambda1, hR[j], cR[j]));
      d2p[j] = change(j, __builtin_fma(-C.lambda2, hR[j], cR[j]));
      d2n[j] = change(j, __builtin_fma(C.lambda2, hR[j], cR[j]));
    }
    // sets 0-2 and the fourth difference, Sample.cuh:176-218
    const double f0 = gated(0u, S0);
    const double base = (f0 * jac) * 2 * (1 - C.ratio);
    double s1 = 0.0, s2 = 0.0, maxdiff = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const double fa = gated(1u << j, Op::join(S0, d1p[j]));
      const double fb = gated(1u << j, Op::join(S0, d1n[j]));
      const double fc = gated(1u << j, Op::join(S0, d2p[j]));
      const double fd = gated(1u << j, Op::join(S0, d2n[j]));
      s1 += fa + fb;
      s2 += fc + fd;
      const double diff =
          __builtin_fabs(base + C.ratio * (fa * jac + fb * jac) - (fc * jac + fd * jac));
      if (diff > maxdiff) {
        maxdiff = diff;
        bisect = j;
      }
    }
#pragma unroll
    for (int r = 0; r < PAGANI_NRULES; ++r)
      rs[r] = C.w[0][r] * f0 + C.w[1][r] * s1 + C.w[2][r] * s2;
    // sets 3-4: +-lambda3, +-lambda4 on one axis
    const double lam34[2] = {C.lambda3, C.lambda4};
#pragma unroll
    for (int s = 3; s <= 4; ++s) {
      const double lam = lam34[s - 3];
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        acc += gated(1u << j, Op::join(S0, change(j, __builtin_fma(-lam, hR[j], cR[j]))));
        acc += gated(1u << j, Op::join(S0, change(j, __builtin_fma(lam, hR[j], cR[j]))));
      }
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[s][r], acc, rs[r]);
    }
    // set 5: (+-lambda1, +-lambda1) on axes j < k
    {
      double acc = 0.0;
#pragma unroll 1
      for (int j = 0; j < D; ++j) {
        const double aj[2] = {Op::join(S0, d1p[j]), Op::join(S0, d1n[j])};
#pragma unroll 1
        for (int k = j + 1; k < D; ++k) {
          const std::uint32_t ch = (1u << j) | (1u << k);
          acc += gated(ch, Op::join(aj[0], d1p[k]));
          acc += gated(ch, Op::join(aj[1], d1p[k]));
          acc += gated(ch, Op::join(aj[0], d1n[k]));
          acc += gated(ch, Op::join(aj[1], d1n[k]));
        }
      }
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[5][r], acc, rs[r]);
    }
    // set 6: (+-lambda1 on axis j, +-lambda2 on axis k != j)
    {
      double acc = 0.0;
#pragma unroll 1
      for (int j = 0; j < D; ++j) {
        const double aj[2] = {Op::join(S0, d1p[j]), Op::join(S0, d1n[j])};
#pragma unroll 1
        for (int k = 0; k < D; ++k) {
          if (k == j) continue;
          const std::uint32_t ch = (1u << j) | (1u << k);
          acc += gated(ch, Op::join(aj[0], d2p[k]));
          acc += gated(ch, Op::join(aj[1], d2p[k]));
          acc += gated(ch, Op::join(aj[0], d2n[k]));
          acc += gated(ch, Op::join(aj[1], d2n[k]));
        }
      }
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[6][r], acc, rs[r]);
    }
    // set 7: (+-lambda1, +-lambda1, +-lambda1) on axes j < k < l
    {
      double acc = 0.0;
#pragma unroll 1
      for (int j = 0; j < D; ++j) {
        const double aj[2] = {Op::join(S0, d1p[j]), Op::join(S0, d1n[j])};
#pragma unroll 1
        for (int k = j + 1; k < D; ++k) {
          const double ajk[4] = {Op::join(aj[0], d1p[k]), Op::join(aj[1], d1p[k]),
                                 Op::join(aj[0], d1n[k]), Op::join(aj[1], d1n[k])};
#pragma unroll 1
          for (int l = k + 1; l < D; ++l) {
            const std::uint32_t ch = (1u << j) | (1u << k) | (1u << l);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc += gated(ch, Op::join(ajk[q], d1p[l]));
#pragma unroll
            for (int q = 0; q < 4; ++q) acc += gated(ch, Op::join(ajk[q], d1n[l]));
          }
        }
      }
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[7][r], acc, rs[r]);
    }
    // set 8: the 2^D corners x = c - g h, g_j = +lambda5 where bit j of the
    // corner index is 0 and -lambda5 where it is 1 (the table's order).  The
    // low LB axes' combined terms are tabulated once; each group of 2^LB
    // corners adds the high axes' combination.  An integrand whose outer
    // function is long (heavy_outer) takes the table loop instead: unrolled
    // 2^LB times, its outer would not fit the registers.
    if constexpr (heavy_outer<F>::value) {
      const int p0 = C.set_begin[3];
      const double acc = set_sum<D, F, double>(
          f, C.gtab + static_cast<std::size_t>(C.set_begin[8] - p0) * D, C.set_begin[8],
          C.set_begin[9], cR, hR);
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[8][r], acc, rs[r]);
    } else {
      constexpr int LB = D < 4 ? D : 4;
      constexpr int HB = D - LB;
      double tp[D], tn[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        // the corners change every axis: plain terms, -infinity where invalid
        const double vp = __builtin_fma(-C.lambda5, hR[j], cR[j]);
        const double vn = __builtin_fma(C.lambda5, hR[j], cR[j]);
        tp[j] = f.term(j, vp);
        tn[j] = f.term(j, vn);
        if constexpr (kMasked) {
          if (!f.valid(j, vp)) tp[j] = -__builtin_inf();
          if (!f.valid(j, vn)) tn[j] = -__builtin_inf();
        }
      }
      double lo[1 << LB];
      lo[0] = tp[0];
      lo[1] = tn[0];
#pragma unroll
      for (int j = 1; j < LB; ++j) {
#pragma unroll
        for (int q = 0; q < (1 << j); ++q) {
          lo[q + (1 << j)] = Op::join(lo[q], tn[j]);
          lo[q] = Op::join(lo[q], tp[j]);
        }
      }
      double acc = 0.0;
#pragma unroll 1
      for (int hq = 0; hq < (1 << HB); ++hq) {
        double hs = 0.0;
        if constexpr (HB > 0) {
          hs = (hq & 1) ? tn[LB] : tp[LB];
#pragma unroll
          for (int j = LB + 1; j < D; ++j)
            hs = Op::join(hs, ((hq >> (j - LB)) & 1) ? tn[j] : tp[j]);
        }
#pragma unroll
        for (int q = 0; q < (1 << LB); ++q) {
          const double t = HB > 0 ? Op::join(lo[q], hs) : lo[q];
          acc += f.outer(t);
        }
      }
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r) rs[r] = __builtin_fma(C.w[8][r], acc, rs[r]);
    }
    return true;
  }
}

// Rule sums and split axis of one region for a cosine integrand (slices ==
// 1).  A point that moves v_j by -g_j h_j on a set A of axes has the angle
// phi0 - sum_A g_j h_j; the rule takes every sign combination of the g_j,
// and the sum over them is Re(e^(i phi0) prod_A (e^(i l h_j) + e^(-i l h_j)))
// = cos(phi0) prod_A 2 cos(l h_j).  So a set's sum is cos(phi0) times a
// symmetric sum of per-axis cosines: per region, one sincos at the centre,
// a sincos at +-lambda1 and +-lambda2 per axis (the fourth difference needs
// those points one by one) and a cosine at lambda3..5 per axis.
__device__ __forceinline__ void sincos_real(double x, double* s, double* c)
{
  math::sincos_bounded(x, *s, *c);
}
__device__ __forceinline__ void sincos_real(float x, float* s, float* c) { ::sincosf(x, s, c); }
__device__ __forceinline__ double cos_real(double x) { return math::cos_bounded(x); }
__device__ __forceinline__ float cos_real(float x) { return ::cosf(x); }

// Returns false, having done nothing, when an angle could leave the range of
// the FP64 sine and cosine (math::kTrigRange).
template <int D, class F, class Real>
__device__ __forceinline__ bool cosine_sums(F& f, const double (&cR)[D], const double (&hR)[D],
                                            double jac, const pagani_eval_consts& C,
                                            double (&rs)[PAGANI_NRULES], int& bisect)
{
  double phi0 = f.phase_angle();
  double hmax = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    phi0 += cR[j];
    hmax = __builtin_fmax(hmax, __builtin_fabs(hR[j]));
  }
  // every lambda is below 1: the axis angles are at most |h_j|
  if (!(__builtin_fabs(phi0) <= math::kTrigRange && hmax <= math::kTrigRange)) return false;
  Real sr, cr;
  sincos_real(static_cast<Real>(phi0), &sr, &cr);
  const double c0 = cr, s0 = sr;
  const double base = (c0 * jac) * 2 * (1 - C.ratio);
  double s1 = 0.0, s2 = 0.0, maxdiff = 0.0;
  // x^m and xy coefficients of prod_j (1 + x A1_j + y A2_j), A = 2 cos(l h_j)
  double e1 = 0.0, e2 = 0.0, e3 = 0.0, g01 = 0.0, g11 = 0.0;
  double a3 = 0.0, a4 = 0.0, p5 = 1.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    Real sa, ca, sb, cb;
    sincos_real(static_cast<Real>(C.lambda1 * hR[j]), &sa, &ca);
    sincos_real(static_cast<Real>(C.lambda2 * hR[j]), &sb, &cb);
    // cos(phi0 -+ l h_j) = c0 cos(l h_j) +- s0 sin(l h_j)
    const double fa = __builtin_fma(s0, static_cast<double>(sa), c0 * ca);
    const double fb = __builtin_fma(-s0, static_cast<double>(sa), c0 * ca);
    const double fc = __builtin_fma(s0, static_cast<double>(sb), c0 * cb);
    const double fd = __builtin_fma(-s0, static_cast<double>(sb), c0 * cb);
    s1 += fa + fb;
    s2 += fc + fd;
    const double diff =
        __builtin_fabs(base + C.ratio * (fa * jac + fb * jac) - (fc * jac + fd * jac));
    if (diff > maxdiff) {
      maxdiff = diff;
      bisect = j;
    }
    const double A1 = 2.0 * static_cast<double>(ca), A2 = 2.0 * static_cast<double>(cb);
    e3 = __builtin_fma(e2, A1, e3);
    e2 = __builtin_fma(e1, A1, e2);
    g11 = __builtin_fma(g01, A1, __builtin_fma(e1, A2, g11));
    e1 += A1;
    g01 += A2;
    a3 += 2.0 * static_cast<double>(cos_real(static_cast<Real>(C.lambda3 * hR[j])));
    a4 += 2.0 * static_cast<double>(cos_real(static_cast<Real>(C.lambda4 * hR[j])));
    p5 *= 2.0 * static_cast<double>(cos_real(static_cast<Real>(C.lambda5 * hR[j])));
  }
  const double v[PAGANI_NSETS] = {c0,      s1,      s2,      c0 * a3, c0 * a4,
                                  c0 * e2, c0 * g11, c0 * e3, c0 * p5};
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r) {
    double acc = C.w[0][r] * v[0] + C.w[1][r] * v[1] + C.w[2][r] * v[2];
#pragma unroll
    for (int s = 3; s < PAGANI_NSETS; ++s) acc = __builtin_fma(C.w[s][r], v[s], acc);
    rs[r] = acc;
  }
  return true;
}

// Region t's left coordinate and length along axis j.  Flat list (pdim
// NULL): row t of n.  Split form: row t mod n/2 of the n/2 parents, halved
// along the parent's axis b = pdim[row] -- lo, h / 2 for child 0 and
// lo + h / 2, h / 2 for child 1, the operations Sub_region_splitter.cuh:
// 8-46 performs, so the bounds are those a written-out split would hold.
struct RegionRow {
  std::size_t row, stride;
  int axis;  // the halved axis, -1 for a flat list
  bool right;
  const double* h0;  // level form: the widths at level 0 (H holds levels); else NULL
  __device__ __forceinline__ RegionRow(const std::uint8_t* __restrict__ pdim, std::size_t n,
                                       std::size_t t, const double* level0 = nullptr)
      : h0(level0)
  {
    if (pdim) {
      const std::size_t m = n / 2;
      right = t >= m;
      row = right ? t - m : t;
      stride = m;
      axis = pdim[row];
    } else {
      right = false;
      row = t;
      stride = n;
      axis = -1;
    }
  }
  __device__ __forceinline__ void bounds(const double* __restrict__ L,
                                         const double* __restrict__ H, int j, double& lo,
                                         double& len) const
  {
    const std::size_t o = static_cast<std::size_t>(j) * stride + row;
    lo = L[o];
    if (h0)
      len = __builtin_ldexp(h0[j], -static_cast<int>(reinterpret_cast<const std::uint16_t*>(H)[o]));
    else
      len = H[o];
    if (j == axis) {
      const double half = len / 2;
      if (right) lo = lo + half;
      len = half;
    }
  }
};

// Region setup, Phases.cuh:323-349: volume, widest axis, and the centre /
// width of region t in the integrand's frame (user coordinates, moved to the
// origin and scaled for a centred integrand); a rule point is x = c - g w.
template <int D, class F>
__device__ __forceinline__ void region_frame(F& f, const double* __restrict__ L,
                                             const double* __restrict__ H,
                                             const std::uint8_t* __restrict__ pdim, std::size_t n,
                                             std::size_t t, const pagani_eval_consts& C,
                                             double (&cR)[D], double (&hR)[D], double& vol,
                                             int& maxdim)
{
  const double* lows = C.lows;
  const double* range = C.range;
  double maxr = 0.0;
  vol = 1.0;
  maxdim = 0;
  const RegionRow rr(pdim, n, t, C.levels ? C.h0 : nullptr);
#pragma unroll
  for (int j = 0; j < D; ++j) {
    double lo, len;
    rr.bounds(L, H, j, lo, len);
    const double up = lo + len;
    const double r = up - lo;
    vol *= r;
    if (r > maxr) {
      maxr = r;
      maxdim = j;
    }
    // one rounding in the map to user coordinates (an FMA), so the points
    // are a fixed function of the region's bounds whatever the compiler
    // contracts
    cR[j] = __builtin_fma(0.5 * (lo + up), range[j], lows[j]);
    hR[j] = r * range[j];
    if constexpr (has_origin<F>::value) cR[j] -= f.origin(j);
    if constexpr (has_origin<F>::value && has_axis_scale<F>::value) {
      cR[j] *= f.axis_scale(j);
      hR[j] *= f.axis_scale(j);
    }
  }
}

// The tile of workgroup b of a grid of g: workgroups are dealt to the eight
// XCDs round robin (b to XCD b mod 8), so each XCD takes a run of
// consecutive tiles, one after another, and the cache lines that
// neighbouring tiles share (the level, axis and split-axis arrays hold less
// than a line per tile) are fetched into one XCD's L2 once.  A bijection of
// [0, g).
__device__ __forceinline__ std::size_t xcd_tile(unsigned b, unsigned g)
{
  const unsigned x = b & 7u, q = g >> 3, r = g & 7u;
  const unsigned first = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return static_cast<std::size_t>(first) + (b >> 3);
}

// Generator coordinate j of rule point p (the order of pagani_rule_point):
// the centre, +-lambda1 and +-lambda2 per axis, then the table.
__device__ __forceinline__ double point_generator(const pagani_eval_consts& C, int p, int j)
{
  if (p == 0) return 0.0;
  if (p < C.set_begin[3]) {
    const bool two = p >= C.set_begin[2];
    const int k = p - (two ? C.set_begin[2] : C.set_begin[1]);
    if (k / 2 != j) return 0.0;
    const double lam = two ? C.lambda2 : C.lambda1;
    return (k & 1) ? -lam : lam;
  }
  return C.gtab[static_cast<std::size_t>(p - C.set_begin[3]) * C.ndim + j];
}

// One wave's share of the evaluation of up to 64 regions, one lane per
// region t of an n-region SoA list (split form when pdim is not NULL), with
// the results left in registers: avg, er and the split axis.  `wave`
// indexes the `slices` waves that split the batch's points (0 when slices
// == 1); with more than one, every wave of the workgroup passes one barrier,
// and only wave 0 returns true (holding the results).  `lows`, `range`,
// `jac` describe the integration box.
template <int D, class F, class Real>
__device__ __forceinline__ bool eval_region_values(
    F& f, const double* __restrict__ L, const double* __restrict__ H,
    const std::uint8_t* __restrict__ pdim, std::size_t n, std::size_t t,
    const double* lows, const double* range, double jac, const pagani_eval_consts& C, int slices,
    int wave, int lane, double* part, double& avg_out, double& er_out, int& axis_out)
{
  double cR[D], hR[D];
  double vol;
  int maxdim;
  region_frame<D>(f, L, H, pdim, n, t, C, cR, hR, vol, maxdim);

  double rs[PAGANI_NRULES] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int bisect = maxdim;

  bool done = false;
  if constexpr (is_cosine<F>::value && has_origin<F>::value && has_axis_scale<F>::value) {
    if (slices == 1) done = cosine_sums<D, F, Real>(f, cR, hR, jac, C, rs, bisect);
  } else if constexpr (is_separable<F>::value) {
    if (slices == 1) done = separable_sums<D, F>(f, cR, hR, jac, C, rs, bisect);
  }
  if (!done) {
    if (wave == 0) {
      // Centre and the +-lambda1, +-lambda2 axis points (sets 0-2), whose
      // values give the fourth difference per axis, Sample.cuh:176-218.
      Real xc[D];
#pragma unroll
      for (int j = 0; j < D; ++j) xc[j] = static_cast<Real>(cR[j]);
      const double f0 = detail::call(f, xc);
      const double base = (f0 * jac) * 2 * (1 - C.ratio);
      double s1 = 0.0, s2 = 0.0, maxdiff = 0.0;
      // f at the centre moved by -g h_j on axis j
      auto moved = [&](int j, double g) {
        Real x[D];
#pragma unroll
        for (int k = 0; k < D; ++k)
          x[k] = k == j ? static_cast<Real>(__builtin_fma(-g, hR[k], cR[k])) : xc[k];
        return detail::call(f, x);
      };
      auto axis = [&](int j) {
        const double fa = moved(j, C.lambda1);
        const double fb = moved(j, -C.lambda1);
        const double fc = moved(j, C.lambda2);
        const double fd = moved(j, -C.lambda2);
        s1 += fa + fb;
        s2 += fc + fd;
        const double diff = __builtin_fabs(
            base + C.ratio * (fa * jac + fb * jac) - (fc * jac + fd * jac));
        if (diff > maxdiff) {
          maxdiff = diff;
          bisect = j;
        }
      };
      if constexpr (heavy_outer<F>::value) {
        // a long integrand: one copy of the four evaluations, the axis
        // chosen by selects
#pragma unroll 1
        for (int j = 0; j < D; ++j) axis(j);
      } else {
#pragma unroll
        for (int j = 0; j < D; ++j) axis(j);
      }
#pragma unroll
      for (int r = 0; r < PAGANI_NRULES; ++r)
        rs[r] = C.w[0][r] * f0 + C.w[1][r] * s1 + C.w[2][r] * s2;
    }

    // Sets 3..8 from the generator table, this wave's slice of the points.
    const int p0 = C.set_begin[3];
    int pa, pb;
    slice_range(slices, wave, p0, C.feval, pa, pb);
#pragma unroll
    for (int s = 3; s < PAGANI_NSETS; ++s) {
      const int a = pa > C.set_begin[s] ? pa : C.set_begin[s];
      const int b = pb < C.set_begin[s + 1] ? pb : C.set_begin[s + 1];
      if (a < b) {
        const double acc = detail::set_sum<D, F, Real>(
            f, C.gtab + static_cast<std::size_t>(a - p0) * D, a, b, cR, hR);
#pragma unroll
        for (int r = 0; r < PAGANI_NRULES; ++r)
          rs[r] = __builtin_fma(C.w[s][r], acc, rs[r]);
      }
    }

    if (slices > 1) {
      if (wave > 0) {
#pragma unroll
        for (int r = 0; r < PAGANI_NRULES; ++r)
          part[((wave - 1) * PAGANI_NRULES + r) * 64 + lane] = rs[r];
      }
      __syncthreads();
      if (wave == 0) {
        for (int w = 1; w < slices; ++w) {
#pragma unroll
          for (int r = 0; r < PAGANI_NRULES; ++r)
            rs[r] += part[((w - 1) * PAGANI_NRULES + r) * 64 + lane];
        }
      }
      if (wave > 0) return false;
    }
  }

  // Null-rule error estimate, Sample.cuh:261-289.
  double s[PAGANI_NRULES];
#pragma unroll
  for (int r = 0; r < PAGANI_NRULES; ++r) s[r] = rs[r] * jac;
  double e1 = 0.0, e2 = 0.0, e3 = 0.0;
#pragma unroll
  for (int k = 0; k < PAGANI_NSETS; ++k) {
    e1 = __builtin_fmax(e1, __builtin_fabs(s[2] + C.scale[k][1] * s[1]) * C.norm[k][1]);
    e2 = __builtin_fmax(e2, __builtin_fabs(s[3] + C.scale[k][2] * s[2]) * C.norm[k][2]);
    e3 = __builtin_fmax(e3, __builtin_fabs(s[4] + C.scale[k][3] * s[3]) * C.norm[k][3]);
  }
  avg_out = vol * s[0];
  er_out = vol * ((5.0 * e1 <= e2 && 5.0 * e2 <= e3)
                      ? e1
                      : 5.0 * __builtin_fmax(__builtin_fmax(e1, e2), e3));
  axis_out = bisect;
  return true;
}

// eval_region_values, stored: region t's estimate, error and split axis.
template <int D, class F, class Real>
__device__ __forceinline__ void eval_regions_wave(
    F& f, const double* __restrict__ L, const double* __restrict__ H,
    const std::uint8_t* __restrict__ pdim, std::size_t n, std::size_t t, bool live,
    const double* lows, const double* range, double jac, const pagani_eval_consts& C, int slices,
    int wave, int lane, double* part, double* __restrict__ est, double* __restrict__ err,
    std::uint8_t* __restrict__ dim)
{
  double avg, er;
  int axis;
  if (!eval_region_values<D, F, Real>(f, L, H, pdim, n, t, lows, range, jac, C, slices, wave,
                                      lane, part, avg, er, axis))
    return;
  if (live) {
    est[t] = avg;
    err[t] = er;
    dim[t] = static_cast<std::uint8_t>(axis);
  }
}

}  // namespace detail

// Kernel: workgroup b evaluates a batch of 64 regions of an n-region SoA
// list (flat: [64 b, 64 b + 64); split form when P is given: the children
// of parents [32 b, 32 b + 32)) with S = slices waves sharing the batch's
// points.  For a
// host-driven launch `it` is NULL and n, the pool (L0, H0) and S = C.slices
// come from the arguments.  In the device-resident loop they come from *it
// (nothing at all if !it->run), and the launch is sized for the round's
// largest list: workgroups past n return at once, and waves past the
// iteration's slice count only keep the barrier company.  Both forms run
// this one kernel, so a given n and slice count give the same bits either
// way.  Dynamic LDS: (waves per workgroup - 1)*5*64 doubles.
template <int D, class F, class Real>
__global__ void __launch_bounds__(64 * PAGANI_MAX_SLICES)
    eval_kernel(F f, const pagani_dev_iter* __restrict__ it, std::size_t n_arg,
                const double* __restrict__ L0, const double* __restrict__ H0,
                const double* __restrict__ L1, const double* __restrict__ H1,
                const std::uint8_t* __restrict__ P0, const std::uint8_t* __restrict__ P1,
                double* __restrict__ est, double* __restrict__ err,
                std::uint8_t* __restrict__ dim, const pagani_eval_consts C)
{
  extern __shared__ double part[];
  std::size_t n = n_arg;
  const double* L = L0;
  const double* H = H0;
  const std::uint8_t* P = P0;
  int slices = C.slices;
  if (it) {
    if (!it->run) return;
    n = it->n;
    if (it->cur) {
      L = L1;
      H = H1;
      P = P1;
    }
    if (!it->split) P = nullptr;
    slices = static_cast<int>(it->slices);
  }
  // the list's tiles of 64 regions; workgroups past them return at once
  const std::size_t tiles = (n + 63) / 64;
  if (blockIdx.x >= tiles) return;
  const std::size_t tile = detail::xcd_tile(blockIdx.x, static_cast<unsigned>(tiles));
  const int lane = threadIdx.x & 63;
  // wave index is uniform across the wave: keep it (and everything derived
  // from it, the slice bounds and the point loop) in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  if (wave >= slices) {
    if (slices > 1) __syncthreads();
    return;
  }
  // A flat list gives tile b the regions 64 b + lane.  A split list (n = 2m
  // children, child t of parent t mod m) gives it both children of the
  // parents 32 b .. 32 b + 31: lanes l and l + 32 read the same parent row
  // in one load instruction, so each row leaves HBM once, and the writes are
  // two runs of 32 consecutive regions.
  std::size_t t;
  bool live;
  if (P) {
    const std::size_t m = n / 2;
    const std::size_t p0 = tile * 32 + (lane & 31);
    live = p0 < m;
    t = (live ? p0 : m - 1) + (lane >= 32 ? m : 0);
  } else {
    const std::size_t t0 = tile * 64 + lane;
    live = t0 < n;
    t = live ? t0 : n - 1;
  }
  detail::eval_regions_wave<D, F, Real>(f, L, H, P, n, t, live, C.lows, C.range, C.jac, C,
                                        slices, wave, lane, part, est, err, dim);
}

// The debug recorder's point dump: thread i writes the value at rule point
// i mod feval of region i / feval, formed as the evaluation kernel forms it.
template <int D, class F, class Real>
__global__ void __launch_bounds__(256)
    point_values_kernel(F f, std::size_t n, const double* __restrict__ L,
                        const double* __restrict__ H, const std::uint8_t* __restrict__ P,
                        double* __restrict__ out, const pagani_eval_consts C)
{
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * 256 + threadIdx.x;
  const std::size_t F_ = static_cast<std::size_t>(C.feval);
  const std::size_t t = i / F_;
  if (t >= n) return;
  const int p = static_cast<int>(i % F_);
  double cR[D], hR[D], vol;
  int maxdim;
  detail::region_frame<D>(f, L, H, P, n, t, C, cR, hR, vol, maxdim);
  Real x[D];
#pragma unroll
  for (int j = 0; j < D; ++j)
    x[j] = static_cast<Real>(__builtin_fma(-detail::point_generator(C, p, j), hR[j], cR[j]));
  out[i] = detail::call(f, x);
}

// Launch the evaluation of args->n regions for integrand f.  Returns 0 or
// PAGANI_EHIP.  Real is the integrand's arithmetic type (double or float).
template <int D, class F, class Real = double>
int eval_regions(const F& f, const pagani_eval_args* args)
{
  static_assert(D >= PAGANI_MIN_DIM && D <= PAGANI_MAX_DIM, "unsupported ndim");
  // a library built from other headers lays its lists out differently
  if (args->consts.abi_version != PAGANI_EVAL_ABI_VERSION) return PAGANI_EUNSUPPORTED;
  if (args->consts.ndim != D) return PAGANI_EBADARG;
  if (args->n == 0) return PAGANI_OK;
  if (!args->dev && args->pdim && (args->n % 2)) return PAGANI_EBADARG;  // pairs of children
  // waves per workgroup: the slice count (an upper bound of the device
  // loop's counts for its launches)
  const int waves = args->consts.slices;
  if (waves < 1 || waves > PAGANI_MAX_SLICES) return PAGANI_EBADARG;
  const std::size_t blocks = (args->n + 63) / 64;
  const std::size_t lds =
      waves > 1 ? static_cast<std::size_t>(waves - 1) * PAGANI_NRULES * 64 * sizeof(double) : 0;
  hipLaunchKernelGGL((eval_kernel<D, F, Real>), dim3(static_cast<unsigned>(blocks)),
                     dim3(64 * waves), lds, static_cast<hipStream_t>(args->stream), f, args->dev,
                     args->n, args->L, args->H, args->dev ? args->L2 : args->L,
                     args->dev ? args->H2 : args->H, args->pdim,
                     args->dev ? args->pdim2 : args->pdim, args->est, args->err, args->dim,
                     args->consts);
  if (args->fevals && !args->dev) {
    const std::size_t threads = args->n * static_cast<std::size_t>(args->consts.feval);
    hipLaunchKernelGGL((point_values_kernel<D, F, Real>), dim3(static_cast<unsigned>((threads + 255) / 256)),
                       dim3(256), 0, static_cast<hipStream_t>(args->stream), f, args->n, args->L,
                       args->H, args->pdim, args->fevals, args->consts);
  }
  return hipGetLastError() == hipSuccess ? PAGANI_OK : PAGANI_EHIP;
}

}  // namespace pagani
