// genz.hpp -- the six Genz test families as device functors.
//
// The reference's fixed-parameter functors F_1..F_6 (common/cuda/
// integrands.cuh:1303-1934) are instances of these families with given
// (a, u).  Real = double for the FP64 path, float for the mixed-precision
// path (FP32 integrand, FP64 accumulation).  Parameters are pre-processed on
// the host (a^2, 1/a^2, 2*pi*u_0) so the device does no divisions by a.
//
// The families that depend on x only through x - u (product peak, Gaussian,
// C0) also offer the "centred" form: origin(i) = u_i and centered(t...) =
// f(u + t).  The evaluation kernel then folds u into each region's centre
// once, so a rule point costs one FMA per axis fewer.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "pagani.h"

namespace pagani {
namespace math {

// e^x for x <= 709 (the exponents of the Gaussian and C0 families are minus
// sums of non-negative terms, the discontinuous family's are bounded by
// sum a_i).  x = n ln2 + r with n the integer nearest x / ln2 and |r| <=
// ln2/2, r from a two-part ln2 (Cody-Waite); e^r by a degree-11 polynomial;
// 2^n by ldexp, which also rounds subnormal results, gives 0 below about
// -745.1 and infinity above about 709.78.  n comes from one FMA: adding
// 1.5 * 2^52 to x log2e (|x log2e| < 2^51) leaves a double whose unit is 1,
// so the sum is rounded to the nearest integer, which its low 32 bits hold
// in two's complement; subtracting the shift again gives n as a double (no
// separate multiply, round or conversion instruction).  The coefficients
// interpolate e^r at the 12 Chebyshev nodes of [-ln2/2, ln2/2] (computed in
// extended precision, then rounded to double; the check that recomputes
// them is in the tests): max polynomial error 0.14 ulp.  The range handling
// is one clamp, which also maps -inf to 0 (and NaN to 0: the families never
// produce NaN from finite input).
__device__ __forceinline__ double exp_bounded(double x)
{
  constexpr double kLog2e = 0x1.71547652b82fep+0;
  constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
  constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;
  constexpr double kShift = 0x1.8p+52;
  constexpr double c[12] = {
      0x1.0000000000000p+0,  0x1.0000000000000p+0,  0x1.0000000000011p-1, 0x1.555555555555fp-3,
      0x1.555555554f0c7p-5,  0x1.111111110e636p-7,  0x1.6c16c187face9p-10, 0x1.a01a01b4a99f9p-13,
      0x1.a01991b378c3ap-16, 0x1.71ddeea4df7bbp-19, 0x1.28b3fce70e98ep-22, 0x1.af6d18013a587p-26};
  x = __builtin_fmax(x, -1.0e4);
  const double t = __builtin_fma(x, kLog2e, kShift);
  const double n = t - kShift;
  double r = __builtin_fma(-n, kLn2Hi, x);
  r = __builtin_fma(-n, kLn2Lo, r);
  double p = c[11];
#pragma unroll
  for (int k = 10; k >= 0; --k) p = __builtin_fma(p, r, c[k]);
  return __builtin_ldexp(p, static_cast<int>(__builtin_bit_cast(unsigned long long, t)));
}
__device__ __forceinline__ double exp_neg(double x) { return exp_bounded(x); }
__device__ __forceinline__ float exp_neg(float x) { return ::expf(x); }
__device__ __forceinline__ float exp_bounded(float x) { return ::expf(x); }

__device__ __forceinline__ double exp_(double x) { return ::exp(x); }
__device__ __forceinline__ float exp_(float x) { return ::expf(x); }
__device__ __forceinline__ double cos_(double x) { return ::cos(x); }
__device__ __forceinline__ float cos_(float x) { return ::cosf(x); }
__device__ __forceinline__ double fabs_(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float fabs_(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double fmax_(double x, double y) { return __builtin_fmax(x, y); }
__device__ __forceinline__ float fmax_(float x, float y) { return __builtin_fmaxf(x, y); }
}  // namespace math

// cos(2 pi u_0 + sum a_i x_i)
template <class Real, int D>
struct Oscillatory {
  Real a[D];
  Real phase;  // 2 pi u_0
  template <class... X>
  __device__ Real operator()(X... xs) const
  {
    const Real x[D] = {static_cast<Real>(xs)...};
    Real s = phase;
#pragma unroll
    for (int i = 0; i < D; ++i) s += a[i] * x[i];
    return math::cos_(s);
  }
  // v = a x (origin 0): f = cos(phase + sum v), additively separable
  __device__ double origin(int) const { return 0.0; }
  __device__ double axis_scale(int i) const { return static_cast<double>(a[i]); }
  template <class... T>
  __device__ Real centered(T... vs) const
  {
    const Real v[D] = {static_cast<Real>(vs)...};
    Real s = phase;
#pragma unroll
    for (int i = 0; i < D; ++i) s += v[i];
    return math::cos_(s);
  }
  static constexpr bool heavy_outer = true;  // cos: argument reduction and two polynomials
  // f = cos(phase_angle() + sum_j v_j): the kernel sums each set's sign
  // combinations as cos of the centre's angle times per-axis cosines
  static constexpr bool cosine = true;
  __device__ double phase_angle() const { return static_cast<double>(phase); }
  __device__ double term(int, double v) const { return v; }
  __device__ double outer(double s) const
  {
    return static_cast<double>(math::cos_(static_cast<Real>(static_cast<double>(phase) + s)));
  }
};

// prod 1 / (a_i^-2 + (x_i - u_i)^2), evaluated as one division
template <class Real, int D>
struct ProductPeak {
  Real ia2[D];  // 1 / a_i^2
  Real u[D];
  Real a[D];
  double k2;    // prod a_i^2
  template <class... X>
  __device__ Real operator()(X... xs) const
  {
    const Real x[D] = {static_cast<Real>(xs)...};
    Real den = 1;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const Real t = x[i] - u[i];
      den *= ia2[i] + t * t;
    }
    return Real(1) / den;
  }
  // f(u + v / a) = prod a_i^2 / (1 + v_i^2): the kernel hands over
  // v = a (x - u); multiplicatively separable, term 1 / (1 + v^2)
  __device__ double origin(int i) const { return static_cast<double>(u[i]); }
  __device__ double axis_scale(int i) const { return static_cast<double>(a[i]); }
  template <class... T>
  __device__ Real centered(T... vs) const
  {
    const Real v[D] = {static_cast<Real>(vs)...};
    Real den = 1;
#pragma unroll
    for (int i = 0; i < D; ++i) den *= Real(1) + v[i] * v[i];
    return static_cast<Real>(k2) / den;
  }
  static constexpr bool multiplicative = true;
  static constexpr bool linear_outer = true;
  __device__ double term(int, double v) const { return 1.0 / (1.0 + v * v); }
  __device__ double outer(double p) const { return static_cast<double>(static_cast<Real>(k2 * p)); }
};

// (1 + sum a_i x_i)^-(d+1), the power by repeated multiplication
template <class Real, int D>
struct CornerPeak {
  Real a[D];
  template <class... X>
  __device__ Real operator()(X... xs) const
  {
    const Real x[D] = {static_cast<Real>(xs)...};
    Real s = 1;
#pragma unroll
    for (int i = 0; i < D; ++i) s += a[i] * x[i];
    return power(s);
  }
  static __device__ Real power(Real s)
  {
    Real p = s;
#pragma unroll
    for (int i = 0; i < D; ++i) p *= s;
    return Real(1) / p;
  }
  // v = a x (origin 0): f = (1 + sum v)^-(d+1), additively separable
  __device__ double origin(int) const { return 0.0; }
  __device__ double axis_scale(int i) const { return static_cast<double>(a[i]); }
  template <class... T>
  __device__ Real centered(T... vs) const
  {
    const Real v[D] = {static_cast<Real>(vs)...};
    Real s = 1;
#pragma unroll
    for (int i = 0; i < D; ++i) s += v[i];
    return power(s);
  }
  __device__ double term(int, double v) const { return v; }
  __device__ double outer(double s) const
  {
    return static_cast<double>(power(static_cast<Real>(1.0 + s)));
  }
};

// The largest centre sum s0 for which an exponential family evaluates a
// region as products: e^s0 and e^-s0 stay normal with room to spare (FP64
// 2^+-1022 is e^+-708; FP32 2^+-126 is e^+-87).
template <class Real>
constexpr double kFactorRange = sizeof(Real) == 8 ? 650.0 : 80.0;

namespace math {
// e^x (x <= 709) for the product forms of the exponential families: at
// the integrand's precision where it holds the value (FP32: |x| <= 80),
// else in FP64, so that an FP32 integrand's centre values and factors stay
// normal over the FP64 range of the products.
__device__ __forceinline__ double exp_factor(double x, double) { return exp_bounded(x); }

// 1 / x for finite, normal x: the hardware reciprocal refined by two
// Newton steps (each doubles the correct bits: r' = r + r (1 - x r)).
__device__ __forceinline__ double recip(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double exp_factor(double x, float)
{
  return __builtin_fabs(x) <= 80.0 ? static_cast<double>(::expf(static_cast<float>(x)))
                                   : exp_bounded(x);
}
}  // namespace math

// exp(-sum a_i^2 (x_i - u_i)^2)
template <class Real, int D>
struct Gaussian {
  Real a[D];   // a_i
  Real a2[D];  // a_i^2
  Real u[D];
  template <class... X>
  __device__ Real operator()(X... xs) const
  {
    const Real x[D] = {static_cast<Real>(xs)...};
    const Real t0 = x[0] - u[0];
    Real s = a2[0] * (t0 * t0);
#pragma unroll
    for (int i = 1; i < D; ++i) {
      const Real t = x[i] - u[i];
      s += a2[i] * (t * t);
    }
    return math::exp_neg(-s);
  }
  // f(u + v / a): the evaluation kernel hands over v = a (x - u), the
  // region's centre and half-widths scaled by a once per region, so a point
  // costs one FMA per axis for v and one for the sum of squares
  __device__ double origin(int i) const { return static_cast<double>(u[i]); }
  __device__ double axis_scale(int i) const { return static_cast<double>(a[i]); }
  template <class... T>
  __device__ Real centered(T... vs) const
  {
    const Real v[D] = {static_cast<Real>(vs)...};
    Real s = v[0] * v[0];
#pragma unroll
    for (int i = 1; i < D; ++i) s += v[i] * v[i];  // one FMA per axis (contracted)
    return math::exp_neg(-s);
  }
  // additively separable in v: f = outer(sum_j term(j, v_j)); the terms
  // and their sums in FP64, outer in the integrand's precision
  __device__ double term(int, double v) const { return v * v; }
  __device__ double outer(double s) const
  {
    return static_cast<double>(math::exp_neg(static_cast<Real>(-s)));
  }
  // outer(s + d) = outer(s) e^-d.  The terms are non-negative, so in a
  // region whose centre sums to s0 every factor is at most e^s0 and every
  // point's value at most 1: s0 bounded well inside the FP64 exponent range
  // keeps the centre's value, each factor and every partial product normal
  // (an FP32 integrand takes those FP32 cannot hold in FP64).
  static constexpr bool exponential = true;
  __device__ double factor(double d) const { return math::exp_factor(-d, Real()); }
  __device__ double centre_value(double s0) const { return math::exp_factor(-s0, Real()); }
  __device__ bool factor_safe(double s0) const { return s0 <= kFactorRange<double>; }
};

// exp(-sum a_i |x_i - u_i|)
template <class Real, int D>
struct C0 {
  Real a[D];
  Real u[D];
  template <class... X>
  __device__ Real operator()(X... xs) const
  {
    const Real x[D] = {static_cast<Real>(xs)...};
    Real s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += a[i] * math::fabs_(x[i] - u[i]);
    return math::exp_neg(-s);
  }
  // f(u + v / a): the kernel hands over v = a (x - u) (see Gaussian)
  __device__ double origin(int i) const { return static_cast<double>(u[i]); }
  __device__ double axis_scale(int i) const { return static_cast<double>(a[i]); }
  template <class... T>
  __device__ Real centered(T... vs) const
  {
    const Real v[D] = {static_cast<Real>(vs)...};
    Real s = math::fabs_(v[0]);
#pragma unroll
    for (int i = 1; i < D; ++i) s += math::fabs_(v[i]);
    return math::exp_neg(-s);
  }
  __device__ double term(int, double v) const { return __builtin_fabs(v); }
  __device__ double outer(double s) const
  {
    return static_cast<double>(math::exp_neg(static_cast<Real>(-s)));
  }
  // The points v -+ delta of one axis change |v| by -+delta when they stay
  // on v's side of the kink: factors e^delta (toward 0) and e^-delta (away),
  // one exponential and its reciprocal.  A point across the kink (delta >
  // |v|) has e^(2|v| - delta) = (e^-delta e^|v|) e^|v|, with aux = e^|v|
  // (|v| <= 650 where the factored form is taken, so aux stays finite and
  // the first product is at most 1).
  // (FP64 integrands only: for an FP32 one the extra registers cost more
  // than the exponentials they save)
  template <class R = Real, std::enable_if_t<std::is_same_v<R, double>, int> = 0>
  __device__ double axis_aux(int, double v) const
  {
    return math::exp_factor(__builtin_fabs(v), Real());
  }
  template <class R = Real, std::enable_if_t<std::is_same_v<R, double>, int> = 0>
  __device__ void factor_pair(int, double v, double delta, double aux, double& fm,
                              double& fp) const
  {
    const double e = math::exp_factor(-delta, Real());
    const double r = math::recip(e);
    const double across = (e * aux) * aux;
    const bool cross = delta > __builtin_fabs(v);
    if (v >= 0.0) {
      fm = cross ? across : r;
      fp = e;
    } else {
      fm = e;
      fp = cross ? across : r;
    }
  }
  // outer(s + d) = outer(s) e^-d, non-negative terms (see Gaussian)
  static constexpr bool exponential = true;
  __device__ double factor(double d) const { return math::exp_factor(-d, Real()); }
  __device__ double centre_value(double s0) const { return math::exp_factor(-s0, Real()); }
  __device__ bool factor_safe(double s0) const { return s0 <= kFactorRange<double>; }
};

// 0 if any x_i > u_i, else exp(sum a_i x_i)
template <class Real, int D>
struct Discontinuous {
  Real a[D];
  Real u[D];
  double c;  // sum a_i u_i
  template <class... X>
  __device__ Real operator()(X... xs) const
  {
    const Real x[D] = {static_cast<Real>(xs)...};
    // "some x_i > u_i" as max_i (x_i - u_i) > 0: one compare instead of a
    // chain of lane masks
    Real m = x[0] - u[0];
    Real s = a[0] * x[0];
#pragma unroll
    for (int i = 1; i < D; ++i) {
      m = math::fmax_(m, x[i] - u[i]);
      s += a[i] * x[i];
    }
    const Real v = math::exp_(s);
    return m > Real(0) ? Real(0) : v;
  }
  // t = x - u (origin u, unit scale): f = [all t_i <= 0] exp(c + sum a_i t_i),
  // additively separable with a validity mask per axis
  __device__ double origin(int i) const { return static_cast<double>(u[i]); }
  __device__ double axis_scale(int) const { return 1.0; }
  template <class... T>
  __device__ Real centered(T... ts) const
  {
    const Real t[D] = {static_cast<Real>(ts)...};
    Real m = t[0];
    Real s = a[0] * t[0];
#pragma unroll
    for (int i = 1; i < D; ++i) {
      m = math::fmax_(m, t[i]);
      s += a[i] * t[i];
    }
    const Real v = math::exp_bounded(static_cast<Real>(c) + s);
    return m > Real(0) ? Real(0) : v;
  }
  static constexpr bool masked = true;
  __device__ bool valid(int, double t) const { return t <= 0.0; }
  __device__ double term(int i, double t) const { return static_cast<double>(a[i]) * t; }
  __device__ double outer(double s) const
  {
    return static_cast<double>(math::exp_bounded(static_cast<Real>(c + s)));
  }
  // outer(s + d) = outer(s) e^d.  The centre's value e^(c + s0) is kept
  // well inside the exponent range; the factors are bounded by e^(a_j h_j)
  // (the kernel sends a region whose sums overflow anyway to the generic
  // loop).
  static constexpr bool exponential = true;
  __device__ double factor(double d) const
  {
    return static_cast<double>(math::exp_bounded(static_cast<Real>(d)));
  }
  // the points t -+ delta change a_j t by -+a_j delta: factors e^(-a_j
  // delta) and its reciprocal
  __device__ void factor_pair(int j, double, double delta, double, double& fm, double& fp) const
  {
    fm = factor(-static_cast<double>(a[j]) * delta);
    fp = math::recip(fm);
  }
  __device__ bool factor_safe(double s0) const
  {
    return __builtin_fabs(c + s0) <= kFactorRange<Real> - (sizeof(Real) == 8 ? 100.0 : 20.0);
  }
};

}  // namespace pagani
