// Sine and cosine for moderate arguments, used by the evaluation kernel's
// factored sums for cosine integrands.  Inlined many times per kernel, the
// device library's versions carry their large-argument reduction into every
// copy; these need |x| <= 2^20, which the kernel checks per region (a region
// outside it takes the generic loop).
#ifndef PAGANI_TRIG_HPP
#define PAGANI_TRIG_HPP

#include <hip/hip_runtime.h>

namespace pagani {
namespace math {

// Largest |x| sincos_bounded accepts.
constexpr double kTrigRange = 1048576.0;  // 2^20

// x = n pi/2 + r with n = rint(x 2/pi) and |r| <= pi/4, r from a two-part
// pi/2 with FMAs (Cody-Waite): x - n Pio2Hi is exact for |n| < 2^20, and the
// part of pi/2 left out of Pio2Hi + Pio2Lo is below 2^-107.  With t = r^2,
// sin r = r + r t S(t) and cos r = 1 - t/2 + t^2 C(t); S and C interpolate
// (sin(sqrt t) - sqrt t) / t^(3/2) and (cos(sqrt t) - 1 + t/2) / t^2 at the
// 6 Chebyshev nodes of [0, (pi/4)^2] (computed from their series in
// extended precision, then rounded to double; the check that recomputes
// them is in the tests): polynomial errors 0.19 ulp (sin) and 0.01 ulp (cos).
// n mod 4 picks the quadrant.
__device__ __forceinline__ void sincos_bounded(double x, double& s, double& c)
{
  constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
  constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
  constexpr double kPio2Lo = 0x1.1a62633145c07p-54;
  constexpr double ks[6] = {-0x1.5555555555555p-3, 0x1.1111111110bb2p-7,
                            -0x1.a01a019e83aaep-13, 0x1.71de37968a100p-19,
                            -0x1.ae600b02b6262p-26, 0x1.5e0b19f8b1451p-33};
  constexpr double kc[6] = {0x1.5555555555555p-5,  -0x1.6c16c16c16967p-10,
                            0x1.a01a019f4eb01p-16, -0x1.27e4fa17da09ep-22,
                            0x1.1eeb68e93b64cp-29, -0x1.907da367a37cbp-37};
  const double n = __builtin_rint(x * kTwoOverPi);
  double r = __builtin_fma(-n, kPio2Hi, x);
  r = __builtin_fma(-n, kPio2Lo, r);
  const double t = r * r;
  double ps = ks[5], pc = kc[5];
#pragma unroll
  for (int k = 4; k >= 0; --k) {
    ps = __builtin_fma(ps, t, ks[k]);
    pc = __builtin_fma(pc, t, kc[k]);
  }
  const double sr = __builtin_fma(r * t, ps, r);
  const double cr = __builtin_fma(t * t, pc, __builtin_fma(-0.5, t, 1.0));
  const int q = static_cast<int>(n) & 3;
  const double ss = (q & 1) ? cr : sr;
  const double cc = (q & 1) ? sr : cr;
  s = (q & 2) ? -ss : ss;
  c = ((q + 1) & 2) ? -cc : cc;
}

__device__ __forceinline__ double cos_bounded(double x)
{
  double s, c;
  sincos_bounded(x, s, c);
  return c;
}

}  // namespace math
}  // namespace pagani

#endif
