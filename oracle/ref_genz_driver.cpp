// extern "C" driver around the reference's closed-form Genz integrals
// (cuda/pagani/demos/compute_genz_integrals.cuh:241-340), which are plain
// host C++.  Built only into oracle/_ref/ by `make ref`; test infrastructure.
#include <array>
#include <cmath>
#include <cstddef>
#include "cuda/pagani/demos/compute_genz_integrals.cuh"

namespace {
template <size_t d>
void load(const double* a, const double* b, std::array<double, d>& A,
          std::array<double, d>& B)
{
  for (size_t i = 0; i < d; ++i) {
    A[i] = a[i];
    B[i] = b[i];
  }
}

template <size_t d>
double call(int which, const double* a, const double* b)
{
  std::array<double, d> A, B;
  load<d>(a, b, A, B);
  switch (which) {
  case 2: return compute_product_peak<d>(A, B);
  case 4: return compute_gaussian<d>(A, B);
  case 5: return compute_c_zero<d>(A, B);
  case 6: return compute_discontinuous<d>(A, B);
  default: return NAN;
  }
}
} // namespace

extern "C" double ref_genz_closed_form(int family, int d, const double* a,
                                       const double* b)
{
  switch (d) {
  case 2: return call<2>(family, a, b);
  case 3: return call<3>(family, a, b);
  case 4: return call<4>(family, a, b);
  case 5: return call<5>(family, a, b);
  case 6: return call<6>(family, a, b);
  case 7: return call<7>(family, a, b);
  case 8: return call<8>(family, a, b);
  default: return NAN;
  }
}

extern "C" double ref_genz_phi(double z) { return genz_phi(z); }
