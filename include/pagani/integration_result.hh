// integration_result.hh -- numint::integration_result, the value returned by
// every integrate call (common/integration_result.hh:11-32 of the reference:
// same members, defaults and stream format).
#pragma once

#include <cstddef>
#include <ostream>

namespace numint {

struct integration_result {
  double estimate = 0.;
  double errorest = 0.;
  std::size_t neval = 0;
  std::size_t nregions = 0;
  std::size_t nFinishedRegions = 0;
  int status = -1;
  int lastPhase = -1;
  double chi_sq = 0.;
  std::size_t iters = 0;
};

inline std::ostream& operator<<(std::ostream& os, integration_result const& res)
{
  os << res.estimate << "," << res.errorest << "," << res.nregions << "," << res.chi_sq << ","
     << res.status;
  return os;
}

}  // namespace numint
