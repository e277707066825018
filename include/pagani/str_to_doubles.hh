// str_to_doubles.hh -- whitespace-separated numbers (C99 hexfloat or
// decimal) from one line of text.
//
// Replaces str_to_doubles (common/cuda/str_to_doubles.hh:9-21), the reader
// of the reference's interpolation-table dumps (three hexfloat lines per
// Interp2D: x knots, y knots, values).  Host-only C++, no HIP dependency.
#pragma once

#include <cerrno>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace pagani {

inline std::vector<double> str_to_doubles(const std::string& txt)
{
  std::vector<double> vals;
  const char* p = txt.c_str();
  while (*p) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') ++p;
    if (!*p) break;
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(p, &end);
    // a token must start with a number; like std::stod, characters after the
    // number are not part of it, so they end the token
    if (end == p) throw std::invalid_argument("str_to_doubles: not a number");
    vals.push_back(v);
    p = end;
    while (*p && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' ||
                   *p == '\f'))
      ++p;
  }
  return vals;
}

}  // namespace pagani

using pagani::str_to_doubles;
