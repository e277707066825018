// rule.hpp -- host-side tables of the degree-9 cubature rule.
//
// Restates quad::Rule<T>::Rule9Generate / Init (cuda/pagani/quad/GPUquad/
// Rule.cuh:24-270, 370-510): nine fully symmetric point sets, five weight
// columns (column 0 the degree-9 rule, columns 1-4 null rules), and the
// null-rule scale/norm constants.  The points are laid out for the device:
// sets 0-2 are generated in registers by the evaluation kernel, sets 3-8 go
// into a point-major generator table.
#pragma once

#include <vector>

#include "pagani.h"

namespace pagani {

struct RuleTables {
  int ndim = 0;
  int feval = 0;
  double w[PAGANI_NSETS][PAGANI_NRULES] = {};
  double scale[PAGANI_NSETS][PAGANI_NRULES] = {};
  double norm[PAGANI_NSETS][PAGANI_NRULES] = {};
  double lambda[5] = {};
  int count[PAGANI_NSETS] = {};
  int set_begin[PAGANI_NSETS + 1] = {};
  // generator coordinates of points set_begin[3] .. feval-1, ndim each
  std::vector<double> gtab;
};

// Points per region for ndim (quad.h:68-75).
int feval_points(int ndim);

// Generator coordinate `axis` of rule point p in the library's point order:
// the centre, then +-lambda1 (set 1) and +-lambda2 (set 2) per axis,
// axis-major with + before -, then the table of sets 3-8.  The point of a
// region with centre c and width w is c - g w.
double point_generator(const RuleTables& R, int p, int axis);

// Build the tables; returns false for an unsupported ndim.
bool build_rule(int ndim, RuleTables& out);

}  // namespace pagani
