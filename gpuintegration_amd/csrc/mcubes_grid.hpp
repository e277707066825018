// mcubes_grid.hpp -- host-side arithmetic of the m-Cubes driver: run
// constants, the importance grid and its refinement, the stopping rule.
// Compiled by the host compiler without floating-point contraction
// (mcubes_grid.cpp), in the reference's operation order.
#pragma once

#include <cstdint>

#include "pagani.h"

namespace pagani {
namespace mc {

// Constants of one vegas() run (vegasT.cuh:686-716).
struct Setup {
  std::uint64_t ng = 0, npg = 0;
  double ncubes = 0, calls = 0, dv2g = 0, dxg = 0, xnd = 0, xjac = 0;
  double lo[PAGANI_MC_MAX_DIM] = {}, dx[PAGANI_MC_MAX_DIM] = {};
};

void setup(Setup& s, int ndim, double ncall, const double* lows, const double* highs);
// the uniform starting grid of every axis (vegasT.cuh:677-680, :718-723)
void grid_init(int ndim, double* xi);
// refinement after an adjusting iteration (vegasT.cuh:880-925; hist_dims 1:
// vegasT1D.cuh:606-636); hist is smoothed in place
void refine(int ndim, int hist_dims, double* hist, double* xi);
// one iteration into the running weighted mean (vegasT.cuh:847-859,
// :968-980): updates si, schi, swgt and writes the estimate, chi^2 per
// degree of freedom and standard deviation
void combine(double ti, double tsi, int iteration, double& si, double& schi, double& swgt,
             double& estimate, double& chi_sq, double& errorest);
// GetStatus (vegas_utils.cuh:233-254)
int get_status(double estimate, double errorest, int iteration, double epsrel, double epsabs);
// AdjustParams (vegas_utils.cuh:283-298)
bool adjust_params(double& ncall, int& total_iters);

}  // namespace mc
}  // namespace pagani
