// Helpers shared by the C-ABI front ends (ps_solver.hip, ps_server.hip):
// argument validation, setup inputs and statistics.
#pragma once

#include <string>

#include "../../include/suli_ps.h"
#include "ps_engine.h"
#include "ps_internal.h"

namespace sps {

// 0, or SULI_PS_EINVAL with the reason in err.
// Sizes, spacings and dt only: what the hydrostatic column sweeps need.
int validate_shape(const suli_ps_grid *g, std::string &err);
int validate_grid(const suli_ps_grid *g, std::string &err);
int validate_problem(const suli_ps_grid *g, const double *u, const double *v, const double *w, const double *rho,
                     const double *ox, const double *oy, const double *oz, const double *dz, const double *dzz,
                     const double *prd1, std::string &err);
// Setup inputs of graddin's problem (pointers as given: host or device).
SetupIn make_setup(const suli_ps_grid *g, const double *u, const double *v, const double *w, const double *rho,
                   const double *prd1, const double *ox, const double *oy, const double *oz, const double *dz,
                   const double *dzz, const uint8_t *solid = nullptr);
SolveOpts solve_opts(const suli_ps_opts &o);
// Statistics of a finished problem from its final scalars.
void fill_stats(const Slot &sl, suli_ps_stats *st);

}  // namespace sps
