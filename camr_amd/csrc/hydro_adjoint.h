// hydro_adjoint.h -- adjoint (reverse mode) of the fused level step.
#pragma once

#include <vector>

#include "hydro_fused.h"

namespace camr {

struct AdjArgs {
    const DevFab *sbar_in;   // per box id: adjoint of S_new (valid cells);
                             // overwritten with the adjoint before clean_state
    const DevFab *sbar_out;  // per box id: adjoint of the raw S_old (accumulated)
    const DevFab *sold_tab;  // per box id: the raw S_old of the step
    const int *owner;
    DomInfo dom;
    HydroConst hc;
    const double *dtp;
    int add_grav;
    int *status;
};

// Launch the adjoint kernels of one step, every box per launch.  The
// engine's scratch must hold the step's primal intermediates (fill + hydro
// of the step just recomputed) and its adjoint scratch must be zero.
int adjoint_step_kernels(const FusedEngine &eng, int parity, const AdjArgs &a, hipStream_t st,
                         Prof *prof);
void zero_doubles(double *p, size_t n, hipStream_t st);
// zero the ranges [start[r], start[r] + len[r]) of p, in one launch for up
// to 64 ranges
void zero_ranges(double *p, const std::vector<size_t> &start, const std::vector<size_t> &len,
                 hipStream_t st);

} // namespace camr
