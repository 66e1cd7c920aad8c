// hydro_eb_adjoint.h -- reverse-mode adjoint of the embedded-boundary level
// step (PLM, FluxRedist): the batched cut-cell stages in reverse order and
// the level's update, clean_state and FillPatch.
#pragma once

#include "hydro_eb.h"
#include "hydro_state.h"

namespace camr {

// The components of the state whose adjoints the EB kernels write
// (URHO..UEINT); the others stay zero.
constexpr int kEbAdjComps = UEINT + 1;

// One box of the adjoint batch: the forward arrays of the recomputed step
// (every pair's and direction's states kept) and their adjoints in arrays
// of the same layouts.  In A: q / qaux hold the adjoints of the primitives,
// uin that of Sborder (grow(bx, 6)), dsdt that of S_new after the update
// and before the cleaning (bx); the geometry and the box fields repeat F's.
struct EbAdjEnt {
    EbBox F, A;
};

// The adjoint stages of hydro_umdrv_eb over the n boxes of tab (device) /
// htab (host copy, for the grid sizes), from A.dsdt (dL/dS_new before the
// cleaning) down to A.uin (dL/dSborder).  Every adjoint array must be zero
// on entry except A.dsdt and the valid cells of A.uin.
// cells (device, ncells): the cut cells of grow(bx, 2) of the table's boxes
// (CutCell), over which the cut-cell divergence and the redistribution's
// source side then run.
int eb_adjoint_batch(const EbAdjEnt *tab, const EbAdjEnt *htab, int n, const HydroConst &hc,
                     const double *dtp, hipStream_t st, Prof *prof = nullptr,
                     const CutCell *cells = nullptr, int ncells = 0);

// Level part, over every box of the tables (grid.y = box):
// the update S_new = Sborder + dt dS/dt, the clean_state passes after it
// (two, three with add_grav), from sbar_in = dL/dS_{n+1}: ub = dL/dS_new
// before the cleaning (valid cells), sbb = the same on the valid cells of
// the Sborder adjoint.  Sets kFloorFlag where the step takes a density
// floor and the adjoint error bit where a reset takes a branch the adjoint
// does not model.
__global__ void kb_ebadj_post(const DevFab *sb, const DevFab *dsdt, const DevFab *sbar_in,
                              const DevFab *ub, const DevFab *sbb, const double *dtp,
                              HydroConst hc, int add_grav, int *status, EbFlags eb);
// FillPatch (fill_raw, resync, the energy reset of Sborder) in reverse: each
// Sborder cell's adjoint to the valid cell it was copied from (sign flips of
// the reflecting walls) in gout, a table of the state layout (atomics).
// sclean: S_old after its clean_state.
__global__ void kb_ebadj_fill(const DevFab *sbb, int g, const DevFab *sclean, const DevFab *gout,
                              const int *owner, DomInfo D, HydroConst hc, int *status);
// clean_state(S_old) in reverse, in place on gout: dL/dS_old.
__global__ void kb_ebadj_old(const DevFab *gout, const DevFab *sclean, HydroConst hc,
                             int *status);
// Seed: gout = d(w . F)/dS for the force F on the embedded body from the
// cut cells of the state S (the functional of eb_force_box).
// geoms / kinds: each box's cut-cell data and kind (1: cut cells).
__global__ void kb_ebadj_force(const DevFab *S, const DevFab *gout, const EbGeom *geoms,
                               const int *kinds, double w0, double w1, double w2, double dx2,
                               HydroConst hc);

} // namespace camr
