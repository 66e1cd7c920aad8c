// level_fused.hip -- the pointwise parts of CAMR_advance fused into two
// gfx950 kernels per box.
//
// k_fill_fused, one thread per cell of grow(box, 4):
//   clean_state(S_old)         (CAMR_advance.cpp:83; CAMR.cpp:1253-1262)
//   FillPatch into Sborder     (expand_state, CAMR.cpp:1243-1250)
//   clean_state(Sborder)       incl. the FillBoundary re-sync of interior
//                              ghosts after normalize_species
//   hydro_ctoprim              (CAMR_construct_hydro_source.cpp:108-120)
//   S_new = Sborder + dt * old source (CAMR_advance.cpp:112-118)
// k_post_fused, one thread per valid cell, after S_new += dt dS/dt
// (k_update_post: the update itself and then the same):
//   clean_state(S_new), new-source Saxpy, clean_state, post_timestep's
//   clean_state (CAMR_advance.cpp:146-154; CAMR.cpp:698-699), then the
//   CFL estimate of the next computeNewDt (Timestep.H:18-61).
//
// Each launch covers a set of boxes (all boxes of the level, or one); ghost
// cells that no box of the level owns (coarse-fine ghosts of a refined
// level) take their values from the level's exterior array when one is set.
//
// Both are exact compositions of the reference's passes as long as no
// density floor fires (enforce_min_density is the only non-pointwise piece;
// normalize_species and the energy reset never change rho, rho*u):
// k_fill_fused flags rho < small_dens in raw S_old, k_update_c flags it in
// S_new, and the level driver then recomputes the step with the unfused
// sequence from the untouched raw S_old.
#include "hydro_fused.h"
#include "state_math.h"

namespace camr {

namespace {

__global__ void __launch_bounds__(kBlock)
k_fill_fused(Batch B, FillArgs a, int *status) {
    BATCH_POINT(ST_G4)
    int m[3];
    unsigned odd;
    const bool phys = map_cell(a.dom, i, j, k, m, &odd);
    const int ob = owner_of(a.dom, a.owner, m);
    double v[NVAR];
    double rho_old = 0.0, rhow_old = 0.0;
    const bool valid = E.bx.contains(i, j, k);
    if (ob >= 0) {
        const Arr s = a.sold_tab[ob].a;
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] = s(m[0], m[1], m[2], n);
        if (valid) floor_detect(v[URHO], a.hc.small_dens, status);
        rho_old = v[URHO];
        rhow_old = v[UMZ];
        // clean_state(S_old) on the source cell
        normalize_species(v);
        reset_e_and_T(v, a.hc);
    } else if (a.has_ext) {
        // exterior data (the FillPatch values of a coarse-fine ghost)
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] = a.ext(m[0], m[1], m[2], n);
    } else {
        atomicOr(status, 1 << CAMR_ERR_SHAPE);
        return;
    }
    // FillPatch copy with the reflect signs of physical boundaries
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
        if ((odd >> n) & 1u) v[n] = -v[n];
    // clean_state(Sborder): normalize_species on valid cells, propagated to
    // interior / periodic ghosts by FillBoundary; computeTemp everywhere
    if (!phys && ob >= 0) normalize_species(v);
    reset_e_and_T(v, a.hc);
    const unsigned v0 = L.off(i, j, k);
    const CView u = cview(L, slot_ptr<S_U>(E), NVAR);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) u.st(v0, n, v[n]);
    double o[NQC], e;
    if (prim_compact(v, o, e, a.hc, status)) {
        const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
#pragma unroll
        for (int n = 0; n < NQC; ++n) q.st(v0, n, o[n]);
    }
    if (valid) {
        const double dt = *a.dtp;
        const Arr snew = E.snew;
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            double x = v[n];
            if (a.add_grav) {
                double src = 0.0;
                if (n == UMZ) src = a.g * rho_old;
                else if (n == UEDEN) src = a.g * rhow_old;
                x += dt * src;
            }
            snew(i, j, k, n) = x;
        }
    }
}

__device__ __forceinline__ void clean_point(double *v, const HydroConst &hc) {
    normalize_species(v);
    reset_e_and_T(v, hc);
}

// The post pass of one valid cell on v (S_new after the update): the
// cleaning passes and new-source Saxpy; returns the cell's CFL estimate.
__device__ __forceinline__ double post_values(const BoxEnt &E, int i, int j, int k,
                                              const PostArgs &a, double *v) {
    const Arr sold = E.sold;
    const double dt = *a.dtp;
    clean_point(v, a.hc);
    if (a.add_grav) {
        const double s_umz = 0.5 * a.g * (v[URHO] - sold(i, j, k, URHO));
        const double s_ued = 0.5 * a.g * (v[UMZ] - sold(i, j, k, UMZ));
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            const double src = (n == UMZ) ? s_umz : ((n == UEDEN) ? s_ued : 0.0);
            v[n] += dt * src;
        }
        clean_point(v, a.hc);
    }
    // post_timestep's clean_state (CAMR.cpp:698-699), unless a finer level's
    // steps come first (then the caller cleans after reflux and avgDown)
    if (!a.defer_post) clean_point(v, a.hc);
    // estTimeStep on the new state
    const double rho = v[URHO];
    const double rhoInv = 1.0 / rho;
    const double eint = v[UEINT] / rho;
    const double pres = (kGamma - 1.0) * rho * eint;
    const double c = sqrt(kGamma * pres / rho);
    double dtv = 1.7976931348623157e308;
    dtv = vmin(dtv, a.hc.dx[0] / (c + fabs(v[UMX] * rhoInv)));
    dtv = vmin(dtv, a.hc.dx[1] / (c + fabs(v[UMY] * rhoInv)));
    dtv = vmin(dtv, a.hc.dx[2] / (c + fabs(v[UMZ] * rhoInv)));
    return dtv;
}

__global__ void __launch_bounds__(kBlock)
k_post_fused(Batch B, PostArgs a, int *status) {
    if (B.halt && *B.halt) return;  // uniform: the whole block leaves
    int lblk;
    const BoxEnt &E = B.tab[batch_box(B, lblk)];
    int i, j, k;
    double dtv = 1.7976931348623157e308;
    // every thread reaches the block reduction below
    if (batch_point(B, E.bx, lblk, i, j, k)) {
        const Arr snew = E.snew;
        double v[NVAR];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] = snew(i, j, k, n);
        dtv = post_values(E, i, j, k, a, v);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) snew(i, j, k, n) = v[n];
    }
    block_min_commit(dtv, a.dtmin, status);
}

// k_update_c followed by k_post_fused, with S_new kept in registers in
// between (the same arithmetic: S_new is rounded once by the update either
// way).
#ifdef CAMR_UPD_MINW
#define UPD_BOUNDS __launch_bounds__(kBlock, CAMR_UPD_MINW)
#else
#define UPD_BOUNDS __launch_bounds__(kBlock)
#endif
__global__ void UPD_BOUNDS
k_update_post(Batch B, HydroConst hc, const double *dtp, PostArgs a, int *status) {
    if (B.halt && *B.halt) return;  // uniform: the whole block leaves
    int lblk;
    const BoxEnt &E = B.tab[batch_box(B, lblk)];
    int i, j, k;
    double dtv = 1.7976931348623157e308;
    if (batch_point(B, E.bx, lblk, i, j, k)) {
        const Arr snew = E.snew;
        double v[NVAR];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] = snew(i, j, k, n);
        update_values(E, i, j, k, hc, *dtp, v);
        const double rho = v[URHO];
        if (rho == 0.) atomicOr(status, 1 << CAMR_ERR_ZERO_DENSITY);
        else if (rho < hc.small_dens) atomicOr(status, kFloorFlag);
        dtv = post_values(E, i, j, k, a, v);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) snew(i, j, k, n) = v[n];
    }
    block_min_commit(dtv, a.dtmin, status);
}

} // namespace

void launch_fill_fused(const LaunchCfg &c, const FillArgs &a, hipStream_t st, int *status) {
    if (c.grid == 0) return;
    hipLaunchKernelGGL(k_fill_fused, dim3(c.grid), dim3(c.block), 0, st, c.B, a, status);
}

void launch_post_fused(const LaunchCfg &c, const PostArgs &a, hipStream_t st, int *status) {
    if (c.grid == 0) return;
    hipLaunchKernelGGL(k_post_fused, dim3(c.grid), dim3(c.block), 0, st, c.B, a, status);
}

void launch_update_post_fused(const LaunchCfg &c, const HydroConst &hc, const double *dtp,
                              const PostArgs &a, hipStream_t st, int *status) {
    if (c.grid == 0) return;
    hipLaunchKernelGGL(k_update_post, dim3(c.grid), dim3(c.block), 0, st, c.B, hc, dtp, a,
                       status);
}

} // namespace camr
