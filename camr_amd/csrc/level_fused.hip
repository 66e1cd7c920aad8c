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
// k_post_fused, one thread per valid cell, after S_new += dt dS/dt:
//   clean_state(S_new), new-source Saxpy, clean_state, post_timestep's
//   clean_state (CAMR_advance.cpp:146-154; CAMR.cpp:698-699), then the
//   CFL estimate of the next computeNewDt (Timestep.H:18-61).
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
k_fill_fused(Range r, FillArgs a, int *status) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    int m[3];
    unsigned odd;
    const bool phys = map_cell(a.dom, i, j, k, m, &odd);
    const int ob = owner_of(a.dom, a.owner, m);
    if (ob < 0) { atomicOr(status, 1 << CAMR_ERR_SHAPE); return; }
    const Arr s = a.sold_tab[ob].a;
    double v[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) v[n] = s(m[0], m[1], m[2], n);
    const bool valid = a.valid.contains(i, j, k);
    if (valid) floor_detect(v[URHO], a.hc.small_dens, status);
    const double rho_old = v[URHO], rhow_old = v[UMZ];
    // clean_state(S_old) on the source cell
    normalize_species(v);
    reset_e_and_T(v, a.hc);
    // FillPatch copy with the reflect signs of physical boundaries
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
        if ((odd >> n) & 1u) v[n] = -v[n];
    // clean_state(Sborder): normalize_species on valid cells, propagated to
    // interior / periodic ghosts by FillBoundary; computeTemp everywhere
    if (!phys) normalize_species(v);
    reset_e_and_T(v, a.hc);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) a.sborder(i, j, k, n) = v[n];
    prim_compact_point(v, a.q, i, j, k, a.hc, status);
    if (valid) {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            double x = v[n];
            if (a.add_grav) {
                double src = 0.0;
                if (n == UMZ) src = a.g * rho_old;
                else if (n == UEDEN) src = a.g * rhow_old;
                x += a.dt * src;
            }
            a.snew(i, j, k, n) = x;
        }
    }
}

__device__ __forceinline__ void clean_point(double *v, const HydroConst &hc) {
    normalize_species(v);
    reset_e_and_T(v, hc);
}

__global__ void __launch_bounds__(kBlock)
k_post_fused(Range r, PostArgs a, int *status) {
    int i, j, k;
    double dtv = 1.7976931348623157e308;
    if (range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) {
        double v[NVAR];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] = a.snew(i, j, k, n);
        clean_point(v, a.hc);
        if (a.add_grav) {
            const double s_umz = 0.5 * a.g * (v[URHO] - a.sold(i, j, k, URHO));
            const double s_ued = 0.5 * a.g * (v[UMZ] - a.sold(i, j, k, UMZ));
#pragma unroll
            for (int n = 0; n < NVAR; ++n) {
                const double src = (n == UMZ) ? s_umz : ((n == UEDEN) ? s_ued : 0.0);
                v[n] += a.dt * src;
            }
            clean_point(v, a.hc);
        }
        clean_point(v, a.hc);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) a.snew(i, j, k, n) = v[n];
        // estTimeStep on the new state
        const double rho = v[URHO];
        const double rhoInv = 1.0 / rho;
        const double eint = v[UEINT] / rho;
        const double pres = (kGamma - 1.0) * rho * eint;
        const double c = sqrt(kGamma * pres / rho);
        dtv = vmin(dtv, a.hc.dx[0] / (c + fabs(v[UMX] * rhoInv)));
        dtv = vmin(dtv, a.hc.dx[1] / (c + fabs(v[UMY] * rhoInv)));
        dtv = vmin(dtv, a.hc.dx[2] / (c + fabs(v[UMZ] * rhoInv)));
    }
    block_min_commit(dtv, a.dtmin, status);
}

} // namespace

void launch_fill_fused(const FillArgs &a, hipStream_t st, int *status, Prof *prof) {
    const Box G = grow(a.valid, CAMR_NGROW);
    Range r = range_of(G);
    Region reg(prof, "fill_clean_prim", st);
    hipLaunchKernelGGL(k_fill_fused, dim3(grid_for(r.n)), dim3(kBlock), 0, st, r, a, status);
}

void launch_post_fused(const PostArgs &a, hipStream_t st, int *status, Prof *prof) {
    Range r = range_of(a.valid);
    Region reg(prof, "clean_estdt", st);
    hipLaunchKernelGGL(k_post_fused, dim3(grid_for(r.n)), dim3(kBlock), 0, st, r, a, status);
}

} // namespace camr
