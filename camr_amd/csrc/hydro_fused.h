// hydro_fused.h -- the face-fused gfx950 engine of the hydro advance.
#pragma once

#include <vector>

#include "hydro_state.h"

namespace camr {

// compact primitive layout: SR..SY of hydro_math.h, then sound speed and
// the small-sound-speed floor of QAUX (QC, QCSML)
enum { SC = NS, SCS = NS + 1, NQC = NS + 2 };

// hydro_ctoprim (Source/Hydro/Hydro_ctoprim.H:11-101) of the conserved
// state v (NVAR values) into the compact primitives at (i,j,k); returns the
// specific internal energy e (needed by hydro_srctoprim).  Cells with
// rho <= 0 are left untouched as in the reference.
__device__ __forceinline__ double prim_compact_point(const double *v, const Arr &q,
                                                     int i, int j, int k,
                                                     const HydroConst &hc, int *status) {
    const double rho = v[URHO];
    if (!(rho > 0.)) return 0.0;
    const double rhoinv = 1.0 / rho;
    const double vx = v[UMX] * rhoinv;
    const double vy = v[UMY] * rhoinv;
    const double vz = v[UMZ] * rhoinv;
    const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    const double ueden = v[UEDEN];
    const double ueint = v[UEINT];
    if (!hc.allow_negative_energy && !(ueden > 0. && ueint > 0.))
        atomicOr(status, 1 << CAMR_ERR_NEG_ENERGY);
    const double e = ((ueden - kineng) / ueden > hc.eta1) ? (ueden - kineng) * rhoinv
                                                         : ueint * rhoinv;
    const double p = (kGamma - 1.0) * rho * e;
    const double cs = sqrt(kGamma * p / rho);
    q(i, j, k, SR) = rho;
    q(i, j, k, SU) = vx;
    q(i, j, k, SV) = vy;
    q(i, j, k, SW) = vz;
    q(i, j, k, SP) = p;
    q(i, j, k, SE) = e * rho;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) q(i, j, k, SY + s) = v[UFS + s] * rhoinv;
    q(i, j, k, SC) = cs;
    q(i, j, k, SCS) = vmax(hc.small_num, hc.small_num * cs);
    return e;
}

// Level-step kernels fusing the pointwise bookkeeping around the hydro
// source (level_fused.hip).
struct FillArgs {
    Box valid;      // box b
    Arr sborder;    // grow(box, 4)
    Arr q;          // compact primitives over grow(box, 4)
    Arr snew;       // valid cells
    const DevFab *sold_tab;  // raw S_old of every box (valid cells)
    const int *owner;
    DomInfo dom;
    HydroConst hc;
    double dt, g;
    int add_grav;
};
void launch_fill_fused(const FillArgs &a, hipStream_t st, int *status, Prof *prof);

struct PostArgs {
    Box valid;
    Arr snew, sold;
    HydroConst hc;
    double dt, g;
    int add_grav;
    unsigned long long *dtmin;
};
void launch_post_fused(const PostArgs &a, hipStream_t st, int *status, Prof *prof);

class FusedEngine {
  public:
    static size_t scratch_doubles(const Box &bx);
    int reserve(const std::vector<Box> &boxes);
    // construct_hydro_source for one box + the S_new Saxpy:
    // snew += dt * dS/dt(sborder); floor / energy checks into *status.
    // have_prim: the compact primitives were already written to prim_view(bx)
    int advance_box(const Box &bx, const Arr &sborder, const Arr &srchyd, int has_src,
                    const Arr &snew, const HydroConst &hc, double dt, hipStream_t st,
                    int *status, Prof *prof, bool have_prim = false);
    // where advance_box expects / writes the compact primitives of bx
    Arr prim_view(const Box &bx) const {
        return dense(scratch_.base(), grow(bx, CAMR_NGROW));
    }

  private:
    Scratch scratch_;
};

} // namespace camr
