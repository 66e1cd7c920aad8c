// hydro_kernels.h -- shared declarations of the gfx950 hydro kernels.
#pragma once

#include "array.h"
#include "hydro_math.h"
#include "profiler.h"

namespace camr {

// Arguments of one hydro_umdrv call (Source/Hydro/Hydro.H:11-35).
struct UmdrvArgs {
    Box bx;        // valid region
    Box qbox;      // box of q / qaux / srcq (flatten's q.contains test)
    HydroConst hc;
    const double *dtp;  // device: the step's dt (read by the kernels)
    Arr uin, dsdt, q, qaux, srcq;
    int has_src;   // 0: identically +0 primitive sources
    Arr flx[3];
    Arr area[3];
    int has_area;
    double area_c[3];
    Arr vol;
    int has_vol;
    double vol_c;
    // 1: dsdt holds nothing yet; hydro_update's += onto its zero fill is
    // written as 0 + the update (the same values, one pass fewer)
    int dsdt_zero = 0;
};

HydroConst make_const(const camr_hydro_params &p);

size_t umdrv_staged_scratch_doubles(const Box &bx);
int umdrv_staged(const UmdrvArgs &A, double *scratch, hipStream_t st,
                 Prof *prof = nullptr);
// hydro_umdrv with do_mol (MOL_umeth): q / qaux must cover grow(bx, 3);
// srcq, has_src and dtp are not read.
size_t umdrv_mol_scratch_doubles(const Box &bx);
int umdrv_mol(const UmdrvArgs &A, double *scratch, hipStream_t st, Prof *prof = nullptr);

// ---- hydro_divu (Source/Hydro/Hydro_utils_K.H:12-169), node (i,j,k) -------
template <class QA>
__device__ __forceinline__ double divu_node(const QA &q, int i, int j, int k,
                                            const HydroConst &hc) {
    const double dx = hc.dx[0], dy = hc.dx[1], dz = hc.dx[2];
    const int NSW = BC_NOSLIPWALL;
#define CQ(a, b, c, n) q(a, b, c, n)
    double ux = 0.25 * (CQ(i, j, k, QU) - CQ(i - 1, j, k, QU) + CQ(i, j, k - 1, QU) -
                        CQ(i - 1, j, k - 1, QU) + CQ(i, j - 1, k, QU) - CQ(i - 1, j - 1, k, QU) +
                        CQ(i, j - 1, k - 1, QU) - CQ(i - 1, j - 1, k - 1, QU)) / dx;
    double vy = 0.25 * (CQ(i, j, k, QV) - CQ(i, j - 1, k, QV) + CQ(i, j, k - 1, QV) -
                        CQ(i, j - 1, k - 1, QV) + CQ(i - 1, j, k, QV) - CQ(i - 1, j - 1, k, QV) +
                        CQ(i - 1, j, k - 1, QV) - CQ(i - 1, j - 1, k - 1, QV)) / dy;
    double wz = 0.25 * (CQ(i, j, k, QW) - CQ(i, j, k - 1, QW) + CQ(i - 1, j, k, QW) -
                        CQ(i - 1, j, k - 1, QW) + CQ(i, j - 1, k, QW) - CQ(i, j - 1, k - 1, QW) +
                        CQ(i - 1, j - 1, k, QW) - CQ(i - 1, j - 1, k - 1, QW)) / dz;
    const bool xl = (i == hc.domlo[0] && hc.bclo[0] == NSW);
    const bool xh = (i == hc.domhi[0] + 1 && hc.bchi[0] == NSW);
    const bool yl = (j == hc.domlo[1] && hc.bclo[1] == NSW);
    const bool yh = (j == hc.domhi[1] + 1 && hc.bchi[1] == NSW);
    const bool zl = (k == hc.domlo[2] && hc.bclo[2] == NSW);
    const bool zh = (k == hc.domhi[2] + 1 && hc.bchi[2] == NSW);
    if (xl | xh | yl | yh | zl | zh) {
        if (xl) {
            vy = 0.5 * (CQ(i, j, k, QV) - CQ(i, j - 1, k, QV) + CQ(i, j, k - 1, QV) - CQ(i, j - 1, k - 1, QV)) / dy;
            wz = 0.5 * (CQ(i, j, k, QW) - CQ(i, j, k - 1, QW) + CQ(i, j - 1, k, QW) - CQ(i, j - 1, k - 1, QW)) / dz;
        } else if (xh) {
            vy = 0.5 * (CQ(i - 1, j, k, QV) - CQ(i - 1, j - 1, k, QV) + CQ(i - 1, j, k - 1, QV) - CQ(i - 1, j - 1, k - 1, QV)) / dy;
            wz = 0.5 * (CQ(i - 1, j, k, QW) - CQ(i - 1, j, k - 1, QW) + CQ(i - 1, j - 1, k, QW) - CQ(i - 1, j - 1, k - 1, QW)) / dz;
        }
        if (yl) {
            ux = 0.5 * (CQ(i, j, k, QU) - CQ(i - 1, j, k, QU) + CQ(i, j, k - 1, QU) - CQ(i - 1, j, k - 1, QU)) / dx;
            wz = 0.5 * (CQ(i, j, k, QW) - CQ(i, j, k - 1, QW) + CQ(i - 1, j, k, QW) - CQ(i - 1, j, k - 1, QW)) / dz;
        } else if (yh) {
            ux = 0.5 * (CQ(i, j - 1, k, QU) - CQ(i - 1, j - 1, k, QU) + CQ(i, j - 1, k - 1, QU) - CQ(i - 1, j - 1, k - 1, QU)) / dx;
            wz = 0.5 * (CQ(i, j - 1, k, QW) - CQ(i, j - 1, k - 1, QW) + CQ(i - 1, j - 1, k, QW) - CQ(i - 1, j - 1, k - 1, QW)) / dz;
        }
        if (zl) {
            ux = 0.5 * (CQ(i, j, k, QU) - CQ(i - 1, j, k, QU) + CQ(i, j - 1, k, QU) - CQ(i - 1, j - 1, k, QU)) / dx;
            vy = 0.5 * (CQ(i, j, k, QV) - CQ(i, j - 1, k, QV) + CQ(i - 1, j, k, QV) - CQ(i - 1, j - 1, k, QV)) / dy;
        } else if (zh) {
            ux = 0.5 * (CQ(i, j, k - 1, QU) - CQ(i - 1, j, k - 1, QU) + CQ(i, j - 1, k - 1, QU) - CQ(i - 1, j - 1, k - 1, QU)) / dx;
            vy = 0.5 * (CQ(i, j, k - 1, QV) - CQ(i, j - 1, k - 1, QV) + CQ(i - 1, j, k - 1, QV) - CQ(i - 1, j - 1, k - 1, QV)) / dy;
        }
        if (yl && zl) ux = (CQ(i, j, k, QU) - CQ(i - 1, j, k, QU)) / dx;
        else if (yh && zl) ux = (CQ(i, j - 1, k, QU) - CQ(i - 1, j - 1, k, QU)) / dx;
        else if (yl && zh) ux = (CQ(i, j, k - 1, QU) - CQ(i - 1, j, k - 1, QU)) / dx;
        else if (yh && zh) ux = (CQ(i, j - 1, k - 1, QU) - CQ(i - 1, j - 1, k - 1, QU)) / dx;
        if (xl && zl) vy = (CQ(i, j, k, QV) - CQ(i, j - 1, k, QV)) / dy;
        else if (xh && zl) vy = (CQ(i - 1, j, k, QV) - CQ(i - 1, j - 1, k, QV)) / dy;
        else if (xl && zh) vy = (CQ(i, j, k - 1, QV) - CQ(i, j - 1, k - 1, QV)) / dy;
        else if (xh && zh) vy = (CQ(i - 1, j, k - 1, QV) - CQ(i - 1, j - 1, k - 1, QV)) / dy;
        if (xl && yl) wz = (CQ(i, j, k, QW) - CQ(i, j, k - 1, QW)) / dz;
        else if (xh && yl) wz = (CQ(i - 1, j, k, QW) - CQ(i - 1, j, k - 1, QW)) / dz;
        else if (xl && yh) wz = (CQ(i, j - 1, k, QW) - CQ(i, j - 1, k - 1, QW)) / dz;
        else if (xh && yh) wz = (CQ(i - 1, j - 1, k, QW) - CQ(i - 1, j - 1, k - 1, QW)) / dz;
    }
#undef CQ
    return ux + vy + wz;
}

// ---- adjust_fluxes for one face: hydro_artif_visc (hydro_artif_visc.H:7-64),
// hydro_norm_spec_flx (Hydro_utils_K.H:207-219), hydro_ext_flx (:224-234).
// f holds the NVAR fluxes of the face, u the conserved state (Sborder).
// open == false (a closed cut-cell face, adjust_fluxes_eb): the species
// normalisation and the area factor are skipped.
template <int D, class DA, class UA>
CAMR_HD void adjust_face(double *f, const DA &divu,
                                            const UA &u, int i, int j, int k,
                                            double area, const HydroConst &hc,
                                            bool open = true) {
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    constexpr int L1 = (D == 0) ? 1 : 0;
    constexpr int L2 = (D == 2) ? 1 : 2;
    constexpr int ax = (L1 == 0), ay = (L1 == 1), az = (L1 == 2);
    constexpr int bx = (L2 == 0), by = (L2 == 1), bz = (L2 == 2);
    const double div =
        hc.difmag * vmin(0.0, 0.25 * (divu(i, j, k, 0) + divu(i + ax, j + ay, k + az, 0) +
                                      divu(i + bx, j + by, k + bz, 0) +
                                      divu(i + ax + bx, j + ay + by, k + az + bz, 0)));
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    const bool at_bndry = (idx == hc.domlo[D] && hc.bclo[D] == BC_NOSLIPWALL) ||
                          (idx == hc.domhi[D] + 1 && hc.bchi[D] == BC_NOSLIPWALL);
    const double dxd = hc.dx[D];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        if (n == UTEMP) continue;
        const bool tang = (n == UMX || n == UMY || n == UMZ) && (n != UMX + D);
        if (!(at_bndry && tang))
            f[n] += dxd * div * (u(i, j, k, n) - u(i - ex, j - ey, k - ez, n));
    }
    f[UTEMP] = 0.0;
    if (!open) return;
    double summ = 0.0;
    summ += f[UFS];
    summ += f[UFS + 1];
    summ += f[UFS + 2];
    const double fac = (summ != 0.0) ? f[URHO] / summ : 1.0;
    f[UFS] *= fac;
    f[UFS + 1] *= fac;
    f[UFS + 2] *= fac;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) f[n] *= area;
}

} // namespace camr
