// state_math.h -- pointwise device arithmetic of the level step: the CFL
// estimate, clean_state pieces and the single-level ghost-cell map.
#pragma once

#include "hydro_state.h"

namespace camr {

// ---------------------------------------------------------------- estdt
// CAMR_estdt_hydro (Source/Utils/Timestep.H:18-61) for one cell.  The min
// over cells is order independent; non-negative doubles order like their
// bit patterns, so block minima are combined with a 64-bit atomicMin.
// the same on a register copy of the cell
__device__ __forceinline__ double estdt_vals(const double *v, double dx, double dy, double dz) {
    const double rho = v[URHO];
    const double rhoInv = 1.0 / rho;
    const double eint = v[UEINT] / rho;
    const double pres = (kGamma - 1.0) * rho * eint;
    const double c = sqrt(kGamma * pres / rho);
    double dt = 1.7976931348623157e308;
    const double ux = v[UMX] * rhoInv;
    dt = vmin(dt, dx / (c + fabs(ux)));
    const double uy = v[UMY] * rhoInv;
    dt = vmin(dt, dy / (c + fabs(uy)));
    const double uz = v[UMZ] * rhoInv;
    dt = vmin(dt, dz / (c + fabs(uz)));
    return dt;
}

__device__ __forceinline__ double estdt_point(const Arr &u, int i, int j, int k,
                                              double dx, double dy, double dz) {
    double w[NVAR] = {};
    w[URHO] = u(i, j, k, URHO);
    w[UMX] = u(i, j, k, UMX);
    w[UMY] = u(i, j, k, UMY);
    w[UMZ] = u(i, j, k, UMZ);
    w[UEINT] = u(i, j, k, UEINT);
    return estdt_vals(w, dx, dy, dz);
}

__device__ __forceinline__ void block_min_commit(double v, unsigned long long *out,
                                                 int *status) {
    __shared__ double red[kBlock / 64];
    // non-positive (or negative-signed) dt is an error (Timestep.H:59)
    if (!(v > 0.0) && v == v) atomicOr(status, 1 << CAMR_ERR_DT);
    if (!(v == v)) v = 1.7976931348623157e308;  // std::min never selects NaN
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = vmin(v, o);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = vmin(m, red[w]);
        if (m > 0.0) atomicMin(out, (unsigned long long)__double_as_longlong(m));
    }
}


// --------------------------------------------------------- clean_state
// CAMR_reset_internal_e (Source/Utils/CAMR_reset_internal_e.H:13-129) then
// the temperature of computeTemp (CAMR.cpp:1141-1185), on a register copy
// of one cell.
CAMR_HD void reset_e_and_T(double *s, const HydroConst &hc) {
    const double rho = s[URHO];
    // EB build: covered cells (rho = 0) keep their zeros, T = 0
    // (CAMR_reset_internal_e.H:36-40, CAMR.cpp:1166-1169)
    if (hc.eb && !(rho > 0.)) { s[UTEMP] = 0.0; return; }
    const double rhoInv = 1.0 / rho;
    const double Up = s[UMX] * rhoInv;
    const double Vp = s[UMY] * rhoInv;
    const double Wp = s[UMZ] * rhoInv;
    const double ke = 0.5 * (Up * Up + Vp * Vp + Wp * Wp);
    const double eden = s[UEDEN] * rhoInv;
    const double y0 = s[UFS] * rhoInv, y1 = s[UFS + 1] * rhoInv, y2 = s[UFS + 2] * rhoInv;
    const double small_T = hc.small_temp;
    if (hc.allow_small_energy == 0) {
        const double small_e = eos_E_from_T(small_T, y0, y1, y2);
        if (eden < small_e) {
            if (s[UEINT] * rhoInv < small_e) {
                const double T = vmax(s[UTEMP], small_T);
                s[UEINT] = rho * eos_E_from_T(T, y0, y1, y2);
            }
            s[UEDEN] = s[UEINT] + rho * ke;
        } else {
            const double rho_eint = s[UEDEN] - s[URHO] * ke;
            if (rho_eint > 0. && rho_eint / s[UEDEN] > hc.eta2) s[UEINT] = rho_eint;
            if (s[UEINT] * rhoInv < small_e) {
                const double T = vmax(s[UTEMP], small_T);
                const double e_out = eos_E_from_T(T, y0, y1, y2);
                if (hc.update_E_from_e == 1) s[UEDEN] += (rho * e_out - s[UEINT]);
                s[UEINT] = s[URHO] * e_out;
            }
        }
    } else if (hc.allow_negative_energy == 0) {
        if (s[UEDEN] < 0.) {
            if (s[UEINT] < 0.) s[UEINT] = rho * eos_E_from_T(small_T, y0, y1, y2);
            s[UEDEN] = s[UEINT] + rho * ke;
        } else {
            const double rho_eint = s[UEDEN] - rho * ke;
            if (rho_eint > 0. && rho_eint / s[UEDEN] > hc.eta2) {
                s[UEINT] = rho_eint;
            } else if (s[UEINT] > 0 && hc.update_E_from_e == 1) {
                s[UEDEN] = s[UEINT] + rho * ke;
            } else if (s[UEINT] <= 0.) {
                const double e_out = eos_E_from_T(small_T, y0, y1, y2);
                s[UEINT] = rho * e_out;
                if (hc.update_E_from_e == 1) s[UEDEN] += (rho * e_out - s[UEINT]);
            }
        }
    } else {
        s[UEINT] = s[UEDEN] - rho * ke;
    }
    const double rinv = 1.0 / s[URHO];
    s[UTEMP] = eos_T_from_e(s[UEINT] * rinv);
}

// normalize_species (CAMR.cpp:866-908) on a register copy of one cell
CAMR_HD void normalize_species(double *s) {
    double sum = 0.;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) {
        s[UFS + n] = vmax(vmin(s[URHO], s[UFS + n]), 0.0);
        sum += s[UFS + n];
    }
    if (s[URHO] > 0.) {
        sum /= s[URHO];
#pragma unroll
        for (int n = 0; n < NSPEC; ++n) s[UFS + n] /= sum;
    }
}

__device__ __forceinline__ void floor_detect(double rho, double small_dens, int *flags) {
    if (rho == 0.) atomicOr(flags, 1 << CAMR_ERR_ZERO_DENSITY);
    else if (rho < small_dens) atomicOr(flags, kFloorFlag);
}


// ---------------------------------------------------------------- ctoprim
// hydro_ctoprim (Source/Hydro/Hydro_ctoprim.H:11-101), all QVAR / NQAUX
// components of cell (i,j,k).  Cells with rho <= 0 are left untouched as in
// the reference (q = 0 in an EB build: covered cells).
__device__ __forceinline__ void pt_ctoprim(const Arr &u, const Arr &q, const Arr &qa,
                                           const HydroConst &hc, int *status, int i, int j,
                                           int k) {
    const double rho = u(i, j, k, URHO);
    if (!(rho > 0.)) {
        if (hc.eb) {  // covered cell: q = 0 (CAMR_construct_hydro_source.cpp:116-118)
            for (int n = 0; n < QVAR; ++n) q(i, j, k, n) = 0.;
            for (int n = 0; n < NQAUX; ++n) qa(i, j, k, n) = 0.;
        }
        return;
    }
    const double rhoinv = 1.0 / rho;
    const double vx = u(i, j, k, UMX) * rhoinv;
    const double vy = u(i, j, k, UMY) * rhoinv;
    const double vz = u(i, j, k, UMZ) * rhoinv;
    const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    q(i, j, k, QRHO) = rho;
    q(i, j, k, QU) = vx;
    q(i, j, k, QV) = vy;
    q(i, j, k, QW) = vz;
    const double ueden = u(i, j, k, UEDEN);
    const double ueint = u(i, j, k, UEINT);
    if (!hc.allow_negative_energy && !(ueden > 0. && ueint > 0.))
        atomicOr(status, 1 << CAMR_ERR_NEG_ENERGY);
    const double e = ((ueden - kineng) / ueden > hc.eta1) ? (ueden - kineng) * rhoinv
                                                         : ueint * rhoinv;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) q(i, j, k, QFS + s) = u(i, j, k, UFS + s) * rhoinv;
    const double p = (kGamma - 1.0) * rho * e;
    q(i, j, k, QTEMP) = eos_T_from_e(e);
    q(i, j, k, QREINT) = e * rho;
    q(i, j, k, QPRES) = p;
    q(i, j, k, QGAME) = p / (e * rho) + 1.0;
    qa(i, j, k, QDPDR) = (kGamma - 1.0) * e;
    qa(i, j, k, QDPDE) = (kGamma - 1.0) * rho;
    qa(i, j, k, QGAMC) = kGamma;
    const double cs = sqrt(kGamma * p / rho);
    qa(i, j, k, QC) = cs;
    qa(i, j, k, QCSML) = vmax(hc.small_num, hc.small_num * cs);
}

// ---------------------------------------------------------- ghost filling
// Source index and sign of component n for a cell c of the grown box:
// periodic wrap, foextrap (Outflow; Inflow with the identity bcnormal of
// the Sod problems), reflect_even / reflect_odd (Symmetry, SlipWall,
// NoSlipWall) per Source/Utils/CAMR_setup.cpp:19-21.  AMReX fills faces,
// edges, then corners from already-filled cells, which composes the
// per-direction maps; returns true if c lies outside the domain in a
// non-periodic direction.
__device__ __forceinline__ bool map_cell(const DomInfo &D, int c0, int c1, int c2,
                                         int *m, unsigned *oddmask) {
    const int c[3] = {c0, c1, c2};
    bool phys = false;
    unsigned odd = 0;  // bit n set -> component n flips sign
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int x = c[d], lo = D.domlo[d], hi = D.domhi[d];
        if (x >= lo && x <= hi) { m[d] = x; continue; }
        if (D.periodic[d]) {
            const int n = hi - lo + 1;
            int r = (x - lo) % n;
            if (r < 0) r += n;
            m[d] = lo + r;
            continue;
        }
        phys = true;
        const int bc = (x < lo) ? D.bclo[d] : D.bchi[d];
        if (bc == BC_OUTFLOW || bc == BC_INFLOW || bc == BC_INTERIOR) {
            m[d] = (x < lo) ? lo : hi;
        } else {
            m[d] = (x < lo) ? (2 * lo - x - 1) : (2 * hi - x + 1);
            odd ^= 1u << (UMX + d);
            if (bc == BC_NOSLIPWALL)
                odd ^= ((1u << UMX) | (1u << UMY) | (1u << UMZ)) & ~(1u << (UMX + d));
        }
    }
    *oddmask = odd;
    return phys;
}

__device__ __forceinline__ int owner_of(const DomInfo &D, const int *owner, const int *m) {
    const int64_t n0 = D.domhi[0] - D.domlo[0] + 1, n1 = D.domhi[1] - D.domlo[1] + 1;
    return owner[(int64_t)(m[0] - D.domlo[0]) + (int64_t)(m[1] - D.domlo[1]) * n0 +
                 (int64_t)(m[2] - D.domlo[2]) * n0 * n1];
}


} // namespace camr
