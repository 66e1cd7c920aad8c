// hydro_eb.h -- embedded-boundary (cut-cell) hydro_umdrv_eb on gfx950.
#pragma once

#include "hydro_kernels.h"

namespace camr {

// Cut-cell flag word: bits 0..26 connectivity to the neighbour at offset
// (ii,jj,kk), bit (ii+1) + 3 (jj+1) + 9 (kk+1); bits 27..28 the type.
constexpr int kEbRegular = 0, kEbSingle = 1, kEbCovered = 2;
__host__ __device__ inline int eb_type(int32_t w) { return (w >> 27) & 3; }
__host__ __device__ inline bool eb_conn(int32_t w, int ii, int jj, int kk) {
    return (w >> ((ii + 1) + 3 * (jj + 1) + 9 * (kk + 1))) & 1;
}

// Flags over a box (the level's domain grown by 7 cells).  p == nullptr:
// no embedded boundary, every cell regular.
struct EbFlags {
    const int32_t *p;
    int lo0, lo1, lo2;
    int64_t js, ks;
    __host__ __device__ int32_t operator()(int i, int j, int k) const {
        return p[(int64_t)(i - lo0) + (int64_t)(j - lo1) * js + (int64_t)(k - lo2) * ks];
    }
    __host__ __device__ bool covered(int i, int j, int k) const {
        return p && eb_type((*this)(i, j, k)) == kEbCovered;
    }
};

// Cut-cell data of one box (EBFArrayBoxFactory arrays of an MFIter): vf on
// grow(bx, 7), ap[d] (apertures) and fc[d] (two face-centroid offsets) on
// the faces of direction d of grow(bx, 7).
struct EbGeom {
    EbFlags f;
    Arr vf, ap[3], fc[3];
};

// Arguments of one hydro_umdrv_eb call (Source/Hydro/Hydro_umdrv_eb.cpp:
// 14-191, Godunov branch, FluxRedist).  q / qaux / srcq and uin cover
// grow(bx, 6); dsdt (bx) must hold zeros on entry; flx[d] on the faces of bx.
struct UmdrvEbArgs {
    Box bx, qbox;
    HydroConst hc;
    int periodic[3];
    const double *dtp;  // device: the step's dt
    Arr uin, dsdt, q, qaux, srcq;
    int has_src;
    Arr flx[3];
    EbGeom g;
};

size_t umdrv_eb_scratch_doubles(const Box &bx, const Box &qbox);
int umdrv_eb(const UmdrvEbArgs &A, double *scratch, hipStream_t st, Prof *prof = nullptr);
// Per valid cell of bx: the wall momentum flux of its cut cell times dx2
// (3 components, 0 for other cells) into out; q / qaux cover bx.
int eb_force_box(const Box &bx, Arr q, Arr qaux, const EbGeom &g, const HydroConst &hc, Arr out,
                 double dx2, hipStream_t st);

} // namespace camr
