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

// One box of a cut-cell batch: its inputs and outputs (those of
// UmdrvEbArgs) and its scratch arrays, carved by eb_box_setup.
struct EbBox {
    Box bx, qbox, nb, dbox2;  // nb: divu nodes clamped to the domain; dbox2: redistribution domain
    EbGeom g;
    Arr q, qaux, srcq, uin, dsdt, flx[3];
    Arr QM[3], QP[3], F1[3], G1[3], F2[3][3], G2[3][3], FT[3], QE[3];
    Arr qmT[3], qpT[3], qmF, qpF, divu, divc, wt, optmp0, delm, winv;
};

// Doubles of scratch one box needs (the one-box call's EbBox record included).
size_t umdrv_eb_scratch_doubles(const Box &bx, const Box &qbox);
// Carves the scratch arrays of box A.bx out of `scratch` (device memory of
// umdrv_eb_scratch_doubles(A.bx, A.qbox) doubles) into *out (host memory).
void eb_box_setup(const UmdrvEbArgs &A, double *scratch, EbBox *out);
// hydro_umdrv_eb over the n boxes of tab (device copy) / htab (the same
// records in host memory, read for the grid sizes): one launch per stage,
// every box of the table in the launch.  All boxes share hc, dtp and has_src
// (srcq holds primitive sources).
int umdrv_eb_batch(const EbBox *tab, const EbBox *htab, int n, const HydroConst &hc,
                   const double *dtp, int has_src, hipStream_t st, Prof *prof = nullptr);
// One box: its EbBox record goes into the head of scratch (written by a
// kernel, so the call can be captured into a graph).
int umdrv_eb(const UmdrvEbArgs &A, double *scratch, hipStream_t st, Prof *prof = nullptr);
// Per valid cell of bx: the wall momentum flux of its cut cell times dx2
// (3 components, 0 for other cells) into out; q / qaux cover bx.
int eb_force_box(const Box &bx, Arr q, Arr qaux, const EbGeom &g, const HydroConst &hc, Arr out,
                 double dx2, hipStream_t st);

} // namespace camr
