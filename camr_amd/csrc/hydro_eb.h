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

// Value of face array `u` (component n) of direction d at face (i,j,k),
// moved to the face centroid (eb_compute_div, Hydro_utils_eb_K.H:410-520).
template <int D>
CAMR_HD double centroid_value(const Arr &u, int n, const EbGeom &g, int i,
                                                 int j, int k) {
    double v = u(i, j, k, n);
    const double a = g.ap[D](i, j, k, 0);
    if (a != 0.0 && a != 1.0) {
        constexpr int T0 = (D == 0) ? 1 : 0, T1 = (D == 2) ? 1 : 2;
        const double c0 = g.fc[D](i, j, k, 0), c1 = g.fc[D](i, j, k, 1);
        const int s0 = (int)copysign(1.0, c0), s1 = (int)copysign(1.0, c1);
        const int o0x = T0 == 0 ? s0 : 0, o0y = T0 == 1 ? s0 : 0, o0z = T0 == 2 ? s0 : 0;
        const int o1x = T1 == 0 ? s1 : 0, o1y = T1 == 1 ? s1 : 0, o1z = T1 == 2 ? s1 : 0;
        const int32_t w = g.f(i, j, k);
        const double f0 = eb_conn(w, o0x, o0y, o0z) ? fabs(c0) : 0.0;
        const double f1 = eb_conn(w, o1x, o1y, o1z) ? fabs(c1) : 0.0;
        v = (1.0 - f0) * (1.0 - f1) * v + f0 * (1.0 - f1) * u(i + o0x, j + o0y, k + o0z, n) +
            f1 * (1.0 - f0) * u(i + o1x, j + o1y, k + o1z, n) +
            f0 * f1 * u(i + o0x + o1x, j + o0y + o1y, k + o0z + o1z, n);
    }
    return v;
}

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
    // 1: every cell of the geometry is regular (constant rows), so the
    // cut-cell divergence and the redistribution's source side have nothing
    // to do and its receiving side is dsdt = divc + 0
    int regular = 0;
    // 1: leave flx unwritten (a recompute that needs only dS/dt and the
    // stages' arrays)
    int noflx = 0;
};

// The index boxes of one box's stages.
struct Layout {
    Box bx, btf, bxg1, bxg2, g1;
    Box qm[3], f1[3], f2[3], ft[3], qe[3], nb, mbx[3];
};

__host__ __device__ inline Layout layout_of(const Box &bx, const Box &qbox) {
    Layout L;
    L.bx = bx;
    L.btf = grow(bx, 2);
    L.bxg1 = grow(L.btf, 1);
    L.bxg2 = grow(L.btf, 2);
    L.g1 = grow(bx, 1);
    for (int d = 0; d < 3; ++d) {
        L.qm[d] = grow_hi(L.bxg2, d, 1);
        L.f1[d] = nodes(grow_dir(L.bxg2, d, -1), d);
        L.f2[d] = nodes(L.bxg2, d);
        L.ft[d] = nodes(grow(bx, 3), d);
        L.qe[d] = nodes(grow(bx, 4), d);
        L.mbx[d] = grow_hi(L.bxg2, d, 1);
    }
    // divu nodes: those of grow(qbox, -1) (Hydro_adjust_fluxes.cpp:73), no
    // further out than the nodes of grow(bx, 5) that the face stencils reach
    const Box a = nodes(nodes(nodes(grow(qbox, -1), 0), 1), 2);
    const Box b = nodes(nodes(nodes(grow(bx, 5), 0), 1), 2);
    for (int d = 0; d < 3; ++d) {
        L.nb.lo[d] = a.lo[d] > b.lo[d] ? a.lo[d] : b.lo[d];
        L.nb.hi[d] = a.hi[d] < b.hi[d] ? a.hi[d] : b.hi[d];
    }
    return L;
}

// pair p: (D, T, O) with O the third direction
__host__ __device__ constexpr int pair_d(int p) { return p / 2; }
__host__ __device__ constexpr int pair_t(int p) {
    return p / 2 == 0 ? (p % 2 ? 2 : 1) : p / 2 == 1 ? (p % 2 ? 2 : 0) : (p % 2 ? 1 : 0);
}
__host__ __device__ constexpr int pair_o(int p) { return 3 - pair_d(p) - pair_t(p); }
__host__ __device__ constexpr int t0_of(int d) { return d == 0 ? 1 : 0; }
__host__ __device__ constexpr int t1_of(int d) { return d == 2 ? 1 : 2; }

// One box of a cut-cell batch: its inputs and outputs (those of
// UmdrvEbArgs) and its scratch arrays, carved by eb_box_setup.
struct EbBox {
    Box bx, qbox, nb, dbox2;  // nb: divu nodes clamped to the domain; dbox2: redistribution domain
    EbGeom g;
    Arr q, qaux, srcq, uin, dsdt, flx[3];
    Arr QM[3], QP[3], F1[3], G1[3], F2[3][3], G2[3][3], FT[3], QE[3];
    // the corrected states of transverse pair p (D|T in the order 0|1, 0|2,
    // 1|0, 1|2, 2|0, 2|1) and the final states of direction d: one buffer
    // for all pairs / all directions, or one each when the adjoint reads
    // them back (`keep`)
    Arr qmT[6], qpT[6], qmF[3], qpF[3];
    Arr divu, divc, wt, optmp0, delm, winv;
    int regular;  // UmdrvEbArgs::regular
    int noflx;    // 1: the face fluxes of bx are not copied into flx (a recompute)
};

// Doubles of scratch one box needs (the one-box call's EbBox record
// included); keep: every pair's and direction's states in their own arrays.
size_t umdrv_eb_scratch_doubles(const Box &bx, const Box &qbox, bool keep = false);
// Carves the scratch arrays of box A.bx out of `scratch` (device memory of
// umdrv_eb_scratch_doubles(A.bx, A.qbox, keep) doubles) into *out (host
// memory).
void eb_box_setup(const UmdrvEbArgs &A, double *scratch, EbBox *out, bool keep = false);
// hydro_umdrv_eb over the n boxes of tab (device copy) / htab (the same
// records in host memory, read for the grid sizes): one launch per stage,
// every box of the table in the launch.  All boxes share hc, dtp and has_src
// (srcq holds primitive sources).
// A cut cell (eb_type kEbSingle) of grow(bx, 2) of entry e of a table.
struct CutCell {
    int e, i, j, k;
};
// cells (device, ncells): every cut cell of grow(bx, 2) of the table's
// boxes; the cut-cell divergence and the redistribution's source side then
// run over them alone (one thread per cut cell) instead of over grow(bx, 2)
// and grow(bx, 1).
int umdrv_eb_batch(const EbBox *tab, const EbBox *htab, int n, const HydroConst &hc,
                   const double *dtp, int has_src, hipStream_t st, Prof *prof = nullptr,
                   const CutCell *cells = nullptr, int ncells = 0);
// hydro_ctoprim of every box of the table over its qbox (Sborder in uin) and
// the zero fill of its dS/dt, in one launch.
int eb_prim_batch(const EbBox *tab, const EbBox *htab, int n, const HydroConst &hc, int *status,
                  hipStream_t st, Prof *prof = nullptr);
// One box: its EbBox record goes into the head of scratch (written by a
// kernel, so the call can be captured into a graph).
int umdrv_eb(const UmdrvEbArgs &A, double *scratch, hipStream_t st, Prof *prof = nullptr);
// Per valid cell of bx: the wall momentum flux of its cut cell times dx2
// (3 components, 0 for other cells) into out; q / qaux cover bx.
int eb_force_box(const Box &bx, Arr q, Arr qaux, const EbGeom &g, const HydroConst &hc, Arr out,
                 double dx2, hipStream_t st);

} // namespace camr
