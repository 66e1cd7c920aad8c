// hydro_state.h -- declarations of the level-step state kernels.
#pragma once

#include "hydro_eb.h"

namespace camr {

constexpr int kFloorFlag = 1 << 30;  // a density below small_dens was seen

struct DevFab {  // device-side box table entry
    Arr a;
    Box b;
};

struct DomInfo {
    int domlo[3], domhi[3], periodic[3], bclo[3], bchi[3];
};

// Device-resident step control of captured level-step graphs.
struct StepCtl {
    double dt;      // dt of the step in progress (read by the step kernels)
    double time;    // level time
    int first;      // next dt is computeInitialDt's, not computeNewDt's
    int halt;       // set when a step must be redone on the host
    int nstep;      // steps taken since the host last reset it
    int cap;        // capacity of dts
    double *dts;    // dt of each step taken
};
struct DtParams {
    double fixed_dt, max_dt, cfl, change_max, init_shrink, stop_time;
};
__global__ void k_step_dt(StepCtl *c, unsigned long long *dtmin, int *status, DtParams P);

__global__ void k_ctoprim(Range r, Arr u, Arr q, Arr qa, HydroConst hc, int *status);
__global__ void k_srctoprim(Range r, Arr q, Arr qa, Arr src, Arr srcq);
__global__ void k_estdt(Range r, Arr u, double dx, double dy, double dz,
                        unsigned long long *out, int *status, EbFlags eb);
__global__ void k_detect(Range r, Arr S, double small_dens, int *flags, EbFlags eb);
__global__ void k_floor_seq(Box bx, Arr S, HydroConst hc, int *flags, EbFlags eb);
__global__ void k_clean_point(Range r, Arr S, HydroConst hc);
__global__ void k_normalize(Range r, Arr S);
__global__ void k_resetT(Range r, Arr S, HydroConst hc);
__global__ void k_init_post(Range r, Arr S, HydroConst hc);
__global__ void k_fill_raw(Range r, Arr dst, const DevFab *src, const int *owner,
                           DomInfo D, Arr ext, int has_ext, int *status);
__global__ void k_resync(Range r, Box valid, Arr dst, const DevFab *src,
                         const int *owner, DomInfo D);
__global__ void k_copy_oldsrc(Range r, Arr snew, Arr sb, Arr sold, double dt,
                              double g, int add_grav);
__global__ void k_add_hydro(Range r, Arr snew, Arr dsdt, double dt,
                            double small_dens, int *flags, EbFlags eb);
__global__ void k_add_newsrc(Range r, Arr snew, Arr sold, double dt, double g,
                             double small_dens, int *flags, EbFlags eb);
__global__ void k_srchyd(Range r, Box valid, Arr dst, Arr sold, double g);
__global__ void k_fill_value(Range r, Arr a, int nc, double v);
__global__ void k_fill_flat(double *p, int64_t n, double v);
__global__ void k_fill_u64(Range r, unsigned long long *p, unsigned long long v);
// the passes above over every box of a table in one launch (grid.y = box)
__global__ void kb_estdt(const DevFab *S, double dx, double dy, double dz,
                         unsigned long long *out, int *status, EbFlags eb);
__global__ void kb_detect(const DevFab *S, double small_dens, int *flags, EbFlags eb);
__global__ void kb_state_flat(const DevFab *S, double *flat, const int64_t *off, int to_flat);
__global__ void kb_zero_state(const DevFab *S);
__global__ void kb_floor_seq(const DevFab *S, int nbox, HydroConst hc, int *flags, EbFlags eb);
__global__ void kb_clean_point(const DevFab *S, HydroConst hc);
__global__ void kb_normalize(const DevFab *S);
__global__ void kb_resetT(const DevFab *S, int g, HydroConst hc);
__global__ void kb_fill_raw(const DevFab *dst, int g, const DevFab *src, const int *owner,
                            DomInfo D, Arr ext, int has_ext, int *status);
__global__ void kb_resync(const DevFab *dst, int g, const DevFab *src, const int *owner, DomInfo D);
__global__ void kb_copy_add_hydro(const DevFab *snew, const DevFab *sb, const DevFab *sold,
                                  const DevFab *dsdt, const double *dtp, double g, int add_grav,
                                  double small_dens, int *flags, EbFlags eb);
// A virtual box: s[0] x s[1] x s[2] boxes of a level, each len[0] x len[1]
// x len[2] cells, that tile the box v, and an array over grow(v, g) (its
// Sborder, or the adjoint of it).  ids[off + (c s[1] + r) s[0] + a]: the
// level index of the box at block position (a, r, c).  The Sborder of v
// holds at each point what the Sborder of the box holding the point
// clamped to v holds there (a point beyond v lies in grow(that box, 6)).
struct VGather {
    Box v;
    Arr sb;
    int len[3], s[3];
    int off;
};
// dL/dSborder of each virtual box of vt: components 0..nc-1 of every point
// of grow(v, g) added into sb_tab's array of the box that holds the point
// clamped to v.
__global__ void kb_vscatter(const VGather *vt, const int *ids, const DevFab *sb_tab, int g,
                            int nc, FlatBlocks fb);
// expand_state of every box of dst (grow g) from the cleaned states of src:
// kb_fill_raw, the detection, normalize_species, kb_resync and kb_resetT
// in one pass (src must have been through clean_state).
// With redo (a source that did not go through clean_state): where a valid
// cell is below small_dens it sets *redo, and kb_expand_redo repeats the
// passes one by one (one workgroup of up to 1024 threads).
__global__ void kb_expand_cleaned(const DevFab *dst, int g, const DevFab *src, const int *owner,
                                  DomInfo D, Arr ext, int has_ext, HydroConst hc, int *status,
                                  EbFlags eb, FlatBlocks fb, int *redo);
__global__ void kb_expand_redo(const DevFab *dst, int nbox, int g, const DevFab *src,
                               const int *owner, DomInfo D, Arr ext, int has_ext, HydroConst hc,
                               int *status, EbFlags eb, int *redo);
// The staged step's pointwise tail after the hydro source in one pass
// (S_new = Sborder + sources + dt dS/dt, or with mol the second stage's
// 0.5 Sborder + 0.5 S_old + 0.5 dt dS/dt, and the clean_state passes), and
// its pass-by-pass repeat where a density floor would act (one workgroup of
// up to 1024 threads; a no-op unless kb_tail_fused set *redo).
__global__ void kb_tail_fused(const DevFab *snew, const DevFab *sb, const DevFab *sold,
                              const DevFab *dsdt, const double *dtp, double g, int add_grav,
                              int post, int mol, HydroConst hc, int *redo, EbFlags eb,
                              unsigned long long *dtmin, int *status);
// With dtmin: the CFL estimate of the new state into *dtmin (kb_estdt's),
// set by kb_tail_fused (which expects DBL_MAX there) and, after a repeat,
// by kb_tail_redo.
__global__ void kb_tail_redo(const DevFab *snew, const DevFab *sb, const DevFab *sold,
                             const DevFab *dsdt, int nbox, const double *dtp, double g,
                             int add_grav, int post, int mol, HydroConst hc, int *flags,
                             int *redo, EbFlags eb, unsigned long long *dtmin);
__global__ void k_coarse_stage(Range r, Arr st, const DevFab *ctab, const int *cowner, DomInfo CD,
                               const DevFab *otab, int tmode, double ta, double tb);
__global__ void k_coarse_stage_cells(const int *cells, int64_t ncells, Arr st, const DevFab *ctab,
                                     const int *cowner, DomInfo CD, const DevFab *otab, int tmode,
                                     double ta, double tb);
__global__ void k_ext_from_coarse(Range r, Arr ext, Arr st, DomInfo CD, int ratio, int flo0,
                                  int flo1, int flo2, const int *fowner, DomInfo FD);
__global__ void k_ext_cells(const int *cells, int64_t ncells, Arr ext, Arr st, DomInfo CD,
                            int ratio, int flo0, int flo1, int flo2);
// coarse-fine synchronisation (flux register, reflux, average_down)
constexpr int kRegBnd = 1, kRegFine = 2;
__global__ void k_reg_crse_add(Range r, Arr d, Arr f0, Arr f1, Arr f2, const uint8_t *flag,
                               DomInfo CD, const double *dtp, double fac, double dx1);
template <int D>
__global__ void k_reg_fine_add(Range r, const DevFab *cfp, const int *cowner, DomInfo CD,
                               const uint8_t *flag, Arr f, int side, int ratio,
                               const double *dtp, double fac, double dx1f);
// flux-register records of one-launch adds (fused-engine levels)
struct RegCrseRec {
    Range r;   // the coarse box
    Arr d;     // its register data
    Arr f[3];  // its final fluxes
};
struct RegFineRec {
    Range r;   // a slab of coarse cells beside one side of a fine box
    Arr f;     // that box's final fluxes of direction d
    int d, side;
};
__global__ void kb_reg_crse_add(const RegCrseRec *recs, const uint8_t *flag, DomInfo CD,
                                const double *dtp, double fac, double dx1);
__global__ void kb_reg_fine_add(const RegFineRec *recs, const DevFab *cfp, const int *cowner,
                                DomInfo CD, const uint8_t *flag, int ratio, const double *dtp,
                                double fac, double dx1f);
__global__ void kb_reflux(const DevFab *snew, const DevFab *crse, const DevFab *cfp);
__global__ void kb_avgdown(const DevFab *csnew, const DevFab *ftab, const int *fowner, DomInfo FD,
                           int ratio);
inline Box make_box1() { Box b; for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = 0; } return b; }

template <typename K, typename... A>
inline void launch_on(K kern, const Box &b, hipStream_t s, A... args) {
    Range r = range_of(b);
    hipLaunchKernelGGL(kern, dim3(grid_for(r.n)), dim3(kBlock), 0, s, r, args...);
}

} // namespace camr
