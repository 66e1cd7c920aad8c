/*
 * camr_hydro.h -- C-ABI of the MI355X (gfx950) single-level hydro advance.
 *
 * Drop-in boundary for CAMR's per-box hydro call chain.  Every entry point
 * takes plain pointers and sizes (no AMReX or torch types); arrays are
 * device pointers described by camr_array4, which has the memory layout of
 * amrex::Array4<Real>: x fastest, then y, then z, component outermost,
 *     p[(i-lo[0]) + (j-lo[1])*jstride + (k-lo[2])*kstride + n*nstride].
 *
 * All functions return a camr_status (0 = success).  Fatal conditions that
 * abort the reference (amrex::Error / AMREX_ALWAYS_ASSERT) become non-zero
 * status codes here; camr_last_error() returns the message of the last
 * failure on the calling thread.  Per-box functions are asynchronous on the
 * given hipStream_t (NULL = default stream) unless stated otherwise.
 *
 * Component maps are CAMR's 3-D maps with NUM_SPECIES = 3
 * (Source/Utils/IndexDefines.H:17-92):
 *   U  (NVAR=10):  URHO UMX UMY UMZ UEDEN UEINT UTEMP UFS..UFS+2
 *   Q  (QVAR=11):  QRHO QU QV QW QGAME QPRES QREINT QTEMP QFS..QFS+2
 *   QAUX (NQAUX=5): QGAMC QC QCSML QDPDR QDPDE
 */
#ifndef CAMR_HYDRO_H
#define CAMR_HYDRO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAMR_NVAR 10
#define CAMR_QVAR 11
#define CAMR_NQAUX 5
#define CAMR_NGROW 4 /* CAMR::numGrow() without EB, Source/CAMR.H:455-483 */

/* Physical boundary codes, Source/Utils/IndexDefines.H:8-13. */
enum camr_bc {
    CAMR_BC_INTERIOR = 0,
    CAMR_BC_INFLOW = 1,
    CAMR_BC_OUTFLOW = 2,
    CAMR_BC_SYMMETRY = 3,
    CAMR_BC_SLIPWALL = 4,
    CAMR_BC_NOSLIPWALL = 5
};

typedef enum camr_status {
    CAMR_OK = 0,
    CAMR_ERR_PARAM = 1,        /* invalid parameter (e.g. ppm_type not 0/1) */
    CAMR_ERR_SHAPE = 2,        /* array does not cover the region required */
    CAMR_ERR_NEG_ENERGY = 3,   /* rho E or rho e <= 0 (Hydro_ctoprim.H:56,68) */
    CAMR_ERR_ZERO_DENSITY = 4, /* rho == 0 in enforce_min_density */
    CAMR_ERR_DT = 5,           /* dt <= 0 (Timestep.H:59) */
    CAMR_ERR_HIP = 6,          /* HIP runtime failure */
    CAMR_ERR_ALLOC = 7,        /* device allocation failure */
    CAMR_ERR_ADJOINT = 8       /* outside the adjoint's scope (settings, or a state
                                  taking a branch it does not differentiate) */
} camr_status;

typedef struct camr_box {
    int32_t lo[3], hi[3]; /* inclusive cell-index bounds */
} camr_box;

typedef struct camr_array4 {
    double *p; /* device pointer to element (lo, comp 0) */
    int64_t jstride, kstride, nstride;
    int32_t lo[3], hi[3];
    int32_t ncomp;
} camr_array4;

/* Hydro runtime parameters passed by CAMR::construct_hydro_source to
 * hydro_umdrv (Source/Hydro/Hydro.H:11-35), plus the ctoprim / clean_state
 * controls (Source/Params/_cpp_parameters). */
typedef struct camr_hydro_params {
    int32_t ppm_type;                 /* 0 PLM, 1 PPM */
    int32_t plm_iorder;               /* 1, 2 or 4 */
    int32_t use_pslope;
    int32_t use_flattening;
    int32_t transverse_reset_density;
    int32_t allow_negative_energy;
    int32_t allow_small_energy;
    int32_t dual_energy_update_E_from_e;
    double small, small_dens, small_pres, small_temp;
    double smallu, small_num, difmag;
    double dual_energy_eta1, dual_energy_eta2;
    int32_t bclo[3], bchi[3];   /* physical BC codes (camr_bc) */
    int32_t domlo[3], domhi[3]; /* problem-domain cell bounds */
    double dx[3];
    int32_t do_mol;             /* 0 Godunov (CTU), 1 method of lines (MOL_umeth) */
} camr_hydro_params;

/* Defaults of Source/Params/_cpp_parameters and CAMR_Constants.H. */
void camr_default_params(camr_hydro_params *p);
const char *camr_last_error(void);
int camr_version(void);

/*
 * Per-box entry points (the MFIter loop body of
 * CAMR::construct_hydro_source, Source/Hydro/CAMR_construct_hydro_source.cpp).
 */

/* hydro_ctoprim over qbx: the ParallelFor at
 * CAMR_construct_hydro_source.cpp:108-120 (Hydro_ctoprim.H:11-101).
 * Writes all QVAR components of q and NQAUX of qaux. */
int camr_hydro_ctoprim(const camr_box *qbx, const camr_array4 *u,
                       const camr_array4 *q, const camr_array4 *qaux,
                       const camr_hydro_params *p, void *stream);

/* hydro_srctoprim over qbx: CAMR_construct_hydro_source.cpp:129-143
 * (Hydro_ctoprim.H:106-157). */
int camr_hydro_srctoprim(const camr_box *qbx, const camr_array4 *q,
                         const camr_array4 *qaux, const camr_array4 *src,
                         const camr_array4 *srcq, void *stream);

/* hydro_umdrv, Source/Hydro/Hydro.H:11-35 / Hydro_umdrv.cpp:11-93, its
 * do_mol argument carried by p->do_mol (Godunov_umeth or MOL_umeth; the
 * method of lines reads neither srcq nor dt's value beyond dt > 0).
 * Same contract: q, qaux, srcq cover grow(bx,4);
 * uin covers grow(bx,1); dsdt is ACCUMULATED (+=) on bx; flx[d] covers the
 * d-faces of bx and is overwritten with extensive (area-weighted) fluxes.
 * area[d].p / vol->p may be NULL for a uniform Cartesian grid (then
 * area = product of the two other dx, vol = dx*dy*dz as geom.GetFaceArea /
 * GetVolume give); srcq->p may be NULL for identically +0 sources.
 * Godunov boxes on a uniform grid (area and vol NULL) whose q covers
 * exactly grow(bx,4), as construct_hydro_source allocates it, run on the
 * fused engine's kernels; other calls on the stage-by-stage kernels; both
 * give the same results.
 * Scratch memory (the reference's The_Async_Arena temporaries,
 * Godunov_3D.cpp:79-98) comes from a stream-ordered memory pool per device
 * owned by the library: freed in stream order at the end of the call and
 * kept in the pool, so repeated calls do not go back to the driver.  Calls
 * on distinct streams may overlap. */
int camr_hydro_umdrv(const camr_box *bx, const camr_array4 *uin,
                     const camr_array4 *dsdt, const camr_array4 *q,
                     const camr_array4 *qaux, const camr_array4 *srcq,
                     const camr_array4 flx[3], const camr_array4 area[3],
                     const camr_array4 *vol, double dt,
                     const camr_hydro_params *p, void *stream);

/* camr_hydro_umdrv with no allocation of any kind: the scratch is a
 * file-scope device array of the library (96 MiB per GPU, held until the
 * process ends), enough for boxes up to 32^3 cells; larger boxes
 * (camr_scratch_bytes above 96 MiB) are refused with CAMR_ERR_SHAPE.  No
 * counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Calls share their GPU's array: calls that overlap on one GPU
 * (on different streams) race. */
int camr_hydro_umdrv_noalloc(const camr_box *bx, const camr_array4 *uin,
                             const camr_array4 *dsdt, const camr_array4 *q,
                             const camr_array4 *qaux, const camr_array4 *srcq,
                             const camr_array4 flx[3], const camr_array4 area[3],
                             const camr_array4 *vol, double dt,
                             const camr_hydro_params *p, void *stream);

/* Bytes of scratch one camr_hydro_umdrv call on box bx takes at most with
 * params p (the Godunov or MOL temporaries of Godunov_3D.cpp:79-98 /
 * MOL_umeth.cpp, plus the device word of dt). */
int camr_scratch_bytes(const camr_box *bx, const camr_hydro_params *p, size_t *bytes);

/* Cut-cell geometry of one box: the EBFArrayBoxFactory arrays that
 * hydro_umdrv_eb reads (Source/Hydro/Hydro_umdrv_eb.cpp:55-64) with the
 * EBCellFlag array replaced by an int32 word per cell: bits 0..26 say
 * whether the neighbour at offset (ii,jj,kk) is connected, bit
 * (ii+1) + 3 (jj+1) + 9 (kk+1); bits 27..28 hold the type (0 regular, 1
 * single-valued cut cell, 2 covered).  flags (device) covers flo..fhi;
 * vfrac, ap[d] (area fractions of the d-faces) and fc[d] (their centroid
 * offsets, 2 components: the two tangential directions in increasing
 * order, in [-1/2, 1/2]) cover grow(bx, 7) and its faces. */
typedef struct camr_eb_geom {
    const int32_t *flags;
    int32_t flo[3], fhi[3];
    camr_array4 vfrac;
    camr_array4 ap[3];
    camr_array4 fc[3];
} camr_eb_geom;

/* hydro_umdrv_eb, Godunov branch with redistribution_type "FluxRedist" and
 * eb_weights_type 2 on a single level (Source/Hydro/Hydro.H:56-91,
 * Hydro_umdrv_eb.cpp:14-191; replaces the call at
 * CAMR_construct_hydro_source.cpp:195-204).  q, qaux, srcq and uin cover
 * grow(bx,6); dsdt (zeroed by the caller, as CAMR_construct_hydro_source.cpp:14
 * does) receives the redistributed update on the non-covered cells of bx;
 * flx[d] receives the face-centroid fluxes of the d-faces of bx (for flux
 * registers).  Needs ppm_type 0 and dx == dy == dz.  No flux-register
 * arguments: the single-level redistribution has none. */
int camr_hydro_umdrv_eb(const camr_box *bx, const camr_array4 *uin,
                        const camr_array4 *dsdt, const camr_array4 *q,
                        const camr_array4 *qaux, const camr_array4 *srcq,
                        const camr_array4 flx[3], const camr_eb_geom *geom,
                        double dt, const camr_hydro_params *p,
                        const int32_t is_periodic[3], void *stream);

/* CAMR_estdt_hydro (Source/Utils/Timestep.H:18-61) min-reduced over bx.
 * Synchronous: *dt_out receives the minimum on return. */
int camr_estdt_hydro(const camr_box *bx, const camr_array4 *u,
                     const double dx[3], double *dt_out, void *stream);

/*
 * Level driver: CAMR::CAMR_advance (Source/CAMR_advance.cpp:38-164) plus
 * post_timestep's clean_state and computeNewDt (Source/CAMR.cpp:566-718) for
 * one AMR level made of a list of boxes.  Device memory for the level
 * (S_new, S_old, Sborder, sources, scratch) is owned by the level object.
 */
typedef struct camr_level camr_level;

typedef struct camr_level_desc {
    camr_hydro_params hp;
    int32_t is_periodic[3];
    int32_t add_grav_src; /* default 1 (_cpp_parameters:52) */
    double const_grav;    /* default 0.0 (_cpp_parameters:55) */
    double cfl, change_max, init_shrink, max_dt, fixed_dt, stop_time;
    int32_t nbox;
    const int32_t *box_lohi; /* host array of nbox*6 ints: lo[3], hi[3] */
    int32_t engine;          /* 0 = fused, or staged for boxes too large for the fused
                                engine's 32-bit layout (default); 1 = staged per-stage
                                kernels; 2 = fused only */
    int32_t batched;         /* fused engine: 1 = each stage is one launch over every
                                box of the level (default); 0 = one launch per box,
                                the per-box loop of construct_hydro_source
                                (CAMR_construct_hydro_source.cpp:59-71) */
    int32_t graph;           /* 1: camr_level_run replays one captured hipGraph per
                                step (device-side dt, no host round trip per step);
                                used with the batched fused engine and zero gravity,
                                otherwise camr_level_run steps on the host */
    int32_t worker_streams;  /* staged engine: streams the per-box hydro of a
                                multi-box level is spread over (default 4, at most
                                16); more than the process's hardware queues
                                (GPU_MAX_HW_QUEUES) gain nothing */
} camr_level_desc;

void camr_default_level_desc(camr_level_desc *d);
int camr_level_create(const camr_level_desc *d, camr_level **out);
int camr_level_destroy(camr_level *L);
int camr_level_set_stream(camr_level *L, void *stream);
/* Valid data of box b in Array4 order (NVAR components), host memory. */
int camr_level_upload_box(camr_level *L, int32_t b, const double *host);
int camr_level_download_box(camr_level *L, int32_t b, double *host);
/* Device view of the new-time state of box b (valid cells, NVAR comps). */
int camr_level_state_array(camr_level *L, int32_t b, camr_array4 *out);
/* CAMR::initData tail: enforce_consistent_e + computeTemp (CAMR.cpp:423-425). */
int camr_level_init_post(camr_level *L);
/* computeInitialDt (CAMR.cpp:638-676) / computeNewDt (CAMR.cpp:566-635). */
int camr_level_initial_dt(camr_level *L, double *dt);
int camr_level_new_dt(camr_level *L, double dt_old, double *dt);
/* One CAMR_advance + post_timestep clean_state, asynchronous. */
int camr_level_advance(camr_level *L, double dt);
/* nsteps with the reference dt sequence; dts (host, may be NULL). */
int camr_level_run(camr_level *L, int32_t nsteps, double *dts);
int camr_level_sync(camr_level *L);
double camr_level_time(const camr_level *L);

/* Exterior data of a level that does not cover its domain: NVAR components
 * over the whole domain (host, Array4 order).  Ghost cells that no box of
 * the level owns take these values where FillPatch would interpolate from
 * the coarser level (FillPatch, CAMR.cpp:1247; CAMR_setup.cpp:141-144).
 * Without it such ghost cells are an error (CAMR_ERR_SHAPE). */
int camr_level_set_exterior(camr_level *L, const double *host);

/* The same exterior data computed on the device from the coarser level's
 * current state: FillPatch's coarse-to-fine interpolation (AmrLevel::
 * FillPatch, CAMR.cpp:1247, with the State_Type interpolater cell_cons_interp,
 * Source/Utils/CAMR_setup.cpp:141-144: conservative linear, MC-limited,
 * linearly limited) at every cell of the fine domain.  The fine domain must
 * be the coarse domain refined by `ratio`.  Single time level: the coarse
 * data of the moment (no time interpolation between coarse old and new). */
int camr_level_exterior_from_coarse(camr_level *fine, camr_level *coarse, int32_t ratio);

/*
 * Coarse-fine synchronisation of level hierarchies (either engine, no
 * embedded boundary).
 */
typedef struct camr_flux_register camr_flux_register;

/* The flux register of the pair: AMReX's YAFluxRegister as CAMR holds it
 * on the fine level (CAMR.H:310-313, defined at CAMR.cpp:287-290) and
 * fills it from the per-box fluxes
 * (CrseAdd / FineAdd, CAMR_construct_hydro_source.cpp:243-254, with
 * fac_for_reflux * dt, :24).  Attaching it makes both levels add their
 * fluxes on every step.  Fine boxes must be aligned to `ratio`, and no
 * coarse cell may border two different fine boxes (CAMR_ERR_SHAPE). */
int camr_flux_register_create(camr_level *coarse, camr_level *fine, int32_t ratio,
                              camr_flux_register **out);
/* Detaches the register from its levels and frees it. */
int camr_flux_register_destroy(camr_flux_register *r);

/* CAMR::avgDown (CAMR.cpp:1047-1066, non-EB amrex::average_down): every
 * coarse cell covered by the fine level becomes the mean of its children. */
int camr_level_average_down(camr_level *coarse, camr_level *fine, int32_t ratio);

/* One coarse step of the hierarchy with subcycling, Amr::timeStep as CAMR
 * drives it: reset the register (CAMR_advance.cpp:64-66), CAMR_advance on
 * the coarse level, `ratio` steps of dt_coarse / ratio on the fine level,
 * each with its exterior ghost data interpolated from the coarse old and
 * new states in time (StateData's linear interpolation) and then in space
 * (as camr_level_exterior_from_coarse), then the coarse post_timestep
 * (CAMR.cpp:679-699): reflux (CAMR.cpp:803-851), avgDown, clean_state.
 * Both levels must be at the same time. */
int camr_amr_step(camr_level *coarse, camr_level *fine, camr_flux_register *r,
                  double dt_coarse);

/* The same for a hierarchy of nlev levels (coarse first, regs[l] joining
 * levels l and l+1): one step of level 0 and, recursively, `ratio` steps
 * of each finer level per step of its coarser one, every level's
 * post_timestep after its finer levels' steps (Amr::timeStep's
 * subcycling).  A level with a finer and a coarser neighbour feeds both
 * registers. */
int camr_amr_advance(camr_level *const *levels, camr_flux_register *const *regs, int32_t nlev,
                     double dt0);

/*
 * AMReX plotfile of a level hierarchy (CAMR::writePlotFile, Source/Utils/
 * CAMR_io.cpp:407-591, with Amr's header lines): dir/Header,
 * dir/Level_<l>/Cell_H, dir/Level_<l>/Cell_D_00000 and dir/job_info.
 * Variables: density, xmom, ymom, zmom, rho_E, rho_e, Temp, the rho_<species>
 * when plot_rhoy is set (CAMR.plot_rhoy, CAMR_io.cpp:100-111), then the
 * requested derived variables among pressure, kineng, soundspeed,
 * MachNumber, x_velocity, y_velocity, z_velocity, magvel (Derive.cpp), in
 * the order CAMR registers them.  Data are the levels' current S_new.
 */
typedef struct camr_plot_desc {
    int32_t nlev;                  /* levels, coarse first */
    camr_level *const *levels;
    const int32_t *ref_ratio;      /* nlev - 1 ratios */
    const int32_t *level_steps;    /* nlev step counts, or NULL (0) */
    double prob_lo[3], prob_hi[3]; /* geometry.prob_lo / prob_hi */
    int32_t plot_rhoy;
    const char *derive_vars;       /* space-separated amr.derive_plot_vars, or NULL */
} camr_plot_desc;
int camr_write_plotfile(const char *dir, const camr_plot_desc *d);

/*
 * Checkpoint of a level hierarchy for restart (CAMR::checkPoint,
 * Source/Utils/CAMR_io.cpp:45-84: AmrLevel's state data and a CPUtime
 * file), in a directory layout modelled on AMReX checkpoints: dir/Header
 * (time, ratios, per-level dt and steps, domains and boxes), the new-time
 * state of each level in VisMF layout (dir/Level_<l>/SD_0_New_H and
 * _D_00000, all NVAR components bit for bit) and dir/CPUtime (written as 0:
 * no CPU-time accounting).  The Header is this library's own, without
 * Geometry, Amr's cycle counts or StateData records: only camr_amd reads it
 * back (camr_amd.plotfile.restart), not an AMReX build.  Restart: upload each
 * box's state (camr_level_upload_box) and the level time
 * (camr_level_set_time).
 */
typedef struct camr_checkpoint_desc {
    int32_t nlev;
    camr_level *const *levels;     /* coarse first */
    const int32_t *ref_ratio;      /* nlev - 1 ratios */
    const int32_t *level_steps;    /* nlev, or NULL */
    const double *dt_level;        /* nlev, the last dt of each level, or NULL */
} camr_checkpoint_desc;
int camr_write_checkpoint(const char *dir, const camr_checkpoint_desc *d);
/* Set a level's time (restart). */
int camr_level_set_time(camr_level *L, double t);

/* The embedded-boundary build of a level (AMREX_USE_EB: CAMR.cpp:303-329,
 * 419-421; CAMR_construct_hydro_source.cpp:77-204): ghost width 6, covered
 * cells skipped by clean_state and the dt estimate, and per box the path of
 * CAMR_construct_hydro_source.cpp:80-83,150-151 -- kind 2 (every valid cell
 * covered) no hydro, kind 1 (a cut cell within 2 cells) hydro_umdrv_eb,
 * kind 0 hydro_umdrv.  Call once, before the first step, on a level created
 * with the staged engine whose boxes cover the domain; flags (host, same
 * word as camr_eb_geom) cover flo..fhi, which must hold the domain grown by
 * 7 cells.  Then give every kind-1 box its cut-cell data (host arrays on
 * grow(bx,7), faces nodal, fc with 2 components). */
int camr_level_set_eb(camr_level *L, const int32_t *flags, const int32_t flo[3],
                      const int32_t fhi[3]);
int camr_level_box_kind(camr_level *L, int32_t b, int32_t *kind);
/* Launch shape of the regular boxes of an EB level, which run as one batch
 * of the fused kernels (the per-box loop of
 * CAMR_construct_hydro_source.cpp:146-238): with merging on (the default),
 * equal regular boxes that tile a box in blocks of up to 4x4x4 run as that
 * one box, with fewer ghost cells; results are identical either way.
 * Set before camr_level_set_eb.  _info: the number of such virtual boxes
 * and of the level boxes they hold. */
int camr_level_set_eb_merge(camr_level *L, int32_t on);
int camr_level_eb_merge_info(camr_level *L, int32_t *nvirtual, int32_t *nmerged);

/* The force on the embedded body from the level's current state: the sum
 * over all cut cells of the wall momentum flux (compute_hyp_wallflux's
 * fluxw[UMX], fluxw[UMY], fluxw[UMZ], Hydro_hypflux_eb_K.H:65) times dx^2,
 * summed box by box and cell by cell (x fastest).  CAMR computes no force;
 * this is the x-force functional of the Sod-past-a-sphere workload (and
 * its y and z analogues). */
int camr_level_eb_force(camr_level *L, double force[3]);
int camr_level_set_box_eb(camr_level *L, int32_t b, const double *vfrac,
                          const double *apx, const double *apy, const double *apz,
                          const double *fcx, const double *fcy, const double *fcz);

/* Launch configuration of the fused engine: block size (64, 128 or 256) of
 * each of the CAMR_NSTAGE stage ranges (grow(bx,4), grow(bx,2), first-flux
 * faces x/y/z, transdo faces x/y/z, nodes, final faces x/y/z, bx) and the
 * box order of the batched launches (0 level order, 1 largest first).  No
 * reference counterpart: set by the autotuner (camr_amd/autotune.py). */
#define CAMR_NSTAGE 13
int camr_level_set_tuning(camr_level *L, const int32_t block[CAMR_NSTAGE], int32_t order);
int camr_level_get_tuning(camr_level *L, int32_t block[CAMR_NSTAGE], int32_t *order);
/* Block shapes of the stage ranges: tile[3*s .. 3*s+2] = (tx, ty, tz) points
 * per block of stage range s, powers of two with tx >= 8 and tx*ty*tz in
 * {64, 128, 256} (that product becomes the stage's block size), or (0, 0, 0)
 * for blocks of consecutive points in x-fastest order.  A tile keeps the
 * y and z stencil neighbours of a block's points inside the block.
 * camr_level_set_tuning resets every stage to (0, 0, 0).  No reference
 * counterpart: set by the autotuner. */
int camr_level_set_tiles(camr_level *L, const int32_t tile[3 * CAMR_NSTAGE]);
int camr_level_get_tiles(camr_level *L, int32_t tile[3 * CAMR_NSTAGE]);
/* XCD chunk of each stage range: 0 keeps the hardware's round-robin of
 * workgroups over the eight XCDs; K > 0 permutes each launch's blocks so
 * that every XCD takes K consecutive blocks of each window of 8K, keeping a
 * block's neighbouring rows in its own XCD's L2.  Results do not depend on
 * it.  camr_level_set_tuning resets it to 16.  No reference counterpart:
 * set by the autotuner. */
int camr_level_set_xcd(camr_level *L, const int32_t chunk[CAMR_NSTAGE]);
int camr_level_get_xcd(camr_level *L, int32_t chunk[CAMR_NSTAGE]);
/* z-march of each stage range: 0 = one point per thread; P > 0 = each
 * thread walks P consecutive z planes of its (x, y) column, carrying what a
 * face shares with the face below in registers.  Only the z-face transverse
 * stages (transdo pairs and transdd, stage ranges 7 and 11) with linear
 * blocks take it; camr_level_set_tiles clears it for tiled stages.  Results
 * do not depend on it.  camr_level_set_tuning resets it to 2 for those two
 * stages and 0 elsewhere.  No reference counterpart: set by the autotuner. */
int camr_level_set_zmarch(camr_level *L, const int32_t planes[CAMR_NSTAGE]);
int camr_level_get_zmarch(camr_level *L, int32_t planes[CAMR_NSTAGE]);
/* LDS-staged CTU face kernels: on = 1 runs the transverse stages of each
 * face direction as one kernel that keeps the transdo pair fluxes in LDS
 * and registers (tiles of 32 x 8 points of the x-y plane, each block
 * walking `planes` z planes, 1..1024); on = 0 (the default, planes 8) runs
 * the staged transdo and transdd launches through memory, which measure
 * faster on MI355X.  Results do not depend on it.  camr_level_set_tuning
 * resets it to the default.  No reference counterpart: set by the
 * autotuner. */
int camr_level_set_ctu(camr_level *L, int32_t on, int32_t planes);
int camr_level_get_ctu(camr_level *L, int32_t *on, int32_t *planes);

/* Adjoint (reverse mode) of level steps, for gradients of a loss on the
 * final state with respect to the initial state.  No reference counterpart
 * (CAMR has no adjoint); for inverse and sensitivity problems.  Scope: the
 * fused engine, periodic (Interior) domain, PPM, const_grav = 0, the
 * default energy branches; or a level with an embedded boundary (staged
 * engine, PLM, FluxRedist; any boundary condition but NoSlipWall).  rho,
 * rho u, rho E, rho e carry the adjoint, species and temperature are
 * passive (zero gradient).  dt is frozen.
 *
 * camr_level_adjoint_forward: from the level's state S_0, nsteps steps with
 * the given dts, keeping device checkpoints every ckpt_every steps; the
 * level ends at S_T.  Then set dL/dS_T box by box (camr_level_adj_upload_box)
 * and call camr_level_adjoint_backward, which recomputes the steps from the
 * checkpoints, leaves dL/dS_0 in the adjoint arrays
 * (camr_level_adj_download_box) and the level at S_0. */
int camr_level_adjoint_forward(camr_level *L, int32_t nsteps, const double *dts,
                               int32_t ckpt_every);
int camr_level_adj_upload_box(camr_level *L, int32_t b, const double *host);
int camr_level_adj_download_box(camr_level *L, int32_t b, double *host);
int camr_level_adjoint_backward(camr_level *L);
/* After camr_level_adjoint_forward on a level with an embedded boundary:
 * set dL/dS_T to the gradient of L = w . F, F the force on the body at the
 * final state (camr_level_eb_force: the wall momentum flux of every cut cell
 * times dx^2, Hydro_hypflux_eb_K.H:65), in place of
 * camr_level_adj_upload_box.  No reference counterpart. */
int camr_level_adj_seed_eb_force(camr_level *L, const double w[3]);

/* Named hipEvent timers around kernel launches on the level's stream -- the
 * counterpart of the reference's BL_PROFILE regions (TinyProfiler).  With a
 * non-NULL filter only launches of that name are timed.  The report is a
 * JSON object {"name": {"count": n, "ms": total}, ...}. */
int camr_level_profile(camr_level *L, int32_t enable, const char *kernel_filter);
int camr_level_profile_report(camr_level *L, char *buf, int64_t buflen);

#ifdef __cplusplus
}
#endif
#endif /* CAMR_HYDRO_H */
