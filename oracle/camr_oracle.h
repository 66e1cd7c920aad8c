/*
 * camr_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C99 CPU restatement of CAMR's single-level Godunov hydro advance,
 * written to follow the reference sources line by line (each function in
 * camr_oracle.c cites the file:line it restates).  It is the parity oracle for
 * the HIP path and the timed "CPU reference path" (cpu_baseline kind "port").
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library.  The product (camr_amd/, include/) never links it.
 *
 * Layout: every array is an AMReX-Array4-compatible view: x fastest, then y,
 * then z, component outermost:
 *     p[(i-lo0) + (j-lo1)*js + (k-lo2)*ks + n*ns]
 */
#ifndef CAMR_ORACLE_H
#define CAMR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Component maps (Source/Utils/IndexDefines.H:17-92, 3-D, NUM_SPECIES=3). */
enum { O_URHO = 0, O_UMX = 1, O_UMY = 2, O_UMZ = 3, O_UEDEN = 4, O_UEINT = 5,
       O_UTEMP = 6, O_UFS = 7, O_NVAR = 10 };
enum { O_QRHO = 0, O_QU = 1, O_QV = 2, O_QW = 3, O_QGAME = 4, O_QPRES = 5,
       O_QREINT = 6, O_QTEMP = 7, O_QFS = 8, O_QVAR = 11 };
enum { O_QGAMC = 0, O_QC = 1, O_QCSML = 2, O_QDPDR = 3, O_QDPDE = 4,
       O_NQAUX = 5 };
enum { O_GDRHO = 0, O_GDU = 1, O_GDV = 2, O_GDW = 3, O_GDPRES = 4,
       O_GDGAME = 5, O_NGDNV = 6 };
/* Physical BC codes (IndexDefines.H:8-13). */
enum { O_Interior = 0, O_Inflow = 1, O_Outflow = 2, O_Symmetry = 3,
       O_SlipWall = 4, O_NoSlipWall = 5 };

typedef struct ora_fab {
    double *p;
    int lo[3], hi[3];
    int64_t js, ks, ns;
    int nc;
} ora_fab;

/* Runtime parameters on the path (Source/Params/_cpp_parameters). */
typedef struct ora_params {
    int ppm_type, plm_iorder, use_pslope, use_flattening;
    int transverse_reset_density;
    int allow_negative_energy, allow_small_energy;
    int dual_energy_update_E_from_e;
    int add_grav_src;
    double small, small_dens, small_pres, small_temp, smallu, small_num;
    double difmag, dual_energy_eta1, dual_energy_eta2, const_grav;
    double cfl, change_max, init_shrink, max_dt, fixed_dt, stop_time;
    int bclo[3], bchi[3];
    int domlo[3], domhi[3];
    int is_periodic[3];
    double dx[3];
    double prob_lo[3];
    int tile[3]; /* hydro_tile_size used by construct_hydro_source */
    int eb;      /* 1: the guards of an AMREX_USE_EB build (covered cells) */
    int do_mol;  /* 1: method-of-lines integrator (two-stage RK, MOL_umeth) */
} ora_params;

/* Cut-cell geometry of one box (the EBFArrayBoxFactory arrays of an MFIter):
 * flags over flo..fhi (a view; see camr_oracle.c for the word layout),
 * vfrac on grow(bx, 7), apertures ap[d] and face centroids fc[d] (2
 * components, the tangential offsets in [-1/2, 1/2]) on its faces. */
typedef struct ora_ebgeom {
    const int32_t *flags;
    int flo[3], fhi[3];
    int64_t fjs, fks;
    ora_fab vfrac, ap[3], fc[3];
} ora_ebgeom;

void ora_default_params(ora_params *p);

/* ---- fab helpers ---- */
int  ora_fab_alloc(ora_fab *f, const int lo[3], const int hi[3], int nc);
void ora_fab_free(ora_fab *f);

/* ---- pointwise kernels (for known-answer tests) ---- */
void ora_riemannus(double rl, double ul, double vl, double v2l, double pl,
                   double rel, double rr, double ur, double vr, double v2r,
                   double pr, double rer, double cav, double gamcl,
                   double gamcr, double small, double smallc,
                   double small_dens, double small_pres, double smallu,
                   double out[12]);
void ora_ppm_reconstruct(const double *s, double flatn, double *sm,
                         double *sp);
void ora_ppm_int_profile(double sm, double sp, double sc, double u, double c,
                         double dtdx, double *Ip, double *Im);
double ora_plm_slope(const ora_fab *q, int i, int j, int k, int n, int dir,
                     double flat, int order);
double ora_flatten(const ora_fab *q, int i, int j, int k, int dir);

/* ---- per-box stages ---- */
void ora_ctoprim_box(const int lo[3], const int hi[3], const ora_fab *u,
                     ora_fab *q, ora_fab *qa, const ora_params *p, int *err);
void ora_srctoprim_box(const int lo[3], const int hi[3], const ora_fab *q,
                       const ora_fab *qa, const ora_fab *src, ora_fab *srcq);
/* hydro_umdrv: dsdt += ..., flx[d] overwritten (extensive). */
void ora_hydro_umdrv(const int lo[3], const int hi[3], const ora_fab *uin,
                     ora_fab *dsdt, const ora_fab *q, const ora_fab *qa,
                     const ora_fab *srcq, ora_fab flx[3], double dt,
                     const ora_params *p);
double ora_estdt_box(const int lo[3], const int hi[3], const ora_fab *u,
                     const double dx[3]);
/* hydro_umdrv_eb (Godunov, FluxRedist): q/qa/srcq and uin on grow(bx, 6);
 * dsdt on bx overwritten on non-covered cells; flx[d] on faces of bx. */
void ora_hydro_umdrv_eb(const int lo[3], const int hi[3], const ora_fab *uin,
                        ora_fab *dsdt, const ora_fab *q, const ora_fab *qa,
                        const ora_fab *srcq, ora_fab flx[3], double dt,
                        const ora_params *p, const ora_ebgeom *geom);

/* ---- level driver (single level, boxes covering the domain) ---- */
typedef struct ora_level ora_level;

ora_level *ora_level_create(const ora_params *p, int nbox,
                            const int *box_lohi /* nbox*6 */);
void   ora_level_destroy(ora_level *L);
/* Values of the domain cells no box owns (NVAR comps over the domain,
 * Array4 order): the FillPatch result for such ghost cells. */
void   ora_level_set_exterior(ora_level *L, const double *data);
int    ora_level_nbox(const ora_level *L);
/* Copy a box's valid data in/out (Array4 order, NVAR components). */
void   ora_level_set_box(ora_level *L, int b, const double *data);
void   ora_level_get_box(const ora_level *L, int b, double *data);
/* CAMR::initData post-processing: enforce_consistent_e + computeTemp. */
void   ora_level_init_post(ora_level *L);
double ora_level_initial_dt(ora_level *L);
double ora_level_new_dt(ora_level *L, double dt_old);
/* One CAMR_advance + post_timestep; returns 0 or an error code. */
int    ora_level_advance(ora_level *L, double dt);
double ora_level_time(const ora_level *L);
/* Run nsteps with the reference's dt sequence; dts (may be NULL) receives
 * the dt of every step. Returns 0 or error code. */
int    ora_level_run(ora_level *L, int nsteps, double *dts);
/* EB build of the level (call right after create, before set_box): flags
 * over flo..fhi = the domain grown by 7 cells; then the cut-cell data of
 * every box whose kind is 1.  Kinds: 0 regular, 1 hydro_umdrv_eb, 2 covered. */
int    ora_level_set_eb(ora_level *L, const int32_t *flags, const int flo[3],
                        const int fhi[3]);
int    ora_level_box_kind(const ora_level *L, int b);
void   ora_level_set_box_eb(ora_level *L, int b, const double *vfrac,
                            const double *apx, const double *apy,
                            const double *apz, const double *fcx,
                            const double *fcy, const double *fcz);
/* Exterior data of fine level F from coarse level C (refinement `ratio`):
 * FillPatch's conservative linear interpolation (cell_cons_interp) at every
 * cell of F's domain.  Returns 1 if the domains do not nest. */
int    ora_level_exterior_from_coarse(ora_level *F, const ora_level *C, int ratio);
void   ora_level_get_exterior(const ora_level *L, double *out);

/* ---- coarse-fine synchronisation (flux register, reflux, average_down) -- */
typedef struct ora_fluxreg ora_fluxreg;
ora_fluxreg *ora_fluxreg_create(ora_level *C, ora_level *F, int ratio);
void   ora_fluxreg_destroy(ora_fluxreg *R);
int    ora_level_average_down(ora_level *C, const ora_level *F, int ratio);
int    ora_amr_step(ora_level *C, ora_level *F, ora_fluxreg *R, double dt_coarse);
int    ora_amr_advance(ora_level **L, ora_fluxreg **R, int nlev, double dt0);
int    ora_level_eb_force(ora_level *L, double f[3]);
int    ora_level_exterior_from_coarse_time(ora_level *F, const ora_level *C, int ratio,
                                           double t_old, double t_new, double t);
void   ora_level_set_old_box(ora_level *L, int b, const double *data);
/* Sum of each component over valid cells (for conservation checks). */
void   ora_level_sums(const ora_level *L, double sums[O_NVAR]);

/* ---- single stage functions at one point (known-answer tests) ---- */
void   ora_transdo_point(int i, int j, int k, int dir, int other_dir, ora_fab *qm,
                         ora_fab *qp, const ora_fab *qnormm, const ora_fab *qnormp,
                         const ora_fab *flxx, const ora_fab *qa, const ora_fab *qint,
                         double cdtdx, int trd, double small_pres);
void   ora_transdd_point(int i, int j, int k, int dir, ora_fab *qm, ora_fab *qp,
                         const ora_fab *qnormm, const ora_fab *qnormp, const ora_fab *flxx,
                         const ora_fab *flxy, const ora_fab *qx, const ora_fab *qy,
                         const ora_fab *qa, const ora_fab *srcq, double hdt, double cdtdx0,
                         double cdtdx1, int trd, double small_pres);
void   ora_trace_ppm_point(int i, int j, int k, int idir, const ora_fab *q, const ora_fab *qa,
                           ora_fab *qm, ora_fab *qp, double dt, const double dx[3],
                           int use_flattening, double small_dens, double small_pres);
void   ora_artif_visc_point(int i, int j, int k, ora_fab *flx, const ora_fab *divu,
                            const ora_fab *u, double dx, double difmag, int dir, int domlo,
                            int domhi, int bclo, int bchi);
void   ora_divu_node_point(int i, int j, int k, const ora_fab *q, ora_fab *divu,
                           const ora_params *P);
/* enforce_min_density over lo..hi of S in place; returns the error bits */
int    ora_enforce_min_density_box(ora_fab *S, const int lo[3], const int hi[3],
                                   const ora_params *P);

/* OpenMP thread count used by the tiled loops (0 = library default). */
void   ora_set_threads(int n);
int    ora_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
