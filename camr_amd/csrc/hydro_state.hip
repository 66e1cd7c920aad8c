// hydro_state.hip -- pointwise state kernels of the level step on gfx950:
// conserved -> primitive conversion, primitive sources, clean_state
// (density floor, species normalisation, internal-energy reset and
// temperature), single-level ghost fill, source Saxpys and the CFL dt
// min-reduction.
#include "hydro_state.h"
#include "state_math.h"

namespace camr {

// ---------------------------------------------------------------- ctoprim
// hydro_ctoprim (Source/Hydro/Hydro_ctoprim.H:11-101), all QVAR / NQAUX
// components.  Cells with rho <= 0 are left untouched as in the reference.
__global__ void __launch_bounds__(kBlock)
k_ctoprim(Range r, Arr u, Arr q, Arr qa, HydroConst hc, int *status) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_ctoprim(u, q, qa, hc, status, i, j, k);
}

// hydro_srctoprim (Hydro_ctoprim.H:106-157)
__global__ void __launch_bounds__(kBlock)
k_srctoprim(Range r, Arr q, Arr qa, Arr src, Arr srcq) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const double srcrho = src(i, j, k, URHO);
    if (q(i, j, k, QRHO) > 0.) {
        const double rhoinv = 1.0 / q(i, j, k, QRHO);
        srcq(i, j, k, QRHO) = srcrho;
        srcq(i, j, k, QU) = (src(i, j, k, UMX) - q(i, j, k, QU) * srcrho) * rhoinv;
        srcq(i, j, k, QV) = (src(i, j, k, UMY) - q(i, j, k, QV) * srcrho) * rhoinv;
        srcq(i, j, k, QW) = (src(i, j, k, UMZ) - q(i, j, k, QW) * srcrho) * rhoinv;
        const double sre = src(i, j, k, UEINT);
        srcq(i, j, k, QREINT) = sre;
#pragma unroll
        for (int s = 0; s < NSPEC; ++s)
            srcq(i, j, k, QFS + s) = (src(i, j, k, UFS + s) - q(i, j, k, QFS + s) * srcrho) * rhoinv;
        srcq(i, j, k, QPRES) =
            qa(i, j, k, QDPDE) * (sre - q(i, j, k, QREINT) * srcrho * rhoinv) * rhoinv +
            qa(i, j, k, QDPDR) * srcrho;
    } else {
        srcq(i, j, k, QRHO) = 0.;
        srcq(i, j, k, QU) = 0.;
        srcq(i, j, k, QV) = 0.;
        srcq(i, j, k, QW) = 0.;
        srcq(i, j, k, QPRES) = 0.;
        srcq(i, j, k, QREINT) = 0.;
    }
}

// ---------------------------------------------------------------- estdt
// CAMR_estdt_hydro (Source/Utils/Timestep.H:18-61).  The min over cells is
// order independent; non-negative doubles order like their bit patterns,
// so the block minimum is combined with a 64-bit integer atomicMin.
__global__ void __launch_bounds__(kBlock)
k_estdt(Range r, Arr u, double dx, double dy, double dz, unsigned long long *out,
        int *status, EbFlags eb) {
    int i, j, k;
    double v = 1.7976931348623157e308;
    if (range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k) &&
        !eb.covered(i, j, k))  // Timestep.H:32-33
        v = estdt_point(u, i, j, k, dx, dy, dz);
    block_min_commit(v, out, status);
}

__device__ __forceinline__ void pt_detect(Arr S, double small_dens, int *flags, EbFlags eb, int i, int j, int k) {
    if (eb.covered(i, j, k)) return;  // CAMR.cpp:987-988
    floor_detect(S(i, j, k, URHO), small_dens, flags);
}

__global__ void __launch_bounds__(kBlock)
k_detect(Range r, Arr S, double small_dens, int *flags, EbFlags eb) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_detect(S, small_dens, flags, eb, i, j, k);
}

// enforce_min_density (CAMR.cpp:941-1045), exact sequential order of the
// reference CPU loop over one box; runs only if a floor was detected.
__device__ void floor_seq_box(Box bx, Arr S, HydroConst hc, EbFlags eb) {
    const double sd = hc.small_dens;
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        if (eb.covered(i, j, k)) continue;
        const double r0 = S(i, j, k, URHO);
        if (r0 == 0. || !(r0 < sd)) continue;
        double max_dens = r0;
        int is = i, js = j, ks = k;
        for (int kk = -1; kk <= 1; kk++)
        for (int jj = -1; jj <= 1; jj++)
        for (int ii = -1; ii <= 1; ii++) {
            if (bx.contains(i + ii, j + jj, k + kk) && S(i + ii, j + jj, k + kk, URHO) > max_dens) {
                is = i + ii; js = j + jj; ks = k + kk;
                max_dens = S(is, js, ks, URHO);
            }
        }
        if (max_dens < sd) {
            const double rho = S(i, j, k, URHO);
            const double small_e = eos_E_from_T(hc.small_temp, S(i, j, k, UFS) / rho,
                                                S(i, j, k, UFS + 1) / rho, S(i, j, k, UFS + 2) / rho);
            S(i, j, k, URHO) = sd;
            S(i, j, k, UTEMP) = hc.small_temp;
            S(i, j, k, UMX) = 0.;
            S(i, j, k, UMY) = 0.;
            S(i, j, k, UMZ) = 0.;
            S(i, j, k, UEINT) = sd * small_e;
            S(i, j, k, UEDEN) = S(i, j, k, UEINT);
        } else {
            for (int n = 0; n < NVAR; n++) S(i, j, k, n) = S(is, js, ks, n);
        }
    }
}

__global__ void k_floor_seq(Box bx, Arr S, HydroConst hc, int *flags, EbFlags eb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!(*flags & kFloorFlag)) return;
    floor_seq_box(bx, S, hc, eb);
}

// normalize_species + computeTemp on valid cells (after the floor pass).
__device__ __forceinline__ void pt_clean_point(Arr S, HydroConst hc, int i, int j, int k) {
    double s[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) s[n] = S(i, j, k, n);
    normalize_species(s);
    reset_e_and_T(s, hc);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) S(i, j, k, n) = s[n];
}

__global__ void __launch_bounds__(kBlock)
k_clean_point(Range r, Arr S, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_clean_point(S, hc, i, j, k);
}

__device__ __forceinline__ void pt_normalize(Arr S, int i, int j, int k) {
    double s[NVAR];
    s[URHO] = S(i, j, k, URHO);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s[UFS + n] = S(i, j, k, UFS + n);
    normalize_species(s);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) S(i, j, k, UFS + n) = s[UFS + n];
}

__global__ void __launch_bounds__(kBlock)
k_normalize(Range r, Arr S) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_normalize(S, i, j, k);
}

__device__ __forceinline__ void pt_resetT(Arr S, HydroConst hc, int i, int j, int k) {
    double s[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) s[n] = S(i, j, k, n);
    reset_e_and_T(s, hc);
    S(i, j, k, UEDEN) = s[UEDEN];
    S(i, j, k, UEINT) = s[UEINT];
    S(i, j, k, UTEMP) = s[UTEMP];
}

__global__ void __launch_bounds__(kBlock)
k_resetT(Range r, Arr S, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_resetT(S, hc, i, j, k);
}

// enforce_consistent_e (CAMR.cpp:911-938) + computeTemp
__global__ void __launch_bounds__(kBlock)
k_init_post(Range r, Arr S, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    double s[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) s[n] = S(i, j, k, n);
    if (!hc.eb || s[URHO] > 0.) {  // EB build: CAMR.cpp:921-923
        const double rhoInv = 1.0 / s[URHO];
        const double u = s[UMX] * rhoInv, v = s[UMY] * rhoInv, w = s[UMZ] * rhoInv;
        s[UEDEN] = s[UEINT] + 0.5 * s[URHO] * (u * u + v * v + w * w);
    }
    reset_e_and_T(s, hc);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) S(i, j, k, n) = s[n];
}

// FillPatchSingleLevel of every cell of the grown box of `dst` from the
// valid data of the level's boxes (`src` table).
__device__ __forceinline__ void pt_fill_raw(Arr dst, const DevFab *src, const int *owner, DomInfo D, Arr ext,
           int has_ext, int *status, int i, int j, int k) {
    int m[3];
    unsigned odd;
    map_cell(D, i, j, k, m, &odd);
    const int ob = owner_of(D, owner, m);
    if (ob < 0 && !has_ext) { atomicOr(status, 1 << CAMR_ERR_SHAPE); return; }
    // cells no box owns take the level's exterior data
    const Arr s = ob >= 0 ? src[ob].a : ext;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        const double v = s(m[0], m[1], m[2], n);
        dst(i, j, k, n) = ((odd >> n) & 1u) ? -v : v;
    }
}

__global__ void __launch_bounds__(kBlock)
k_fill_raw(Range r, Arr dst, const DevFab *src, const int *owner, DomInfo D, Arr ext,
           int has_ext, int *status) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_fill_raw(dst, src, owner, D, ext, has_ext, status, i, j, k);
}

// FillBoundary: ghost cells of box b that are periodic / interior images of
// valid cells get those cells' values (physical-boundary ghosts untouched).
__device__ __forceinline__ void pt_resync(Box valid, Arr dst, const DevFab *src, const int *owner, DomInfo D, int i, int j, int k) {
    if (valid.contains(i, j, k)) return;
    int m[3];
    unsigned odd;
    if (map_cell(D, i, j, k, m, &odd)) return;
    const int ob = owner_of(D, owner, m);
    if (ob < 0) return;
    const Arr s = src[ob].a;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) dst(i, j, k, n) = s(m[0], m[1], m[2], n);
}

__global__ void __launch_bounds__(kBlock)
k_resync(Range r, Box valid, Arr dst, const DevFab *src, const int *owner, DomInfo D) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_resync(valid, dst, src, owner, D, i, j, k);
}

// ----------------------------------------------------- Saxpy with sources
// S_new = Sborder (MultiFab::Copy, CAMR_advance.cpp:112), then the old
// gravity source Saxpy (:115-118; GravitySrc.cpp:36-78).  Zero-valued
// sources still change -0.0 into +0.0, so they are applied literally.
__device__ __forceinline__ void pt_copy_oldsrc(Arr snew, Arr sb, Arr sold, double dt, double g, int add_grav, int i, int j, int k) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        double v = sb(i, j, k, n);
        if (add_grav) {
            double src = 0.0;
            if (n == UMZ) src = g * sold(i, j, k, URHO);
            else if (n == UEDEN) src = g * sold(i, j, k, UMZ);
            v += dt * src;
        }
        snew(i, j, k, n) = v;
    }
}

__global__ void __launch_bounds__(kBlock)
k_copy_oldsrc(Range r, Arr snew, Arr sb, Arr sold, double dt, double g, int add_grav) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_copy_oldsrc(snew, sb, sold, dt, g, add_grav, i, j, k);
}

// S_new += dt * dSdt (CAMR_advance.cpp:141) + floor detection for the
// following clean_state.
__device__ __forceinline__ void pt_add_hydro(Arr snew, Arr dsdt, double dt, double small_dens, int *flags,
            EbFlags eb, int i, int j, int k) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n) snew(i, j, k, n) += dt * dsdt(i, j, k, n);
    if (!eb.covered(i, j, k)) floor_detect(snew(i, j, k, URHO), small_dens, flags);
}

__global__ void __launch_bounds__(kBlock)
k_add_hydro(Range r, Arr snew, Arr dsdt, double dt, double small_dens, int *flags,
            EbFlags eb) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_add_hydro(snew, dsdt, dt, small_dens, flags, eb, i, j, k);
}

// New-time gravity correction source (GravitySrc.cpp:81-125) Saxpy
// (CAMR_advance.cpp:150-152) + floor detection.
__device__ __forceinline__ void pt_add_newsrc(Arr snew, Arr sold, double dt, double g, double small_dens, int *flags,
             EbFlags eb, int i, int j, int k) {
    const double s_umz = 0.5 * g * (snew(i, j, k, URHO) - sold(i, j, k, URHO));
    const double s_ued = 0.5 * g * (snew(i, j, k, UMZ) - sold(i, j, k, UMZ));
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        const double src = (n == UMZ) ? s_umz : ((n == UEDEN) ? s_ued : 0.0);
        snew(i, j, k, n) += dt * src;
    }
    if (!eb.covered(i, j, k)) floor_detect(snew(i, j, k, URHO), small_dens, flags);
}

__global__ void __launch_bounds__(kBlock)
k_add_newsrc(Range r, Arr snew, Arr sold, double dt, double g, double small_dens, int *flags,
             EbFlags eb) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    pt_add_newsrc(snew, sold, dt, g, small_dens, flags, eb, i, j, k);
}

// sources_for_hydro = 0 + 1.0 * old source on valid cells
// (CAMR_construct_hydro_source.cpp:31-34), ghosts zero; FillBoundary is a
// separate k_resync pass.
__global__ void __launch_bounds__(kBlock)
k_srchyd(Range r, Box valid, Arr dst, Arr sold, double g) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const bool v = valid.contains(i, j, k);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        double src = 0.0;
        if (v) {
            if (n == UMZ) src = g * sold(i, j, k, URHO);
            else if (n == UEDEN) src = g * sold(i, j, k, UMZ);
        }
        dst(i, j, k, n) = v ? (0.0 + 1.0 * src) : 0.0;
    }
}

__global__ void __launch_bounds__(kBlock)
k_fill_value(Range r, Arr a, int nc, double v) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    for (int n = 0; n < nc; ++n) a(i, j, k, n) = v;
}

__global__ void __launch_bounds__(kBlock)
k_fill_flat(double *p, int64_t n, double v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// one-element initialiser (a kernel rather than a memset so that the step
// stays capturable in a HIP graph)
__global__ void k_fill_u64(Range r, unsigned long long *p, unsigned long long v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// ----------------------------------------------------- device dt control
// First node of a captured level step: estTimeStep + computeNewDt (or
// computeInitialDt on the first step) of CAMR.cpp:474-676 from the CFL
// minimum of the previous step, in the host driver's operation order; then
// re-arms the minimum for this step's estimate.  A density floor flagged by
// the previous step (or a bad dt) halts the graph: every later kernel
// returns at once and the host finishes the step.
__global__ void k_step_dt(StepCtl *c, unsigned long long *dtmin, int *status, DtParams P) {
    if (threadIdx.x != 0 || c->halt) return;
    if (*status & kFloorFlag) { c->halt = 1; return; }
    double est;
    if (P.fixed_dt > 0.0) {
        est = P.fixed_dt;
    } else {
        double estdt = P.max_dt;
        const double max_dt_over_cfl = P.max_dt / P.cfl;
        double estdt_hydro = max_dt_over_cfl;
        const double red = __longlong_as_double((long long)*dtmin);
        estdt_hydro = vmin(estdt_hydro, red);
        estdt_hydro *= P.cfl;
        if (estdt_hydro < estdt) estdt = estdt_hydro;
        est = estdt;
    }
    double dt_0 = 1.7976931348623157e308;
    if (c->first) {
        dt_0 = vmin(dt_0, 1 * (P.init_shrink * est));
    } else {
        double dt_min = est;
        if (P.fixed_dt <= 0.0) dt_min = vmin(dt_min, P.change_max * c->dt);
        dt_0 = vmin(dt_0, 1 * dt_min);
    }
    const double dt_eps = 0.001 * dt_0;
    if (P.stop_time >= 0.0 && (c->time + dt_0) > (P.stop_time - dt_eps)) dt_0 = P.stop_time - c->time;
    const double dt = dt_0 / 1;
    if (!(dt > 0.0)) {
        atomicOr(status, 1 << CAMR_ERR_DT);
        c->halt = 1;
        return;
    }
    c->first = 0;
    c->dt = dt;
    if (c->nstep < c->cap) c->dts[c->nstep] = dt;
    c->nstep += 1;
    c->time += dt;
    *dtmin = 0x7fefffffffffffffULL;  // DBL_MAX for this step's estimate
}


// ---------------------------------------------- one launch over all boxes
// The same passes over every box of a table in one launch: blockIdx.y is
// the box, blockIdx.x walks the points of the box grown by g.
__device__ __forceinline__ bool batch_cell(const DevFab *tab, int g, int &b, int &i, int &j,
                                           int &k) {
    b = blockIdx.y;
    const Range r = range_of(grow(tab[b].b, g));
    return range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k);
}

// k_estdt over every box of the table in one launch (the minimum is exact,
// so the box order does not matter)
// The state of every box (NVAR dense components) to or from a flat buffer
// that holds the boxes one after another (off[b]: box b's first double).
__global__ void __launch_bounds__(kBlock)
kb_state_flat(const DevFab *S, double *flat, const int64_t *off, int to_flat) {
    const int b = blockIdx.y;
    double *p = S[b].a.p;
    double *f = flat + off[b];
    const int64_t n = (int64_t)NVAR * S[b].b.npts();
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (to_flat) f[t] = p[t];
        else p[t] = f[t];
    }
}

__global__ void __launch_bounds__(kBlock)
kb_zero_state(const DevFab *S) {
    const int b = blockIdx.y;
    double *p = S[b].a.p;
    const int64_t n = (int64_t)NVAR * S[b].b.npts();
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

__global__ void __launch_bounds__(kBlock)
kb_estdt(const DevFab *S, double dx, double dy, double dz, unsigned long long *out, int *status,
         EbFlags eb) {
    int b, i, j, k;
    double v = 1.7976931348623157e308;
    if (batch_cell(S, 0, b, i, j, k) && !eb.covered(i, j, k))
        v = estdt_point(S[b].a, i, j, k, dx, dy, dz);
    block_min_commit(v, out, status);
}

__global__ void __launch_bounds__(kBlock)
kb_detect(const DevFab *S, double small_dens, int *flags, EbFlags eb) {
    int b, i, j, k;
    if (batch_cell(S, 0, b, i, j, k)) pt_detect(S[b].a, small_dens, flags, eb, i, j, k);
}

// enforce_min_density box after box (each box independent of the others)
__global__ void kb_floor_seq(const DevFab *S, int nbox, HydroConst hc, int *flags, EbFlags eb) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || blockIdx.y != 0) return;
    if (!(*flags & kFloorFlag)) return;
    for (int b = 0; b < nbox; ++b) floor_seq_box(S[b].b, S[b].a, hc, eb);
}

__global__ void __launch_bounds__(kBlock)
kb_clean_point(const DevFab *S, HydroConst hc) {
    int b, i, j, k;
    if (batch_cell(S, 0, b, i, j, k)) pt_clean_point(S[b].a, hc, i, j, k);
}

__global__ void __launch_bounds__(kBlock)
kb_normalize(const DevFab *S) {
    int b, i, j, k;
    if (batch_cell(S, 0, b, i, j, k)) pt_normalize(S[b].a, i, j, k);
}

__global__ void __launch_bounds__(kBlock)
kb_resetT(const DevFab *S, int g, HydroConst hc) {
    int b, i, j, k;
    if (batch_cell(S, g, b, i, j, k)) pt_resetT(S[b].a, hc, i, j, k);
}

__global__ void __launch_bounds__(kBlock)
kb_fill_raw(const DevFab *dst, int g, const DevFab *src, const int *owner, DomInfo D, Arr ext,
            int has_ext, int *status) {
    int b, i, j, k;
    if (batch_cell(dst, g, b, i, j, k))
        pt_fill_raw(dst[b].a, src, owner, D, ext, has_ext, status, i, j, k);
}

__global__ void __launch_bounds__(kBlock)
kb_resync(const DevFab *dst, int g, const DevFab *src, const int *owner, DomInfo D) {
    int b, i, j, k;
    if (batch_cell(dst, g, b, i, j, k)) pt_resync(dst[b].b, dst[b].a, src, owner, D, i, j, k);
}

// MultiFab::Copy(S_new, Sborder) + old-source Saxpy (pt_copy_oldsrc),
// then S_new += dt dS/dt with floor detection (pt_add_hydro), on one cell
__global__ void __launch_bounds__(kBlock)
kb_copy_add_hydro(const DevFab *snew, const DevFab *sb, const DevFab *sold, const DevFab *dsdt,
                  const double *dtp, double g, int add_grav, double small_dens, int *flags,
                  EbFlags eb) {
    int b, i, j, k;
    if (!batch_cell(snew, 0, b, i, j, k)) return;
    pt_copy_oldsrc(snew[b].a, sb[b].a, sold[b].a, *dtp, g, add_grav, i, j, k);
    pt_add_hydro(snew[b].a, dsdt[b].a, *dtp, small_dens, flags, eb, i, j, k);
}

// Every point of grow(b, g) is the image of at most one point of the
// virtual box (the clamp owner is unique), so the adds do not collide.
__global__ void __launch_bounds__(kBlock)
kb_vscatter(const VGather *vt, const int *ids, const DevFab *sb_tab, int g, int nc,
            FlatBlocks fb) {
    int u, lb;
    flat_entry(fb, u, lb);
    const VGather &V = vt[u];
    int i, j, k;
    if (!range_point(range_of(grow(V.v, g)), (int64_t)lb * blockDim.x + threadIdx.x, i, j, k))
        return;
    const int ci = min(max(i, V.v.lo[0]), V.v.hi[0]);
    const int cj = min(max(j, V.v.lo[1]), V.v.hi[1]);
    const int ck = min(max(k, V.v.lo[2]), V.v.hi[2]);
    const int a = (ci - V.v.lo[0]) / V.len[0], r = (cj - V.v.lo[1]) / V.len[1],
              c = (ck - V.v.lo[2]) / V.len[2];
    const Arr s = sb_tab[ids[V.off + (c * V.s[1] + r) * V.s[0] + a]].a;
    for (int n = 0; n < nc; ++n) s(i, j, k, n) += V.sb(i, j, k, n);
}

// ------------------------------------- expand_state after clean_state(S_old)
// FillPatch of every point of grow(b, g) from the level's cleaned S_old
// (pt_fill_raw), the floor detection and normalize_species of the valid
// cells, FillBoundary (pt_resync: interior and periodic ghost cells take
// their owner's normalized valid cell) and computeTemp (pt_resetT) on one
// point in registers.  Exact when src was cleaned: enforce_min_density
// leaves no cell with 0 < rho < small_dens (a fixed cell takes a
// neighbour's state of density >= small_dens, or the small state), so the
// floor pass between the detection and normalize_species finds nothing.
__global__ void __launch_bounds__(kBlock)
kb_expand_cleaned(const DevFab *dst, int g, const DevFab *src, const int *owner, DomInfo D,
                  Arr ext, int has_ext, HydroConst hc, int *status, EbFlags eb, FlatBlocks fb,
                  int *redo) {
    int b, lb, i, j, k;
    flat_entry(fb, b, lb);
    if (!range_point(range_of(grow(dst[b].b, g)), (int64_t)lb * blockDim.x + threadIdx.x, i, j, k))
        return;
    const Arr S = dst[b].a;
    int m[3];
    unsigned odd;
    const bool phys = map_cell(D, i, j, k, m, &odd);
    const int ob = owner_of(D, owner, m);
    double v[NVAR];
    if (ob < 0 && !has_ext) {
        // pt_fill_raw leaves the point as it was; computeTemp still acts on it
        atomicOr(status, 1 << CAMR_ERR_SHAPE);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] = S(i, j, k, n);
    } else {
        const Arr s = ob >= 0 ? src[ob].a : ext;
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            const double x = s(m[0], m[1], m[2], n);
            v[n] = ((odd >> n) & 1u) ? -x : x;
        }
        if (dst[b].b.contains(i, j, k)) {
            if (!eb.covered(i, j, k)) {
                floor_detect(v[URHO], hc.small_dens, status);
                // an uncleaned source: the floor pass may act here
                if (redo && (v[URHO] == 0. || v[URHO] < hc.small_dens)) atomicOr(redo, 1);
            }
            normalize_species(v);
        } else if (!phys && ob >= 0) {
            normalize_species(v);
        }
    }
    reset_e_and_T(v, hc);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) S(i, j, k, n) = v[n];
}

// expand_state pass by pass (kb_fill_raw, kb_detect, kb_floor_seq,
// kb_normalize, kb_resync, kb_resetT) by one workgroup whose barriers order
// the passes; returns at once unless kb_expand_cleaned set *redo (which it
// clears).  For a source that did not go through clean_state.
__global__ void __launch_bounds__(1024)
kb_expand_redo(const DevFab *dst, int nbox, int g, const DevFab *src, const int *owner,
               DomInfo D, Arr ext, int has_ext, HydroConst hc, int *status, EbFlags eb,
               int *redo) {
    if (!*redo) return;
    auto each = [&](int gg, auto f) {
        for (int b = 0; b < nbox; ++b) {
            const Range r = range_of(grow(dst[b].b, gg));
            for (int64_t t = threadIdx.x; t < r.n; t += blockDim.x) {
                int i, j, k;
                range_point(r, t, i, j, k);
                f(b, i, j, k);
            }
        }
        __syncthreads();
    };
    each(g, [&](int b, int i, int j, int k) {
        pt_fill_raw(dst[b].a, src, owner, D, ext, has_ext, status, i, j, k);
    });
    each(0, [&](int b, int i, int j, int k) {
        pt_detect(dst[b].a, hc.small_dens, status, eb, i, j, k);
    });
    if (threadIdx.x == 0 &&
        (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kFloorFlag))
        for (int b = 0; b < nbox; ++b) floor_seq_box(dst[b].b, dst[b].a, hc, eb);
    __syncthreads();
    each(0, [&](int b, int i, int j, int k) { pt_normalize(dst[b].a, i, j, k); });
    each(g, [&](int b, int i, int j, int k) {
        pt_resync(dst[b].b, dst[b].a, dst, owner, D, i, j, k);
    });
    each(g, [&](int b, int i, int j, int k) { pt_resetT(dst[b].a, hc, i, j, k); });
    if (threadIdx.x == 0) *redo = 0;
}

// ------------------------------------------- the staged step's tail fused
// S_new = Sborder + old source (CAMR_advance.cpp:112-118), += dt dS/dt
// (:141), clean_state (:146), the new-time source and its clean_state
// (:150-154) and post_timestep's clean_state (CAMR.cpp:698-699, when
// `post`) on one cell in registers, the arithmetic of pt_copy_oldsrc,
// pt_add_hydro, pt_clean_point and pt_add_newsrc in their order.  The
// enforce_min_density passes between them act only on cells below
// small_dens; where a floor check of the sequence would fire, the cell
// sets *redo and kb_tail_redo repeats the tail pass by pass.
__device__ __forceinline__ bool below_floor(double rho, double small_dens) {
    return rho == 0. || rho < small_dens;
}

__device__ __forceinline__ double tail_point(const Arr sn, const Arr u, const Arr so,
                                             const Arr ds, double dt, double g, int add_grav,
                                             int post, int mol, const HydroConst &hc, int *redo,
                                             const EbFlags &eb, bool want_dt, int i, int j,
                                             int k) {
    const bool cov = eb.covered(i, j, k);
    double v[NVAR];
    if (mol) {
        // the second stage of the method of lines (CAMR_advance.cpp:134):
        // 0.5 Sborder + 0.5 S_old + 0.5 dt dS/dt
        const double a = 0.5 * dt;
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            double x = 0.5 * u(i, j, k, n) + 0.5 * so(i, j, k, n);
            x += a * ds(i, j, k, n);
            v[n] = x;
        }
    } else {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            double x = u(i, j, k, n);
            if (add_grav) {
                double src = 0.0;
                if (n == UMZ) src = g * so(i, j, k, URHO);
                else if (n == UEDEN) src = g * so(i, j, k, UMZ);
                x += dt * src;
            }
            v[n] = x;
        }
#pragma unroll
        for (int n = 0; n < NVAR; ++n) v[n] += dt * ds(i, j, k, n);
    }
    bool fire = !cov && below_floor(v[URHO], hc.small_dens);
    normalize_species(v);
    reset_e_and_T(v, hc);
    if (add_grav) {
        const double s_umz = 0.5 * g * (v[URHO] - so(i, j, k, URHO));
        const double s_ued = 0.5 * g * (v[UMZ] - so(i, j, k, UMZ));
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            const double src = (n == UMZ) ? s_umz : ((n == UEDEN) ? s_ued : 0.0);
            v[n] += dt * src;
        }
        fire = fire || (!cov && below_floor(v[URHO], hc.small_dens));
        normalize_species(v);
        reset_e_and_T(v, hc);
    }
    if (post) {
        fire = fire || (!cov && below_floor(v[URHO], hc.small_dens));
        normalize_species(v);
        reset_e_and_T(v, hc);
    }
#pragma unroll
    for (int n = 0; n < NVAR; ++n) sn(i, j, k, n) = v[n];
    if (fire) atomicOr(redo, 1);
    // a cell no floor acts on ends as the pass-by-pass tail leaves it (the
    // floors change only the cells they act on); kb_tail_redo estimates
    // the others
    return (want_dt && !cov && !fire) ? estdt_vals(v, hc.dx[0], hc.dx[1], hc.dx[2])
                                      : 1.7976931348623157e308;
}

// With dtmin (post only), also the CFL estimate of the new state
// (kb_estdt's minimum, exact: the minimum does not depend on the order).
__global__ void __launch_bounds__(kBlock)
kb_tail_fused(const DevFab *snew, const DevFab *sb, const DevFab *sold, const DevFab *dsdt,
              const double *dtp, double g, int add_grav, int post, int mol, HydroConst hc,
              int *redo, EbFlags eb, unsigned long long *dtmin, int *status) {
    int b, i, j, k;
    const bool in = batch_cell(snew, 0, b, i, j, k);
    double est = 1.7976931348623157e308;
    if (in) est = tail_point(snew[b].a, sb[b].a, sold[b].a, dsdt[b].a, *dtp, g, add_grav, post,
                             mol, hc, redo, eb, dtmin != nullptr, i, j, k);
    if (dtmin) block_min_commit(est, dtmin, status);
}


// The same tail pass by pass (floor detection and enforce_min_density
// included), by one workgroup whose barriers order the passes; returns at
// once unless kb_tail_fused set *redo, which it clears.
__global__ void __launch_bounds__(1024)
kb_tail_redo(const DevFab *snew, const DevFab *sb, const DevFab *sold, const DevFab *dsdt,
             int nbox, const double *dtp, double g, int add_grav, int post, int mol,
             HydroConst hc, int *flags, int *redo, EbFlags eb, unsigned long long *dtmin) {
    if (!*redo) return;
    const double dt = *dtp;
    auto each = [&](auto f) {
        for (int b = 0; b < nbox; ++b) {
            const Range r = range_of(snew[b].b);
            for (int64_t t = threadIdx.x; t < r.n; t += blockDim.x) {
                int i, j, k;
                range_point(r, t, i, j, k);
                f(b, i, j, k);
            }
        }
        __syncthreads();
    };
    auto floor = [&]() {
        if (threadIdx.x == 0 &&
            (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kFloorFlag))
            for (int b = 0; b < nbox; ++b) floor_seq_box(snew[b].b, snew[b].a, hc, eb);
        __syncthreads();
    };
    if (mol) {
        each([&](int b, int i, int j, int k) {
            const Arr S = snew[b].a, B = sb[b].a, O = sold[b].a, D = dsdt[b].a;
            const double a = 0.5 * dt;
            for (int n = 0; n < NVAR; ++n) {
                double v = 0.5 * B(i, j, k, n) + 0.5 * O(i, j, k, n);
                v += a * D(i, j, k, n);
                S(i, j, k, n) = v;
            }
            if (!eb.covered(i, j, k)) floor_detect(S(i, j, k, URHO), hc.small_dens, flags);
        });
    } else {
        each([&](int b, int i, int j, int k) {
            pt_copy_oldsrc(snew[b].a, sb[b].a, sold[b].a, dt, g, add_grav, i, j, k);
        });
        each([&](int b, int i, int j, int k) {
            pt_add_hydro(snew[b].a, dsdt[b].a, dt, hc.small_dens, flags, eb, i, j, k);
        });
    }
    floor();
    each([&](int b, int i, int j, int k) { pt_clean_point(snew[b].a, hc, i, j, k); });
    if (add_grav) {
        each([&](int b, int i, int j, int k) {
            pt_add_newsrc(snew[b].a, sold[b].a, dt, g, hc.small_dens, flags, eb, i, j, k);
        });
        floor();
        each([&](int b, int i, int j, int k) { pt_clean_point(snew[b].a, hc, i, j, k); });
    }
    if (post) {
        each([&](int b, int i, int j, int k) {
            pt_detect(snew[b].a, hc.small_dens, flags, eb, i, j, k);
        });
        floor();
        each([&](int b, int i, int j, int k) { pt_clean_point(snew[b].a, hc, i, j, k); });
    }
    if (dtmin) {
        // the CFL estimate of the repeated state, as kb_estdt
        if (threadIdx.x == 0) *dtmin = (unsigned long long)0x7fefffffffffffffULL;
        __syncthreads();
        double m = 1.7976931348623157e308;
        each([&](int b, int i, int j, int k) {
            if (eb.covered(i, j, k)) return;
            double v = estdt_point(snew[b].a, i, j, k, hc.dx[0], hc.dx[1], hc.dx[2]);
            if (!(v > 0.0) && v == v) atomicOr(flags, 1 << CAMR_ERR_DT);
            if (!(v == v)) v = 1.7976931348623157e308;
            m = vmin(m, v);
        });
        for (int off = 32; off > 0; off >>= 1) m = vmin(m, __shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0 && m > 0.0)
            atomicMin(dtmin, (unsigned long long)__double_as_longlong(m));
    }
    if (threadIdx.x == 0) *redo = 0;
}

// periodic image of an index in domain D (false outside a non-periodic
// domain) and the coarse-fine flag of a cell
__device__ __forceinline__ bool reg_map(const DomInfo &D, const int *c, int *m) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int lo = D.domlo[d], n = D.domhi[d] - lo + 1;
        int x = c[d];
        if (x < lo || x > lo + n - 1) {
            if (!D.periodic[d]) return false;
            x = lo + ((x - lo) % n + n) % n;
        }
        m[d] = x;
    }
    return true;
}
__device__ __forceinline__ int reg_flag(const DomInfo &D, const uint8_t *flag, const int *c) {
    int m[3];
    if (!reg_map(D, c, m)) return 0;
    const int64_t n0 = D.domhi[0] - D.domlo[0] + 1, n1 = D.domhi[1] - D.domlo[1] + 1;
    return flag[(int64_t)(m[0] - D.domlo[0]) + (int64_t)(m[1] - D.domlo[1]) * n0 +
                (int64_t)(m[2] - D.domlo[2]) * n0 * n1];
}

// ------------------------------------------- coarse-fine synchronisation
// AMReX's YAFluxRegister as CAMR uses it (CAMR.H:310; CrseAdd / FineAdd at
// CAMR_construct_hydro_source.cpp:243-254; Reflux at CAMR.cpp:803-851) and
// average_down (CAMR.cpp:1047-1066), restated.  Coarse cells carry a flag
// over the coarse domain: kRegFine (covered by the fine level), kRegBnd
// (not covered, within one cell of a covered cell) or 0.
// CrseAdd over the coarse box of `d`: a boundary cell takes -/+ dt/vol_c
// times the coarse flux through each face it shares with a covered cell
// (x, then y, then z).
__global__ void __launch_bounds__(kBlock)
k_reg_crse_add(Range r, Arr d, Arr f0, Arr f1, Arr f2, const uint8_t *flag, DomInfo CD,
               const double *dtp, double fac, double dx1) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    if (reg_flag(CD, flag, c) != kRegBnd) return;
    const double dtdx = (fac * *dtp) / dx1;
    const Arr f[3] = {f0, f1, f2};
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
        int cm[3] = {i, j, k}, cp[3] = {i, j, k};
        cm[dd] -= 1;
        cp[dd] += 1;
        if (reg_flag(CD, flag, cm) == kRegFine) {
            for (int n = 0; n < NVAR; ++n) d(i, j, k, n) -= dtdx * f[dd](i, j, k, n);
        } else if (reg_flag(CD, flag, cp) == kRegFine) {
            for (int n = 0; n < NVAR; ++n) d(i, j, k, n) += dtdx * f[dd](cp[0], cp[1], cp[2], n);
        }
    }
}

// FineAdd of one fine box's D-faces on one side: r is the slab of coarse
// cells (unwrapped index) next to the coarsened box; each boundary cell
// takes +/- dt_f/(vol_f r^3) times every fine flux on its face, one
// addition per fine face, (outer, inner) = (k, j) / (k, i) / (j, i).
template <int D>
__global__ void __launch_bounds__(kBlock)
k_reg_fine_add(Range r, const DevFab *cfp, const int *cowner, DomInfo CD, const uint8_t *flag,
               Arr f, int side, int ratio, const double *dtp, double fac, double dx1f) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    int m[3];
    if (!reg_map(CD, c, m)) return;
    if (reg_flag(CD, flag, c) != kRegBnd) return;
    const int cb = owner_of(CD, cowner, m);
    if (cb < 0) return;
    const Arr d = cfp[cb].a;
    const double dtdx = (fac * *dtp) / (dx1f * (double)(ratio * ratio * ratio));
    constexpr int T0 = (D == 0) ? 1 : 0, T1 = (D == 2) ? 1 : 2;  // inner, outer
    for (int n = 0; n < NVAR; ++n) {
        double acc = d(m[0], m[1], m[2], n);
        for (int o1 = 0; o1 < ratio; ++o1)
            for (int o0 = 0; o0 < ratio; ++o0) {
                int fi[3];
                fi[D] = side ? c[D] * ratio : (c[D] + 1) * ratio;
                fi[T1] = c[T1] * ratio + o1;
                fi[T0] = c[T0] * ratio + o0;
                const double v = f(fi[0], fi[1], fi[2], n);
                acc += side ? dtdx * v : -dtdx * v;
            }
        d(m[0], m[1], m[2], n) = acc;
    }
}
template __global__ void k_reg_fine_add<0>(Range, const DevFab *, const int *, DomInfo, const uint8_t *, Arr, int, int, const double *, double, double);
template __global__ void k_reg_fine_add<1>(Range, const DevFab *, const int *, DomInfo, const uint8_t *, Arr, int, int, const double *, double, double);
template __global__ void k_reg_fine_add<2>(Range, const DevFab *, const int *, DomInfo, const uint8_t *, Arr, int, int, const double *, double, double);

// The same two adds over a table of records, one launch for every box
// (grid.y = record): CrseAdd per coarse box, FineAdd per bordering side of
// every fine box.  Each coarse cell still takes its contributions from one
// record, in the same order.
__global__ void __launch_bounds__(kBlock)
kb_reg_crse_add(const RegCrseRec *recs, const uint8_t *flag, DomInfo CD, const double *dtp,
                double fac, double dx1) {
    const RegCrseRec &R = recs[blockIdx.y];
    int i, j, k;
    if (!range_point(R.r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    if (reg_flag(CD, flag, c) != kRegBnd) return;
    const double dtdx = (fac * *dtp) / dx1;
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
        int cm[3] = {i, j, k}, cp[3] = {i, j, k};
        cm[dd] -= 1;
        cp[dd] += 1;
        if (reg_flag(CD, flag, cm) == kRegFine) {
            for (int n = 0; n < NVAR; ++n) R.d(i, j, k, n) -= dtdx * R.f[dd](i, j, k, n);
        } else if (reg_flag(CD, flag, cp) == kRegFine) {
            for (int n = 0; n < NVAR; ++n) R.d(i, j, k, n) += dtdx * R.f[dd](cp[0], cp[1], cp[2], n);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
kb_reg_fine_add(const RegFineRec *recs, const DevFab *cfp, const int *cowner, DomInfo CD,
                const uint8_t *flag, int ratio, const double *dtp, double fac, double dx1f) {
    const RegFineRec &R = recs[blockIdx.y];
    int i, j, k;
    if (!range_point(R.r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    int m[3];
    if (!reg_map(CD, c, m)) return;
    if (reg_flag(CD, flag, c) != kRegBnd) return;
    const int cb = owner_of(CD, cowner, m);
    if (cb < 0) return;
    const Arr d = cfp[cb].a;
    const double dtdx = (fac * *dtp) / (dx1f * (double)(ratio * ratio * ratio));
    const int D = R.d, side = R.side;
    const int T0 = (D == 0) ? 1 : 0, T1 = (D == 2) ? 1 : 2;  // inner, outer
    for (int n = 0; n < NVAR; ++n) {
        double acc = d(m[0], m[1], m[2], n);
        for (int o1 = 0; o1 < ratio; ++o1)
            for (int o0 = 0; o0 < ratio; ++o0) {
                int fi[3];
                fi[D] = side ? c[D] * ratio : (c[D] + 1) * ratio;
                fi[T1] = c[T1] * ratio + o1;
                fi[T0] = c[T0] * ratio + o0;
                const double v = R.f(fi[0], fi[1], fi[2], n);
                acc += side ? dtdx * v : -dtdx * v;
            }
        d(m[0], m[1], m[2], n) = acc;
    }
}

// Reflux: crse_data += cfp, then S_new += crse_data, every valid cell.
__global__ void __launch_bounds__(kBlock)
kb_reflux(const DevFab *snew, const DevFab *crse, const DevFab *cfp) {
    int b, i, j, k;
    if (!batch_cell(snew, 0, b, i, j, k)) return;
    const Arr S = snew[b].a, Cr = crse[b].a, Fp = cfp[b].a;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        const double dd = Cr(i, j, k, n) + Fp(i, j, k, n);
        Cr(i, j, k, n) = dd;
        S(i, j, k, n) += dd;
    }
}

// average_down: a covered coarse cell = (1/r^3) * the sum of its children
// in (k, j, i) order.  Fine boxes are aligned to the ratio (checked by the
// host), so all children of a coarse cell lie in the box owning the first.
__global__ void __launch_bounds__(kBlock)
kb_avgdown(const DevFab *csnew, const DevFab *ftab, const int *fowner, DomInfo FD, int ratio) {
    int b, i, j, k;
    if (!batch_cell(csnew, 0, b, i, j, k)) return;
    const int f0[3] = {i * ratio, j * ratio, k * ratio};
    const int fo = owner_of(FD, fowner, f0);
    if (fo < 0) return;
    const Arr F = ftab[fo].a;
    const double volfrac = 1.0 / (double)(ratio * ratio * ratio);
    for (int n = 0; n < NVAR; ++n) {
        double c = 0.0;
        for (int kr = 0; kr < ratio; ++kr)
            for (int jr = 0; jr < ratio; ++jr)
                for (int ir = 0; ir < ratio; ++ir)
                    c += F(f0[0] + ir, f0[1] + jr, f0[2] + kr, n);
        csnew[b].a(i, j, k, n) = volfrac * c;
    }
}

// ------------------------------------------------ coarse-to-fine ghost data
// FillPatch's interpolation from the coarser level (AmrLevel::FillPatch,
// CAMR.cpp:1247; interpolater cell_cons_interp, Source/Utils/
// CAMR_setup.cpp:141-144): AMReX's CellConservativeLinear with linear
// limiting, restated -- per direction the monotonized-central slope of the
// coarse data, then one factor alpha in [0,1] for the slope vector so that
// every child lies within the min / max of the 27 coarse neighbours, into
// the fine level's exterior array.  Two passes: the coarse values staged
// densely (k_coarse_stage*), then one thread per fine cell and component
// (k_ext_*).

// The coarse value FillPatch sees at (i, j, k), components [nlo, nhi):
// its source cell through the domain's periodicity / BCs, at the fill
// time -- new (tmode 0), old (1) or ta * old + tb * new
// (StateData::InterpAddBox) -- with the reflection sign; 0 where no box
// owns the source.
__device__ __forceinline__ void stage_point(int i, int j, int k, const Arr &st, const DevFab *ctab,
                                            const int *cowner, const DomInfo &CD,
                                            const DevFab *otab, int tmode, double ta, double tb,
                                            int nlo, int nhi) {
    int m[3];
    unsigned odd;
    map_cell(CD, i, j, k, m, &odd);
    const int ob = owner_of(CD, cowner, m);
    for (int n = nlo; n < nhi; ++n) {
        double v = 0.0;
        if (ob >= 0) {
            const double vn = ctab[ob].a(m[0], m[1], m[2], n);
            if (tmode == 0) v = vn;
            else {
                const double vo = otab[ob].a(m[0], m[1], m[2], n);
                v = (tmode == 1) ? vo : ta * vo + tb * vn;
            }
        }
        st(i, j, k, n) = ((odd >> n) & 1u) ? -v : v;
    }
}

__global__ void __launch_bounds__(kBlock)
k_coarse_stage(Range r, Arr st, const DevFab *ctab, const int *cowner, DomInfo CD,
               const DevFab *otab, int tmode, double ta, double tb) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int nlo = gridDim.y > 1 ? (int)blockIdx.y : 0;
    const int nhi = gridDim.y > 1 ? (int)blockIdx.y + 1 : NVAR;
    stage_point(i, j, k, st, ctab, cowner, CD, otab, tmode, ta, tb, nlo, nhi);
}

// The same staging over a list of coarse cells (x, y, z triples), one
// component per block row.
__global__ void __launch_bounds__(kBlock)
k_coarse_stage_cells(const int *cells, int64_t ncells, Arr st, const DevFab *ctab,
                     const int *cowner, DomInfo CD, const DevFab *otab, int tmode, double ta,
                     double tb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const int n = (int)blockIdx.y;
    stage_point(cells[3 * t], cells[3 * t + 1], cells[3 * t + 2], st, ctab, cowner, CD, otab,
                tmode, ta, tb, n, n + 1);
}

// FillPatch's conservative linear interpolation of fine cell (i,j,k),
// components [nlo, nhi), from the staged coarse values.
__device__ __forceinline__ void ext_point(int i, int j, int k, const Arr &ext, const Arr &st,
                                          const DomInfo &CD, int ratio, const int *flo, int nlo,
                                          int nhi);

__global__ void __launch_bounds__(kBlock)
k_ext_from_coarse(Range r, Arr ext, Arr st, DomInfo CD, int ratio, int flo0, int flo1, int flo2,
                  const int *fowner, DomInfo FD) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    if (fowner) {
        // ghost-shell mode: r is a grown fine box; take the cell's domain
        // image and skip it when outside a non-periodic domain or owned
        int c[3] = {i, j, k}, m[3];
        if (!reg_map(FD, c, m) || owner_of(FD, fowner, m) >= 0) return;
        i = m[0];
        j = m[1];
        k = m[2];
    }
    const int flo[3] = {flo0, flo1, flo2};
    // grid.y > 1: one component per block row
    const int nlo = gridDim.y > 1 ? (int)blockIdx.y : 0;
    const int nhi = gridDim.y > 1 ? (int)blockIdx.y + 1 : NVAR;
    ext_point(i, j, k, ext, st, CD, ratio, flo, nlo, nhi);
}

// The same interpolation over a list of fine cells (x, y, z triples, each
// a domain cell no box owns), one component per block row.
__global__ void __launch_bounds__(kBlock)
k_ext_cells(const int *cells, int64_t ncells, Arr ext, Arr st, DomInfo CD, int ratio, int flo0,
            int flo1, int flo2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const int flo[3] = {flo0, flo1, flo2};
    const int n = (int)blockIdx.y;
    ext_point(cells[3 * t], cells[3 * t + 1], cells[3 * t + 2], ext, st, CD, ratio, flo, n, n + 1);
}

__device__ __forceinline__ void ext_point(int i, int j, int k, const Arr &ext, const Arr &st,
                                          const DomInfo &CD, int ratio, const int *flo, int nlo,
                                          int nhi) {
    const int f[3] = {i, j, k};
    const double half = 0.5 * (ratio - 1);
    int cc[3];
    double off[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        cc[d] = (f[d] - flo[d]) / ratio + CD.domlo[d];
        const int sub = (f[d] - flo[d]) - (cc[d] - CD.domlo[d]) * ratio;
        off[d] = (sub - half) / ratio;
    }
    const double edge = half / ratio;
#pragma unroll 1
    for (int n = nlo; n < nhi; ++n) {
        double u[27];
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int d = -1; d <= 1; ++d)
                    u[(a + 1) * 9 + (b + 1) * 3 + (d + 1)] = st(cc[0] + d, cc[1] + b, cc[2] + a, n);
        const double u0 = u[13];
        const double um[3] = {u[12], u[10], u[4]};
        const double up[3] = {u[14], u[16], u[22]};
        double s[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double cen = 0.5 * (up[d] - um[d]);
            const double forw = 2.0 * (up[d] - u0);
            const double back = 2.0 * (u0 - um[d]);
            const double slp = (forw * back >= 0.0) ? vmin(fabs(forw), fabs(back)) : 0.0;
            s[d] = copysign(1.0, cen) * vmin(slp, fabs(cen));
        }
        double cmn = u0, cmx = u0;
        for (int t = 0; t < 27; ++t) {
            cmn = vmin(cmn, u[t]);
            cmx = vmax(cmx, u[t]);
        }
        const double dmn = cmn - u0, dmx = cmx - u0;
        double alpha = 1.0;
        for (int cz = -1; cz <= 1; cz += 2)
            for (int cy = -1; cy <= 1; cy += 2)
                for (int cx = -1; cx <= 1; cx += 2) {
                    const double dd = cx * edge * s[0] + cy * edge * s[1] + cz * edge * s[2];
                    if (dd > dmx) alpha = vmin(alpha, dmx / dd);
                    if (dd < dmn) alpha = vmin(alpha, dmn / dd);
                }
        ext(i, j, k, n) = u0 + off[0] * (alpha * s[0]) + off[1] * (alpha * s[1]) + off[2] * (alpha * s[2]);
    }
}

} // namespace camr

