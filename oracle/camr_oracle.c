/*
 * camr_oracle.c -- TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline).
 *
 * Plain-C99 restatement of CAMR's single-level Godunov hydro advance.  Every
 * function follows the reference file:line cited above it, keeping the
 * reference's stage order, per-tile temporaries and floating-point operation
 * order.  Build with -O3 -ffp-contract=off (no FMA contraction) so that
 * results are those of the reference's x86-64 CPU build.
 *
 * std::min(a,b) == (b < a) ? b : a   and   std::max(a,b) == (a < b) ? b : a
 * are reproduced exactly by smin/smax (NaN and signed-zero behaviour of the
 * C++ library templates, which differs from fmin/fmax).
 */
#include "camr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* constants                                                                 */
/* ------------------------------------------------------------------------ */
#define EOS_GAMMA 1.4                /* Source/EOS/GammaLaw/EOS.H:16       */
#define PC_RU 8.31446261815324e7     /* Source/EOS/PhysicsConstants.H:8    */
#define PC_AIRMW 26.93               /* Source/EOS/PhysicsConstants.H:11   */
#define NSPEC 3
#define NGROW 4                      /* CAMR::numGrow, Source/CAMR.H:455-483 */
#define EB_NGROW 6 /* numGrow for Godunov + FluxRedist, Source/CAMR.H:455-483 */

#define ERR_NEG_ENERGY 1
#define ERR_ZERO_DENS  2
#define ERR_DT         4
#define ERR_UNOWNED    8  /* a ghost cell no box owns and no exterior data */

static const int UPASS[NSPEC] = {7, 8, 9};  /* IndexDefines.cpp:4-22 */
static const int QPASS[NSPEC] = {8, 9, 10};

static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* Planted mistakes for the mutation tests (tests/test_oracle_plants.py): a
 * build with -DORA_PLANT=n carries slip n, so that the tests can show that
 * their pins catch it.  The default build (ORA_PLANT 0) carries none. */
#ifndef ORA_PLANT
#define ORA_PLANT 0
#endif

/* Stage markers of the operation-counting build (oracle/opcount.cpp), which
 * attributes every FP64 operation to the stage running; no-ops otherwise. */
#ifdef ORA_OPCOUNT
#define ORA_STAGE(s) (ora_stage = (s))
#else
#define ORA_STAGE(s) ((void)0)
#endif

#define FA(f, i, j, k, n)                                                   \
    ((f)->p[(int64_t)((i) - (f)->lo[0]) + (int64_t)((j) - (f)->lo[1]) * (f)->js + \
            (int64_t)((k) - (f)->lo[2]) * (f)->ks + (int64_t)(n) * (f)->ns])
#define FV(f, iv, n) FA(f, (iv)[0], (iv)[1], (iv)[2], n)

typedef struct { int lo[3], hi[3]; } obox;

/* Embedded-boundary flag word: bits 0..26 connectivity to the neighbour at
 * offset (ii,jj,kk), bit (ii+1)+3(jj+1)+9(kk+1); bits 27..28 the type (0
 * regular, 1 single-valued, 2 covered).  AMReX's EBCellFlag meaning in an
 * encoding of our own. */
#define FLG(G, i, j, k)                                                      \
    ((G)->flags[(int64_t)((i) - (G)->flo[0]) +                                \
                (int64_t)((j) - (G)->flo[1]) * (G)->fjs +                     \
                (int64_t)((k) - (G)->flo[2]) * (G)->fks])
static inline int fl_type(int32_t w) { return (w >> 27) & 3; }
static inline int fl_covered(int32_t w) { return fl_type(w) == 2; }
static inline int fl_regular(int32_t w) { return fl_type(w) == 0; }
static inline int fl_single(int32_t w) { return fl_type(w) == 1; }
static inline int fl_conn(int32_t w, int ii, int jj, int kk) {
    return (w >> ((ii + 1) + 3 * (jj + 1) + 9 * (kk + 1))) & 1;
}

static obox mkbox(const int lo[3], const int hi[3]) {
    obox b; for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.hi[d] = hi[d]; }
    return b;
}
static obox bgrow(obox b, int n) {
    for (int d = 0; d < 3; ++d) { b.lo[d] -= n; b.hi[d] += n; }
    return b;
}
static obox bgrowd(obox b, int d, int n) { b.lo[d] -= n; b.hi[d] += n; return b; }
static obox bgrowhi(obox b, int d, int n) { b.hi[d] += n; return b; }
static obox bnodes(obox b, int d) { b.hi[d] += 1; return b; }

static int g_threads = 0;
void ora_set_threads(int n) { g_threads = n; }
int ora_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

void ora_default_params(ora_params *p) {
    memset(p, 0, sizeof(*p));
    /* Source/Params/_cpp_parameters */
    p->ppm_type = 0; p->plm_iorder = 4; p->use_pslope = 0;
    p->use_flattening = 1; p->transverse_reset_density = 1;
    p->allow_negative_energy = 0; p->allow_small_energy = 1;
    p->dual_energy_update_E_from_e = 1; p->add_grav_src = 1;
    p->small = 1.e-8; p->small_dens = 1.e-7; p->small_pres = 1.e-200;
    p->small_temp = 1.e2;
    p->smallu = 1.e-12; p->small_num = 1.e-8; /* CAMR_Constants.H:9-10 */
    p->difmag = 0.1; p->dual_energy_eta1 = 1.0e0; p->dual_energy_eta2 = 1.0e-4;
    p->const_grav = 0.0;
    p->cfl = 0.8; p->change_max = 1.1; p->init_shrink = 1.0;
    p->max_dt = 1.e200; p->fixed_dt = -1.0; p->stop_time = -1.0;
    p->tile[0] = 1024; p->tile[1] = 16; p->tile[2] = 16; /* CAMR.cpp:29-30 */
}

int ora_fab_alloc(ora_fab *f, const int lo[3], const int hi[3], int nc) {
    int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1,
            nz = hi[2] - lo[2] + 1;
    for (int d = 0; d < 3; ++d) { f->lo[d] = lo[d]; f->hi[d] = hi[d]; }
    f->js = nx; f->ks = nx * ny; f->ns = nx * ny * nz; f->nc = nc;
    f->p = (double *)calloc((size_t)(f->ns * nc > 0 ? f->ns * nc : 1),
                            sizeof(double));
    return f->p ? 0 : -1;
}
static int fab_alloc_box(ora_fab *f, obox b, int nc) {
    return ora_fab_alloc(f, b.lo, b.hi, nc);
}
void ora_fab_free(ora_fab *f) { free(f->p); f->p = NULL; }

static int fab_contains(const ora_fab *f, const int iv[3]) {
    for (int d = 0; d < 3; ++d)
        if (iv[d] < f->lo[d] || iv[d] > f->hi[d]) return 0;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* EOS (Source/EOS/GammaLaw/EOS.H)                                           */
/* ------------------------------------------------------------------------ */
/* EY2T / REY2T, EOS.H:118-131 */
static inline double eos_rey2t(double e) {
    double poverrho = (EOS_GAMMA - 1.0) * e;
    return poverrho * PC_AIRMW / PC_RU;
}
/* RTY2E -> TY2E -> T2Ei, EOS.H:27-33,266-281 */
static inline double eos_rty2e(double T, const double Y[NSPEC]) {
    double wbar = PC_AIRMW;
    const double Cv = PC_RU / (wbar * (EOS_GAMMA - 1.0));
    double ei[NSPEC];
    for (int n = 0; n < NSPEC; n++) ei[n] = Cv * T;
    double E = 0.0;
    for (int n = 0; n < NSPEC; n++) E += Y[n] * ei[n];
    return E;
}

/* ------------------------------------------------------------------------ */
/* hydro_ctoprim, Source/Hydro/Hydro_ctoprim.H:11-101                        */
/* ------------------------------------------------------------------------ */
static void ctoprim_cell(int i, int j, int k, const ora_fab *u, ora_fab *q,
                         ora_fab *qa, double small_num, double eta1,
                         int allow_neg, int *err) {
    const double rho = FA(u, i, j, k, O_URHO);
    if (rho > 0.) {
        const double rhoinv = 1.0 / rho;
        const double vx = FA(u, i, j, k, O_UMX) * rhoinv;
        const double vy = FA(u, i, j, k, O_UMY) * rhoinv;
        const double vz = FA(u, i, j, k, O_UMZ) * rhoinv;
        const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
        FA(q, i, j, k, O_QRHO) = rho;
        FA(q, i, j, k, O_QU) = vx;
        FA(q, i, j, k, O_QV) = vy;
        FA(q, i, j, k, O_QW) = vz;
        for (int ip = 0; ip < NSPEC; ++ip)
            FA(q, i, j, k, QPASS[ip]) = FA(u, i, j, k, UPASS[ip]) * rhoinv;
        const double ueden = FA(u, i, j, k, O_UEDEN);
        if (!allow_neg && !(ueden > 0.)) *err |= ERR_NEG_ENERGY;
        if ((ueden - kineng) / ueden > eta1)
            FA(q, i, j, k, O_QREINT) = (ueden - kineng) * rhoinv;
        else
            FA(q, i, j, k, O_QREINT) = FA(u, i, j, k, O_UEINT) * rhoinv;
        const double e = FA(q, i, j, k, O_QREINT);
        if (!allow_neg && !(FA(u, i, j, k, O_UEINT) > 0.)) *err |= ERR_NEG_ENERGY;
        for (int sp = 0; sp < NSPEC; ++sp)
            FA(q, i, j, k, sp + O_QFS) = FA(u, i, j, k, O_UFS + sp) * rhoinv;
        /* EOS::REY2T, REY2P, REY2Gam, REY2dpde, REY2dpdr_e */
        const double T = eos_rey2t(e);
        const double p = (EOS_GAMMA - 1.0) * rho * e;
        const double gam1 = EOS_GAMMA;
        const double dpde = (EOS_GAMMA - 1.0) * rho;
        const double dpdr_e = (EOS_GAMMA - 1.0) * e;
        FA(q, i, j, k, O_QTEMP) = T;
        FA(q, i, j, k, O_QREINT) = e * rho;
        FA(q, i, j, k, O_QPRES) = p;
        FA(q, i, j, k, O_QGAME) = p / (e * rho) + 1.0;
        FA(qa, i, j, k, O_QDPDR) = dpdr_e;
        FA(qa, i, j, k, O_QDPDE) = dpde;
        FA(qa, i, j, k, O_QGAMC) = gam1;
        const double cs = sqrt(gam1 * p / rho);
        FA(qa, i, j, k, O_QC) = cs;
        FA(qa, i, j, k, O_QCSML) = smax(small_num, small_num * cs);
    }
}

void ora_ctoprim_box(const int lo[3], const int hi[3], const ora_fab *u,
                     ora_fab *q, ora_fab *qa, const ora_params *p, int *err) {
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i)
                ctoprim_cell(i, j, k, u, q, qa, p->small_num,
                             p->dual_energy_eta1, p->allow_negative_energy, err);
}

/* hydro_srctoprim, Hydro_ctoprim.H:106-157 */
static void srctoprim_cell(int i, int j, int k, const ora_fab *q,
                           const ora_fab *qa, const ora_fab *src,
                           ora_fab *srcq) {
    const double srcrho = FA(src, i, j, k, O_URHO);
    if (FA(q, i, j, k, O_QRHO) > 0.) {
        const double rhoinv = 1.0 / FA(q, i, j, k, O_QRHO);
        const double vx = FA(q, i, j, k, O_QU);
        const double vy = FA(q, i, j, k, O_QV);
        const double vz = FA(q, i, j, k, O_QW);
        const double smx = FA(src, i, j, k, O_UMX);
        const double smy = FA(src, i, j, k, O_UMY);
        const double smz = FA(src, i, j, k, O_UMZ);
        FA(srcq, i, j, k, O_QRHO) = srcrho;
        FA(srcq, i, j, k, O_QU) = (smx - vx * srcrho) * rhoinv;
        FA(srcq, i, j, k, O_QV) = (smy - vy * srcrho) * rhoinv;
        FA(srcq, i, j, k, O_QW) = (smz - vz * srcrho) * rhoinv;
        FA(srcq, i, j, k, O_QREINT) = FA(src, i, j, k, O_UEINT);
        for (int ip = 0; ip < NSPEC; ++ip) {
            const int n = UPASS[ip], nq = QPASS[ip];
            FA(srcq, i, j, k, nq) =
                (FA(src, i, j, k, n) - FA(q, i, j, k, nq) * srcrho) * rhoinv;
        }
        FA(srcq, i, j, k, O_QPRES) =
            FA(qa, i, j, k, O_QDPDE) *
                (FA(srcq, i, j, k, O_QREINT) -
                 FA(q, i, j, k, O_QREINT) * srcrho * rhoinv) * rhoinv +
            FA(qa, i, j, k, O_QDPDR) * srcrho;
    } else {
        FA(srcq, i, j, k, O_QRHO) = 0.;
        FA(srcq, i, j, k, O_QU) = 0.;
        FA(srcq, i, j, k, O_QV) = 0.;
        FA(srcq, i, j, k, O_QW) = 0.;
        FA(srcq, i, j, k, O_QPRES) = 0.;
        FA(srcq, i, j, k, O_QREINT) = 0.;
    }
}

void ora_srctoprim_box(const int lo[3], const int hi[3], const ora_fab *q,
                       const ora_fab *qa, const ora_fab *src, ora_fab *srcq) {
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i)
                srctoprim_cell(i, j, k, q, qa, src, srcq);
}

/* ------------------------------------------------------------------------ */
/* riemannus, Source/Hydro/Hydro_riemann.H:8-173 (3-D)                       */
/* out: ustar, f_rho, f_u, f_v, f_w, f_eden, f_eint, gd_iu, gd_iv1, gd_iv2,  */
/*      gd_pres, gd_game                                                     */
/* ------------------------------------------------------------------------ */
void ora_riemannus(double rl_in, double ul, double vl, double v2l,
                   double pl_in, double rel, double rr_in, double ur,
                   double vr, double v2r, double pr_in, double rer,
                   double cav, double gamcl, double gamcr, double small,
                   double smallc, double small_dens, double small_pres,
                   double smallu, double out[12]) {
    double wsmall = smallc * small_dens;
    double rl = smax(rl_in, small_dens);
    double pl = smax(pl_in, small_pres);
    double rr = smax(rr_in, small_dens);
    double pr = smax(pr_in, small_pres);

    double wl = smax(wsmall, sqrt(fabs(gamcl * pl * rl)));
    double wr = smax(wsmall, sqrt(fabs(gamcr * pr * rr)));

    double pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr);
    pstar = smax(pstar, small_pres);
    double ustar = ((wl * ul + wr * ur) + (pl - pr)) / (wl + wr);

    double ro, uo, po, reo, gamco;

    int mask = fabs(ustar) < smallu * 0.5 * (fabs(ul) + fabs(ur));
    ustar = mask ? 0.0 : ustar;

    if (ustar > 0.) {
        ro = rl; uo = ul; po = pl; reo = rel; gamco = gamcl;
    } else if (ustar < 0) {
        ro = rr; uo = ur; po = pr; reo = rer; gamco = gamcr;
    } else {
        ro = 0.5 * (rl + rr);
        uo = 0.5 * (ul + ur);
        po = 0.5 * (pl + pr);
        reo = 0.5 * (rel + rer);
        gamco = 0.5 * (gamcl + gamcr);
    }

    ro = smax(ro, small_dens);
    double roinv = 1.0 / ro;

    double co, co2inv, rstar, cstar, entho, estar, sgnm, spout, spin, ushock,
        frac, drho;
    co = sqrt(fabs(gamco * po * roinv));
    co = smax(smallc, co);
    co2inv = 1.0 / (co * co);

    drho = (pstar - po) * co2inv;
    rstar = ro + drho;
    rstar = smax(small_dens, rstar);

    entho = (reo + po) * roinv * co2inv;
    estar = reo + (pstar - po) * entho;
    cstar = sqrt(fabs(gamco * pstar / rstar));
    cstar = smax(cstar, smallc);

    sgnm = (ustar < 0.) ? -1.0 : 1.0;
    spout = co - sgnm * uo;
    spin = cstar - sgnm * ustar;
    ushock = 0.5 * (spin + spout);

    if (pstar - po > 0.0) {
        spin = ushock;
        spout = ushock;
    }

    double scr;
    if (spout - spin == 0.0) scr = small * cav;
    else scr = spout - spin;

    frac = (1.0 + (spout + spin) / scr) * 0.5;
    frac = smax(0.0, smin(1.0, frac));

    double qint_iv1, qint_iv2, qint_iu, qint_gdpres, qint_gdgame;
    mask = ustar > 0.0;
    qint_iv1 = mask ? vl : vr;
    qint_iv2 = mask ? v2l : v2r;

    mask = (ustar == 0.0);
    qint_iv1 = mask ? 0.5 * (vl + vr) : qint_iv1;
    qint_iv2 = mask ? 0.5 * (v2l + v2r) : qint_iv2;

    double rho_gd = frac * rstar + (1.0 - frac) * ro;
    qint_iu = frac * ustar + (1.0 - frac) * uo;
    qint_gdpres = frac * pstar + (1.0 - frac) * po;
    double re_gd = frac * estar + (1.0 - frac) * reo;

    mask = (spout < 0.0);
    rho_gd = mask ? ro : rho_gd;
    qint_iu = mask ? uo : qint_iu;
    qint_gdpres = mask ? po : qint_gdpres;
    re_gd = mask ? reo : re_gd;

    mask = (spin >= 0.0);
    rho_gd = mask ? rstar : rho_gd;
    qint_iu = mask ? ustar : qint_iu;
    qint_gdpres = mask ? pstar : qint_gdpres;
    re_gd = mask ? estar : re_gd;

    qint_gdgame = qint_gdpres / re_gd + 1.0;
    qint_gdpres = smax(qint_gdpres, small_pres);

    double uflx_rho = rho_gd * qint_iu;
    double uflx_u = uflx_rho * qint_iu + qint_gdpres;
    double uflx_v = uflx_rho * qint_iv1;
    double uflx_w = uflx_rho * qint_iv2;
    const double rhoetot =
        re_gd + 0.5 * rho_gd *
                    (qint_iu * qint_iu + qint_iv1 * qint_iv1 +
                     qint_iv2 * qint_iv2);
    double uflx_eden = qint_iu * (rhoetot + qint_gdpres);
    double uflx_eint = qint_iu * re_gd;

    out[0] = ustar; out[1] = uflx_rho; out[2] = uflx_u; out[3] = uflx_v;
    out[4] = uflx_w; out[5] = uflx_eden; out[6] = uflx_eint;
    out[7] = qint_iu; out[8] = qint_iv1; out[9] = qint_iv2;
    out[10] = qint_gdpres; out[11] = qint_gdgame;
}

/* ------------------------------------------------------------------------ */
/* hydro_cmpflx (3-D), Source/Hydro/Hydro_cmpflx.H:153-310                   */
/* Mutates ql/qr at domain faces exactly as the reference does.              */
/* ------------------------------------------------------------------------ */
static void cmpflx(int i, int j, int k, int bclo, int bchi, int domlo,
                   int domhi, ora_fab *ql, ora_fab *qr, ora_fab *flx,
                   ora_fab *qg, const ora_fab *qa, int dir, double small,
                   double small_dens, double small_pres, double smallu) {
    double cav, smallc, gamcl, gamcr;
    double ul, ur, vl, vr, v2l, v2r;
    int IU, IV, IV2, GU, GV, GV2, f_idx[3];
    int im[3] = {i, j, k};
    im[dir] -= 1;
    if (dir == 0) {
        IU = O_QU; IV = O_QV; IV2 = O_QW;
        GU = O_GDU; GV = O_GDV; GV2 = O_GDW;
        f_idx[0] = O_UMX; f_idx[1] = O_UMY; f_idx[2] = O_UMZ;
    } else if (dir == 1) {
        IU = O_QV; IV = O_QU; IV2 = O_QW;
        GU = O_GDV; GV = O_GDU; GV2 = O_GDW;
        f_idx[0] = O_UMY; f_idx[1] = O_UMX; f_idx[2] = O_UMZ;
    } else {
        IU = O_QW; IV = O_QU; IV2 = O_QV;
        GU = O_GDW; GV = O_GDU; GV2 = O_GDV;
        f_idx[0] = O_UMZ; f_idx[1] = O_UMX; f_idx[2] = O_UMY;
    }
    cav = 0.5 * (FA(qa, i, j, k, O_QC) + FV(qa, im, O_QC));
    smallc = smax(FA(qa, i, j, k, O_QCSML), FV(qa, im, O_QCSML));
    gamcl = FV(qa, im, O_QGAMC);
    gamcr = FA(qa, i, j, k, O_QGAMC);

    ul = FA(ql, i, j, k, IU);
    vl = FA(ql, i, j, k, IV);
    v2l = FA(ql, i, j, k, IV2);
    ur = FA(qr, i, j, k, IU);
    vr = FA(qr, i, j, k, IV);
    v2r = FA(qr, i, j, k, IV2);

    int idx = (dir == 2) ? k : ((dir == 0) ? i : j);

    if (idx == domlo) {
        if (bclo == O_Symmetry || bclo == O_SlipWall || bclo == O_NoSlipWall) {
            FA(ql, i, j, k, O_QRHO) = FA(qr, i, j, k, O_QRHO);
            FA(ql, i, j, k, O_QPRES) = FA(qr, i, j, k, O_QPRES);
            FA(ql, i, j, k, O_QREINT) = FA(qr, i, j, k, O_QREINT);
            gamcl = gamcr;
            ul = -ur;
            vl = vr;
            v2l = v2r;
        } else if (bclo == O_Outflow) {
            ul = ur;
            vl = vr;
            v2l = v2r;
            gamcl = gamcr;
            FA(ql, i, j, k, O_QRHO) = FA(qr, i, j, k, O_QRHO);
            FA(ql, i, j, k, O_QPRES) = FA(qr, i, j, k, O_QPRES);
            FA(ql, i, j, k, O_QREINT) = FA(qr, i, j, k, O_QREINT);
        }
    }
    if (idx == domhi + 1) {
        if (bchi == O_Symmetry || bchi == O_SlipWall || bchi == O_NoSlipWall) {
            FA(qr, i, j, k, O_QRHO) = FA(ql, i, j, k, O_QRHO);
            FA(qr, i, j, k, O_QPRES) = FA(ql, i, j, k, O_QPRES);
            FA(qr, i, j, k, O_QREINT) = FA(ql, i, j, k, O_QREINT);
            gamcr = gamcl;
            ur = -ul;
            vr = vl;
            v2r = v2l;
        } else if (bchi == O_Outflow) {
            FA(qr, i, j, k, O_QRHO) = FA(ql, i, j, k, O_QRHO);
            FA(qr, i, j, k, O_QPRES) = FA(ql, i, j, k, O_QPRES);
            FA(qr, i, j, k, O_QREINT) = FA(ql, i, j, k, O_QREINT);
            gamcr = gamcl;
            ur = ul;
            vr = vl;
            v2r = v2l;
        }
    }

    double o[12];
    ora_riemannus(FA(ql, i, j, k, O_QRHO), ul, vl, v2l,
                  FA(ql, i, j, k, O_QPRES), FA(ql, i, j, k, O_QREINT),
                  FA(qr, i, j, k, O_QRHO), ur, vr, v2r,
                  FA(qr, i, j, k, O_QPRES), FA(qr, i, j, k, O_QREINT), cav,
                  gamcl, gamcr, small, smallc, small_dens, small_pres, smallu,
                  o);
    const double ustar = o[0];
    FA(flx, i, j, k, O_URHO) = o[1];
    FA(flx, i, j, k, f_idx[0]) = o[2];
    FA(flx, i, j, k, f_idx[1]) = o[3];
    FA(flx, i, j, k, f_idx[2]) = o[4];
    FA(flx, i, j, k, O_UEDEN) = o[5];
    FA(flx, i, j, k, O_UEINT) = o[6];
    FA(qg, i, j, k, GU) = o[7];
    FA(qg, i, j, k, GV) = o[8];
    FA(qg, i, j, k, GV2) = o[9];
    FA(qg, i, j, k, O_GDPRES) = o[10];
    FA(qg, i, j, k, O_GDGAME) = o[11];

    const double flxrho = FA(flx, i, j, k, O_URHO);
    for (int ip = 0; ip < NSPEC; ++ip) {
        const int n = UPASS[ip], nqp = QPASS[ip];
        if (ustar > 0.0)
            FA(flx, i, j, k, n) = flxrho * FA(ql, i, j, k, nqp);
        else if (ustar < 0.0)
            FA(flx, i, j, k, n) = flxrho * FA(qr, i, j, k, nqp);
        else
            FA(flx, i, j, k, n) =
                flxrho * 0.50 * (FA(ql, i, j, k, nqp) + FA(qr, i, j, k, nqp));
    }
}

/* ------------------------------------------------------------------------ */
/* flatten, Source/Hydro/Godunov/flatten.H:19-99                             */
/* ------------------------------------------------------------------------ */
double ora_flatten(const ora_fab *q, int i, int j, int k, int dir) {
    int iv[3] = {i, j, k};
    int ivm3[3] = {i, j, k}, ivp3[3] = {i, j, k};
    ivm3[dir] -= 3; ivp3[dir] += 3;
    if (!fab_contains(q, ivp3) || !fab_contains(q, ivm3)) return 1.0;

    const int n = O_QPRES;
    const double small_pres = 1.e-200;
    const double shktst = 0.33e0;
    const double zcut1 = 0.75e0;
    const double zcut2 = 0.85e0;
    const double dzcut = 1.0 / (zcut2 - zcut1);

#define QOFF(off, comp) \
    FA(q, iv[0] + (dir == 0) * (off), iv[1] + (dir == 1) * (off), \
       iv[2] + (dir == 2) * (off), comp)
    const double qm = QOFF(-1, n);
    const double qp = QOFF(+1, n);
    const double qmun = QOFF(-1, O_QU + dir);
    const double qpun = QOFF(+1, O_QU + dir);
    const double qmm = QOFF(-2, n);
    const double qpp = QOFF(+2, n);

    double dp = qp - qm;
    const int ishft = (dp > 0.0) ? 1 : -1;

    double denom = smax(small_pres, fabs(qpp - qmm));
    double zeta = fabs(dp) / denom;
    const double z = smin(1.0, smax(0.0, dzcut * (zeta - zcut1)));

    double tst = (qmun - qpun >= 0.0) ? 1.0 : 0.0;
    double tmp = smin(qp, qm);
    const double chi = (fabs(dp) / tmp) > shktst ? tst : 0.0;

    const double qpshft = QOFF(1 - ishft, n);
    const double qmshft = QOFF(-(1 + ishft), n);
    const double qpunshft = QOFF(1 - ishft, O_QU + dir);
    const double qmunshft = QOFF(-(1 + ishft), O_QU + dir);
    const double qppshft = QOFF(2 - ishft, n);
    const double qmmshft = QOFF(-(2 + ishft), n);
#undef QOFF

    dp = qpshft - qmshft;
    denom = smax(small_pres, fabs(qppshft - qmmshft));
    zeta = fabs(dp) / denom;
    const double z2 = smin(1.0, smax(0.0, dzcut * (zeta - zcut1)));

    tst = (qmunshft - qpunshft >= 0.0) ? 1.0 : 0.0;
    tmp = smin(qpshft, qmshft);
    const double chi2 = (fabs(dp) / tmp) > shktst ? tst : 0.0;

    return 1.0 - smax(chi2 * z2, chi * z);
}

/* ------------------------------------------------------------------------ */
/* plm_slope / plm_pslope / hydro_plm_d, Source/Hydro/Godunov/PLM.H          */
/* ------------------------------------------------------------------------ */
#define QD(qf, iv, dir, off, comp)                                  \
    FA(qf, (iv)[0] + ((dir) == 0) * (off), (iv)[1] + ((dir) == 1) * (off), \
       (iv)[2] + ((dir) == 2) * (off), comp)

/* PLM.H:19-79 */
double ora_plm_slope(const ora_fab *q, int i, int j, int k, int n, int dir,
                     double flat, int order) {
    if (order == 1) return 0.;
    const int iv[3] = {i, j, k};
    double dlft, drgt, dcen, dfm = 0.0, dfp = 0.0, dlim, dsgn, dtemp;
    const double qc = QD(q, iv, dir, 0, n);
    const double qm2 = QD(q, iv, dir, -2, n);
    const double qm = QD(q, iv, dir, -1, n);
    const double qp = QD(q, iv, dir, +1, n);
    const double qp2 = QD(q, iv, dir, +2, n);
    if (order == 4) {
        dlft = qm - qm2;
        drgt = qc - qm;
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * smin(fabs(dlft), fabs(drgt)) : 0.0;
        dfm = dsgn * smin(dlim, fabs(dcen));

        dlft = qp - qc;
        drgt = (order == 4) ? qp2 - qp : 0.0;
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * smin(fabs(dlft), fabs(drgt)) : 0.0;
        dfp = dsgn * smin(dlim, fabs(dcen));
    }
    dlft = qc - qm;
    drgt = qp - qc;
    dcen = 0.5 * (dlft + drgt);
    dsgn = copysign(1.0, dcen);
    dlim = (dlft * drgt >= 0.0) ? 2.0 * smin(fabs(dlft), fabs(drgt)) : 0.0;
    dtemp = 4.0 / 3.0 * dcen - 1.0 / 6.0 * (dfp + dfm);
    return flat * dsgn * smin(dlim, fabs(dtemp));
}

/* PLM.H:84-165 */
static double plm_pslope(const ora_fab *q, int i, int j, int k, int n,
                         int dir, double dx, const ora_fab *srcQ, double flat,
                         int order) {
    if (order == 1) return 0.;
    const int iv[3] = {i, j, k};
    double dlft, drgt, dcen, dfm = 0.0, dfp = 0.0, dlim, dsgn, dtemp;
    const double qc = QD(q, iv, dir, 0, n);
    const double qm2 = QD(q, iv, dir, -2, n);
    const double qm = QD(q, iv, dir, -1, n);
    const double qp = QD(q, iv, dir, +1, n);
    const double qp2 = QD(q, iv, dir, +2, n);
    const int QIU = (dir == 0) ? O_QU : ((dir == 1) ? O_QV : O_QW);
#define RH(off) QD(q, iv, dir, off, O_QRHO)
#define SU(off) QD(srcQ, iv, dir, off, QIU)
    if (order == 4) {
        dlft = qm - qm2 - 0.25 * dx * (RH(-1) + RH(-2)) * (SU(-1) + SU(-2));
        drgt = qc - qm - 0.25 * dx * (RH(-1) + RH(0)) * (SU(-1) + SU(0));
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * smin(fabs(dlft), fabs(drgt)) : 0.0;
        dfm = dsgn * smin(dlim, fabs(dcen));

        dlft = qp - qc - 0.25 * dx * (RH(1) + RH(0)) * (SU(1) + SU(0));
        drgt = qp2 - qp - 0.25 * dx * (RH(1) + RH(2)) * (SU(1) + SU(2));
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * smin(fabs(dlft), fabs(drgt)) : 0.0;
        dfp = dsgn * smin(dlim, fabs(dcen));
    }
    dlft = qc - qm - 0.25 * dx * (RH(0) + RH(-1)) * (SU(0) + SU(-1));
    drgt = qp - qc - 0.25 * dx * (RH(0) + RH(1)) * (SU(0) + SU(1));
    dcen = 0.5 * (dlft + drgt);
    dsgn = copysign(1.0, dcen);
    dlim = (dlft * drgt >= 0.0) ? 2.0 * smin(fabs(dlft), fabs(drgt)) : 0.0;
    dtemp = 4.0 / 3.0 * dcen - 1.0 / 6.0 * (dfp + dfm);
    double dp = flat * dsgn * smin(dlim, fabs(dtemp));
    dp += RH(0) * SU(0) * dx;
#undef RH
#undef SU
    return dp;
}

/* PLM.H:170-323; with `area` (face apertures of direction dir) it is
 * hydro_plm_d_eb, PLM_eb.H:103-287: the state of a closed face is zero. */
static void plm_d(int i, int j, int k, int dir, ora_fab *qm, ora_fab *qp,
                  const double slope[O_QVAR], const ora_fab *q, double cc,
                  double dx, double dt, double small_dens, double small_pres,
                  const ora_fab *area) {
    const int iv[3] = {i, j, k};
    int ivp[3] = {i, j, k};
    ivp[dir] += 1;
    const int open_lo = area ? (FV(area, iv, 0) > 0.0) : 1;
    const int open_hi = area ? (FV(area, ivp, 0) > 0.0) : 1;
    int l_idx[3];
    l_idx[0] = dir;
    l_idx[1] = (dir == 0) ? 1 : 0;
    l_idx[2] = (dir == 2) ? 1 : 2;

    const double dtdx = dt / dx;
    const double cs = cc * cc;
    const double rho = FV(q, iv, O_QRHO);
    const double vel[3] = {FV(q, iv, O_QU), FV(q, iv, O_QV), FV(q, iv, O_QW)};
    const double p = FV(q, iv, O_QPRES);
    const double rhoe = FV(q, iv, O_QREINT);
    const double enth = ((rhoe + p) / rho) / cs;
    const double drho = slope[O_QRHO];
    const double dvel[3] = {slope[O_QU], slope[O_QV], slope[O_QW]};
    const double dp = slope[O_QPRES];
    const double drhoe = slope[O_QREINT];
    const double alpham = 0.5 * (dp / (rho * cc) - dvel[dir]) * rho / cc;
    const double alphap = 0.5 * (dp / (rho * cc) + dvel[dir]) * rho / cc;
    const double alpha0r = drho - dp / cs;
    const double alpha0e = drhoe - dp * enth;
    const double alpha0v = dvel[l_idx[1]];
    const double alpha0w = dvel[l_idx[2]];
    const double wv[3] = {vel[dir] - cc, vel[dir], vel[dir] + cc};

    if (!open_lo) {
        for (int n = 0; n < O_QVAR; n++) FV(qp, iv, n) = 0.0;
    } else { /* right state on the i-1/2 interface */
        const double rho_ref =
            rho - 0.5 * (1.0 + dtdx * smin(wv[0], 0.0)) * drho;
        double vel_ref[3];
        for (int dim = 0; dim < 3; dim++)
            vel_ref[dim] =
                vel[dim] - 0.5 * (1.0 + dtdx * smin(wv[0], 0.0)) * dvel[dim];
        const double p_ref = p - 0.5 * (1.0 + dtdx * smin(wv[0], 0.0)) * dp;
        const double rhoe_ref =
            rhoe - 0.5 * (1.0 + dtdx * smin(wv[0], 0.0)) * drhoe;
        const double apright = 0.25 * dtdx * (wv[0] - wv[2]) *
                               (1.0 - copysign(1.0, wv[2])) * alphap;
        const double amright = 0.0;
        const double azrright = 0.25 * dtdx * (wv[0] - wv[1]) *
                                (1.0 - copysign(1.0, wv[1])) * alpha0r;
        const double azeright = 0.25 * dtdx * (wv[0] - wv[1]) *
                                (1.0 - copysign(1.0, wv[1])) * alpha0e;
        const double azv1rght = 0.25 * dtdx * (wv[0] - wv[1]) *
                                (1.0 - copysign(1.0, wv[1])) * alpha0v;
        const double azw1rght = 0.25 * dtdx * (wv[0] - wv[1]) *
                                (1.0 - copysign(1.0, wv[1])) * alpha0w;
        FV(qp, iv, O_QRHO) = rho_ref + apright + amright + azrright;
        FV(qp, iv, O_QRHO) = smax(FV(qp, iv, O_QRHO), small_dens);
        FV(qp, iv, O_QU + l_idx[0]) =
            vel_ref[l_idx[0]] + (apright - amright) * cc / rho;
        FV(qp, iv, O_QU + l_idx[1]) = 0.;
        FV(qp, iv, O_QU + l_idx[2]) = 0.;
        FV(qp, iv, O_QU + l_idx[1]) = vel_ref[l_idx[1]] + azv1rght;
        FV(qp, iv, O_QU + l_idx[2]) = vel_ref[l_idx[2]] + azw1rght;
        FV(qp, iv, O_QPRES) = p_ref + (apright + amright) * cs;
        FV(qp, iv, O_QPRES) = smax(FV(qp, iv, O_QPRES), small_pres);
        FV(qp, iv, O_QREINT) =
            rhoe_ref + (apright + amright) * enth * cs + azeright;
    }
    if (!open_hi) {
        for (int n = 0; n < O_QVAR; n++) FV(qm, ivp, n) = 0.0;
    } else { /* left state on the i+1/2 interface */
        const double rho_ref =
            rho + 0.5 * (1.0 - dtdx * smax(wv[2], 0.0)) * drho;
        double vel_ref[3];
        for (int dim = 0; dim < 3; dim++)
            vel_ref[dim] =
                vel[dim] + 0.5 * (1.0 - dtdx * smax(wv[2], 0.0)) * dvel[dim];
        const double p_ref = p + 0.5 * (1.0 - dtdx * smax(wv[2], 0.0)) * dp;
        const double rhoe_ref =
            rhoe + 0.5 * (1.0 - dtdx * smax(wv[2], 0.0)) * drhoe;
        const double apleft = 0.0;
        const double amleft = 0.25 * dtdx * (wv[2] - wv[0]) *
                              (1.0 + copysign(1.0, wv[0])) * alpham;
        const double azrleft = 0.25 * dtdx * (wv[2] - wv[1]) *
                               (1.0 + copysign(1.0, wv[1])) * alpha0r;
        const double azeleft = 0.25 * dtdx * (wv[2] - wv[1]) *
                               (1.0 + copysign(1.0, wv[1])) * alpha0e;
        const double azv1left = 0.25 * dtdx * (wv[2] - wv[1]) *
                                (1.0 + copysign(1.0, wv[1])) * alpha0v;
        const double azw1left = 0.25 * dtdx * (wv[2] - wv[1]) *
                                (1.0 + copysign(1.0, wv[1])) * alpha0w;
        FV(qm, ivp, O_QRHO) = rho_ref + apleft + amleft + azrleft;
        FV(qm, ivp, O_QRHO) = smax(FV(qm, ivp, O_QRHO), small_dens);
        FV(qm, ivp, O_QU + l_idx[0]) =
            vel_ref[l_idx[0]] + (apleft - amleft) * cc / rho;
        FV(qm, ivp, O_QU + l_idx[1]) = 0.;
        FV(qm, ivp, O_QU + l_idx[2]) = 0.;
        FV(qm, ivp, O_QU + l_idx[1]) = vel_ref[l_idx[1]] + azv1left;
        FV(qm, ivp, O_QU + l_idx[2]) = vel_ref[l_idx[2]] + azw1left;
        FV(qm, ivp, O_QPRES) = p_ref + (apleft + amleft) * cs;
        FV(qm, ivp, O_QPRES) = smax(FV(qm, ivp, O_QPRES), small_pres);
        FV(qm, ivp, O_QREINT) =
            rhoe_ref + (apleft + amleft) * enth * cs + azeleft;
    }
    /* upwind the passive variables */
    for (int ip = 0; ip < NSPEC; ++ip) {
        const int n = QPASS[ip];
        const double vel_ad = FV(q, iv, O_QU + l_idx[0]);
        if (open_lo) {
            const double spzero = (vel_ad > 0) ? -1.0 : vel_ad * dtdx;
            const double acmprght = 0.5 * (-1.0 - spzero) * slope[n];
            FV(qp, iv, n) = FV(q, iv, n) + acmprght;
        } else {
            FV(qp, iv, n) = 0.0;
        }
        if (open_hi) {
            const double spzero = vel_ad >= 0 ? vel_ad * dtdx : 1.0;
            const double acmpleft = 0.5 * (1.0 - spzero) * slope[n];
            FV(qm, ivp, n) = FV(q, iv, n) + acmpleft;
        } else {
            FV(qm, ivp, n) = 0.0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* PPM, Source/Hydro/Godunov/PPM.H:25-183 and PPM.cpp:11-411                 */
/* ------------------------------------------------------------------------ */
void ora_ppm_reconstruct(const double *s, double flatn, double *smo,
                         double *spo) {
    enum { im2 = 0, im1 = 1, i0 = 2, ip1 = 3, ip2 = 4 };
    double sm, sp;
    double dsl = 2.0 * (s[im1] - s[im2]);
    double dsr = 2.0 * (s[i0] - s[im1]);
    double dsvl_l = 0.0;
    if (dsl * dsr > 0.0) {
        double dsc = 0.5 * (s[i0] - s[im2]);
        dsvl_l = copysign(1.0, dsc) * smin(fabs(dsc), smin(fabs(dsl), fabs(dsr)));
    }
    dsl = 2.0 * (s[i0] - s[im1]);
    dsr = 2.0 * (s[ip1] - s[i0]);
    double dsvl_r = 0.0;
    if (dsl * dsr > 0.0) {
        double dsc = 0.5 * (s[ip1] - s[im1]);
        dsvl_r = copysign(1.0, dsc) * smin(fabs(dsc), smin(fabs(dsl), fabs(dsr)));
    }
    sm = 0.5 * (s[i0] + s[im1]) - (1.0 / 6.0) * (dsvl_r - dsvl_l);
    sm = smax(sm, smin(s[i0], s[im1]));
    sm = smin(sm, smax(s[i0], s[im1]));

    dsl = 2.0 * (s[i0] - s[im1]);
    dsr = 2.0 * (s[ip1] - s[i0]);
    dsvl_l = 0.0;
    if (dsl * dsr > 0.0) {
        double dsc = 0.5 * (s[ip1] - s[im1]);
        dsvl_l = copysign(1.0, dsc) * smin(fabs(dsc), smin(fabs(dsl), fabs(dsr)));
    }
    dsl = 2.0 * (s[ip1] - s[i0]);
    dsr = 2.0 * (s[ip2] - s[ip1]);
    dsvl_r = 0.0;
    if (dsl * dsr > 0.0) {
        double dsc = 0.5 * (s[ip2] - s[i0]);
        dsvl_r = copysign(1.0, dsc) * smin(fabs(dsc), smin(fabs(dsl), fabs(dsr)));
    }
    sp = 0.5 * (s[ip1] + s[i0]) - (1.0 / 6.0) * (dsvl_r - dsvl_l);
    sp = smax(sp, smin(s[ip1], s[i0]));
    sp = smin(sp, smax(s[ip1], s[i0]));

    sm = flatn * sm + (1.0 - flatn) * s[i0];
    sp = flatn * sp + (1.0 - flatn) * s[i0];

    if ((sp - s[i0]) * (s[i0] - sm) <= 0.0) {
        sp = s[i0];
        sm = s[i0];
    } else if (fabs(sp - s[i0]) >= 2.0 * fabs(sm - s[i0])) {
        sp = 3.0 * s[i0] - 2.0 * sm;
    } else if (fabs(sm - s[i0]) >= 2.0 * fabs(sp - s[i0])) {
        sm = 3.0 * s[i0] - 2.0 * sp;
    }
    *smo = sm;
    *spo = sp;
}

void ora_ppm_int_profile(double sm, double sp, double sc, double u, double c,
                         double dtdx, double *Ip, double *Im) {
    double s6 = 6.0 * sc - 3.0 * (sm + sp);
    double speed = u - c;
    double sigma = fabs(speed) * dtdx;
    if (speed <= 0.0) {
        Ip[0] = sp;
        Im[0] = sm + 0.5 * sigma * (sp - sm + (1.0 - (2.0 / 3.0) * sigma) * s6);
    } else {
        Ip[0] = sp - 0.5 * sigma * (sp - sm - (1.0 - (2.0 / 3.0) * sigma) * s6);
        Im[0] = sm;
    }
    speed = u;
    sigma = fabs(speed) * dtdx;
    if (speed <= 0.0) {
        Ip[1] = sp;
        Im[1] = sm + 0.5 * sigma * (sp - sm + (1.0 - (2.0 / 3.0) * sigma) * s6);
    } else {
        Ip[1] = sp - 0.5 * sigma * (sp - sm - (1.0 - (2.0 / 3.0) * sigma) * s6);
        Im[1] = sm;
    }
    speed = u + c;
    sigma = fabs(speed) * dtdx;
    if (speed <= 0.0) {
        Ip[2] = sp;
        Im[2] = sm + 0.5 * sigma * (sp - sm + (1.0 - (2.0 / 3.0) * sigma) * s6);
    } else {
        Ip[2] = sp - 0.5 * sigma * (sp - sm - (1.0 - (2.0 / 3.0) * sigma) * s6);
        Im[2] = sm;
    }
}

/* trace_ppm, PPM.cpp:11-411 */
static void trace_ppm(obox bx, int idir, const ora_fab *q, const ora_fab *qa,
                      ora_fab *qm, ora_fab *qp, obox vbx, double dt,
                      const double dx[3], int use_flattening,
                      double small_dens, double small_pres) {
    const double dtdx = dt / dx[idir];
    int QUN = 0, QUT = 0, QUTT = 0;
    if (idir == 0) { QUN = O_QU; QUT = O_QV; QUTT = O_QW; }
    else if (idir == 1) { QUN = O_QV; QUT = O_QW; QUTT = O_QU; }
    else { QUN = O_QW; QUT = O_QU; QUTT = O_QV; }

    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        const int iv[3] = {i, j, k};
        int ivp1[3] = {i, j, k};
        ivp1[idir] += 1;
        double cc = sqrt(FV(qa, iv, O_QGAMC) * FV(q, iv, O_QPRES) /
                         FV(q, iv, O_QRHO));
        double un = FV(q, iv, QUN);
        double flat = 1.0;
        ORA_STAGE(8);
        if (use_flattening == 1)
            for (int dir_flat = 0; dir_flat < 3; dir_flat++)
                flat = smin(flat, ora_flatten(q, i, j, k, dir_flat));
        ORA_STAGE(1);

        double Ip[O_QVAR + 1][3];
        double Im[O_QVAR + 1][3];
        for (int n = 0; n < O_QVAR; n++) {
            double s[5];
            for (int o = -2; o <= 2; ++o) s[o + 2] = QD(q, iv, idir, o, n);
            double sm, sp;
            ora_ppm_reconstruct(s, flat, &sm, &sp);
            ora_ppm_int_profile(sm, sp, s[2], un, cc, dtdx, Ip[n], Im[n]);
        }
        {
            double s[5];
            for (int o = -2; o <= 2; ++o) s[o + 2] = QD(qa, iv, idir, o, O_QGAMC);
            double sm, sp;
            ora_ppm_reconstruct(s, flat, &sm, &sp);
            ora_ppm_int_profile(sm, sp, s[2], un, cc, dtdx, Ip[O_QVAR],
                                Im[O_QVAR]);
        }
        for (int ip = 0; ip < NSPEC; ++ip) {
            const int n = QPASS[ip];
            if (iv[idir] >= vbx.lo[idir]) FV(qp, iv, n) = Im[n][1];
            if (iv[idir] <= vbx.hi[idir]) FV(qm, ivp1, n) = Ip[n][1];
        }
        if (iv[idir] >= vbx.lo[idir]) {
            double rho_ref = Im[O_QRHO][0];
            double un_ref = Im[QUN][0];
            double p_ref = Im[O_QPRES][0];
            double rhoe_g_ref = Im[O_QREINT][0];
            double gam_ref = Im[O_QVAR][0];
            rho_ref = smax(rho_ref, small_dens);
            double rho_ref_inv = 1.0 / rho_ref;
            p_ref = smax(p_ref, small_pres);
            double cc_ref = sqrt(gam_ref * p_ref / rho_ref);
            double csq_ref = cc_ref * cc_ref;
            double cc_ref_inv = 1.0 / cc_ref;
            double h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv / csq_ref;
            double dum = un_ref - Im[QUN][0];
            double dptotm = p_ref - Im[O_QPRES][0];
            double drho = rho_ref - Im[O_QRHO][1];
            double dptot = p_ref - Im[O_QPRES][1];
            double drhoe_g = rhoe_g_ref - Im[O_QREINT][1];
            double dup = un_ref - Im[QUN][2];
            double dptotp = p_ref - Im[O_QPRES][2];
            double alpham =
                0.5 * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
            double alphap =
                0.5 * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
            double alpha0r = drho - dptot / csq_ref;
            double alpha0e_g = drhoe_g - dptot * h_g_ref;
            if (un - cc > 0.) alpham = 0.;
            else if (un - cc < 0.) alpham *= -1.0;
            else alpham *= -0.5;
            if (un + cc > 0.) alphap = 0.;
            else if (un + cc < 0.) alphap *= -1.0;
            else alphap *= -0.5;
            if (un > 0.) { alpha0r = 0.; alpha0e_g = 0.; }
            else if (un < 0.) { alpha0r *= -1.0; alpha0e_g *= -1.0; }
            else { alpha0r *= -0.5; alpha0e_g *= -0.5; }
            FV(qp, iv, O_QRHO) = rho_ref + alphap + alpham + alpha0r;
#if ORA_PLANT == 4
            FV(qp, iv, QUN) = un_ref + (alpham - alphap) * cc_ref * rho_ref_inv;
#else
            FV(qp, iv, QUN) = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
#endif
            FV(qp, iv, O_QPRES) = p_ref + (alphap + alpham) * csq_ref;
            FV(qp, iv, O_QRHO) = smax(FV(qp, iv, O_QRHO), small_dens);
            FV(qp, iv, O_QPRES) = smax(FV(qp, iv, O_QPRES), small_pres);
            FV(qp, iv, QUT) = Im[QUT][1];
            FV(qp, iv, QUTT) = Im[QUTT][1];
            FV(qp, iv, O_QREINT) =
                rhoe_g_ref + (alphap + alpham) * h_g_ref * csq_ref + alpha0e_g;
        }
        if (iv[idir] <= vbx.hi[idir]) {
            double rho_ref = Ip[O_QRHO][2];
            double un_ref = Ip[QUN][2];
            double p_ref = Ip[O_QPRES][2];
            double rhoe_g_ref = Ip[O_QREINT][2];
            double gam_ref = Ip[O_QVAR][2];
            rho_ref = smax(rho_ref, small_dens);
            double rho_ref_inv = 1.0 / rho_ref;
            p_ref = smax(p_ref, small_pres);
            double cc_ref = sqrt(gam_ref * p_ref / rho_ref);
            double csq_ref = cc_ref * cc_ref;
            double cc_ref_inv = 1.0 / cc_ref;
            double h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv / csq_ref;
            double dum = un_ref - Ip[QUN][0];
            double dptotm = p_ref - Ip[O_QPRES][0];
            double drho = rho_ref - Ip[O_QRHO][1];
            double dptot = p_ref - Ip[O_QPRES][1];
            double drhoe_g = rhoe_g_ref - Ip[O_QREINT][1];
            double dup = un_ref - Ip[QUN][2];
            double dptotp = p_ref - Ip[O_QPRES][2];
            double alpham =
                0.5 * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
            double alphap =
                0.5 * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
            double alpha0r = drho - dptot / csq_ref;
            double alpha0e_g = drhoe_g - dptot * h_g_ref;
            if (un - cc > 0.) alpham *= -1.0;
            else if (un - cc < 0.) alpham = 0.;
            else alpham *= -0.5;
            if (un + cc > 0.) alphap *= -1.0;
            else if (un + cc < 0.) alphap = 0.;
            else alphap *= -0.5;
            if (un > 0.) { alpha0r *= -1.0; alpha0e_g *= -1.0; }
            else if (un < 0.) { alpha0r = 0.; alpha0e_g = 0.; }
            else { alpha0r *= -0.5; alpha0e_g *= -0.5; }
            FV(qm, ivp1, O_QRHO) = rho_ref + alphap + alpham + alpha0r;
            FV(qm, ivp1, QUN) = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
            FV(qm, ivp1, O_QPRES) = p_ref + (alphap + alpham) * csq_ref;
            FV(qm, ivp1, O_QRHO) = smax(FV(qm, ivp1, O_QRHO), small_dens);
            /* PPM.cpp:398 clamps the minus-state pressure with small_dens */
            FV(qm, ivp1, O_QPRES) = smax(FV(qm, ivp1, O_QPRES), small_dens);
            FV(qm, ivp1, QUT) = Ip[QUT][1];
            FV(qm, ivp1, QUTT) = Ip[QUTT][1];
            FV(qm, ivp1, O_QREINT) =
                rhoe_g_ref + (alphap + alpham) * h_g_ref * csq_ref + alpha0e_g;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* hydro_transdo / hydro_transdd, Godunov_utils_3D.H:22-239, 245-482        */
/* With NULL apertures every face is open (the non-EB build); with
 * norm_area / trans_area (apertures of the faces normal to dir and to
 * other_dir) it is the EB form.                                              */
/* ------------------------------------------------------------------------ */
static void transdo(int i, int j, int k, int dir, int other_dir, ora_fab *qm,
                    ora_fab *qp, const ora_fab *qnormm, const ora_fab *qnormp,
                    const ora_fab *flxx, const ora_fab *qa,
                    const ora_fab *qint, double cdtdx, int trd,
                    double small_pres, const ora_fab *norm_area,
                    const ora_fab *trans_area) {
    const int iv[3] = {i, j, k};
    int ivpn[3] = {i, j, k}, ivpt[3] = {i, j, k};
    ivpn[dir] += 1;
    ivpt[other_dir] += 1;
#if ORA_PLANT == 7
    const int qvidx = (dir == 0) ? O_GDU : ((dir == 1) ? O_GDV : O_GDW);
#else
    const int qvidx = (other_dir == 0) ? O_GDU : ((other_dir == 1) ? O_GDV : O_GDW);
#endif
    int no_cov_face = 1, lo_open = 1, hi_open = 1;
    if (trans_area)
        no_cov_face = (FV(trans_area, ivpt, 0) > 0.0) && (FV(trans_area, iv, 0) > 0.0);
    if (norm_area) {
        lo_open = FV(norm_area, iv, 0) > 0.0;
        hi_open = FV(norm_area, ivpn, 0) > 0.0;
    }

    const double flxrho = no_cov_face ? cdtdx * (FV(flxx, ivpt, O_URHO) - FV(flxx, iv, O_URHO)) : 0.0;
    const double flxu = no_cov_face ? cdtdx * (FV(flxx, ivpt, O_UMX) - FV(flxx, iv, O_UMX)) : 0.0;
    const double flxv = no_cov_face ? cdtdx * (FV(flxx, ivpt, O_UMY) - FV(flxx, iv, O_UMY)) : 0.0;
    const double flxw = no_cov_face ? cdtdx * (FV(flxx, ivpt, O_UMZ) - FV(flxx, iv, O_UMZ)) : 0.0;
    const double flxe = no_cov_face ? cdtdx * (FV(flxx, ivpt, O_UEDEN) - FV(flxx, iv, O_UEDEN)) : 0.0;
    const double c = FV(qa, iv, O_QGAMC);

    for (int ip = 0; ip < NSPEC; ++ip) {
        const int n = UPASS[ip], nqp = QPASS[ip];
        const double compn = no_cov_face ? cdtdx * (FV(flxx, ivpt, n) - FV(flxx, iv, n)) : 0.0;
        double rr, rrnew, compo;
        if (lo_open) {
            rr = FV(qnormp, iv, O_QRHO);
            rrnew = rr - flxrho;
            compo = rr * FV(qnormp, iv, nqp) - compn;
            FV(qp, iv, nqp) = compo / rrnew;
        } else {
            FV(qp, iv, nqp) = FV(qnormp, iv, nqp);
        }
        if (hi_open) {
            rr = FV(qnormm, ivpn, O_QRHO);
            rrnew = rr - flxrho;
            compo = rr * FV(qnormm, ivpn, nqp) - compn;
            FV(qm, ivpn, nqp) = compo / rrnew;
        } else {
            FV(qm, ivpn, nqp) = FV(qnormm, ivpn, nqp);
        }
    }

    double dAup = 0.0, pav = 0.0, dAu = 0.0;
    if (no_cov_face) {
        const double pggp = FV(qint, ivpt, O_GDPRES);
        const double pggm = FV(qint, iv, O_GDPRES);
        const double ugp = FV(qint, ivpt, qvidx);
        const double ugm = FV(qint, iv, qvidx);
        dAup = pggp * ugp - pggm * ugm;
        pav = 0.5 * (pggp + pggm);
        dAu = ugp - ugm;
    }
    if (!lo_open) {
        /* the reference leaves QW unset here; the state of a closed face
         * is never read */
        FV(qp, iv, O_QRHO) = FV(qnormp, iv, O_QRHO);
        FV(qp, iv, O_QU) = FV(qnormp, iv, O_QU);
        FV(qp, iv, O_QV) = FV(qnormp, iv, O_QV);
        FV(qp, iv, O_QW) = FV(qnormp, iv, O_QW);
        FV(qp, iv, O_QREINT) = FV(qnormp, iv, O_QREINT);
        FV(qp, iv, O_QPRES) = FV(qnormp, iv, O_QPRES);
    } else { /* QP */
        const double rrr = FV(qnormp, iv, O_QRHO);
        double rur = FV(qnormp, iv, O_QU);
        double rvr = FV(qnormp, iv, O_QV);
        double rwr = FV(qnormp, iv, O_QW);
        const double ekinr = 0.5 * rrr * (rur * rur + rvr * rvr + rwr * rwr);
        rur *= rrr;
        rvr *= rrr;
        rwr *= rrr;
        const double rer = FV(qnormp, iv, O_QREINT) + ekinr;
        double rrnewr = rrr - flxrho;
        double runewr = rur - flxu;
        double rvnewr = rvr - flxv;
        double rwnewr = rwr - flxw;
        double renewr = rer - flxe;
        int reset_state = 0;
        if (trd == 1 && rrnewr < 0.) {
            rrnewr = rrr; runewr = rur; rvnewr = rvr; rwnewr = rwr;
            renewr = rer; reset_state = 1;
        }
        FV(qp, iv, O_QRHO) = rrnewr;
        FV(qp, iv, O_QU) = runewr / rrnewr;
        FV(qp, iv, O_QV) = rvnewr / rrnewr;
        FV(qp, iv, O_QW) = rwnewr / rrnewr;
        const double rhoekinr =
            0.5 * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr) / rrnewr;
        FV(qp, iv, O_QREINT) = renewr - rhoekinr;
        if (!reset_state) {
#if ORA_PLANT == 1
            double pnewr = FV(qnormp, iv, O_QPRES) +
                           cdtdx * (dAup + pav * dAu * (c - 1.0));
#else
            double pnewr = FV(qnormp, iv, O_QPRES) -
                           cdtdx * (dAup + pav * dAu * (c - 1.0));
#endif
            FV(qp, iv, O_QPRES) = smax(pnewr, small_pres);
        } else {
            FV(qp, iv, O_QPRES) = FV(qnormp, iv, O_QPRES);
        }
    }
    if (!hi_open) {
        FV(qm, ivpn, O_QRHO) = FV(qnormm, ivpn, O_QRHO);
        FV(qm, ivpn, O_QU) = FV(qnormm, ivpn, O_QU);
        FV(qm, ivpn, O_QV) = FV(qnormm, ivpn, O_QV);
        FV(qm, ivpn, O_QW) = FV(qnormm, ivpn, O_QW);
        FV(qm, ivpn, O_QREINT) = FV(qnormm, ivpn, O_QREINT);
        FV(qm, ivpn, O_QPRES) = FV(qnormm, ivpn, O_QPRES);
    } else { /* QM */
        const double rrl = FV(qnormm, ivpn, O_QRHO);
        double rul = FV(qnormm, ivpn, O_QU);
        double rvl = FV(qnormm, ivpn, O_QV);
        double rwl = FV(qnormm, ivpn, O_QW);
        const double ekinl = 0.5 * rrl * (rul * rul + rvl * rvl + rwl * rwl);
        rul *= rrl;
        rvl *= rrl;
        rwl *= rrl;
        const double rel = FV(qnormm, ivpn, O_QREINT) + ekinl;
        double rrnewl = rrl - flxrho;
        double runewl = rul - flxu;
        double rvnewl = rvl - flxv;
        double rwnewl = rwl - flxw;
        double renewl = rel - flxe;
        int reset_state = 0;
        if (trd == 1 && rrnewl < 0.) {
            rrnewl = rrl; runewl = rul; rvnewl = rvl; rwnewl = rwl;
            renewl = rel; reset_state = 1;
        }
        FV(qm, ivpn, O_QRHO) = rrnewl;
        FV(qm, ivpn, O_QU) = runewl / rrnewl;
        FV(qm, ivpn, O_QV) = rvnewl / rrnewl;
        FV(qm, ivpn, O_QW) = rwnewl / rrnewl;
        const double rhoekinl =
            0.5 * (runewl * runewl + rvnewl * rvnewl + rwnewl * rwnewl) / rrnewl;
        FV(qm, ivpn, O_QREINT) = renewl - rhoekinl;
        if (!reset_state) {
#if ORA_PLANT == 1
            double pnewl = FV(qnormm, ivpn, O_QPRES) +
                           cdtdx * (dAup + pav * dAu * (c - 1.0));
#else
            double pnewl = FV(qnormm, ivpn, O_QPRES) -
                           cdtdx * (dAup + pav * dAu * (c - 1.0));
#endif
            FV(qm, ivpn, O_QPRES) = smax(pnewl, small_pres);
        } else {
            FV(qm, ivpn, O_QPRES) = FV(qnormm, ivpn, O_QPRES);
        }
    }
}

static void transdd(int i, int j, int k, int dir, ora_fab *qm, ora_fab *qp,
                    const ora_fab *qnormm, const ora_fab *qnormp,
                    const ora_fab *flxx, const ora_fab *flxy,
                    const ora_fab *qx, const ora_fab *qy, const ora_fab *qa,
                    const ora_fab *srcq, double hdt, double cdtdx0,
                    double cdtdx1, int trd, double small_pres,
                    const ora_fab *norm_area, const ora_fab *trans0_area,
                    const ora_fab *trans1_area) {
    const int iv[3] = {i, j, k};
    const int qvidx0 = (dir == 0) ? O_GDV : O_GDU;
    const int qvidx1 = (dir == 2) ? O_GDV : O_GDW;
    const int l1 = (dir == 0) ? 1 : 0;
    const int l2 = (dir == 2) ? 1 : 2;
    int ivpt0[3] = {i, j, k}, ivpt1[3] = {i, j, k}, ivpn[3] = {i, j, k};
    ivpt0[l1] += 1;
    ivpt1[l2] += 1;
    ivpn[dir] += 1;
    int no_cov_face = 1, lo_open = 1, hi_open = 1;
    if (trans0_area && trans1_area)
        no_cov_face = (FV(trans0_area, ivpt0, 0) > 0.0) && (FV(trans1_area, ivpt1, 0) > 0.0) &&
                      (FV(trans0_area, iv, 0) > 0.0) && (FV(trans1_area, iv, 0) > 0.0);
    if (norm_area) {
        lo_open = FV(norm_area, iv, 0) > 0.0;
        hi_open = FV(norm_area, ivpn, 0) > 0.0;
    }

#define TF(comp) (no_cov_face ? cdtdx0 * (FV(flxx, ivpt0, comp) - FV(flxx, iv, comp)) + \
                  cdtdx1 * (FV(flxy, ivpt1, comp) - FV(flxy, iv, comp)) : 0.0)
    const double flxrho = TF(O_URHO);
    const double flxu = TF(O_UMX);
    const double flxv = TF(O_UMY);
    const double flxw = TF(O_UMZ);
    const double flxe = TF(O_UEDEN);
    const double c = FV(qa, iv, O_QGAMC);
    const double srcrho = FV(srcq, iv, O_QRHO);
    const double srcu = FV(srcq, iv, O_QU);
    const double srcv = FV(srcq, iv, O_QV);
    const double srcw = FV(srcq, iv, O_QW);
    const double srce = FV(srcq, iv, O_QREINT);
    const double srcp = FV(srcq, iv, O_QPRES);

    const double rrr = FV(qnormp, iv, O_QRHO);
    const double rrl = FV(qnormm, ivpn, O_QRHO);
    double rrnewl = rrl - flxrho;
    double rrnewr = rrr - flxrho;

    for (int ip = 0; ip < NSPEC; ++ip) {
        const int n = UPASS[ip], nqp = QPASS[ip];
        const double srcpass = FV(srcq, iv, nqp);
        const double compn = TF(n);
        double compo;
        if (lo_open) {
            compo = rrr * FV(qnormp, iv, nqp) - compn;
            FV(qp, iv, nqp) = compo / rrnewr + hdt * srcpass;
        } else {
            FV(qp, iv, nqp) = FV(qnormp, iv, nqp);
        }
        if (hi_open) {
            compo = rrl * FV(qnormm, ivpn, nqp) - compn;
            FV(qm, ivpn, nqp) = compo / rrnewl + hdt * srcpass;
        } else {
            FV(qm, ivpn, nqp) = FV(qnormm, ivpn, nqp);
        }
    }
#undef TF

    double dAupx = 0.0, pavx = 0.0, dAux = 0.0, dAupy = 0.0, pavy = 0.0, dAuy = 0.0;
    if (no_cov_face) {
        const double pggpx = FV(qx, ivpt0, O_GDPRES);
        const double pggmx = FV(qx, iv, O_GDPRES);
        const double ugpx = FV(qx, ivpt0, qvidx0);
        const double ugmx = FV(qx, iv, qvidx0);
        dAupx = pggpx * ugpx - pggmx * ugmx;
        pavx = 0.5 * (pggpx + pggmx);
        dAux = ugpx - ugmx;
        const double pggpy = FV(qy, ivpt1, O_GDPRES);
        const double pggmy = FV(qy, iv, O_GDPRES);
        const double ugpy = FV(qy, ivpt1, qvidx1);
        const double ugmy = FV(qy, iv, qvidx1);
        dAupy = pggpy * ugpy - pggmy * ugmy;
        pavy = 0.5 * (pggpy + pggmy);
        dAuy = ugpy - ugmy;
    }
    const double pxnew = cdtdx0 * (dAupx + pavx * dAux * (c - 1.0));
#if ORA_PLANT == 8
    const double pynew = 0.0 * cdtdx1 * (dAupy + pavy * dAuy * (c - 1.0));
#else
    const double pynew = cdtdx1 * (dAupy + pavy * dAuy * (c - 1.0));
#endif

    if (!lo_open) {
        FV(qp, iv, O_QRHO) = FV(qnormp, iv, O_QRHO);
        FV(qp, iv, O_QU) = FV(qnormp, iv, O_QU);
        FV(qp, iv, O_QV) = FV(qnormp, iv, O_QV);
        FV(qp, iv, O_QW) = FV(qnormp, iv, O_QW);
        FV(qp, iv, O_QREINT) = FV(qnormp, iv, O_QREINT);
        FV(qp, iv, O_QPRES) = FV(qnormp, iv, O_QPRES);
    } else { /* qp */
        const double rur = rrr * FV(qnormp, iv, O_QU);
        const double rvr = rrr * FV(qnormp, iv, O_QV);
        const double rwr = rrr * FV(qnormp, iv, O_QW);
        const double ekinr = 0.5 * (rur * rur + rvr * rvr + rwr * rwr) / rrr;
        const double rer = FV(qnormp, iv, O_QREINT) + ekinr;
        double runewr = rur - flxu;
        double rvnewr = rvr - flxv;
        double rwnewr = rwr - flxw;
        double renewr = rer - flxe;
        int reset_state = 0;
        if (trd == 1 && rrnewr < 0.) {
            rrnewr = rrr; runewr = rur; rvnewr = rvr; rwnewr = rwr;
            renewr = rer; reset_state = 1;
        }
        FV(qp, iv, O_QRHO) = rrnewr + hdt * srcrho;
        FV(qp, iv, O_QU) = runewr / rrnewr + hdt * srcu;
        FV(qp, iv, O_QV) = rvnewr / rrnewr + hdt * srcv;
        FV(qp, iv, O_QW) = rwnewr / rrnewr + hdt * srcw;
        const double rhoekinr =
            0.5 * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr) / rrnewr;
        FV(qp, iv, O_QREINT) = renewr - rhoekinr + hdt * srce;
        if (!reset_state)
            FV(qp, iv, O_QPRES) = FV(qnormp, iv, O_QPRES) - pxnew - pynew + hdt * srcp;
        else
            FV(qp, iv, O_QPRES) = FV(qnormp, iv, O_QPRES) + hdt * srcp;
        FV(qp, iv, O_QPRES) = smax(FV(qp, iv, O_QPRES), small_pres);
    }
    if (!hi_open) {
        FV(qm, ivpn, O_QRHO) = FV(qnormm, ivpn, O_QRHO);
        FV(qm, ivpn, O_QU) = FV(qnormm, ivpn, O_QU);
        FV(qm, ivpn, O_QV) = FV(qnormm, ivpn, O_QV);
        FV(qm, ivpn, O_QW) = FV(qnormm, ivpn, O_QW);
        FV(qm, ivpn, O_QREINT) = FV(qnormm, ivpn, O_QREINT);
        FV(qm, ivpn, O_QPRES) = FV(qnormm, ivpn, O_QPRES);
    } else { /* qm */
        const double rul = rrl * FV(qnormm, ivpn, O_QU);
        const double rvl = rrl * FV(qnormm, ivpn, O_QV);
        const double rwl = rrl * FV(qnormm, ivpn, O_QW);
        const double ekinl = 0.5 * (rul * rul + rvl * rvl + rwl * rwl) / rrl;
        const double rel = FV(qnormm, ivpn, O_QREINT) + ekinl;
        double runewl = rul - flxu;
        double rvnewl = rvl - flxv;
        double rwnewl = rwl - flxw;
        double renewl = rel - flxe;
        int reset_state = 0;
        if (trd == 1 && rrnewl < 0.) {
            rrnewl = rrl; runewl = rul; rvnewl = rvl; rwnewl = rwl;
            renewl = rel; reset_state = 1;
        }
        FV(qm, ivpn, O_QRHO) = rrnewl + hdt * srcrho;
        FV(qm, ivpn, O_QU) = runewl / rrnewl + hdt * srcu;
        FV(qm, ivpn, O_QV) = rvnewl / rrnewl + hdt * srcv;
        FV(qm, ivpn, O_QW) = rwnewl / rrnewl + hdt * srcw;
        const double rhoekinl =
            0.5 * (runewl * runewl + rvnewl * rvnewl + rwnewl * rwnewl) / rrnewl;
        FV(qm, ivpn, O_QREINT) = renewl - rhoekinl + hdt * srce;
        if (!reset_state)
            FV(qm, ivpn, O_QPRES) = FV(qnormm, ivpn, O_QPRES) - pxnew - pynew + hdt * srcp;
        else
            FV(qm, ivpn, O_QPRES) = FV(qnormm, ivpn, O_QPRES) + hdt * srcp;
        FV(qm, ivpn, O_QPRES) = smax(FV(qm, ivpn, O_QPRES), small_pres);
    }
}

/* ------------------------------------------------------------------------ */
/* Hydro_utils_K.H: hydro_divu :12-169, hydro_pdivu :175-202,                */
/* hydro_norm_spec_flx :207-219, hydro_ext_flx :224-234, hydro_update        */
/* :239-257; hydro_artif_visc.H:7-64                                          */
/* ------------------------------------------------------------------------ */
static void divu_node(int i, int j, int k, const ora_fab *q, double dx,
                      double dy, double dz, ora_fab *divu, const int domlo[3],
                      const int domhi[3], const int bclo[3],
                      const int bchi[3]) {
#define Q(a, b, c, n) FA(q, a, b, c, n)
    const int NS = O_NoSlipWall;
    double ux, vy, wz;
    ux = 0.25 * (Q(i, j, k, O_QU) - Q(i - 1, j, k, O_QU) +
                 Q(i, j, k - 1, O_QU) - Q(i - 1, j, k - 1, O_QU) +
                 Q(i, j - 1, k, O_QU) - Q(i - 1, j - 1, k, O_QU) +
                 Q(i, j - 1, k - 1, O_QU) - Q(i - 1, j - 1, k - 1, O_QU)) / dx;
    vy = 0.25 * (Q(i, j, k, O_QV) - Q(i, j - 1, k, O_QV) +
                 Q(i, j, k - 1, O_QV) - Q(i, j - 1, k - 1, O_QV) +
                 Q(i - 1, j, k, O_QV) - Q(i - 1, j - 1, k, O_QV) +
                 Q(i - 1, j, k - 1, O_QV) - Q(i - 1, j - 1, k - 1, O_QV)) / dy;
    wz = 0.25 * (Q(i, j, k, O_QW) - Q(i, j, k - 1, O_QW) +
                 Q(i - 1, j, k, O_QW) - Q(i - 1, j, k - 1, O_QW) +
                 Q(i, j - 1, k, O_QW) - Q(i, j - 1, k - 1, O_QW) +
                 Q(i - 1, j - 1, k, O_QW) - Q(i - 1, j - 1, k - 1, O_QW)) / dz;

    if (i == domlo[0] && bclo[0] == NS) {
        vy = 0.5 * (Q(i, j, k, O_QV) - Q(i, j - 1, k, O_QV) +
                    Q(i, j, k - 1, O_QV) - Q(i, j - 1, k - 1, O_QV)) / dy;
        wz = 0.5 * (Q(i, j, k, O_QW) - Q(i, j, k - 1, O_QW) +
                    Q(i, j - 1, k, O_QW) - Q(i, j - 1, k - 1, O_QW)) / dz;
    } else if (i == domhi[0] + 1 && bchi[0] == NS) {
        vy = 0.5 * (Q(i - 1, j, k, O_QV) - Q(i - 1, j - 1, k, O_QV) +
                    Q(i - 1, j, k - 1, O_QV) - Q(i - 1, j - 1, k - 1, O_QV)) / dy;
        wz = 0.5 * (Q(i - 1, j, k, O_QW) - Q(i - 1, j, k - 1, O_QW) +
                    Q(i - 1, j - 1, k, O_QW) - Q(i - 1, j - 1, k - 1, O_QW)) / dz;
    }
    if (j == domlo[1] && bclo[1] == NS) {
        ux = 0.5 * (Q(i, j, k, O_QU) - Q(i - 1, j, k, O_QU) +
                    Q(i, j, k - 1, O_QU) - Q(i - 1, j, k - 1, O_QU)) / dx;
        wz = 0.5 * (Q(i, j, k, O_QW) - Q(i, j, k - 1, O_QW) +
                    Q(i - 1, j, k, O_QW) - Q(i - 1, j, k - 1, O_QW)) / dz;
    } else if (j == domhi[1] + 1 && bchi[1] == NS) {
        ux = 0.5 * (Q(i, j - 1, k, O_QU) - Q(i - 1, j - 1, k, O_QU) +
                    Q(i, j - 1, k - 1, O_QU) - Q(i - 1, j - 1, k - 1, O_QU)) / dx;
        wz = 0.5 * (Q(i, j - 1, k, O_QW) - Q(i, j - 1, k - 1, O_QW) +
                    Q(i - 1, j - 1, k, O_QW) - Q(i - 1, j - 1, k - 1, O_QW)) / dz;
    }
    if (k == domlo[2] && bclo[2] == NS) {
        ux = 0.5 * (Q(i, j, k, O_QU) - Q(i - 1, j, k, O_QU) +
                    Q(i, j - 1, k, O_QU) - Q(i - 1, j - 1, k, O_QU)) / dx;
        vy = 0.5 * (Q(i, j, k, O_QV) - Q(i, j - 1, k, O_QV) +
                    Q(i - 1, j, k, O_QV) - Q(i - 1, j - 1, k, O_QV)) / dy;
    } else if (k == domhi[2] + 1 && bchi[2] == NS) {
        ux = 0.5 * (Q(i, j, k - 1, O_QU) - Q(i - 1, j, k - 1, O_QU) +
                    Q(i, j - 1, k - 1, O_QU) - Q(i - 1, j - 1, k - 1, O_QU)) / dx;
        vy = 0.5 * (Q(i, j, k - 1, O_QV) - Q(i, j - 1, k - 1, O_QV) +
                    Q(i - 1, j, k - 1, O_QV) - Q(i - 1, j - 1, k - 1, O_QV)) / dy;
    }
    /* corners */
    if (j == domlo[1] && bclo[1] == NS && k == domlo[2] && bclo[2] == NS)
        ux = (Q(i, j, k, O_QU) - Q(i - 1, j, k, O_QU)) / dx;
    else if ((j == domhi[1] + 1 && bchi[1] == NS) && (k == domlo[2] && bclo[2] == NS))
        ux = (Q(i, j - 1, k, O_QU) - Q(i - 1, j - 1, k, O_QU)) / dx;
    else if ((j == domlo[1] && bclo[1] == NS) && (k == domhi[2] + 1 && bchi[2] == NS))
        ux = (Q(i, j, k - 1, O_QU) - Q(i - 1, j, k - 1, O_QU)) / dx;
    else if ((j == domhi[1] + 1 && bchi[1] == NS) && (k == domhi[2] + 1 && bchi[2] == NS))
        ux = (Q(i, j - 1, k - 1, O_QU) - Q(i - 1, j - 1, k - 1, O_QU)) / dx;

    if (i == domlo[0] && bclo[0] == NS && k == domlo[2] && bclo[2] == NS)
        vy = (Q(i, j, k, O_QV) - Q(i, j - 1, k, O_QV)) / dy;
    else if ((i == domhi[0] + 1 && bchi[0] == NS) && (k == domlo[2] && bclo[2] == NS))
        vy = (Q(i - 1, j, k, O_QV) - Q(i - 1, j - 1, k, O_QV)) / dy;
    else if ((i == domlo[0] && bclo[0] == NS) && (k == domhi[2] + 1 && bchi[2] == NS))
        vy = (Q(i, j, k - 1, O_QV) - Q(i, j - 1, k - 1, O_QV)) / dy;
    else if ((i == domhi[0] + 1 && bchi[0] == NS) && (k == domhi[2] + 1 && bchi[2] == NS))
        vy = (Q(i - 1, j, k - 1, O_QV) - Q(i - 1, j - 1, k - 1, O_QV)) / dy;

    if (i == domlo[0] && bclo[0] == NS && j == domlo[1] && bclo[1] == NS)
        wz = (Q(i, j, k, O_QW) - Q(i, j, k - 1, O_QW)) / dz;
    else if ((i == domhi[0] + 1 && bchi[0] == NS) && (j == domlo[1] && bclo[1] == NS))
        wz = (Q(i - 1, j, k, O_QW) - Q(i - 1, j, k - 1, O_QW)) / dz;
    else if ((i == domlo[0] && bclo[0] == NS) && (j == domhi[1] + 1 && bchi[1] == NS))
        wz = (Q(i, j - 1, k, O_QW) - Q(i, j - 1, k - 1, O_QW)) / dz;
    else if ((i == domhi[0] + 1 && bchi[0] == NS) && (j == domhi[1] + 1 && bchi[1] == NS))
        wz = (Q(i - 1, j - 1, k, O_QW) - Q(i - 1, j - 1, k - 1, O_QW)) / dz;
#undef Q
    FA(divu, i, j, k, 0) = ux + vy + wz;
}

static void artif_visc(int i, int j, int k, ora_fab *flx, const ora_fab *divu,
                       const ora_fab *u, double dx, double difmag, int dir,
                       int domlo, int domhi, int bclo, int bchi) {
    const int iv[3] = {i, j, k};
    int ivm[3] = {i, j, k};
    ivm[dir] -= 1;
    const int l1 = (dir == 0) ? 1 : 0;
    const int l2 = (dir == 2) ? 1 : 2;
    int ivpj[3] = {i, j, k}, ivpk[3] = {i, j, k}, ivpp[3] = {i, j, k};
    ivpj[l1] += 1;
    ivpk[l2] += 1;
    ivpp[l1] += 1;
    ivpp[l2] += 1;
    const double div =
#if ORA_PLANT == 5
        difmag * smax(0.0, 0.25 * (FV(divu, iv, 0) + FV(divu, ivpj, 0) +
                                   FV(divu, ivpk, 0) + FV(divu, ivpp, 0)));
#else
        difmag * smin(0.0, 0.25 * (FV(divu, iv, 0) + FV(divu, ivpj, 0) +
                                   FV(divu, ivpk, 0) + FV(divu, ivpp, 0)));
#endif
    int at_bndry = ((dir == 0) && (i == domlo) && (bclo == O_NoSlipWall)) ||
                   ((dir == 0) && (i == domhi + 1) && (bchi == O_NoSlipWall)) ||
                   ((dir == 1) && (j == domlo) && (bclo == O_NoSlipWall)) ||
                   ((dir == 1) && (j == domhi + 1) && (bchi == O_NoSlipWall));
    at_bndry = at_bndry ||
               ((dir == 2) && (k == domlo) && (bclo == O_NoSlipWall)) ||
               ((dir == 2) && (k == domhi + 1) && (bchi == O_NoSlipWall));
    for (int n = 0; n < O_NVAR; ++n) {
        if (n != O_UTEMP) {
            int is_tang_vel = (dir == 0 && (n == O_UMY || n == O_UMZ)) ||
                              (dir == 1 && (n == O_UMX || n == O_UMZ)) ||
                              (dir == 2 && (n == O_UMX || n == O_UMY));
            if (!(at_bndry && is_tang_vel))
                FV(flx, iv, n) += dx * div * (FV(u, iv, n) - FV(u, ivm, n));
        }
    }
    FV(flx, iv, O_UTEMP) = 0.0;
}

static void norm_spec_flx(int i, int j, int k, ora_fab *flx) {
    const int num = NSPEC + O_UFS;
    double summ = 0.0, fac;
    for (int n = O_UFS; n < num; ++n) summ += FA(flx, i, j, k, n);
    fac = (summ != 0.0) ? FA(flx, i, j, k, O_URHO) / summ : 1.0;
    for (int n = O_UFS; n < num; ++n) FA(flx, i, j, k, n) *= fac;
}

/* p div(u) from the Godunov face states, hydro_pdivu Hydro_utils_K.H:175-202 */
static void pdivu_box(obox bx, const ora_fab *q1, const ora_fab *q2, const ora_fab *q3,
                      const double area[3], ora_fab *pdivu, double vol) {
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        FA(pdivu, i, j, k, 0) =
            0.5 *
            ((FA(q1, i + 1, j, k, O_GDPRES) + FA(q1, i, j, k, O_GDPRES)) *
                 (FA(q1, i + 1, j, k, O_GDU) * area[0] -
                  FA(q1, i, j, k, O_GDU) * area[0]) +
             (FA(q2, i, j + 1, k, O_GDPRES) + FA(q2, i, j, k, O_GDPRES)) *
                 (FA(q2, i, j + 1, k, O_GDV) * area[1] -
                  FA(q2, i, j, k, O_GDV) * area[1]) +
             (FA(q3, i, j, k + 1, O_GDPRES) + FA(q3, i, j, k, O_GDPRES)) *
                 (FA(q3, i, j, k + 1, O_GDW) * area[2] -
                  FA(q3, i, j, k, O_GDW) * area[2])) /
            vol;
    }
}

/* ------------------------------------------------------------------------ */
/* Godunov_umeth (3-D), Source/Hydro/Godunov/Godunov_3D.cpp:12-461           */
/* ------------------------------------------------------------------------ */
static void godunov_umeth(obox bx, const ora_params *P, const ora_fab *q,
                          const ora_fab *qaux, const ora_fab *srcQ,
                          ora_fab *flx1, ora_fab *flx2, ora_fab *flx3,
                          ora_fab *q1, ora_fab *q2, ora_fab *q3,
                          const double area[3], ora_fab *pdivu, double vol,
                          double dt) {
    const double dx = P->dx[0], dy = P->dx[1], dz = P->dx[2];
    const double hdtdx = 0.5 * dt / dx;
    const double hdtdy = 0.5 * dt / dy;
    const double hdtdz = 0.5 * dt / dz;
    const double cdtdx = 1.0 / 3.0 * dt / dx;
    const double cdtdy = 1.0 / 3.0 * dt / dy;
    const double cdtdz = 1.0 / 3.0 * dt / dz;
    const double hdt = 0.5 * dt;
    const int *bclo = P->bclo, *bchi = P->bchi, *dlo = P->domlo, *dhi = P->domhi;
    const double small = P->small, sd = P->small_dens, sp = P->small_pres,
                 su = P->smallu;
    const int trd = P->transverse_reset_density;
    (void)hdtdx; (void)hdtdy; (void)hdtdz;

    const obox bxg1 = bgrow(bx, 1);
    const obox bxg2 = bgrow(bx, 2);

    ora_fab qdm[3], qdp[3];
    obox flxbx[3];
    for (int d = 0; d < 3; ++d) {
        fab_alloc_box(&qdm[d], bgrowhi(bxg2, d, 1), O_QVAR);
        fab_alloc_box(&qdp[d], bxg2, O_QVAR);
        flxbx[d] = bnodes(bgrowd(bxg2, d, -1), d);
    }

    ORA_STAGE(1);
    if (P->ppm_type == 0) {
        const double del[3] = {dx, dy, dz};
        for (int k = bxg2.lo[2]; k <= bxg2.hi[2]; ++k)
        for (int j = bxg2.lo[1]; j <= bxg2.hi[1]; ++j)
        for (int i = bxg2.lo[0]; i <= bxg2.hi[0]; ++i) {
            double slope[O_QVAR];
            double flat = 1.0;
            ORA_STAGE(8);
            if (P->use_flattening == 1)
                for (int dir_flat = 0; dir_flat < 3; dir_flat++)
                    flat = smin(flat, ora_flatten(q, i, j, k, dir_flat));
            ORA_STAGE(1);
            for (int idir = 0; idir < 3; ++idir) {
                for (int n = 0; n < O_QVAR; ++n) {
                    if (n == O_QPRES && P->use_pslope)
                        slope[n] = plm_pslope(q, i, j, k, n, idir, del[idir],
                                              srcQ, flat, P->plm_iorder);
                    else
                        slope[n] = ora_plm_slope(q, i, j, k, n, idir, flat,
                                                 P->plm_iorder);
                }
                plm_d(i, j, k, idir, &qdm[idir], &qdp[idir], slope, q,
                      FA(qaux, i, j, k, O_QC), del[idir], dt, sd, sp, NULL);
            }
        }
    } else {
        for (int idir = 0; idir < 3; ++idir)
            trace_ppm(bxg2, idir, q, qaux, &qdm[idir], &qdp[idir], bxg2, dt,
                      P->dx, P->use_flattening, sd, sp);
    }

    /* first flux estimates */
    ORA_STAGE(2);
    ora_fab f[3], qgd[3];
    for (int d = 0; d < 3; ++d) {
        fab_alloc_box(&f[d], flxbx[d], O_NVAR);
        fab_alloc_box(&qgd[d], flxbx[d], O_NGDNV);
        obox b = flxbx[d];
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i)
            cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qdm[d], &qdp[d],
                   &f[d], &qgd[d], qaux, d, small, sd, sp, su);
    }

    /* transverse corrections and Riemann solves for the pairs d|t.
     * Pair order follows the reference: X|Y, X|Z, Y|X, Y|Z, Z|X, Z|Y. */
#if ORA_PLANT == 2
    const double cdt[3] = {0.5 * dt / dx, 0.5 * dt / dy, 0.5 * dt / dz};
#else
    const double cdt[3] = {cdtdx, cdtdy, cdtdz};
#endif
    ORA_STAGE(3);
    ora_fab fl2[3][3], qg2[3][3]; /* fl2[d][t]: flux of direction d corrected by t */
    for (int d = 0; d < 3; ++d) {
        const int t0 = (d == 0) ? 1 : 0;
        const int t1 = (d == 2) ? 1 : 2;
        const obox tbx = bgrowd(bxg1, d, 1);
        const obox tbxm = bgrowhi(tbx, d, 1);
        ora_fab qmA, qpA, qmB, qpB;
        fab_alloc_box(&qmA, tbxm, O_QVAR);
        fab_alloc_box(&qpA, tbx, O_QVAR);
        fab_alloc_box(&qmB, tbxm, O_QVAR);
        fab_alloc_box(&qpB, tbx, O_QVAR);
        for (int k = tbx.lo[2]; k <= tbx.hi[2]; ++k)
        for (int j = tbx.lo[1]; j <= tbx.hi[1]; ++j)
        for (int i = tbx.lo[0]; i <= tbx.hi[0]; ++i) {
            transdo(i, j, k, d, t0, &qmA, &qpA, &qdm[d], &qdp[d], &f[t0],
                    qaux, &qgd[t0], cdt[t0], trd, sp, NULL, NULL);
            transdo(i, j, k, d, t1, &qmB, &qpB, &qdm[d], &qdp[d], &f[t1],
                    qaux, &qgd[t1], cdt[t1], trd, sp, NULL, NULL);
        }
        const obox tfx = bnodes(bxg1, d);
        fab_alloc_box(&fl2[d][t0], tfx, O_NVAR);
        fab_alloc_box(&fl2[d][t1], tfx, O_NVAR);
        fab_alloc_box(&qg2[d][t0], tfx, O_NGDNV);
        fab_alloc_box(&qg2[d][t1], tfx, O_NGDNV);
        for (int k = tfx.lo[2]; k <= tfx.hi[2]; ++k)
        for (int j = tfx.lo[1]; j <= tfx.hi[1]; ++j)
        for (int i = tfx.lo[0]; i <= tfx.hi[0]; ++i) {
            cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qmA, &qpA,
                   &fl2[d][t0], &qg2[d][t0], qaux, d, small, sd, sp, su);
            cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qmB, &qpB,
                   &fl2[d][t1], &qg2[d][t1], qaux, d, small, sd, sp, su);
        }
        ora_fab_free(&qmA); ora_fab_free(&qpA);
        ora_fab_free(&qmB); ora_fab_free(&qpB);
    }
    for (int d = 0; d < 3; ++d) { ora_fab_free(&f[d]); ora_fab_free(&qgd[d]); }

    /* final states and fluxes */
    ORA_STAGE(5);
    ora_fab qmf, qpf;
    fab_alloc_box(&qmf, bxg2, O_QVAR);
    fab_alloc_box(&qpf, bxg1, O_QVAR);
    ora_fab *flxd[3] = {flx1, flx2, flx3};
    ora_fab *qd[3] = {q1, q2, q3};
    for (int d = 0; d < 3; ++d) {
        /* dir d corrected by the fluxes (t0|t1) and (t1|t0) */
        const int t0 = (d == 0) ? 1 : 0;
        const int t1 = (d == 2) ? 1 : 2;
        const double hdtd[3] = {0.5 * dt / dx, 0.5 * dt / dy, 0.5 * dt / dz};
        const obox tb = bgrowd(bx, d, 1);
        for (int k = tb.lo[2]; k <= tb.hi[2]; ++k)
        for (int j = tb.lo[1]; j <= tb.hi[1]; ++j)
        for (int i = tb.lo[0]; i <= tb.hi[0]; ++i)
#if ORA_PLANT == 3
            transdd(i, j, k, d, &qmf, &qpf, &qdm[d], &qdp[d], &fl2[t0][d],
                    &fl2[t1][d], &qg2[t0][d], &qg2[t1][d], qaux, srcQ, hdt,
                    hdtd[t0], hdtd[t1], trd, sp, NULL, NULL, NULL);
#else
            transdd(i, j, k, d, &qmf, &qpf, &qdm[d], &qdp[d], &fl2[t0][t1],
                    &fl2[t1][t0], &qg2[t0][t1], &qg2[t1][t0], qaux, srcQ, hdt,
                    hdtd[t0], hdtd[t1], trd, sp, NULL, NULL, NULL);
#endif
        const obox fb = bnodes(bx, d);
        for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
        for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
        for (int i = fb.lo[0]; i <= fb.hi[0]; ++i)
            cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qmf, &qpf,
                   flxd[d], qd[d], qaux, d, small, sd, sp, su);
    }
    ora_fab_free(&qmf); ora_fab_free(&qpf);
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < 3; ++t)
            if (t != d) { ora_fab_free(&fl2[d][t]); ora_fab_free(&qg2[d][t]); }
    for (int d = 0; d < 3; ++d) { ora_fab_free(&qdm[d]); ora_fab_free(&qdp[d]); }

    ORA_STAGE(6);
    pdivu_box(bx, q1, q2, q3, area, pdivu, vol);
}

/* ------------------------------------------------------------------------ */
/* Method of lines: MOL_umeth (3-D), Source/Hydro/MOL/MOL_umeth.cpp:15-142;  */
/* slopes MOL_hydro_K.H:10-256, face states MOL_riemann_K.H:12-491           */
/* ------------------------------------------------------------------------ */
/* limiter, MOL_hydro_K.H:10-20 (plm_theta 2: van Leer's MC) */
static double mol_limiter(double dlft, double drgt, double plm_theta) {
    const double dcen = 0.5 * (dlft + drgt);
    const double dsgn = copysign(1.0, dcen);
    const double slop = plm_theta * smin(fabs(dlft), fabs(drgt));
    const double dlim = (dlft * drgt >= 0.0) ? slop : 0.0;
    return dsgn * smin(dlim, fabs(dcen));
}

/* Characteristic slopes of cell (i,j,k) along direction d: mol_slope_x/y/z
 * (MOL_hydro_K.H:22-256).  dq[0] / dq[2] the acoustic waves, dq[3] / dq[4]
 * the transverse velocities (x: v, w; y: u, w; z: u, v), dq[QREINT] the
 * entropy-like wave, dq[QFS+n] the species waves.  dq[1] (the density
 * wave) is not computed by the x variant and no face state reads it; it is
 * left 0 here for every direction. */
static void mol_slope(const ora_fab *q, const ora_fab *qa, int i, int j, int k, int d,
                      int plm_iorder, double plm_theta, double small_dens, double dq[O_QVAR]) {
    int m[3] = {i, j, k}, p[3] = {i, j, k};
    m[d] -= 1;
    p[d] += 1;
    const int un = O_QU + d;
    const int t1 = (d == 0) ? O_QV : O_QU;
    const int t2 = (d == 2) ? O_QV : O_QW;
    for (int n = 0; n < O_QVAR; ++n) dq[n] = 0.0;
    if (plm_iorder == 1) return;
#define QC_(n) FA(q, i, j, k, n)
#define QM_(n) FV(q, m, n)
#define QP_(n) FV(q, p, n)
    const double c = FA(qa, i, j, k, O_QC);
    double dlft = 0.5 * (QC_(O_QPRES) - QM_(O_QPRES)) / c - 0.5 * QC_(O_QRHO) * (QC_(un) - QM_(un));
    double drgt = 0.5 * (QP_(O_QPRES) - QC_(O_QPRES)) / c - 0.5 * QC_(O_QRHO) * (QP_(un) - QC_(un));
    dq[0] = mol_limiter(dlft, drgt, plm_theta);
    const double cs2 = c * c;
    dlft = 0.5 * (QC_(O_QPRES) - QM_(O_QPRES)) / c + 0.5 * QC_(O_QRHO) * (QC_(un) - QM_(un));
    drgt = 0.5 * (QP_(O_QPRES) - QC_(O_QPRES)) / c + 0.5 * QC_(O_QRHO) * (QP_(un) - QC_(un));
    dq[2] = mol_limiter(dlft, drgt, plm_theta);
    dlft = QC_(t1) - QM_(t1);
    drgt = QP_(t1) - QC_(t1);
    dq[3] = mol_limiter(dlft, drgt, plm_theta);
    dlft = QC_(t2) - QM_(t2);
    drgt = QP_(t2) - QC_(t2);
    dq[4] = mol_limiter(dlft, drgt, plm_theta);
    const double enth = (QC_(O_QREINT) + QC_(O_QPRES)) / smax(QC_(O_QRHO), small_dens);
    dlft = (QC_(O_QREINT) - QM_(O_QREINT)) - enth * (QC_(O_QPRES) - QM_(O_QPRES)) / cs2;
    drgt = (QP_(O_QREINT) - QC_(O_QREINT)) - enth * (QP_(O_QPRES) - QC_(O_QPRES)) / cs2;
    dq[O_QREINT] = mol_limiter(dlft, drgt, plm_theta);
    for (int n = 0; n < NSPEC; ++n) {
        const int f = O_QFS + n;
        dlft = QC_(O_QRHO) * QC_(f) - QM_(O_QRHO) * QM_(f) - QC_(f) * (QC_(O_QPRES) - QM_(O_QPRES)) / cs2;
        drgt = QP_(O_QRHO) * QP_(f) - QC_(O_QRHO) * QC_(f) - QC_(f) * (QP_(O_QPRES) - QC_(O_QPRES)) / cs2;
        dq[f] = mol_limiter(dlft, drgt, plm_theta);
    }
#undef QC_
#undef QM_
#undef QP_
}

/* Left / right states at the face (i,j,k) of direction d from the slopes
 * dqm of the cell below and dqc of the cell above: mol_riemann_x/y/z before
 * their hydro_cmpflx call (MOL_riemann_K.H:12-168, 171-326, 329-488).
 * Species are reconstructed as partial densities and renormalised; a
 * vanishing total falls back to the mass fractions of the cell ABOVE the
 * face on both sides, as the reference does. */
static void mol_face_states(const ora_fab *q, const ora_fab *qa, int i, int j, int k, int d,
                            const double *dqm, const double *dqc, double small_dens,
                            double small_pres, double *ql, double *qr) {
    int m[3] = {i, j, k};
    m[d] -= 1;
    const int un = O_QU + d;
    const int t1 = (d == 0) ? O_QV : O_QU;
    const int t2 = (d == 2) ? O_QV : O_QW;
    double cspeed = FV(qa, m, O_QC);
    double enth = (FV(q, m, O_QREINT) + FV(q, m, O_QPRES)) / smax(FV(q, m, O_QRHO), small_dens);
    double pl = FV(q, m, O_QPRES) + 0.5 * (dqm[0] + dqm[2]) * cspeed;
    pl = smax(pl, small_pres);
    ql[O_QPRES] = pl;
    ql[un] = FV(q, m, un) + 0.5 * ((dqm[2] - dqm[0]) / FV(q, m, O_QRHO));
    ql[t1] = FV(q, m, t1) + 0.5 * dqm[3];
    ql[t2] = FV(q, m, t2) + 0.5 * dqm[4];
    double sum = 0.;
    for (int n = 0; n < NSPEC; ++n) {
        const int f = O_QFS + n;
        const double specden = FV(q, m, O_QRHO) * FV(q, m, f) +
                               0.5 * (dqm[f] + FV(q, m, f) * (dqm[0] + dqm[2]) / cspeed);
        sum += specden;
        ql[f] = specden;
    }
    if (sum < small_dens) {
        for (int n = 0; n < NSPEC; ++n) ql[O_QFS + n] = FA(q, i, j, k, O_QFS + n);
    } else {
        for (int n = 0; n < NSPEC; ++n) ql[O_QFS + n] = ql[O_QFS + n] / sum;
    }
    ql[O_QRHO] = smax(sum, small_dens);
    ql[O_QREINT] = FV(q, m, O_QREINT) + 0.5 * (dqm[O_QREINT] + enth * (dqm[0] + dqm[2]) / cspeed);

    cspeed = FA(qa, i, j, k, O_QC);
    enth = (FA(q, i, j, k, O_QREINT) + FA(q, i, j, k, O_QPRES)) / smax(FA(q, i, j, k, O_QRHO), small_dens);
    double pr = FA(q, i, j, k, O_QPRES) - 0.5 * (dqc[0] + dqc[2]) * cspeed;
    pr = smax(pr, small_pres);
    qr[O_QPRES] = pr;
    qr[un] = FA(q, i, j, k, un) - 0.5 * ((dqc[2] - dqc[0]) / FA(q, i, j, k, O_QRHO));
    qr[t1] = FA(q, i, j, k, t1) - 0.5 * dqc[3];
    qr[t2] = FA(q, i, j, k, t2) - 0.5 * dqc[4];
    sum = 0.;
    for (int n = 0; n < NSPEC; ++n) {
        const int f = O_QFS + n;
        const double specden = FA(q, i, j, k, O_QRHO) * FA(q, i, j, k, f) -
                               0.5 * (dqc[f] + FA(q, i, j, k, f) * (dqc[0] + dqc[2]) / cspeed);
        sum += specden;
        qr[f] = specden;
    }
    if (sum < small_dens) {
        for (int n = 0; n < NSPEC; ++n) qr[O_QFS + n] = FA(q, i, j, k, O_QFS + n);
    } else {
        for (int n = 0; n < NSPEC; ++n) qr[O_QFS + n] = qr[O_QFS + n] / sum;
    }
    qr[O_QRHO] = smax(sum, small_dens);
    qr[O_QREINT] = FA(q, i, j, k, O_QREINT) - 0.5 * (dqc[O_QREINT] + enth * (dqc[0] + dqc[2]) / cspeed);
}

/* MOL_umeth: per direction, slopes on grow(bx, d, 1), then the face states
 * and hydro_cmpflx on the faces of bx; then p div(u). */
static void mol_umeth(obox bx, const ora_params *P, const ora_fab *q, const ora_fab *qa,
                      ora_fab *flx[3], ora_fab *qg[3], const double area[3], ora_fab *pdivu,
                      double vol) {
    const double plm_theta = 2.0;
    ora_fab slope, ql, qr;
    fab_alloc_box(&slope, bgrow(bx, 1), O_QVAR);
    for (int d = 0; d < 3; ++d) {
        const obox sb = bgrowd(bx, d, 1);
        for (int k = sb.lo[2]; k <= sb.hi[2]; ++k)
        for (int j = sb.lo[1]; j <= sb.hi[1]; ++j)
        for (int i = sb.lo[0]; i <= sb.hi[0]; ++i) {
            double dq[O_QVAR];
            mol_slope(q, qa, i, j, k, d, P->plm_iorder, plm_theta, P->small_dens, dq);
            for (int n = 0; n < O_QVAR; ++n) FA(&slope, i, j, k, n) = dq[n];
        }
        const obox fb = bnodes(bx, d);
        fab_alloc_box(&ql, fb, O_QVAR);
        fab_alloc_box(&qr, fb, O_QVAR);
        for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
        for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
        for (int i = fb.lo[0]; i <= fb.hi[0]; ++i) {
            int m[3] = {i, j, k};
            m[d] -= 1;
            double dqm[O_QVAR], dqc[O_QVAR], l[O_QVAR], r[O_QVAR];
            for (int n = 0; n < O_QVAR; ++n) {
                dqm[n] = FV(&slope, m, n);
                dqc[n] = FA(&slope, i, j, k, n);
                l[n] = r[n] = 0.0;
            }
            mol_face_states(q, qa, i, j, k, d, dqm, dqc, P->small_dens, P->small_pres, l, r);
            for (int n = 0; n < O_QVAR; ++n) {
                FA(&ql, i, j, k, n) = l[n];
                FA(&qr, i, j, k, n) = r[n];
            }
            cmpflx(i, j, k, P->bclo[d], P->bchi[d], P->domlo[d], P->domhi[d], &ql, &qr,
                   flx[d], qg[d], qa, d, P->small, P->small_dens, P->small_pres, P->smallu);
        }
        ora_fab_free(&ql);
        ora_fab_free(&qr);
    }
    ora_fab_free(&slope);
    pdivu_box(bx, qg[0], qg[1], qg[2], area, pdivu, vol);
}

/* ------------------------------------------------------------------------ */
/* hydro_umdrv, Source/Hydro/Hydro_umdrv.cpp:11-93 (Godunov or, with        */
/* do_mol, MOL_umeth)                                                         */
/* + adjust_fluxes Hydro_adjust_fluxes.cpp:13-44 + hydro_consup              */
/* Hydro_consup.cpp:13-25                                                     */
/* ------------------------------------------------------------------------ */
void ora_hydro_umdrv(const int lo[3], const int hi[3], const ora_fab *uin,
                     ora_fab *dsdt, const ora_fab *q, const ora_fab *qa,
                     const ora_fab *srcq, ora_fab flx[3], double dt,
                     const ora_params *P) {
    const obox bx = mkbox(lo, hi);
    const obox bxg2 = bgrow(bx, 2);
    ora_fab qec[3];
    for (int d = 0; d < 3; ++d) fab_alloc_box(&qec[d], bnodes(bxg2, d), O_NGDNV);
    ora_fab divu, pdivu;
    fab_alloc_box(&divu, bxg2, 1);
    fab_alloc_box(&pdivu, bx, 1);

    /* geom.GetVolume / GetFaceArea for a Cartesian grid (CAMR.cpp:337-348) */
    const double *dx = P->dx;
    const double vol = dx[0] * dx[1] * dx[2];
    const double area[3] = {dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1]};

    if (P->do_mol) {
        ora_fab *fl[3] = {&flx[0], &flx[1], &flx[2]}, *qg[3] = {&qec[0], &qec[1], &qec[2]};
        mol_umeth(bx, P, q, qa, fl, qg, area, &pdivu, vol);
    } else {
        godunov_umeth(bx, P, q, qa, srcq, &flx[0], &flx[1], &flx[2], &qec[0],
                      &qec[1], &qec[2], area, &pdivu, vol, dt);
    }

    ORA_STAGE(4);
    for (int k = bxg2.lo[2]; k <= bxg2.hi[2]; ++k)
    for (int j = bxg2.lo[1]; j <= bxg2.hi[1]; ++j)
    for (int i = bxg2.lo[0]; i <= bxg2.hi[0]; ++i)
        divu_node(i, j, k, q, dx[0], dx[1], dx[2], &divu, P->domlo, P->domhi,
                  P->bclo, P->bchi);

    ORA_STAGE(5);
    for (int d = 0; d < 3; ++d) {
        const obox fb = bnodes(bx, d);
        for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
        for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
        for (int i = fb.lo[0]; i <= fb.hi[0]; ++i) {
            artif_visc(i, j, k, &flx[d], &divu, uin, dx[d], P->difmag, d,
                       P->domlo[d], P->domhi[d], P->bclo[d], P->bchi[d]);
            norm_spec_flx(i, j, k, &flx[d]);
            const double a = area[d];
            for (int n = 0; n < O_NVAR; ++n) FA(&flx[d], i, j, k, n) *= a;
        }
    }

    ORA_STAGE(6);
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        double volinv = 1.0 / vol;
        for (int n = 0; n < O_NVAR; ++n) {
            FA(dsdt, i, j, k, n) +=
                (FA(&flx[0], i, j, k, n) - FA(&flx[0], i + 1, j, k, n) +
                 FA(&flx[1], i, j, k, n) - FA(&flx[1], i, j + 1, k, n) +
                 FA(&flx[2], i, j, k, n) - FA(&flx[2], i, j, k + 1, n)) *
                volinv;
        }
        FA(dsdt, i, j, k, O_UEINT) -= FA(&pdivu, i, j, k, 0);
    }

    for (int d = 0; d < 3; ++d) ora_fab_free(&qec[d]);
    ora_fab_free(&divu);
    ora_fab_free(&pdivu);
}

/* ------------------------------------------------------------------------ */
/* CAMR_estdt_hydro, Source/Utils/Timestep.H:18-61                           */
/* ------------------------------------------------------------------------ */
static double estdt_box_eb(const int lo[3], const int hi[3], const ora_fab *u,
                           const double dx[3], const ora_ebgeom *G) {
    double dt = DBL_MAX;
    for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
    for (int i = lo[0]; i <= hi[0]; ++i) {
        if (G && fl_covered(FLG(G, i, j, k))) continue; /* Timestep.H:32-33 */
        const double rho = FA(u, i, j, k, O_URHO);
        const double rhoInv = 1.0 / rho;
        double eint = FA(u, i, j, k, O_UEINT) / rho;
        const double pres = (EOS_GAMMA - 1.0) * rho * eint;
        const double gam = EOS_GAMMA;
        const double c = sqrt(gam * pres / rho);
        const double ux = FA(u, i, j, k, O_UMX) * rhoInv;
        const double dt1 = dx[0] / (c + fabs(ux));
        dt = smin(dt, dt1);
        const double uy = FA(u, i, j, k, O_UMY) * rhoInv;
        const double dt2 = dx[1] / (c + fabs(uy));
        dt = smin(dt, dt2);
        const double uz = FA(u, i, j, k, O_UMZ) * rhoInv;
        const double dt3 = dx[2] / (c + fabs(uz));
        dt = smin(dt, dt3);
    }
    return dt;
}

double ora_estdt_box(const int lo[3], const int hi[3], const ora_fab *u,
                     const double dx[3]) {
    return estdt_box_eb(lo, hi, u, dx, NULL);
}

/* ------------------------------------------------------------------------ */
/* Level driver: CAMR_advance (Source/CAMR_advance.cpp:38-164),              */
/* construct_hydro_source (Hydro/CAMR_construct_hydro_source.cpp:9-267),     */
/* clean_state family (Source/CAMR.cpp:866-1262), estTimeStep /             */
/* computeNewDt / computeInitialDt (CAMR.cpp:474-676).                       */
/* ------------------------------------------------------------------------ */
struct ora_level {
    ora_params P;
    int nbox;
    obox *box;
    ora_fab *Snew, *Sold, *Sborder, *hsrc, *oldsrc, *newsrc, *srchyd;
    int *owner;          /* domain cell -> box index (or -1) */
    double *ext;         /* exterior data over the domain (NVAR comps), or NULL */
    int64_t dn[3];
    double time;
    int err;
    int ng;              /* ghost width of Sborder: 4, or 6 with EB */
    int32_t *flags;      /* EB flags over the domain grown by 7 (or NULL) */
    ora_ebgeom *geom;    /* per box: flags view + cut-cell data on grow(bx,7) */
    int *kind;           /* per box: 0 regular, 1 EB path, 2 all covered */
    int defer_post;      /* 1: the step leaves post_timestep's clean_state to the caller */
    struct ora_fluxreg *reg_crse, *reg_fine; /* flux registers: this level as coarse / fine */
};

static int64_t dom_index(const ora_level *L, const int c[3]) {
    return (int64_t)(c[0] - L->P.domlo[0]) +
           (int64_t)(c[1] - L->P.domlo[1]) * L->dn[0] +
           (int64_t)(c[2] - L->P.domlo[2]) * L->dn[0] * L->dn[1];
}

ora_level *ora_level_create(const ora_params *p, int nbox, const int *lohi) {
    ora_level *L = (ora_level *)calloc(1, sizeof(ora_level));
    L->P = *p;
    L->nbox = nbox;
    L->ng = NGROW;
    L->box = (obox *)calloc(nbox, sizeof(obox));
    L->Snew = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    L->Sold = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    L->Sborder = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    L->hsrc = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    L->oldsrc = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    L->newsrc = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    L->srchyd = (ora_fab *)calloc(nbox, sizeof(ora_fab));
    for (int d = 0; d < 3; ++d)
        L->dn[d] = p->domhi[d] - p->domlo[d] + 1;
    int64_t ncell = L->dn[0] * L->dn[1] * L->dn[2];
    L->owner = (int *)malloc(sizeof(int) * (size_t)ncell);
    for (int64_t c = 0; c < ncell; ++c) L->owner[c] = -1;
    for (int b = 0; b < nbox; ++b) {
        obox B;
        for (int d = 0; d < 3; ++d) { B.lo[d] = lohi[6 * b + d]; B.hi[d] = lohi[6 * b + 3 + d]; }
        L->box[b] = B;
        fab_alloc_box(&L->Snew[b], B, O_NVAR);
        fab_alloc_box(&L->Sold[b], B, O_NVAR);
        fab_alloc_box(&L->Sborder[b], bgrow(B, L->ng), O_NVAR);
        fab_alloc_box(&L->hsrc[b], bgrow(B, L->ng), O_NVAR);
        fab_alloc_box(&L->oldsrc[b], bgrow(B, L->ng), O_NVAR);
        fab_alloc_box(&L->newsrc[b], B, O_NVAR);
        fab_alloc_box(&L->srchyd[b], bgrow(B, L->ng), O_NVAR);
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            int c[3] = {i, j, k};
            L->owner[dom_index(L, c)] = b;
        }
    }
    return L;
}

void ora_level_set_exterior(ora_level *L, const double *data) {
    const int64_t nc = L->dn[0] * L->dn[1] * L->dn[2] * O_NVAR;
    if (!L->ext) L->ext = (double *)malloc(sizeof(double) * (size_t)nc);
    memcpy(L->ext, data, sizeof(double) * (size_t)nc);
}

void ora_level_destroy(ora_level *L) {
    if (!L) return;
    if (L->geom) {
        for (int b = 0; b < L->nbox; ++b) {
            ora_fab_free(&L->geom[b].vfrac);
            for (int d = 0; d < 3; ++d) { ora_fab_free(&L->geom[b].ap[d]); ora_fab_free(&L->geom[b].fc[d]); }
        }
        free(L->geom);
    }
    free(L->kind);
    free(L->flags);
    for (int b = 0; b < L->nbox; ++b) {
        ora_fab_free(&L->Snew[b]); ora_fab_free(&L->Sold[b]);
        ora_fab_free(&L->Sborder[b]); ora_fab_free(&L->hsrc[b]);
        ora_fab_free(&L->oldsrc[b]); ora_fab_free(&L->newsrc[b]);
        ora_fab_free(&L->srchyd[b]);
    }
    free(L->Snew); free(L->Sold); free(L->Sborder); free(L->hsrc);
    free(L->oldsrc); free(L->newsrc); free(L->srchyd); free(L->box);
    free(L->owner); free(L->ext); free(L);
}

int ora_level_nbox(const ora_level *L) { return L->nbox; }
double ora_level_time(const ora_level *L) { return L->time; }

void ora_level_set_box(ora_level *L, int b, const double *data) {
    memcpy(L->Snew[b].p, data, sizeof(double) * (size_t)(L->Snew[b].ns * O_NVAR));
}
void ora_level_get_box(const ora_level *L, int b, double *data) {
    memcpy(data, L->Snew[b].p, sizeof(double) * (size_t)(L->Snew[b].ns * O_NVAR));
}

/* CAMR_reset_internal_e, Source/Utils/CAMR_reset_internal_e.H:13-129 */
static void reset_internal_e(ora_fab *S, int i, int j, int k,
                             const ora_params *P) {
    const double rho = FA(S, i, j, k, O_URHO);
    if (P->eb && !(rho > 0.)) return; /* EB build: CAMR_reset_internal_e.H:36-40 */
    const double rhoInv = 1.0 / rho;
    const double Up = FA(S, i, j, k, O_UMX) * rhoInv;
    const double Vp = FA(S, i, j, k, O_UMY) * rhoInv;
    const double Wp = FA(S, i, j, k, O_UMZ) * rhoInv;
    const double ke = 0.5 * (Up * Up + Vp * Vp + Wp * Wp);
    const double eden = FA(S, i, j, k, O_UEDEN) * rhoInv;
    double massfrac[NSPEC];
    for (int sp = 0; sp < NSPEC; sp++) massfrac[sp] = FA(S, i, j, k, sp + O_UFS) * rhoInv;
    double e_out = 0.;
    double small_e = 0.;
    const double small_T = P->small_temp;
    if (P->allow_small_energy == 0) {
        small_e = eos_rty2e(small_T, massfrac);
        if (eden < small_e) {
            if (FA(S, i, j, k, O_UEINT) * rhoInv < small_e) {
                const double T = smax(FA(S, i, j, k, O_UTEMP), small_T);
                e_out = eos_rty2e(T, massfrac);
                FA(S, i, j, k, O_UEINT) = rho * e_out;
            }
            FA(S, i, j, k, O_UEDEN) = FA(S, i, j, k, O_UEINT) + rho * ke;
        } else {
            const double rho_eint = FA(S, i, j, k, O_UEDEN) - FA(S, i, j, k, O_URHO) * ke;
            if (rho_eint > 0. && rho_eint / FA(S, i, j, k, O_UEDEN) > P->dual_energy_eta2)
                FA(S, i, j, k, O_UEINT) = rho_eint;
            if (FA(S, i, j, k, O_UEINT) * rhoInv < small_e) {
                const double T = smax(FA(S, i, j, k, O_UTEMP), small_T);
                e_out = eos_rty2e(T, massfrac);
                if (P->dual_energy_update_E_from_e == 1)
                    FA(S, i, j, k, O_UEDEN) += (rho * e_out - FA(S, i, j, k, O_UEINT));
                FA(S, i, j, k, O_UEINT) = FA(S, i, j, k, O_URHO) * e_out;
            }
        }
    } else if (P->allow_negative_energy == 0) {
        if (FA(S, i, j, k, O_UEDEN) < 0.) {
            if (FA(S, i, j, k, O_UEINT) < 0.) {
                e_out = eos_rty2e(small_T, massfrac);
                FA(S, i, j, k, O_UEINT) = rho * e_out;
            }
            FA(S, i, j, k, O_UEDEN) = FA(S, i, j, k, O_UEINT) + rho * ke;
        } else {
            const double rho_eint = FA(S, i, j, k, O_UEDEN) - rho * ke;
            if (rho_eint > 0. && rho_eint / FA(S, i, j, k, O_UEDEN) > P->dual_energy_eta2) {
                FA(S, i, j, k, O_UEINT) = rho_eint;
            } else if (FA(S, i, j, k, O_UEINT) > 0 && P->dual_energy_update_E_from_e == 1) {
                FA(S, i, j, k, O_UEDEN) = FA(S, i, j, k, O_UEINT) + rho * ke;
            } else if (FA(S, i, j, k, O_UEINT) <= 0.) {
                e_out = eos_rty2e(small_T, massfrac);
                FA(S, i, j, k, O_UEINT) = rho * e_out;
                if (P->dual_energy_update_E_from_e == 1)
                    FA(S, i, j, k, O_UEDEN) += (rho * e_out - FA(S, i, j, k, O_UEINT));
            }
        }
    } else {
        FA(S, i, j, k, O_UEINT) = FA(S, i, j, k, O_UEDEN) - rho * ke;
    }
}

/* computeTemp, CAMR.cpp:1141-1185 (reset_internal_energy first) */
static void compute_temp(ora_fab *S, obox gb, const ora_params *P) {
    for (int k = gb.lo[2]; k <= gb.hi[2]; ++k)
    for (int j = gb.lo[1]; j <= gb.hi[1]; ++j)
    for (int i = gb.lo[0]; i <= gb.hi[0]; ++i)
        reset_internal_e(S, i, j, k, P);
    for (int k = gb.lo[2]; k <= gb.hi[2]; ++k)
    for (int j = gb.lo[1]; j <= gb.hi[1]; ++j)
    for (int i = gb.lo[0]; i <= gb.hi[0]; ++i) {
        if (P->eb) { /* EB build, CAMR.cpp:1166-1169 */
            FA(S, i, j, k, O_UTEMP) = 0.0;
            if (!(FA(S, i, j, k, O_URHO) > 0.)) continue;
        }
        double rhoInv = 1.0 / FA(S, i, j, k, O_URHO);
        double e = FA(S, i, j, k, O_UEINT) * rhoInv;
        FA(S, i, j, k, O_UTEMP) = eos_rey2t(e);
    }
}

/* enforce_min_density, CAMR.cpp:941-1045 (untiled box, sequential order) */
static void enforce_min_density(ora_fab *S, obox bx, const ora_params *P,
                                int *err, const ora_ebgeom *G) {
    const double sd = P->small_dens;
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        if (G && fl_covered(FLG(G, i, j, k))) continue; /* CAMR.cpp:987-988 */
        if (FA(S, i, j, k, O_URHO) == 0.) {
            *err |= ERR_ZERO_DENS;
        } else if (FA(S, i, j, k, O_URHO) < sd) {
            double max_dens = FA(S, i, j, k, O_URHO);
            int is = i, js = j, ks = k;
            for (int kk = -1; kk <= 1; kk++)
            for (int jj = -1; jj <= 1; jj++)
            for (int ii = -1; ii <= 1; ii++) {
                if ((i + ii >= bx.lo[0]) && (j + jj >= bx.lo[1]) && (k + kk >= bx.lo[2]) &&
                    (i + ii <= bx.hi[0]) && (j + jj <= bx.hi[1]) && (k + kk <= bx.hi[2])) {
#if ORA_PLANT == 6
                    if (FA(S, i + ii, j + jj, k + kk, O_URHO) >= max_dens) {
#else
                    if (FA(S, i + ii, j + jj, k + kk, O_URHO) > max_dens) {
#endif
                        is = i + ii; js = j + jj; ks = k + kk;
                        max_dens = FA(S, is, js, ks, O_URHO);
                    }
                }
            }
            if (max_dens < sd) {
                double massfrac[NSPEC];
                for (int s = 0; s < NSPEC; s++)
                    massfrac[s] = FA(S, i, j, k, O_UFS + s) / FA(S, i, j, k, O_URHO);
                double small_e = eos_rty2e(P->small_temp, massfrac);
                FA(S, i, j, k, O_URHO) = sd;
                FA(S, i, j, k, O_UTEMP) = P->small_temp;
                FA(S, i, j, k, O_UMX) = 0.;
                FA(S, i, j, k, O_UMY) = 0.;
                FA(S, i, j, k, O_UMZ) = 0.;
                FA(S, i, j, k, O_UEINT) = sd * small_e;
                FA(S, i, j, k, O_UEDEN) = FA(S, i, j, k, O_UEINT);
            } else {
                for (int n = 0; n < O_NVAR; n++)
                    FA(S, i, j, k, n) = FA(S, is, js, ks, n);
            }
        }
    }
}

/* normalize_species, CAMR.cpp:866-908 (valid cells) */
static void normalize_species(ora_fab *S, obox bx, const ora_ebgeom *G) {
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        if (G && fl_covered(FLG(G, i, j, k))) continue; /* CAMR.cpp:886-887 */
        double sum = 0.;
        for (int n = 0; n < NSPEC; n++) {
            FA(S, i, j, k, O_UFS + n) =
                smax(smin(FA(S, i, j, k, O_URHO), FA(S, i, j, k, O_UFS + n)), 0.0);
            sum += FA(S, i, j, k, O_UFS + n);
        }
        if (FA(S, i, j, k, O_URHO) > 0.) {
            sum /= FA(S, i, j, k, O_URHO);
            for (int n = 0; n < NSPEC; n++) FA(S, i, j, k, O_UFS + n) /= sum;
        }
    }
}

/* Per-direction ghost-cell source index and sign (AMReX FillPatchSingleLevel
 * + FilccCell semantics with the BC tables of CAMR_setup.cpp:19-21).
 * Returns 1 if the cell is outside the domain in a non-periodic direction. */
static int map_cell(const ora_params *P, const int c[3], int comp, int m[3],
                    double *sign) {
    int phys = 0;
    *sign = 1.0;
    for (int d = 0; d < 3; ++d) {
        int x = c[d];
        const int lo = P->domlo[d], hi = P->domhi[d], n = hi - lo + 1;
        if (x >= lo && x <= hi) { m[d] = x; continue; }
        if (P->is_periodic[d]) {
            int r = (x - lo) % n;
            if (r < 0) r += n;
            m[d] = lo + r;
            continue;
        }
        phys = 1;
        const int bc = (x < lo) ? P->bclo[d] : P->bchi[d];
        const int is_norm = (comp == O_UMX + d);
        const int is_tang = (comp >= O_UMX && comp <= O_UMZ && !is_norm);
        if (bc == O_Outflow || bc == O_Inflow || bc == O_Interior) {
            m[d] = (x < lo) ? lo : hi; /* foextrap (Inflow: bcnormal == copy) */
        } else {
            m[d] = (x < lo) ? (2 * lo - x - 1) : (2 * hi - x + 1);
            int odd = is_norm || (is_tang && bc == O_NoSlipWall);
            if (odd) *sign = -(*sign);
        }
    }
    return phys;
}

/* FillPatch (single level) of all ghost+valid cells of dst from src. */
static void fill_patch(ora_level *L, ora_fab *dst, int b, ora_fab *src) {
    obox g = bgrow(L->box[b], L->ng);
    for (int n = 0; n < O_NVAR; ++n)
    for (int k = g.lo[2]; k <= g.hi[2]; ++k)
    for (int j = g.lo[1]; j <= g.hi[1]; ++j)
    for (int i = g.lo[0]; i <= g.hi[0]; ++i) {
        int c[3] = {i, j, k}, m[3];
        double s;
        map_cell(&L->P, c, n, m, &s);
        const int64_t di = dom_index(L, m);
        int ob = L->owner[di];
        double v;
        if (ob >= 0) {
            v = FA(&src[ob], m[0], m[1], m[2], n);
        } else if (L->ext) {
            /* a cell no box owns: the exterior data stands for the
             * coarse-level FillPatch values */
            v = L->ext[di + (int64_t)n * L->dn[0] * L->dn[1] * L->dn[2]];
        } else {
            L->err |= ERR_UNOWNED;
            v = 0.0;
        }
        FA(dst, i, j, k, n) = (s < 0.0) ? -v : v;
    }
}

/* FillBoundary: re-copy ghost cells that are periodic/interior images of
 * valid cells (physical-boundary ghosts untouched). */
static void fill_boundary(ora_level *L, ora_fab *S, int ng) {
    for (int b = 0; b < L->nbox; ++b) {
        obox g = bgrow(L->box[b], ng);
        obox v = L->box[b];
        for (int k = g.lo[2]; k <= g.hi[2]; ++k)
        for (int j = g.lo[1]; j <= g.hi[1]; ++j)
        for (int i = g.lo[0]; i <= g.hi[0]; ++i) {
            if (i >= v.lo[0] && i <= v.hi[0] && j >= v.lo[1] && j <= v.hi[1] &&
                k >= v.lo[2] && k <= v.hi[2]) continue;
            int c[3] = {i, j, k}, m[3];
            double s;
            if (map_cell(&L->P, c, 0, m, &s)) continue;
            int ob = L->owner[dom_index(L, m)];
            if (ob < 0) continue;
            for (int n = 0; n < O_NVAR; ++n)
                FA(&S[b], i, j, k, n) = FA(&S[ob], m[0], m[1], m[2], n);
        }
    }
}

/* clean_state, CAMR.cpp:1253-1262 */
static void clean_state(ora_level *L, ora_fab *S, int ng) {
    for (int b = 0; b < L->nbox; ++b)
        enforce_min_density(&S[b], L->box[b], &L->P, &L->err, L->geom ? &L->geom[b] : NULL);
    for (int b = 0; b < L->nbox; ++b)
        normalize_species(&S[b], L->box[b], L->geom ? &L->geom[b] : NULL);
    if (ng > 0) fill_boundary(L, S, ng);
    for (int b = 0; b < L->nbox; ++b) compute_temp(&S[b], bgrow(L->box[b], ng), &L->P);
}

/* enforce_consistent_e, CAMR.cpp:911-938 */
void ora_level_init_post(ora_level *L) {
    for (int b = 0; b < L->nbox; ++b) {
        ora_fab *S = &L->Snew[b];
        obox B = L->box[b];
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            if (L->P.eb && !(FA(S, i, j, k, O_URHO) > 0.)) continue; /* CAMR.cpp:921-923 */
            const double rhoInv = 1.0 / FA(S, i, j, k, O_URHO);
            const double u = FA(S, i, j, k, O_UMX) * rhoInv;
            const double v = FA(S, i, j, k, O_UMY) * rhoInv;
            const double w = FA(S, i, j, k, O_UMZ) * rhoInv;
            FA(S, i, j, k, O_UEDEN) =
                FA(S, i, j, k, O_UEINT) + 0.5 * FA(S, i, j, k, O_URHO) * (u * u + v * v + w * w);
        }
        compute_temp(S, B, &L->P);
    }
}

/* estTimeStep, CAMR.cpp:490-563 */
static double est_time_step(ora_level *L) {
    const ora_params *P = &L->P;
    ORA_STAGE(7);
    if (P->fixed_dt > 0.0) return P->fixed_dt;
    double estdt = P->max_dt;
    const double max_dt_over_cfl = P->max_dt / P->cfl;
    double estdt_hydro = max_dt_over_cfl;
    double dt = DBL_MAX;
    for (int b = 0; b < L->nbox; ++b)
        dt = smin(dt, estdt_box_eb(L->box[b].lo, L->box[b].hi, &L->Snew[b], P->dx,
                                   L->geom ? &L->geom[b] : NULL));
    if (!(dt > 0.0)) L->err |= ERR_DT;
    estdt_hydro = smin(estdt_hydro, dt);
    estdt_hydro *= P->cfl;
    if (estdt_hydro < estdt) estdt = estdt_hydro;
    return estdt;
}

/* computeInitialDt / initialTimeStep, CAMR.cpp:474-488, 638-676 */
double ora_level_initial_dt(ora_level *L) {
    double init_dt = L->P.init_shrink * est_time_step(L);
    double dt_0 = DBL_MAX;
    dt_0 = smin(dt_0, 1 * init_dt);
    const double dt_eps = 0.001 * dt_0;
    if (L->P.stop_time >= 0.0)
        if ((L->time + dt_0) > (L->P.stop_time - dt_eps)) dt_0 = L->P.stop_time - L->time;
    return dt_0 / 1;
}

/* computeNewDt, CAMR.cpp:566-635 (single level, no regrid) */
double ora_level_new_dt(ora_level *L, double dt_old) {
    double dt_min = est_time_step(L);
    if (L->P.fixed_dt <= 0.0) dt_min = smin(dt_min, L->P.change_max * dt_old);
    double dt_0 = DBL_MAX;
    dt_0 = smin(dt_0, 1 * dt_min);
    const double dt_eps = 0.001 * dt_0;
    if (L->P.stop_time >= 0.0)
        if ((L->time + dt_0) > (L->P.stop_time - dt_eps)) dt_0 = L->P.stop_time - L->time;
    return dt_0 / 1;
}

typedef struct { int b; obox t; } otile;

/* construct_hydro_source, Hydro/CAMR_construct_hydro_source.cpp:9-267 */
typedef struct ora_fluxreg ora_fluxreg;
static void fluxreg_crse_add(ora_fluxreg *R, int b, obox bx, ora_fab flux[3], double dt);
static void fluxreg_fine_add(ora_fluxreg *R, obox fbx, ora_fab flux[3], double dt);

static void construct_hydro_source(ora_level *L, double dt) {
    const ora_params *P = &L->P;
    for (int b = 0; b < L->nbox; ++b) {
        ora_fab *h = &L->hsrc[b];
        memset(h->p, 0, sizeof(double) * (size_t)(h->ns * h->nc));
    }
    /* sources_for_hydro (zeroed at CAMR_advance.cpp:120) += old sources on
     * valid cells, then FillBoundary */
    for (int b = 0; b < L->nbox; ++b) {
        obox B = L->box[b];
        for (int n = 0; n < O_NVAR; ++n)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
            FA(&L->srchyd[b], i, j, k, n) += 1.0 * FA(&L->oldsrc[b], i, j, k, n);
    }
    fill_boundary(L, L->srchyd, L->ng);

    /* tiles (hydro_tile_size) */
    int ntile = 0, cap = 64;
    otile *tiles = (otile *)malloc(sizeof(otile) * cap);
    for (int b = 0; b < L->nbox; ++b) {
        obox B = L->box[b];
        if (L->kind && L->kind[b] == 2) continue; /* covered box: dsdt stays 0 */
        const int untiled = L->kind && L->kind[b] == 1; /* hydro_umdrv_eb runs per box */
        for (int k0 = B.lo[2]; k0 <= B.hi[2]; k0 += untiled ? (B.hi[2] - B.lo[2] + 1) : P->tile[2])
        for (int j0 = B.lo[1]; j0 <= B.hi[1]; j0 += untiled ? (B.hi[1] - B.lo[1] + 1) : P->tile[1])
        for (int i0 = B.lo[0]; i0 <= B.hi[0]; i0 += untiled ? (B.hi[0] - B.lo[0] + 1) : P->tile[0]) {
            otile t; t.b = b;
            t.t.lo[0] = i0; t.t.lo[1] = j0; t.t.lo[2] = k0;
            t.t.hi[0] = (i0 + P->tile[0] - 1 < B.hi[0]) ? i0 + P->tile[0] - 1 : B.hi[0];
            t.t.hi[1] = (j0 + P->tile[1] - 1 < B.hi[1]) ? j0 + P->tile[1] - 1 : B.hi[1];
            t.t.hi[2] = (k0 + P->tile[2] - 1 < B.hi[2]) ? k0 + P->tile[2] - 1 : B.hi[2];
            if (ntile == cap) { cap *= 2; tiles = (otile *)realloc(tiles, sizeof(otile) * cap); }
            tiles[ntile++] = t;
        }
    }
    int err = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(ora_get_threads()) reduction(| : err)
#endif
    for (int it = 0; it < ntile; ++it) {
        const int b = tiles[it].b;
        const obox bx = tiles[it].t;
        const obox qbx = bgrow(bx, L->ng);
        ora_fab flux[3];
        for (int d = 0; d < 3; ++d) fab_alloc_box(&flux[d], bnodes(bx, d), O_NVAR);
        ora_fab q, qaux, srcq;
        fab_alloc_box(&q, qbx, O_QVAR);
        fab_alloc_box(&qaux, qbx, O_NQAUX);
        fab_alloc_box(&srcq, qbx, O_QVAR);
        ORA_STAGE(0);
        ora_ctoprim_box(qbx.lo, qbx.hi, &L->Sborder[b], &q, &qaux, P, &err);
        if (!P->do_mol) /* CAMR_construct_hydro_source.cpp:129-143 */
            ora_srctoprim_box(qbx.lo, qbx.hi, &q, &qaux, &L->srchyd[b], &srcq);
        if (L->kind && L->kind[b] == 1)
            ora_hydro_umdrv_eb(bx.lo, bx.hi, &L->Sborder[b], &L->hsrc[b], &q, &qaux,
                               &srcq, flux, dt, P, &L->geom[b]);
        else
            ora_hydro_umdrv(bx.lo, bx.hi, &L->Sborder[b], &L->hsrc[b], &q, &qaux,
                            &srcq, flux, dt, P);
        /* do_reflux: the tile's fluxes into the flux registers
         * (CAMR_construct_hydro_source.cpp:243-254), fac_for_reflux * dt */
        const double fdt = (P->do_mol ? 0.5 : 1.0) * dt;
        if (L->reg_crse) fluxreg_crse_add(L->reg_crse, b, bx, flux, fdt);
        if (L->reg_fine) fluxreg_fine_add(L->reg_fine, bx, flux, fdt);
        ora_fab_free(&q); ora_fab_free(&qaux); ora_fab_free(&srcq);
        for (int d = 0; d < 3; ++d) ora_fab_free(&flux[d]);
    }
    L->err |= err;
    free(tiles);
}

/* S_new += a * src on valid cells (MultiFab::Saxpy with separate mul/add) */
static void saxpy_valid(ora_level *L, ora_fab *dst, double a, ora_fab *src) {
    for (int b = 0; b < L->nbox; ++b) {
        obox B = L->box[b];
        for (int n = 0; n < O_NVAR; ++n)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
            FA(&dst[b], i, j, k, n) += a * FA(&src[b], i, j, k, n);
    }
}

static void set_val(ora_fab *f, int nbox, double v) {
    for (int b = 0; b < nbox; ++b) {
        int64_t N = f[b].ns * f[b].nc;
        for (int64_t c = 0; c < N; ++c) f[b].p[c] = v;
    }
}

/* CAMR_advance (Source/CAMR_advance.cpp:38-164) followed by post_timestep's
 * clean_state (CAMR.cpp:698-699). */
int ora_level_advance(ora_level *L, double dt) {
    const ora_params *P = &L->P;
    L->err = 0;
    ORA_STAGE(0);
    /* swapTimeLevels: old <- new */
    for (int b = 0; b < L->nbox; ++b) {
        ora_fab t = L->Sold[b]; L->Sold[b] = L->Snew[b]; L->Snew[b] = t;
    }
    clean_state(L, L->Sold, 0);
    /* expand_state: FillPatch + clean_state(Sborder) */
    for (int b = 0; b < L->nbox; ++b) fill_patch(L, &L->Sborder[b], b, L->Sold);
    clean_state(L, L->Sborder, L->ng);
    const int nsrc = P->add_grav_src ? 1 : 0; /* src_list = {grav_src} */
    if (nsrc) { set_val(L->oldsrc, L->nbox, 0.0); set_val(L->newsrc, L->nbox, 0.0); }
    /* MultiFab::Copy(S_new, Sborder) on valid cells */
    for (int b = 0; b < L->nbox; ++b) {
        obox B = L->box[b];
        for (int n = 0; n < O_NVAR; ++n)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
            FA(&L->Snew[b], i, j, k, n) = FA(&L->Sborder[b], i, j, k, n);
    }
    if (nsrc) {
        /* construct_old_grav_source, GravitySrc.cpp:4-17,36-78 */
        set_val(L->oldsrc, L->nbox, 0.0);
        for (int b = 0; b < L->nbox; ++b) {
            obox B = L->box[b];
            for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
            for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                FA(&L->oldsrc[b], i, j, k, O_UMZ) = P->const_grav * FA(&L->Sold[b], i, j, k, O_URHO);
                FA(&L->oldsrc[b], i, j, k, O_UEDEN) = P->const_grav * FA(&L->Sold[b], i, j, k, O_UMZ);
            }
        }
        fill_boundary(L, L->oldsrc, L->ng);
        saxpy_valid(L, L->Snew, dt, L->oldsrc);
    }
    set_val(L->srchyd, L->nbox, 0.0);
    construct_hydro_source(L, dt);
    saxpy_valid(L, L->Snew, dt, L->hsrc);
    if (P->do_mol) {
        /* second stage (CAMR_advance.cpp:124-135): expand_state(Sborder,
         * cur_time) = FillPatch from S_new + clean_state, the hydro source
         * of that state, LinComb(S_new, 0.5, Sborder, 0.5, S_old) and
         * Saxpy(S_new, 0.5 dt, new_hydro_source) */
        for (int b = 0; b < L->nbox; ++b) fill_patch(L, &L->Sborder[b], b, L->Snew);
        clean_state(L, L->Sborder, L->ng);
        construct_hydro_source(L, dt);
        for (int b = 0; b < L->nbox; ++b) {
            obox B = L->box[b];
            for (int n = 0; n < O_NVAR; ++n)
            for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
            for (int i = B.lo[0]; i <= B.hi[0]; ++i)
                FA(&L->Snew[b], i, j, k, n) = 0.5 * FA(&L->Sborder[b], i, j, k, n) +
                                              0.5 * FA(&L->Sold[b], i, j, k, n);
        }
        saxpy_valid(L, L->Snew, 0.5 * dt, L->hsrc);
    }
    ORA_STAGE(7);
    clean_state(L, L->Snew, 0);
    if (nsrc) {
        /* construct_new_grav_source, GravitySrc.cpp:20-33,81-125 */
        set_val(L->newsrc, L->nbox, 0.0);
        for (int b = 0; b < L->nbox; ++b) {
            obox B = L->box[b];
            for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
            for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                FA(&L->newsrc[b], i, j, k, O_UMZ) =
                    0.5 * P->const_grav *
                    (FA(&L->Snew[b], i, j, k, O_URHO) - FA(&L->Sold[b], i, j, k, O_URHO));
                FA(&L->newsrc[b], i, j, k, O_UEDEN) =
                    0.5 * P->const_grav *
                    (FA(&L->Snew[b], i, j, k, O_UMZ) - FA(&L->Sold[b], i, j, k, O_UMZ));
            }
        }
        saxpy_valid(L, L->Snew, dt, L->newsrc);
        clean_state(L, L->Snew, 0);
    }
    /* post_timestep: clean_state(S_new); a coarse level of a two-level
     * step runs it after reflux and avgDown instead */
    if (!L->defer_post) clean_state(L, L->Snew, 0);
    L->time += dt;
    return L->err;
}

int ora_level_run(ora_level *L, int nsteps, double *dts) {
    double dt = 0.0;
    for (int s = 0; s < nsteps; ++s) {
        dt = (s == 0) ? ora_level_initial_dt(L) : ora_level_new_dt(L, dt);
        if (dts) dts[s] = dt;
        int e = ora_level_advance(L, dt);
        if (e) return e;
    }
    return 0;
}

void ora_level_sums(const ora_level *L, double sums[O_NVAR]) {
    for (int n = 0; n < O_NVAR; ++n) sums[n] = 0.0;
    for (int b = 0; b < L->nbox; ++b) {
        const ora_fab *S = &L->Snew[b];
        obox B = L->box[b];
        for (int n = 0; n < O_NVAR; ++n)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
            sums[n] += FA(S, i, j, k, n);
    }
}

/* ======================================================================== */
/* Embedded boundaries: hydro_umdrv_eb (Source/Hydro/Hydro_umdrv_eb.cpp:14-191)*/
/* ======================================================================== */

/* is_cut_neighborhood, Source/EB/eb_utils.H:7-21, ng = 3 along dir only */
static int cut_along(const ora_ebgeom *G, int i, int j, int k, int dir, int ng) {
    int has = fl_single(FLG(G, i, j, k));
    const int e[3] = {dir == 0, dir == 1, dir == 2};
    for (int n = 1; n <= ng; n++)
        has |= fl_single(FLG(G, i - n * e[0], j - n * e[1], k - n * e[2])) ||
               fl_single(FLG(G, i + n * e[0], j + n * e[1], k + n * e[2]));
    return has;
}

/* flatten_eb, Source/Hydro/Godunov/flatten.H:104-134 */
static double flatten_eb(const ora_fab *q, const ora_ebgeom *G, int i, int j, int k, int dir) {
    int ivp3[3] = {i, j, k}, ivm3[3] = {i, j, k};
    ivp3[dir] += 3; ivm3[dir] -= 3;
    if (!fab_contains(q, ivp3) || !fab_contains(q, ivm3)) return 1.0;
    if (cut_along(G, i, j, k, dir, 3)) return 1.0;
    return ora_flatten(q, i, j, k, dir);
}

/* plm_slope_eb / plm_pslope_eb, Source/Hydro/Godunov/PLM_eb.H:18-100 */
static double plm_slope_eb(const ora_fab *q, const ora_ebgeom *G, int i, int j, int k,
                           int n, int dir, double flat, int order, int pslope,
                           double dx, const ora_fab *srcQ) {
    const int e[3] = {dir == 0, dir == 1, dir == 2};
#define CV(o) fl_covered(FLG(G, i + (o) * e[0], j + (o) * e[1], k + (o) * e[2]))
    int ord = 0;
    if (order == 4 && !CV(0) && !CV(-1) && !CV(-2) && !CV(1) && !CV(2)) ord = order;
    else if (order > 1 && !CV(0) && !CV(-1) && !CV(1)) ord = 2;
#undef CV
    if (!ord) return 0.0;
    return pslope ? plm_pslope(q, i, j, k, n, dir, dx, srcQ, flat, ord)
                  : ora_plm_slope(q, i, j, k, n, dir, flat, ord);
}

/* Godunov_umeth_eb, Source/Hydro/Godunov/Godunov_3D_eb.cpp:16-567.  `btf` is
 * bx_to_fill = grow(bx, 2); flx[d] (zeroed by the caller) and qec[d] receive
 * the final fluxes and Godunov states on faces of btf grown by one
 * tangentially, where the aperture is > 0. */
static void godunov_umeth_eb(obox btf, const ora_params *P, const ora_ebgeom *G,
                             const ora_fab *q, const ora_fab *qaux, const ora_fab *srcQ,
                             ora_fab *flx, ora_fab *qec, double dt) {
    const double dx = P->dx[0], dy = P->dx[1], dz = P->dx[2];
    const double hdt = 0.5 * dt;
    const double hdtd[3] = {hdt / dx, hdt / dy, hdt / dz};
    const double cdt[3] = {1.0 / 3.0 * dt / dx, 1.0 / 3.0 * dt / dy, 1.0 / 3.0 * dt / dz};
    const double del[3] = {dx, dy, dz};
    const int *bclo = P->bclo, *bchi = P->bchi, *dlo = P->domlo, *dhi = P->domhi;
    const double small = P->small, sd = P->small_dens, sp = P->small_pres, su = P->smallu;
    const int trd = P->transverse_reset_density;
    const ora_fab *ap = G->ap;
    const obox bxg1 = bgrow(btf, 1), bxg2 = bgrow(btf, 2);

    ora_fab qdm[3], qdp[3];
    for (int d = 0; d < 3; ++d) {
        fab_alloc_box(&qdm[d], bgrowhi(bxg2, d, 1), O_QVAR);
        fab_alloc_box(&qdp[d], bxg2, O_QVAR);
    }
    /* PLM slopes and traces on bxg2, covered cells skipped (:94-152) */
    for (int k = bxg2.lo[2]; k <= bxg2.hi[2]; ++k)
    for (int j = bxg2.lo[1]; j <= bxg2.hi[1]; ++j)
    for (int i = bxg2.lo[0]; i <= bxg2.hi[0]; ++i) {
        if (fl_covered(FLG(G, i, j, k))) continue;
        double slope[O_QVAR];
        double flat = 1.0;
        if (P->use_flattening == 1)
            for (int df = 0; df < 3; df++) flat = smin(flat, flatten_eb(q, G, i, j, k, df));
        for (int idir = 0; idir < 3; ++idir) {
            for (int n = 0; n < O_QVAR; ++n)
                slope[n] = plm_slope_eb(q, G, i, j, k, n, idir, flat, P->plm_iorder,
                                        n == O_QPRES && P->use_pslope, del[idir], srcQ);
            plm_d(i, j, k, idir, &qdm[idir], &qdp[idir], slope, q,
                  FA(qaux, i, j, k, O_QC), del[idir], dt, sd, sp, &ap[idir]);
        }
    }
    /* first fluxes where the aperture is open (:158-212) */
    ora_fab f[3], qgd[3];
    for (int d = 0; d < 3; ++d) {
        const obox b = bnodes(bgrowd(bxg2, d, -1), d);
        fab_alloc_box(&f[d], b, O_NVAR);
        fab_alloc_box(&qgd[d], b, O_NGDNV);
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i)
            if (FA(&ap[d], i, j, k, 0) > 0.)
                cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qdm[d], &qdp[d],
                       &f[d], &qgd[d], qaux, d, small, sd, sp, su);
    }
    /* transverse corrections d|t and their Riemann solves (:214-455) */
    ora_fab fl2[3][3], qg2[3][3];
    for (int d = 0; d < 3; ++d) {
        const int tt[2] = {(d == 0) ? 1 : 0, (d == 2) ? 1 : 2};
        const obox mbx = bgrowhi(bxg2, d, 1);
        for (int w = 0; w < 2; ++w) {
            const int t = tt[w], o = tt[1 - w];
            ora_fab qmT, qpT;
            fab_alloc_box(&qmT, mbx, O_QVAR);
            fab_alloc_box(&qpT, mbx, O_QVAR);
            const obox tb = bgrowd(bxg2, t, -1);
            for (int k = tb.lo[2]; k <= tb.hi[2]; ++k)
            for (int j = tb.lo[1]; j <= tb.hi[1]; ++j)
            for (int i = tb.lo[0]; i <= tb.hi[0]; ++i)
                if (!fl_covered(FLG(G, i, j, k)))
                    transdo(i, j, k, d, t, &qmT, &qpT, &qdm[d], &qdp[d], &f[t], qaux,
                            &qgd[t], cdt[t], trd, sp, &ap[d], &ap[t]);
            fab_alloc_box(&fl2[d][t], bnodes(bxg2, d), O_NVAR);
            fab_alloc_box(&qg2[d][t], bnodes(bxg2, d), O_NGDNV);
            const obox cb = bgrowd(bnodes(bxg1, d), o, 1);
            for (int k = cb.lo[2]; k <= cb.hi[2]; ++k)
            for (int j = cb.lo[1]; j <= cb.hi[1]; ++j)
            for (int i = cb.lo[0]; i <= cb.hi[0]; ++i)
                if (FA(&ap[d], i, j, k, 0) > 0.)
                    cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qmT, &qpT,
                           &fl2[d][t], &qg2[d][t], qaux, d, small, sd, sp, su);
            ora_fab_free(&qmT); ora_fab_free(&qpT);
        }
    }
    for (int d = 0; d < 3; ++d) { ora_fab_free(&f[d]); ora_fab_free(&qgd[d]); }
    /* final states on bxg1 and fluxes on faces of btf grown tangentially
     * (:457-565) */
    ora_fab qm, qp;
    fab_alloc_box(&qm, bxg2, O_QVAR);
    fab_alloc_box(&qp, bxg1, O_QVAR);
    for (int d = 0; d < 3; ++d) {
        const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2;
        for (int k = bxg1.lo[2]; k <= bxg1.hi[2]; ++k)
        for (int j = bxg1.lo[1]; j <= bxg1.hi[1]; ++j)
        for (int i = bxg1.lo[0]; i <= bxg1.hi[0]; ++i)
            if (!fl_covered(FLG(G, i, j, k)))
                transdd(i, j, k, d, &qm, &qp, &qdm[d], &qdp[d], &fl2[t0][t1], &fl2[t1][t0],
                        &qg2[t0][t1], &qg2[t1][t0], qaux, srcQ, hdt, hdtd[t0], hdtd[t1],
                        trd, sp, &ap[d], &ap[t0], &ap[t1]);
        const obox fb = bgrowd(bgrowd(bnodes(btf, d), t0, 1), t1, 1);
        for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
        for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
        for (int i = fb.lo[0]; i <= fb.hi[0]; ++i)
            if (FA(&ap[d], i, j, k, 0) > 0.)
                cmpflx(i, j, k, bclo[d], bchi[d], dlo[d], dhi[d], &qm, &qp, &flx[d],
                       &qec[d], qaux, d, small, sd, sp, su);
    }
    ora_fab_free(&qm); ora_fab_free(&qp);
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < 3; ++t)
            if (t != d) { ora_fab_free(&fl2[d][t]); ora_fab_free(&qg2[d][t]); }
    for (int d = 0; d < 3; ++d) { ora_fab_free(&qdm[d]); ora_fab_free(&qdp[d]); }
}

/* eb_divu, Source/Hydro/Hydro_utils_eb_K.H:31-78 (node i,j,k) */
static double eb_divu_node(const ora_fab *q, const ora_fab *vf, int i, int j, int k,
                           const double dxinv[3]) {
    const double sv = FA(vf, i, j, k, 0) + FA(vf, i - 1, j, k, 0) + FA(vf, i, j - 1, k, 0) +
                      FA(vf, i - 1, j - 1, k, 0) + FA(vf, i, j, k - 1, 0) +
                      FA(vf, i - 1, j, k - 1, 0) + FA(vf, i, j - 1, k - 1, 0) +
                      FA(vf, i - 1, j - 1, k - 1, 0);
    const double cutoff = 8.0 * (1.0 - 1.e-12);
    if (!(sv >= cutoff)) return 0.0;
#define Q(a, b, c, n) FA(q, a, b, c, n)
    const double ux = 0.25 *
        (Q(i, j, k, O_QU) - Q(i - 1, j, k, O_QU) + Q(i, j, k - 1, O_QU) - Q(i - 1, j, k - 1, O_QU) +
         Q(i, j - 1, k, O_QU) - Q(i - 1, j - 1, k, O_QU) + Q(i, j - 1, k - 1, O_QU) -
         Q(i - 1, j - 1, k - 1, O_QU)) * dxinv[0];
    const double vy = 0.25 *
        (Q(i, j, k, O_QV) - Q(i, j - 1, k, O_QV) + Q(i, j, k - 1, O_QV) - Q(i, j - 1, k - 1, O_QV) +
         Q(i - 1, j, k, O_QV) - Q(i - 1, j - 1, k, O_QV) + Q(i - 1, j, k - 1, O_QV) -
         Q(i - 1, j - 1, k - 1, O_QV)) * dxinv[1];
    const double wz = 0.25 *
        (Q(i, j, k, O_QW) - Q(i, j, k - 1, O_QW) + Q(i - 1, j, k, O_QW) - Q(i - 1, j, k - 1, O_QW) +
         Q(i, j - 1, k, O_QW) - Q(i, j - 1, k - 1, O_QW) + Q(i - 1, j - 1, k, O_QW) -
         Q(i - 1, j - 1, k - 1, O_QW)) * dxinv[2];
#undef Q
    return ux + vy + wz;
}

/* compute_hyp_wallflux, Source/Hydro/Hydro_hypflux_eb_K.H:12-68 */
static void hyp_wallflux(const ora_fab *q, const ora_fab *qa, int i, int j, int k,
                         double axm, double axp, double aym, double ayp, double azm,
                         double azp, const ora_params *P, double fluxw[O_NVAR]) {
    for (int n = 0; n < O_NVAR; ++n) fluxw[n] = 0.;
    const double u = FA(q, i, j, k, O_QU), v = FA(q, i, j, k, O_QV), w = FA(q, i, j, k, O_QW);
    const double apnorm = sqrt((axm - axp) * (axm - axp) + (aym - ayp) * (aym - ayp) +
                               (azm - azp) * (azm - azp));
    const double apnorminv = 1. / apnorm;
    const double un = u * (axm - axp) * apnorminv + v * (aym - ayp) * apnorminv +
                      w * (azm - azp) * apnorminv;
    const double rho = FA(q, i, j, k, O_QRHO), p = FA(q, i, j, k, O_QPRES);
    const double qrei = FA(q, i, j, k, O_QREINT), gamc = FA(qa, i, j, k, O_QGAMC);
    const double cav = FA(qa, i, j, k, O_QC), smallc = FA(qa, i, j, k, O_QCSML);
    double o[12];
    ora_riemannus(rho, un, 0., 0., p, qrei, rho, -un, 0., 0., p, qrei, cav, gamc, gamc,
                  P->small, smallc, P->small_dens, P->small_pres, P->smallu, o);
    fluxw[O_URHO] = o[1];
    fluxw[O_UEDEN] = o[5];
    fluxw[O_UEINT] = o[6];
    fluxw[O_UMX] = (axm - axp) * o[2];
    fluxw[O_UMY] = (aym - ayp) * o[2];
    fluxw[O_UMZ] = (azm - azp) * o[2];
}

/* Face value of direction d at face (i,j,k), moved to the face centroid by
 * the bilinear interpolation of eb_compute_div / eb_pdivu
 * (Hydro_utils_eb_K.H:161-179, 437-447): u is the face array, comp its
 * component, c the cell whose connectivity is consulted. */
static double face_centroid_value(const ora_fab *u, int comp, const ora_fab *fc,
                                  const ora_fab *apd, const ora_ebgeom *G, int d,
                                  int i, int j, int k) {
    double v = FA(u, i, j, k, comp);
    const double a = FA(apd, i, j, k, 0);
    if (a != 0.0 && a != 1.0) {
        const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2;
        int o0[3] = {0, 0, 0}, o1[3] = {0, 0, 0};
        o0[t0] = (int)copysign(1.0, FA(fc, i, j, k, 0));
        o1[t1] = (int)copysign(1.0, FA(fc, i, j, k, 1));
        const int32_t w = FLG(G, i, j, k);
        const double f0 = fl_conn(w, o0[0], o0[1], o0[2]) ? fabs(FA(fc, i, j, k, 0)) : 0.0;
        const double f1 = fl_conn(w, o1[0], o1[1], o1[2]) ? fabs(FA(fc, i, j, k, 1)) : 0.0;
        v = (1.0 - f0) * (1.0 - f1) * v +
            f0 * (1.0 - f1) * FA(u, i + o0[0], j + o0[1], k + o0[2], comp) +
            f1 * (1.0 - f0) * FA(u, i + o1[0], j + o1[1], k + o1[2], comp) +
            f0 * f1 * FA(u, i + o0[0] + o1[0], j + o0[1] + o1[1], k + o0[2] + o1[2], comp);
    }
    return v;
}

/* ApplyMLRedistribution with "FluxRedist" on a single level (AMReX
 * AMReX_EB_Redistribution.cpp, amrex_flux_redistribute; as_crse = as_fine =
 * 0, use_wts_in_divnc = false): the published algorithm, restated.  dsdt =
 * divc + the redistributed mass deficit of the cut cells.  The scatter is
 * done in the sequential source order of the CPU build. */
static void flux_redistribute(obox bx, const ora_params *P, const ora_ebgeom *G,
                              ora_fab *dsdt, const ora_fab *divc, const ora_fab *wt) {
    const ora_fab *vf = &G->vfrac;
    const obox g1 = bgrow(bx, 1), g2 = bgrow(bx, 2);
    obox dbox2;
    for (int d = 0; d < 3; ++d) {
        const int gp = P->is_periodic[d] ? 2 : 0;
        dbox2.lo[d] = P->domlo[d] - gp; dbox2.hi[d] = P->domhi[d] + gp;
    }
#define IN2(a, b, c) ((a) >= dbox2.lo[0] && (a) <= dbox2.hi[0] && (b) >= dbox2.lo[1] && \
                      (b) <= dbox2.hi[1] && (c) >= dbox2.lo[2] && (c) <= dbox2.hi[2])
    ora_fab optmp, delm;
    fab_alloc_box(&optmp, g2, O_NVAR);
    fab_alloc_box(&delm, g1, O_NVAR);
    for (int n = 0; n < O_NVAR; ++n)
    for (int k = g1.lo[2]; k <= g1.hi[2]; ++k)
    for (int j = g1.lo[1]; j <= g1.hi[1]; ++j)
    for (int i = g1.lo[0]; i <= g1.hi[0]; ++i) {
        const int32_t w = FLG(G, i, j, k);
        if (fl_single(w)) {
            double vtot = 0.0, divnc = 0.0;
            for (int kk = -1; kk <= 1; kk++)
            for (int jj = -1; jj <= 1; jj++)
            for (int ii = -1; ii <= 1; ii++)
                if ((ii != 0 || jj != 0 || kk != 0) && fl_conn(w, ii, jj, kk) &&
                    IN2(i + ii, j + jj, k + kk)) {
                    const double wf = FA(vf, i + ii, j + jj, k + kk, 0) * 1.0;
                    vtot += wf;
                    divnc += wf * FA(divc, i + ii, j + jj, k + kk, n);
                }
            divnc /= vtot;
            const double mask = IN2(i, j, k) ? 1.0 : 0.0;
            FA(&optmp, i, j, k, n) = (1 - FA(vf, i, j, k, 0)) * (divnc - FA(divc, i, j, k, n)) * mask;
            FA(&delm, i, j, k, n) = -(FA(vf, i, j, k, 0)) * FA(&optmp, i, j, k, n);
        } else {
            FA(&delm, i, j, k, n) = 0;
        }
    }
    for (int n = 0; n < O_NVAR; ++n)
    for (int k = g1.lo[2]; k <= g1.hi[2]; ++k)
    for (int j = g1.lo[1]; j <= g1.hi[1]; ++j)
    for (int i = g1.lo[0]; i <= g1.hi[0]; ++i) {
        const int32_t w = FLG(G, i, j, k);
        if (!fl_single(w)) continue;
        double wtot = 0.0;
        for (int kk = -1; kk <= 1; kk++)
        for (int jj = -1; jj <= 1; jj++)
        for (int ii = -1; ii <= 1; ii++)
            if ((ii != 0 || jj != 0 || kk != 0) && fl_conn(w, ii, jj, kk))
                wtot += FA(wt, i + ii, j + jj, k + kk, 0) * FA(vf, i + ii, j + jj, k + kk, 0) *
                        (IN2(i + ii, j + jj, k + kk) ? 1.0 : 0.0);
        wtot = 1.0 / wtot;
        for (int kk = -1; kk <= 1; kk++)
        for (int jj = -1; jj <= 1; jj++)
        for (int ii = -1; ii <= 1; ii++)
            if ((ii != 0 || jj != 0 || kk != 0) && fl_conn(w, ii, jj, kk)) {
                const double drho = FA(&delm, i, j, k, n) * wtot * FA(wt, i + ii, j + jj, k + kk, 0) *
                                    (IN2(i + ii, j + jj, k + kk) ? 1.0 : 0.0);
                FA(&optmp, i + ii, j + jj, k + kk, n) += drho;
            }
    }
#undef IN2
    for (int n = 0; n < O_NVAR; ++n)
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i)
        if (!fl_covered(FLG(G, i, j, k)))
            FA(dsdt, i, j, k, n) = FA(divc, i, j, k, n) + FA(&optmp, i, j, k, n);
    ora_fab_free(&optmp);
    ora_fab_free(&delm);
}

/* hydro_umdrv_eb, Source/Hydro/Hydro_umdrv_eb.cpp:14-191 (Godunov branch,
 * FluxRedist): q/qaux/srcq on grow(bx, 6), uin (Sborder) likewise; dsdt on
 * bx is overwritten on non-covered cells; flx[d] (faces of bx) receive the
 * face-centroid fluxes (extensive, for flux registers). */
void ora_hydro_umdrv_eb(const int lo[3], const int hi[3], const ora_fab *uin,
                        ora_fab *dsdt, const ora_fab *q, const ora_fab *qa,
                        const ora_fab *srcq, ora_fab flx[3], double dt,
                        const ora_params *P, const ora_ebgeom *G) {
    const obox bx = mkbox(lo, hi);
    const obox bxg_i = bgrow(bx, 2), bxg_ii = bgrow(bxg_i, 1);
    const double *dx = P->dx;
    const double dxinv[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
    ora_fab ftmp[3], qec[3];
    for (int d = 0; d < 3; ++d) {
        fab_alloc_box(&ftmp[d], bnodes(bxg_ii, d), O_NVAR);     /* setVal(0) */
        /* qec is not initialised in the reference; values of closed faces
         * are never used except through the connectivity-weighted
         * interpolation of eb_pdivu, which can reach them: zero here */
        fab_alloc_box(&qec[d], bnodes(bgrow(bxg_ii, 1), d), O_NGDNV);
    }
    godunov_umeth_eb(bxg_i, P, G, q, qa, srcq, ftmp, qec, dt);

    /* adjust_fluxes_eb, Hydro_adjust_fluxes.cpp:48-170 */
    {
        obox qb = {{q->lo[0], q->lo[1], q->lo[2]}, {q->hi[0], q->hi[1], q->hi[2]}};
        obox nb = bnodes(bnodes(bnodes(bgrow(qb, -1), 0), 1), 2);
        for (int d = 0; d < 3; ++d) {
            const int g = P->is_periodic[d] ? 16 : 0;
            const int dl = P->domlo[d] - g, dh = P->domhi[d] + 1 + g;
            if (nb.lo[d] < dl) nb.lo[d] = dl;
            if (nb.hi[d] > dh) nb.hi[d] = dh;
        }
        ora_fab divu;
        obox db = bnodes(bnodes(bnodes(bgrow(qb, -1), 0), 1), 2);
        fab_alloc_box(&divu, db, 1);
        for (int k = nb.lo[2]; k <= nb.hi[2]; ++k)
        for (int j = nb.lo[1]; j <= nb.hi[1]; ++j)
        for (int i = nb.lo[0]; i <= nb.hi[0]; ++i)
            FA(&divu, i, j, k, 0) = eb_divu_node(q, &G->vfrac, i, j, k, dxinv);
        const double areafac[3] = {dx[1] * dx[2], dx[0] * dx[2], dx[1] * dx[0]};
        for (int d = 0; d < 3; ++d) {
            const obox fb = bnodes(bxg_ii, d);
            for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
            for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
            for (int i = fb.lo[0]; i <= fb.hi[0]; ++i) {
                artif_visc(i, j, k, &ftmp[d], &divu, uin, dx[d], P->difmag, d,
                           P->domlo[d], P->domhi[d], P->bclo[d], P->bchi[d]);
                if (FA(&G->ap[d], i, j, k, 0) > 0.) norm_spec_flx(i, j, k, &ftmp[d]);
                if (FA(&G->ap[d], i, j, k, 0) > 0.)
                    for (int n = 0; n < O_NVAR; ++n) FA(&ftmp[d], i, j, k, n) *= areafac[d];
            }
        }
        ora_fab_free(&divu);
    }

    /* hydro_consup_eb, Hydro_consup.cpp:29-98: eb_compute_div + eb_pdivu */
    ora_fab divc, wt;
    fab_alloc_box(&divc, bxg_i, O_NVAR);
    fab_alloc_box(&wt, bxg_i, 1);
    const double cell_vol_inv = dxinv[0] * dxinv[1] * dxinv[2];
    const ora_fab *ap = G->ap, *fc = G->fc;
    for (int n = 0; n < O_NVAR; ++n)
    for (int k = bxg_i.lo[2]; k <= bxg_i.hi[2]; ++k)
    for (int j = bxg_i.lo[1]; j <= bxg_i.hi[1]; ++j)
    for (int i = bxg_i.lo[0]; i <= bxg_i.hi[0]; ++i) {
        const int32_t w = FLG(G, i, j, k);
        if (fl_covered(w)) continue;
        const int valid = i >= bx.lo[0] && i <= bx.hi[0] && j >= bx.lo[1] && j <= bx.hi[1] &&
                          k >= bx.lo[2] && k <= bx.hi[2];
        const int xh = (i == bx.hi[0]), yh = (j == bx.hi[1]), zh = (k == bx.hi[2]);
        double dv;
        if (fl_regular(w)) {
            dv = cell_vol_inv * (FA(&ftmp[0], i + 1, j, k, n) - FA(&ftmp[0], i, j, k, n)) +
                 cell_vol_inv * (FA(&ftmp[1], i, j + 1, k, n) - FA(&ftmp[1], i, j, k, n)) +
                 cell_vol_inv * (FA(&ftmp[2], i, j, k + 1, n) - FA(&ftmp[2], i, j, k, n));
            if (valid) {
                FA(&flx[0], i, j, k, n) = FA(&ftmp[0], i, j, k, n);
                FA(&flx[1], i, j, k, n) = FA(&ftmp[1], i, j, k, n);
                FA(&flx[2], i, j, k, n) = FA(&ftmp[2], i, j, k, n);
                if (xh) FA(&flx[0], i + 1, j, k, n) = FA(&ftmp[0], i + 1, j, k, n);
                if (yh) FA(&flx[1], i, j + 1, k, n) = FA(&ftmp[1], i, j + 1, k, n);
                if (zh) FA(&flx[2], i, j, k + 1, n) = FA(&ftmp[2], i, j, k + 1, n);
            }
        } else {
            const double fxm = face_centroid_value(&ftmp[0], n, &fc[0], &ap[0], G, 0, i, j, k);
            const double fxp = face_centroid_value(&ftmp[0], n, &fc[0], &ap[0], G, 0, i + 1, j, k);
            const double fym = face_centroid_value(&ftmp[1], n, &fc[1], &ap[1], G, 1, i, j, k);
            const double fyp = face_centroid_value(&ftmp[1], n, &fc[1], &ap[1], G, 1, i, j + 1, k);
            const double fzm = face_centroid_value(&ftmp[2], n, &fc[2], &ap[2], G, 2, i, j, k);
            const double fzp = face_centroid_value(&ftmp[2], n, &fc[2], &ap[2], G, 2, i, j, k + 1);
            if (valid) {
                FA(&flx[0], i, j, k, n) = fxm;
                FA(&flx[1], i, j, k, n) = fym;
                FA(&flx[2], i, j, k, n) = fzm;
                if (xh) FA(&flx[0], i + 1, j, k, n) = fxp;
                if (yh) FA(&flx[1], i, j + 1, k, n) = fyp;
                if (zh) FA(&flx[2], i, j, k + 1, n) = fzp;
            }
            const double axm = FA(&ap[0], i, j, k, 0), axp = FA(&ap[0], i + 1, j, k, 0);
            const double aym = FA(&ap[1], i, j, k, 0), ayp = FA(&ap[1], i, j + 1, k, 0);
            const double azm = FA(&ap[2], i, j, k, 0), azp = FA(&ap[2], i, j, k + 1, 0);
            double fw[O_NVAR];
            hyp_wallflux(q, qa, i, j, k, axm, axp, aym, ayp, azm, azp, P, fw);
            const double vfr = FA(&G->vfrac, i, j, k, 0);
            dv = cell_vol_inv / vfr *
                 (axp * fxp - axm * fxm + ayp * fyp - aym * fym + azp * fzp - azm * fzm);
            dv += fw[n] * dxinv[0] / vfr;
        }
        FA(&divc, i, j, k, n) = dv * -1.0;
        FA(&wt, i, j, k, 0) = FA(q, i, j, k, O_QRHO); /* eb_weights_type 2 */
    }
    for (int k = bxg_i.lo[2]; k <= bxg_i.hi[2]; ++k)
    for (int j = bxg_i.lo[1]; j <= bxg_i.hi[1]; ++j)
    for (int i = bxg_i.lo[0]; i <= bxg_i.hi[0]; ++i) {
        const int32_t w = FLG(G, i, j, k);
        if (fl_covered(w)) continue;
        if (fl_regular(w)) {
            const double pdivu =
                0.5 * ((FA(&qec[0], i + 1, j, k, O_GDPRES) + FA(&qec[0], i, j, k, O_GDPRES)) *
                           (FA(&qec[0], i + 1, j, k, O_GDU) - FA(&qec[0], i, j, k, O_GDU)) * dxinv[0] +
                       (FA(&qec[1], i, j + 1, k, O_GDPRES) + FA(&qec[1], i, j, k, O_GDPRES)) *
                           (FA(&qec[1], i, j + 1, k, O_GDV) - FA(&qec[1], i, j, k, O_GDV)) * dxinv[1] +
                       (FA(&qec[2], i, j, k + 1, O_GDPRES) + FA(&qec[2], i, j, k, O_GDPRES)) *
                           (FA(&qec[2], i, j, k + 1, O_GDW) - FA(&qec[2], i, j, k, O_GDW)) * dxinv[2]);
            FA(&divc, i, j, k, O_UEINT) -= pdivu;
        } else {
            double fm[3], fp[3];
            const int gdn[3] = {O_GDU, O_GDV, O_GDW};
            for (int d = 0; d < 3; ++d) {
                int ip[3] = {i, j, k};
                ip[d] += 1;
                fm[d] = (FA(&ap[d], i, j, k, 0) > 0.0)
                            ? face_centroid_value(&qec[d], gdn[d], &fc[d], &ap[d], G, d, i, j, k)
                            : 0.0;
                fp[d] = (FA(&ap[d], ip[0], ip[1], ip[2], 0) > 0.0)
                            ? face_centroid_value(&qec[d], gdn[d], &fc[d], &ap[d], G, d, ip[0], ip[1], ip[2])
                            : 0.0;
            }
            FA(&divc, i, j, k, O_UEINT) -=
                (1.0 / FA(&G->vfrac, i, j, k, 0)) * FA(q, i, j, k, O_QPRES) *
                (dxinv[0] * (FA(&ap[0], i + 1, j, k, 0) * fp[0] - FA(&ap[0], i, j, k, 0) * fm[0]) +
                 dxinv[1] * (FA(&ap[1], i, j + 1, k, 0) * fp[1] - FA(&ap[1], i, j, k, 0) * fm[1]) +
                 dxinv[2] * (FA(&ap[2], i, j, k + 1, 0) * fp[2] - FA(&ap[2], i, j, k, 0) * fm[2]));
        }
    }
    flux_redistribute(bx, P, G, dsdt, &divc, &wt);
    ora_fab_free(&divc);
    ora_fab_free(&wt);
    for (int d = 0; d < 3; ++d) { ora_fab_free(&ftmp[d]); ora_fab_free(&qec[d]); }
}

static void geom_strides(ora_ebgeom *G) {
    G->fjs = (int64_t)(G->fhi[0] - G->flo[0] + 1);
    G->fks = G->fjs * (int64_t)(G->fhi[1] - G->flo[1] + 1);
}

/* Switch a freshly created level to the EB build: ghost width 6
 * (CAMR::numGrow, CAMR.H:455-483), the flags over flo..fhi (which must hold
 * the domain grown by 7), and the per-box path of
 * CAMR_construct_hydro_source.cpp:80-83,150-151. */
int ora_level_set_eb(ora_level *L, const int32_t *flags, const int flo[3], const int fhi[3]) {
    int64_t n = 1;
    for (int d = 0; d < 3; ++d) n *= (int64_t)(fhi[d] - flo[d] + 1);
    L->flags = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    memcpy(L->flags, flags, sizeof(int32_t) * (size_t)n);
    L->P.eb = 1;
    L->ng = EB_NGROW;
    L->geom = (ora_ebgeom *)calloc(L->nbox, sizeof(ora_ebgeom));
    L->kind = (int *)calloc(L->nbox, sizeof(int));
    for (int b = 0; b < L->nbox; ++b) {
        const obox B = L->box[b];
        ora_ebgeom *G = &L->geom[b];
        G->flags = L->flags;
        for (int d = 0; d < 3; ++d) { G->flo[d] = flo[d]; G->fhi[d] = fhi[d]; }
        geom_strides(G);
        ora_fab_free(&L->Sborder[b]); fab_alloc_box(&L->Sborder[b], bgrow(B, L->ng), O_NVAR);
        ora_fab_free(&L->hsrc[b]); fab_alloc_box(&L->hsrc[b], bgrow(B, L->ng), O_NVAR);
        ora_fab_free(&L->oldsrc[b]); fab_alloc_box(&L->oldsrc[b], bgrow(B, L->ng), O_NVAR);
        ora_fab_free(&L->srchyd[b]); fab_alloc_box(&L->srchyd[b], bgrow(B, L->ng), O_NVAR);
        int all_cov = 1, any_nonreg = 0;
        for (int k = B.lo[2] - 2; k <= B.hi[2] + 2; ++k)
        for (int j = B.lo[1] - 2; j <= B.hi[1] + 2; ++j)
        for (int i = B.lo[0] - 2; i <= B.hi[0] + 2; ++i) {
            const int32_t w = FLG(G, i, j, k);
            const int valid = i >= B.lo[0] && i <= B.hi[0] && j >= B.lo[1] && j <= B.hi[1] &&
                              k >= B.lo[2] && k <= B.hi[2];
            if (valid && !fl_covered(w)) all_cov = 0;
            if (!fl_regular(w)) any_nonreg = 1;
        }
        L->kind[b] = all_cov ? 2 : (any_nonreg ? 1 : 0);
    }
    return 0;
}

int ora_level_box_kind(const ora_level *L, int b) { return L->kind ? L->kind[b] : 0; }

/* Cut-cell data of box b on grow(bx, 7): vfrac (cells), apertures ap[d]
 * (faces of direction d), face centroids fc[d] (2 components). */
void ora_level_set_box_eb(ora_level *L, int b, const double *vfrac, const double *apx,
                          const double *apy, const double *apz, const double *fcx,
                          const double *fcy, const double *fcz) {
    ora_ebgeom *G = &L->geom[b];
    const obox g = bgrow(L->box[b], 7);
    const double *ap[3] = {apx, apy, apz}, *fc[3] = {fcx, fcy, fcz};
    ora_fab_free(&G->vfrac);
    fab_alloc_box(&G->vfrac, g, 1);
    memcpy(G->vfrac.p, vfrac, sizeof(double) * (size_t)G->vfrac.ns);
    for (int d = 0; d < 3; ++d) {
        ora_fab_free(&G->ap[d]); ora_fab_free(&G->fc[d]);
        fab_alloc_box(&G->ap[d], bnodes(g, d), 1);
        fab_alloc_box(&G->fc[d], bnodes(g, d), 2);
        memcpy(G->ap[d].p, ap[d], sizeof(double) * (size_t)G->ap[d].ns);
        memcpy(G->fc[d].p, fc[d], sizeof(double) * (size_t)(2 * G->fc[d].ns));
    }
}

/* ======================================================================== */
/* Coarse-to-fine ghost data: FillPatch's interpolation from the coarser     */
/* level (AmrLevel::FillPatch, CAMR.cpp:1247, with the State_Type            */
/* interpolater cell_cons_interp, Source/Utils/CAMR_setup.cpp:141-144)       */
/* ======================================================================== */
/* Value of component n of the coarse level at cell c (any cell: ghost cells
 * map through the domain BCs as FillPatch does). */
/* Coarse data at the fill time: new only (tw.mode 0), old only (1), or
 * alpha * old + beta * new (2), StateData::InterpAddBox's choice. */
typedef struct { int mode; double alpha, beta; } otimew;

static double level_value(const ora_level *L, const int c[3], int n, otimew tw) {
    int m[3];
    double s;
    map_cell(&L->P, c, n, m, &s);
    const int ob = L->owner[dom_index(L, m)];
    double v = 0.0;
    if (ob >= 0) {
        const double vn = FA(&L->Snew[ob], m[0], m[1], m[2], n);
        const double vo = FA(&L->Sold[ob], m[0], m[1], m[2], n);
        v = tw.mode == 0 ? vn : (tw.mode == 1 ? vo : tw.alpha * vo + tw.beta * vn);
    }
    return (s < 0.0) ? -v : v;
}

/* AMReX's CellConservativeLinear (linear limiting), restated: per
 * direction the monotonized-central slope of the coarse data, then one
 * factor alpha in [0,1] for the slope vector so that every child value lies
 * within the min / max of the 27 coarse neighbours.  The children of a
 * coarse cell average back to it (the offsets are symmetric). */
static double cons_interp(const ora_level *C, const int cc[3], int n, const double off[3],
                          double edge, otimew tw) {
    double u[3][3][3];
    for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
    for (int d = -1; d <= 1; ++d) {
        const int c[3] = {cc[0] + d, cc[1] + b, cc[2] + a};
        u[a + 1][b + 1][d + 1] = level_value(C, c, n, tw);
    }
    const double u0 = u[1][1][1];
    const double um[3] = {u[1][1][0], u[1][0][1], u[0][1][1]};
    const double up[3] = {u[1][1][2], u[1][2][1], u[2][1][1]};
    double s[3];
    for (int d = 0; d < 3; ++d) {
        const double cen = 0.5 * (up[d] - um[d]);
        const double forw = 2.0 * (up[d] - u0);
        const double back = 2.0 * (u0 - um[d]);
        const double slp = (forw * back >= 0.0) ? smin(fabs(forw), fabs(back)) : 0.0;
        s[d] = copysign(1.0, cen) * smin(slp, fabs(cen));
    }
    double cmn = u0, cmx = u0;
    for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
    for (int d = 0; d < 3; ++d) {
        cmn = smin(cmn, u[a][b][d]);
        cmx = smax(cmx, u[a][b][d]);
    }
    const double dmn = cmn - u0, dmx = cmx - u0;
    double alpha = 1.0;
    for (int cz = -1; cz <= 1; cz += 2)
    for (int cy = -1; cy <= 1; cy += 2)
    for (int cx = -1; cx <= 1; cx += 2) {
        /* the extreme children: the outermost fine cells, at +-edge */
        const double dd = cx * edge * s[0] + cy * edge * s[1] + cz * edge * s[2];
        if (dd > dmx) alpha = smin(alpha, dmx / dd);
        if (dd < dmn) alpha = smin(alpha, dmn / dd);
    }
    return u0 + off[0] * (alpha * s[0]) + off[1] * (alpha * s[1]) + off[2] * (alpha * s[2]);
}

/* The exterior data of fine level F (every cell of its domain, NVAR
 * components) interpolated from coarse level C with refinement `ratio`. */
static int exterior_from_coarse_t(ora_level *F, const ora_level *C, int ratio, otimew tw) {
    for (int d = 0; d < 3; ++d)
        if (F->P.domlo[d] != C->P.domlo[d] * ratio ||
            F->P.domhi[d] + 1 != (C->P.domhi[d] + 1) * ratio) return 1;
    const int64_t nc = F->dn[0] * F->dn[1] * F->dn[2];
    if (!F->ext) F->ext = (double *)malloc(sizeof(double) * (size_t)(nc * O_NVAR));
    const double half = 0.5 * (ratio - 1);
    for (int n = 0; n < O_NVAR; ++n)
    for (int k = F->P.domlo[2]; k <= F->P.domhi[2]; ++k)
    for (int j = F->P.domlo[1]; j <= F->P.domhi[1]; ++j)
    for (int i = F->P.domlo[0]; i <= F->P.domhi[0]; ++i) {
        const int f[3] = {i, j, k};
        int cc[3];
        double off[3];
        for (int d = 0; d < 3; ++d) {
            cc[d] = (f[d] - F->P.domlo[d]) / ratio + C->P.domlo[d];
            const int sub = (f[d] - F->P.domlo[d]) - (cc[d] - C->P.domlo[d]) * ratio;
            off[d] = (sub - half) / ratio; /* fine-centre offset in coarse cells */
        }
        F->ext[dom_index(F, f) + (int64_t)n * nc] = cons_interp(C, cc, n, off, half / ratio, tw);
    }
    return 0;
}

int ora_level_exterior_from_coarse(ora_level *F, const ora_level *C, int ratio) {
    const otimew tw = {0, 0.0, 0.0};
    return exterior_from_coarse_t(F, C, ratio, tw);
}

/* The exterior of F from C at time t between C's old state (t_old) and new
 * state (t_new), StateData::InterpAddBox's rule, over F's whole domain. */
int ora_level_exterior_from_coarse_time(ora_level *F, const ora_level *C, int ratio,
                                        double t_old, double t_new, double t) {
    const double teps = (t_new - t_old) * 1.e-3;
    otimew tw = {2, (t_new - t) / (t_new - t_old), (t - t_old) / (t_new - t_old)};
    if (t > t_old - teps && t < t_old + teps) tw.mode = 1;
    else if (t > t_new - teps && t < t_new + teps) tw.mode = 0;
    return exterior_from_coarse_t(F, C, ratio, tw);
}

/* Set a level's old state (the time-interpolation source of its finer
 * level's ghost data). */
void ora_level_set_old_box(ora_level *L, int b, const double *data) {
    memcpy(L->Sold[b].p, data, sizeof(double) * (size_t)(L->Sold[b].ns * O_NVAR));
}

/* The same interpolation at just the cells a fill of F reads from the
 * exterior: the domain images of grow(box, ng) that no box of F owns. */
static int exterior_ghosts_from_coarse_t(ora_level *F, const ora_level *C, int ratio, otimew tw) {
    for (int d = 0; d < 3; ++d)
        if (F->P.domlo[d] != C->P.domlo[d] * ratio ||
            F->P.domhi[d] + 1 != (C->P.domhi[d] + 1) * ratio) return 1;
    const int64_t nc = F->dn[0] * F->dn[1] * F->dn[2];
    if (!F->ext) F->ext = (double *)calloc((size_t)(nc * O_NVAR), sizeof(double));
    const double half = 0.5 * (ratio - 1);
    for (int b = 0; b < F->nbox; ++b) {
        const obox G = bgrow(F->box[b], F->ng);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(ora_get_threads())
#endif
        for (int k = G.lo[2]; k <= G.hi[2]; ++k)
        for (int j = G.lo[1]; j <= G.hi[1]; ++j)
        for (int i = G.lo[0]; i <= G.hi[0]; ++i) {
            int f[3] = {i, j, k};
            int ok = 1;
            for (int d = 0; d < 3; ++d) {
                const int lo = F->P.domlo[d], n = F->P.domhi[d] - lo + 1;
                if (f[d] < lo || f[d] > lo + n - 1) {
                    if (!F->P.is_periodic[d]) { ok = 0; break; }
                    f[d] = lo + ((f[d] - lo) % n + n) % n;
                }
            }
            if (!ok || F->owner[dom_index(F, f)] >= 0) continue;
            int cc[3];
            double off[3];
            for (int d = 0; d < 3; ++d) {
                cc[d] = (f[d] - F->P.domlo[d]) / ratio + C->P.domlo[d];
                const int sub = (f[d] - F->P.domlo[d]) - (cc[d] - C->P.domlo[d]) * ratio;
                off[d] = (sub - half) / ratio;
            }
            for (int n = 0; n < O_NVAR; ++n)
                F->ext[dom_index(F, f) + (int64_t)n * nc] = cons_interp(C, cc, n, off, half / ratio, tw);
        }
    }
    return 0;
}

void ora_level_get_exterior(const ora_level *L, double *out) {
    if (L->ext)
        memcpy(out, L->ext, sizeof(double) * (size_t)(L->dn[0] * L->dn[1] * L->dn[2] * O_NVAR));
}

/* The force on the embedded body from the level's current state: the sum
 * over the cut cells of the wall momentum flux (compute_hyp_wallflux's
 * fluxw[UMX..UMZ], Hydro_hypflux_eb_K.H:65) times dx^2, summed box by box,
 * cell by cell (x fastest).  CAMR computes no force; this is the x-force
 * functional of the Sod-past-a-sphere workload (and its y, z analogues). */
int ora_level_eb_force(ora_level *L, double f[3]) {
    if (!L->kind || !L->geom) return 1;
    f[0] = f[1] = f[2] = 0.0;
    const double dx2 = L->P.dx[0] * L->P.dx[0];
    for (int b = 0; b < L->nbox; ++b) {
        if (L->kind[b] != 1) continue;
        const obox B = L->box[b];
        const ora_ebgeom *G = &L->geom[b];
        ora_fab q, qa;
        fab_alloc_box(&q, B, O_QVAR);
        fab_alloc_box(&qa, B, O_NQAUX);
        int err = 0;
        ora_ctoprim_box(B.lo, B.hi, &L->Snew[b], &q, &qa, &L->P, &err);
        const int64_t np = q.ns;
        double *c = (double *)calloc((size_t)(3 * np), sizeof(double));
        int64_t idx = 0;
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i, ++idx) {
            if (!fl_single(FLG(G, i, j, k))) continue;
            const double axm = FA(&G->ap[0], i, j, k, 0), axp = FA(&G->ap[0], i + 1, j, k, 0);
            const double aym = FA(&G->ap[1], i, j, k, 0), ayp = FA(&G->ap[1], i, j + 1, k, 0);
            const double azm = FA(&G->ap[2], i, j, k, 0), azp = FA(&G->ap[2], i, j, k + 1, 0);
            double fw[O_NVAR];
            hyp_wallflux(&q, &qa, i, j, k, axm, axp, aym, ayp, azm, azp, &L->P, fw);
            for (int d = 0; d < 3; ++d) c[d * np + idx] = fw[O_UMX + d] * dx2;
        }
        for (int d = 0; d < 3; ++d)
            for (int64_t p = 0; p < np; ++p) f[d] += c[d * np + p];
        free(c);
        ora_fab_free(&q);
        ora_fab_free(&qa);
        if (err) return err;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Coarse-fine synchronisation: the flux register, reflux, average_down and */
/* the two-level subcycled step.  CAMR uses AMReX's YAFluxRegister          */
/* (CAMR.H:310, CrseAdd / FineAdd at CAMR_construct_hydro_source.cpp:       */
/* 243-254, Reflux at CAMR.cpp:803-851), average_down (CAMR.cpp:1047-1066)  */
/* and Amr::timeStep; AMReX is absent, so these are restated from their     */
/* published definitions:                                                   */
/*  - coarse cells are flagged fine (covered by the coarsened fine boxes),  */
/*    boundary (not covered, within one cell of a covered one) or neither;  */
/*  - CrseAdd: a boundary cell takes -/+ dt/vol_c times the coarse flux     */
/*    through each face it shares with a covered cell;                      */
/*  - FineAdd: the same cell takes +/- dt_f/(vol_f r^3) times each fine     */
/*    flux on that face, one addition per fine face, in the order (outer,   */
/*    inner) = (k, j) for x faces, (k, i) for y, (j, i) for z;              */
/*  - Reflux: crse_data += cfp, then S_new += crse_data on every valid cell;*/
/*  - average_down: a covered coarse cell = (1/r^3) * the sum of its        */
/*    children in (k, j, i) order.                                          */
/* Coarse boundary cells bordering two different fine boxes are not        */
/* supported (where AMReX's coarse-fine patches overlap it weighs them).    */
/* ------------------------------------------------------------------------ */
enum { FR_NONE = 0, FR_BND = 1, FR_FINE = 2 };
struct ora_fluxreg {
    ora_level *C, *F;
    int ratio;
    unsigned char *flag; /* per coarse domain cell */
    ora_fab *crse, *cfp; /* per coarse box, NVAR components */
};

/* periodic image of a coarse index; 0 outside a non-periodic domain */
static int fr_map(const ora_level *C, const int c[3], int m[3]) {
    for (int d = 0; d < 3; ++d) {
        const int lo = C->P.domlo[d], n = C->P.domhi[d] - lo + 1;
        int x = c[d];
        if (x < lo || x > lo + n - 1) {
            if (!C->P.is_periodic[d]) return 0;
            x = lo + ((x - lo) % n + n) % n;
        }
        m[d] = x;
    }
    return 1;
}

static int fr_flag(const ora_fluxreg *R, const int c[3]) {
    int m[3];
    if (!fr_map(R->C, c, m)) return FR_NONE;
    return R->flag[dom_index(R->C, m)];
}

void ora_fluxreg_destroy(ora_fluxreg *R) {
    if (!R) return;
    if (R->C->reg_crse == R) R->C->reg_crse = NULL;
    if (R->F->reg_fine == R) R->F->reg_fine = NULL;
    for (int b = 0; b < R->C->nbox; ++b) { ora_fab_free(&R->crse[b]); ora_fab_free(&R->cfp[b]); }
    free(R->crse); free(R->cfp); free(R->flag); free(R);
}

ora_fluxreg *ora_fluxreg_create(ora_level *C, ora_level *F, int ratio) {
    if (ratio < 2) return NULL;
    for (int d = 0; d < 3; ++d)
        if (F->P.domlo[d] != C->P.domlo[d] * ratio ||
            F->P.domhi[d] + 1 != (C->P.domhi[d] + 1) * ratio) return NULL;
    for (int b = 0; b < F->nbox; ++b)
        for (int d = 0; d < 3; ++d)
            if (F->box[b].lo[d] % ratio || (F->box[b].hi[d] + 1) % ratio) return NULL;
    ora_fluxreg *R = (ora_fluxreg *)calloc(1, sizeof(ora_fluxreg));
    R->C = C; R->F = F; R->ratio = ratio;
    const int64_t nc = C->dn[0] * C->dn[1] * C->dn[2];
    R->flag = (unsigned char *)calloc((size_t)nc, 1);
    int *cov = (int *)malloc(sizeof(int) * (size_t)nc); /* covering fine box */
    for (int64_t c = 0; c < nc; ++c) cov[c] = -1;
    for (int b = 0; b < F->nbox; ++b) {
        const obox fb = F->box[b];
        for (int k = fb.lo[2] / ratio; k <= fb.hi[2] / ratio; ++k)
        for (int j = fb.lo[1] / ratio; j <= fb.hi[1] / ratio; ++j)
        for (int i = fb.lo[0] / ratio; i <= fb.hi[0] / ratio; ++i) {
            const int c[3] = {i, j, k};
            R->flag[dom_index(C, c)] = FR_FINE;
            cov[dom_index(C, c)] = b;
        }
    }
    int bad = 0;
    for (int b = 0; b < F->nbox; ++b) {
        const obox fb = F->box[b];
        for (int k = fb.lo[2] / ratio - 1; k <= fb.hi[2] / ratio + 1; ++k)
        for (int j = fb.lo[1] / ratio - 1; j <= fb.hi[1] / ratio + 1; ++j)
        for (int i = fb.lo[0] / ratio - 1; i <= fb.hi[0] / ratio + 1; ++i) {
            const int c[3] = {i, j, k};
            int m[3];
            if (!fr_map(C, c, m)) continue;
            if (R->flag[dom_index(C, m)] != FR_FINE) R->flag[dom_index(C, m)] = FR_BND;
        }
    }
    for (int k = C->P.domlo[2]; k <= C->P.domhi[2]; ++k)
    for (int j = C->P.domlo[1]; j <= C->P.domhi[1]; ++j)
    for (int i = C->P.domlo[0]; i <= C->P.domhi[0]; ++i) {
        const int c[3] = {i, j, k};
        if (R->flag[dom_index(C, c)] != FR_BND) continue;
        int seen = -1;
        for (int d = 0; d < 3; ++d)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                int nb[3] = {i, j, k}, m[3];
                nb[d] += sgn;
                if (!fr_map(C, nb, m)) continue;
                const int fbx = cov[dom_index(C, m)];
                if (fbx < 0) continue;
                if (seen >= 0 && fbx != seen) bad = 1;
                seen = fbx;
            }
    }
    free(cov);
    if (bad) { free(R->flag); free(R); return NULL; }
    R->crse = (ora_fab *)calloc(C->nbox, sizeof(ora_fab));
    R->cfp = (ora_fab *)calloc(C->nbox, sizeof(ora_fab));
    for (int b = 0; b < C->nbox; ++b) {
        fab_alloc_box(&R->crse[b], C->box[b], O_NVAR);
        fab_alloc_box(&R->cfp[b], C->box[b], O_NVAR);
    }
    C->reg_crse = R;
    F->reg_fine = R;
    return R;
}

/* YAFluxRegister::CrseAdd over the tile bx of coarse box b */
static void fluxreg_crse_add(ora_fluxreg *R, int b, obox bx, ora_fab flux[3], double dt) {
    const double *dx = R->C->P.dx;
    double dx1 = dx[0];
    dx1 *= dx[1];
    dx1 *= dx[2];
    const double dtdx = dt / dx1;
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
        const int c[3] = {i, j, k};
        if (fr_flag(R, c) != FR_BND) continue;
        for (int d = 0; d < 3; ++d) {
            int cm[3] = {i, j, k}, cp[3] = {i, j, k};
            cm[d] -= 1;
            cp[d] += 1;
            if (fr_flag(R, cm) == FR_FINE) {
                for (int n = 0; n < O_NVAR; ++n)
                    FA(&R->crse[b], i, j, k, n) -= dtdx * FA(&flux[d], i, j, k, n);
            } else if (fr_flag(R, cp) == FR_FINE) {
                for (int n = 0; n < O_NVAR; ++n)
                    FA(&R->crse[b], i, j, k, n) += dtdx * FV(&flux[d], cp, n);
            }
        }
    }
}

/* YAFluxRegister::FineAdd over the fine tile fbx (aligned to the ratio) */
static void fluxreg_fine_add(ora_fluxreg *R, obox fbx, ora_fab flux[3], double dt) {
    const ora_level *C = R->C;
    const int r = R->ratio;
    const double *dx = R->F->P.dx;
    double dx1 = dx[0];
    dx1 *= dx[1];
    dx1 *= dx[2];
    const double dtdx = dt / (dx1 * (double)(r * r * r));
    obox cb;
    for (int d = 0; d < 3; ++d) { cb.lo[d] = fbx.lo[d] / r; cb.hi[d] = fbx.hi[d] / r; }
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            obox sl = cb;
            sl.lo[d] = sl.hi[d] = side ? cb.hi[d] + 1 : cb.lo[d] - 1;
            const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2; /* inner, outer */
            for (int k = sl.lo[2]; k <= sl.hi[2]; ++k)
            for (int j = sl.lo[1]; j <= sl.hi[1]; ++j)
            for (int i = sl.lo[0]; i <= sl.hi[0]; ++i) {
                const int c[3] = {i, j, k};
                int m[3];
                if (!fr_map(C, c, m) || R->flag[dom_index(C, m)] != FR_BND) continue;
                /* the coarse box holding the cell */
                const int cbx = C->owner[dom_index(C, m)];
                if (cbx < 0) continue;
                for (int n = 0; n < O_NVAR; ++n) {
                    double *dp = &FA(&R->cfp[cbx], m[0], m[1], m[2], n);
                    for (int o1 = 0; o1 < r; ++o1)
                        for (int o0 = 0; o0 < r; ++o0) {
                            int f[3];
                            f[d] = side ? c[d] * r : (c[d] + 1) * r;
                            f[t1] = c[t1] * r + o1;
                            f[t0] = c[t0] * r + o0;
                            const double tmp = side ? dtdx * FV(&flux[d], f, n)
                                                    : -dtdx * FV(&flux[d], f, n);
                            *dp += tmp;
                        }
                }
            }
        }
}

/* average_down of fine level F onto coarse level C (non-EB, Cartesian) */
int ora_level_average_down(ora_level *C, const ora_level *F, int ratio) {
    for (int d = 0; d < 3; ++d)
        if (F->P.domlo[d] != C->P.domlo[d] * ratio ||
            F->P.domhi[d] + 1 != (C->P.domhi[d] + 1) * ratio) return 1;
    const double volfrac = 1.0 / (double)(ratio * ratio * ratio);
    for (int b = 0; b < C->nbox; ++b) {
        const obox B = C->box[b];
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            const int f0[3] = {i * ratio, j * ratio, k * ratio};
            const int fo = F->owner[dom_index(F, f0)];
            if (fo < 0) continue; /* not covered (fine boxes are aligned to the ratio) */
            for (int n = 0; n < O_NVAR; ++n) {
                double c = 0.0;
                for (int kr = 0; kr < ratio; ++kr)
                for (int jr = 0; jr < ratio; ++jr)
                for (int ir = 0; ir < ratio; ++ir) {
                    const int fc[3] = {f0[0] + ir, f0[1] + jr, f0[2] + kr};
                    c += FV(&F->Snew[F->owner[dom_index(F, fc)]], fc, n);
                }
                FA(&C->Snew[b], i, j, k, n) = volfrac * c;
            }
        }
    }
    return 0;
}

/* One step of level l of a hierarchy with subcycling (Amr::timeStep as CAMR
 * drives it): reset the register to the finer level and advance level l
 * (CAMR_advance, CrseAdd / FineAdd), then `ratio` steps of level l+1 --
 * recursively, with their own finer levels -- whose ghost data come from
 * level l interpolated in time (StateData::InterpAddBox: old / new within
 * (t_new - t_old) 1e-3 of either end, else alpha old + beta new) and in
 * space (FillPatch); then level l's post_timestep (CAMR.cpp:679-699):
 * reflux, avgDown, clean_state. */
static int amr_advance(ora_level **L, ora_fluxreg **R, int nlev, int l, double dt) {
    ora_level *C = L[l];
    const int has_fine = l + 1 < nlev;
    if (has_fine) {
        ora_fluxreg *Rg = R[l];
        for (int b = 0; b < C->nbox; ++b) {
            memset(Rg->crse[b].p, 0, sizeof(double) * (size_t)(Rg->crse[b].ns * O_NVAR));
            memset(Rg->cfp[b].p, 0, sizeof(double) * (size_t)(Rg->cfp[b].ns * O_NVAR));
        }
    }
    const double t_old = C->time;
    C->defer_post = has_fine;
    int err = ora_level_advance(C, dt);
    C->defer_post = 0;
    if (err || !has_fine) return err;
    ora_fluxreg *Rg = R[l];
    ora_level *F = L[l + 1];
    const int r = Rg->ratio;
    const double t_new = C->time;
    const double dtf = dt / r;
    const double teps = (t_new - t_old) * 1.e-3;
    for (int s = 0; s < r; ++s) {
        const double t = F->time;
        otimew tw = {2, (t_new - t) / (t_new - t_old), (t - t_old) / (t_new - t_old)};
        if (t > t_old - teps && t < t_old + teps) tw.mode = 1;
        else if (t > t_new - teps && t < t_new + teps) tw.mode = 0;
        if (exterior_ghosts_from_coarse_t(F, C, r, tw)) return 1;
        err = amr_advance(L, R, nlev, l + 1, dtf);
        if (err) return err;
    }
    for (int b = 0; b < C->nbox; ++b) {
        const obox B = C->box[b];
        for (int n = 0; n < O_NVAR; ++n)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            double *d = &FA(&Rg->crse[b], i, j, k, n);
            *d += FA(&Rg->cfp[b], i, j, k, n);
            FA(&C->Snew[b], i, j, k, n) += *d;
        }
    }
    ora_level_average_down(C, F, r);
    clean_state(C, C->Snew, 0);
    return C->err;
}

/* One coarse step of an nlev-level hierarchy; R[l] is the register between
 * levels l and l+1. */
int ora_amr_advance(ora_level **L, ora_fluxreg **R, int nlev, double dt0) {
    for (int l = 0; l + 1 < nlev; ++l)
        if (!R[l] || R[l]->C != L[l] || R[l]->F != L[l + 1]) return 1;
    return amr_advance(L, R, nlev, 0, dt0);
}

int ora_amr_step(ora_level *C, ora_level *F, ora_fluxreg *R, double dtc) {
    ora_level *L[2] = {C, F};
    ora_fluxreg *Rs[1] = {R};
    return ora_amr_advance(L, Rs, 2, dtc);
}

/* ------------------------------------------------------------------------ */
/* Single-point entry points of the known-answer tests (tests/               */
/* test_oracle_godunov.py): the stage functions above, on caller arrays.     */
/* ------------------------------------------------------------------------ */
void ora_transdo_point(int i, int j, int k, int dir, int other_dir, ora_fab *qm,
                       ora_fab *qp, const ora_fab *qnormm, const ora_fab *qnormp,
                       const ora_fab *flxx, const ora_fab *qa, const ora_fab *qint,
                       double cdtdx, int trd, double small_pres) {
    transdo(i, j, k, dir, other_dir, qm, qp, qnormm, qnormp, flxx, qa, qint, cdtdx, trd,
            small_pres, NULL, NULL);
}

void ora_transdd_point(int i, int j, int k, int dir, ora_fab *qm, ora_fab *qp,
                       const ora_fab *qnormm, const ora_fab *qnormp, const ora_fab *flxx,
                       const ora_fab *flxy, const ora_fab *qx, const ora_fab *qy,
                       const ora_fab *qa, const ora_fab *srcq, double hdt, double cdtdx0,
                       double cdtdx1, int trd, double small_pres) {
    transdd(i, j, k, dir, qm, qp, qnormm, qnormp, flxx, flxy, qx, qy, qa, srcq, hdt, cdtdx0,
            cdtdx1, trd, small_pres, NULL, NULL, NULL);
}

void ora_trace_ppm_point(int i, int j, int k, int idir, const ora_fab *q, const ora_fab *qa,
                         ora_fab *qm, ora_fab *qp, double dt, const double dx[3],
                         int use_flattening, double small_dens, double small_pres) {
    obox b;
    b.lo[0] = b.hi[0] = i;
    b.lo[1] = b.hi[1] = j;
    b.lo[2] = b.hi[2] = k;
    trace_ppm(b, idir, q, qa, qm, qp, b, dt, dx, use_flattening, small_dens, small_pres);
}

void ora_artif_visc_point(int i, int j, int k, ora_fab *flx, const ora_fab *divu,
                          const ora_fab *u, double dx, double difmag, int dir, int domlo,
                          int domhi, int bclo, int bchi) {
    artif_visc(i, j, k, flx, divu, u, dx, difmag, dir, domlo, domhi, bclo, bchi);
}

void ora_divu_node_point(int i, int j, int k, const ora_fab *q, ora_fab *divu,
                         const ora_params *P) {
    divu_node(i, j, k, q, P->dx[0], P->dx[1], P->dx[2], divu, P->domlo, P->domhi, P->bclo,
              P->bchi);
}

int ora_enforce_min_density_box(ora_fab *S, const int lo[3], const int hi[3],
                                const ora_params *P) {
    int err = 0;
    enforce_min_density(S, mkbox(lo, hi), P, &err, NULL);
    return err;
}
