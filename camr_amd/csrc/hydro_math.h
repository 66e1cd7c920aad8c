// hydro_math.h -- pointwise FP64 arithmetic of CAMR's Godunov hydro advance,
// written for gfx950 device code (and usable on the host for self-checks).
//
// Every function reproduces the floating-point operation order of the
// CAMR routine it cites, so that a build with -ffp-contract=off gives the
// same bits as the reference CPU build.  std::min / std::max semantics
// (NaN and signed-zero behaviour) are reproduced by vmin / vmax below.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CAMR_HD __host__ __device__ __forceinline__

namespace camr {

// ---- component maps (Source/Utils/IndexDefines.H:17-92, 3-D) -------------
enum { URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UEDEN = 4, UEINT = 5, UTEMP = 6,
       UFS = 7, NVAR = 10 };
enum { QRHO = 0, QU = 1, QV = 2, QW = 3, QGAME = 4, QPRES = 5, QREINT = 6,
       QTEMP = 7, QFS = 8, QVAR = 11 };
enum { QGAMC = 0, QC = 1, QCSML = 2, QDPDR = 3, QDPDE = 4, NQAUX = 5 };
enum { NSPEC = 3 };
enum { BC_INTERIOR = 0, BC_INFLOW = 1, BC_OUTFLOW = 2, BC_SYMMETRY = 3,
       BC_SLIPWALL = 4, BC_NOSLIPWALL = 5 };

// Compact face-state layout used between kernels: primitive variables that
// later stages read (QGAME / QTEMP are never read after tracing).
enum { SR = 0, SU = 1, SV = 2, SW = 3, SP = 4, SE = 5, SY = 6, NS = 9 };
// Compact primitives of a cell: the state layout, then the sound speed and
// the small-sound-speed floor of QAUX (QC, QCSML).
enum { SC = NS, SCS = NS + 1, NQC = NS + 2 };
// Compact flux layout: the NVAR fluxes without UTEMP (which is always 0).
enum { FR = 0, FMX = 1, FMY = 2, FMZ = 3, FE = 4, FEI = 5, FY = 6, NF = 9 };
// Godunov (GD) face quantities kept: normal velocity and pressure.
enum { GUN = 0, GP = 1, NG = 2 };

constexpr double kGamma = 1.4;             // EOS.H:16
constexpr double kRU = 8.31446261815324e7; // PhysicsConstants.H:8
constexpr double kAirMW = 26.93;           // PhysicsConstants.H:11

CAMR_HD double vmax(double a, double b) { return (a < b) ? b : a; }
CAMR_HD double vmin(double a, double b) { return (b < a) ? b : a; }

// Hydro parameters in device-friendly form.
struct HydroConst {
    int ppm_type, plm_iorder, use_pslope, use_flattening, trd;
    int allow_negative_energy, allow_small_energy, update_E_from_e;
    double small, small_dens, small_pres, small_temp, smallu, small_num, difmag;
    double eta1, eta2;
    int bclo[3], bchi[3], domlo[3], domhi[3];
    double dx[3];
    int eb;  // 1: the covered-cell guards of an AMREX_USE_EB build
};

// ---- EOS (Source/EOS/GammaLaw/EOS.H) ---------------------------------------
// REY2T / EY2T (EOS.H:118-131)
CAMR_HD double eos_T_from_e(double e) {
    double poverrho = (kGamma - 1.0) * e;
    return poverrho * kAirMW / kRU;
}
// RTY2E (EOS.H:27-33, 266-281): sum_n Y_n * Cv * T
CAMR_HD double eos_E_from_T(double T, double y0, double y1, double y2) {
    const double Cv = kRU / (kAirMW * (kGamma - 1.0));
    const double ei = Cv * T;
    double E = 0.0;
    E += y0 * ei;
    E += y1 * ei;
    E += y2 * ei;
    return E;
}

// ---- riemannus (Source/Hydro/Hydro_riemann.H:8-173) ------------------------
struct RiemannOut {
    double ustar;
    double f_rho, f_un, f_ut1, f_ut2, f_eden, f_eint;
    double g_un, g_ut1, g_ut2, g_p, g_game;
};

CAMR_HD void riemann_two_shock(double rl_in, double ul, double vl, double v2l,
                               double pl_in, double rel, double rr_in,
                               double ur, double vr, double v2r, double pr_in,
                               double rer, double cav, double gamcl,
                               double gamcr, double small, double smallc,
                               double small_dens, double small_pres,
                               double smallu, RiemannOut &o) {
    const double wsmall = smallc * small_dens;
    const double rl = vmax(rl_in, small_dens);
    const double pl = vmax(pl_in, small_pres);
    const double rr = vmax(rr_in, small_dens);
    const double pr = vmax(pr_in, small_pres);
    const double wl = vmax(wsmall, sqrt(fabs(gamcl * pl * rl)));
    const double wr = vmax(wsmall, sqrt(fabs(gamcr * pr * rr)));
    const double wsum = wl + wr;
    double pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / wsum;
    pstar = vmax(pstar, small_pres);
    double ustar = ((wl * ul + wr * ur) + (pl - pr)) / wsum;
    if (fabs(ustar) < smallu * 0.5 * (fabs(ul) + fabs(ur))) ustar = 0.0;

    double ro, uo, po, reo, gamco;
    if (ustar > 0.) {
        ro = rl; uo = ul; po = pl; reo = rel; gamco = gamcl;
    } else if (ustar < 0.) {
        ro = rr; uo = ur; po = pr; reo = rer; gamco = gamcr;
    } else {
        ro = 0.5 * (rl + rr);
        uo = 0.5 * (ul + ur);
        po = 0.5 * (pl + pr);
        reo = 0.5 * (rel + rer);
        gamco = 0.5 * (gamcl + gamcr);
    }
    ro = vmax(ro, small_dens);
    const double roinv = 1.0 / ro;
    double co = sqrt(fabs(gamco * po * roinv));
    co = vmax(smallc, co);
    const double co2inv = 1.0 / (co * co);
    const double dpo = pstar - po;
    double rstar = ro + dpo * co2inv;
    rstar = vmax(small_dens, rstar);
    const double entho = (reo + po) * roinv * co2inv;
    const double estar = reo + dpo * entho;
    double cstar = sqrt(fabs(gamco * pstar / rstar));
    cstar = vmax(cstar, smallc);

    const double sgnm = (ustar < 0.) ? -1.0 : 1.0;
    double spout = co - sgnm * uo;
    double spin = cstar - sgnm * ustar;
    const double ushock = 0.5 * (spin + spout);
    if (dpo > 0.0) {
        spin = ushock;
        spout = ushock;
    }
    const double sdiff = spout - spin;
    const double scr = (sdiff == 0.0) ? small * cav : sdiff;
    double frac = (1.0 + (spout + spin) / scr) * 0.5;
    frac = vmax(0.0, vmin(1.0, frac));

    double t1 = (ustar > 0.0) ? vl : vr;
    double t2 = (ustar > 0.0) ? v2l : v2r;
    if (ustar == 0.0) {
        t1 = 0.5 * (vl + vr);
        t2 = 0.5 * (v2l + v2r);
    }
    const double omf = 1.0 - frac;
    double rho_gd = frac * rstar + omf * ro;
    double un = frac * ustar + omf * uo;
    double pg = frac * pstar + omf * po;
    double re_gd = frac * estar + omf * reo;
    if (spout < 0.0) {
        rho_gd = ro; un = uo; pg = po; re_gd = reo;
    }
    if (spin >= 0.0) {
        rho_gd = rstar; un = ustar; pg = pstar; re_gd = estar;
    }
    o.g_game = pg / re_gd + 1.0;
    pg = vmax(pg, small_pres);

    o.ustar = ustar;
    o.f_rho = rho_gd * un;
    o.f_un = o.f_rho * un + pg;
    o.f_ut1 = o.f_rho * t1;
    o.f_ut2 = o.f_rho * t2;
    const double rhoetot = re_gd + 0.5 * rho_gd * (un * un + t1 * t1 + t2 * t2);
    o.f_eden = un * (rhoetot + pg);
    o.f_eint = un * re_gd;
    o.g_un = un;
    o.g_ut1 = t1;
    o.g_ut2 = t2;
    o.g_p = pg;
}

// Normal / transverse velocity slots of a compact state for face direction d
// (hydro_cmpflx's IU/IV/IV2 choice, Hydro_cmpflx.H:181-223).
CAMR_HD int vel_n(int d) { return SU + d; }
CAMR_HD int vel_t1(int d) { return (d == 0) ? SV : SU; }
CAMR_HD int vel_t2(int d) { return (d == 2) ? SV : SW; }

// hydro_cmpflx (3-D), Hydro_cmpflx.H:153-310.  ql / qr are the compact
// states on either side of the face; at domain faces the reference
// overwrites ql (lo) or qr (hi) density, pressure and rho*e in place, which
// later stages observe -- the caller writes them back when *mutated != 0.
// Fluxes are returned in compact order (FR..FY), Godunov state in g.
struct FaceAux {
    double c_l, c_r, csml_l, csml_r, gam_l, gam_r;
};

CAMR_HD void face_flux(int d, int idx, const HydroConst &hc, double *ql,
                       double *qr, const FaceAux &a, double *f, double *g,
                       int *mutated) {
    const double cav = 0.5 * (a.c_r + a.c_l);
    const double smallc = vmax(a.csml_r, a.csml_l);
    double gamcl = a.gam_l, gamcr = a.gam_r;
    const int IU = vel_n(d), IV = vel_t1(d), IV2 = vel_t2(d);
    double ul = ql[IU], vl = ql[IV], v2l = ql[IV2];
    double ur = qr[IU], vr = qr[IV], v2r = qr[IV2];
    *mutated = 0;
    if (idx == hc.domlo[d]) {
        const int bc = hc.bclo[d];
        if (bc == BC_SYMMETRY || bc == BC_SLIPWALL || bc == BC_NOSLIPWALL) {
            ql[SR] = qr[SR]; ql[SP] = qr[SP]; ql[SE] = qr[SE];
            gamcl = gamcr; ul = -ur; vl = vr; v2l = v2r;
            *mutated = 1;
        } else if (bc == BC_OUTFLOW) {
            ul = ur; vl = vr; v2l = v2r; gamcl = gamcr;
            ql[SR] = qr[SR]; ql[SP] = qr[SP]; ql[SE] = qr[SE];
            *mutated = 1;
        }
    }
    if (idx == hc.domhi[d] + 1) {
        const int bc = hc.bchi[d];
        if (bc == BC_SYMMETRY || bc == BC_SLIPWALL || bc == BC_NOSLIPWALL) {
            qr[SR] = ql[SR]; qr[SP] = ql[SP]; qr[SE] = ql[SE];
            gamcr = gamcl; ur = -ul; vr = vl; v2r = v2l;
            *mutated |= 2;
        } else if (bc == BC_OUTFLOW) {
            qr[SR] = ql[SR]; qr[SP] = ql[SP]; qr[SE] = ql[SE];
            gamcr = gamcl; ur = ul; vr = vl; v2r = v2l;
            *mutated |= 2;
        }
    }
    RiemannOut o;
    riemann_two_shock(ql[SR], ul, vl, v2l, ql[SP], ql[SE], qr[SR], ur, vr, v2r,
                      qr[SP], qr[SE], cav, gamcl, gamcr, hc.small, smallc,
                      hc.small_dens, hc.small_pres, hc.smallu, o);
    f[FR] = o.f_rho;
    f[FMX + d] = o.f_un;
    f[FMX + (IV - SU)] = o.f_ut1;
    f[FMX + (IV2 - SU)] = o.f_ut2;
    f[FE] = o.f_eden;
    f[FEI] = o.f_eint;
    g[GUN] = o.g_un;
    g[GP] = o.g_p;
    const double flxrho = f[FR];
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) {
        if (o.ustar > 0.0)
            f[FY + s] = flxrho * ql[SY + s];
        else if (o.ustar < 0.0)
            f[FY + s] = flxrho * qr[SY + s];
        else
            f[FY + s] = flxrho * 0.50 * (ql[SY + s] + qr[SY + s]);
    }
}

// ---- method of lines (Source/Hydro/MOL/) ------------------------------------
// limiter, MOL_hydro_K.H:10-20 (plm_theta = 2, van Leer's MC).
CAMR_HD double mol_limiter(double dlft, double drgt) {
    const double plm_theta = 2.0;
    const double dcen = 0.5 * (dlft + drgt);
    const double dsgn = copysign(1.0, dcen);
    const double slop = plm_theta * vmin(fabs(dlft), fabs(drgt));
    const double dlim = (dlft * drgt >= 0.0) ? slop : 0.0;
    return dsgn * vmin(dlim, fabs(dcen));
}

// Characteristic slopes of one cell along direction d, mol_slope_x/y/z
// (MOL_hydro_K.H:22-256), from the compact states of the cell (c) and its
// neighbours below (m) and above (p) and the cell's sound speed.  Output
// (compact): [0] / [1] the acoustic waves (reference dq 0 / 2), [2] / [3]
// the transverse velocities (dq 3 / 4), [4] the rho*e wave (dq QREINT),
// [5..7] the species waves.  The density wave (dq 1) is read by no face
// state and is not formed.
enum { MS_AM = 0, MS_AP = 1, MS_T1 = 2, MS_T2 = 3, MS_E = 4, MS_Y = 5, NMS = 8 };
CAMR_HD void mol_slopes(int d, const double *m, const double *c, const double *p, double cc,
                        int plm_iorder, double small_dens, double *dq) {
    if (plm_iorder == 1) {
        for (int n = 0; n < NMS; ++n) dq[n] = 0.0;
        return;
    }
    const int un = vel_n(d), t1 = vel_t1(d), t2 = vel_t2(d);
    double dlft = 0.5 * (c[SP] - m[SP]) / cc - 0.5 * c[SR] * (c[un] - m[un]);
    double drgt = 0.5 * (p[SP] - c[SP]) / cc - 0.5 * c[SR] * (p[un] - c[un]);
    dq[MS_AM] = mol_limiter(dlft, drgt);
    const double cs2 = cc * cc;
    dlft = 0.5 * (c[SP] - m[SP]) / cc + 0.5 * c[SR] * (c[un] - m[un]);
    drgt = 0.5 * (p[SP] - c[SP]) / cc + 0.5 * c[SR] * (p[un] - c[un]);
    dq[MS_AP] = mol_limiter(dlft, drgt);
    dq[MS_T1] = mol_limiter(c[t1] - m[t1], p[t1] - c[t1]);
    dq[MS_T2] = mol_limiter(c[t2] - m[t2], p[t2] - c[t2]);
    const double enth = (c[SE] + c[SP]) / vmax(c[SR], small_dens);
    dlft = (c[SE] - m[SE]) - enth * (c[SP] - m[SP]) / cs2;
    drgt = (p[SE] - c[SE]) - enth * (p[SP] - c[SP]) / cs2;
    dq[MS_E] = mol_limiter(dlft, drgt);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) {
        const int f = SY + n;
        dlft = c[SR] * c[f] - m[SR] * m[f] - c[f] * (c[SP] - m[SP]) / cs2;
        drgt = p[SR] * p[f] - c[SR] * c[f] - c[f] * (p[SP] - c[SP]) / cs2;
        dq[MS_Y + n] = mol_limiter(dlft, drgt);
    }
}

// Face state on one side of a face of direction d, mol_riemann_x/y/z
// (MOL_riemann_K.H:12-168): side = +1 extrapolates cell c (below the face)
// up, side = -1 cell c (above) down.  ya holds the mass fractions of the
// cell above the face, which both sides take when the reconstructed
// partial densities sum below small_dens.
CAMR_HD void mol_face_state(int d, double side, const double *c, const double *dq, double cc,
                            const double *ya, double small_dens, double small_pres,
                            double *out) {
    const int un = vel_n(d), t1 = vel_t1(d), t2 = vel_t2(d);
    const double enth = (c[SE] + c[SP]) / vmax(c[SR], small_dens);
    const double ac = dq[MS_AM] + dq[MS_AP];
    double pv = (side > 0.0) ? c[SP] + 0.5 * ac * cc : c[SP] - 0.5 * ac * cc;
    out[SP] = vmax(pv, small_pres);
    const double du = (dq[MS_AP] - dq[MS_AM]) / c[SR];
    out[un] = (side > 0.0) ? c[un] + 0.5 * du : c[un] - 0.5 * du;
    out[t1] = (side > 0.0) ? c[t1] + 0.5 * dq[MS_T1] : c[t1] - 0.5 * dq[MS_T1];
    out[t2] = (side > 0.0) ? c[t2] + 0.5 * dq[MS_T2] : c[t2] - 0.5 * dq[MS_T2];
    double sum = 0.;
    double sd[NSPEC];
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) {
        const double h = 0.5 * (dq[MS_Y + n] + c[SY + n] * ac / cc);
        sd[n] = (side > 0.0) ? c[SR] * c[SY + n] + h : c[SR] * c[SY + n] - h;
        sum += sd[n];
    }
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) out[SY + n] = (sum < small_dens) ? ya[n] : sd[n] / sum;
    out[SR] = vmax(sum, small_dens);
    const double he = 0.5 * (dq[MS_E] + enth * ac / cc);
    out[SE] = (side > 0.0) ? c[SE] + he : c[SE] - he;
}

// ---- flatten (Source/Hydro/Godunov/flatten.H:19-99) ------------------------
// p[o], un[o] hold pressure / normal velocity at offsets o-3, o = 0..6.
CAMR_HD double flatten_1d(const double *p, const double *un) {
    const double small_pres = 1.e-200;
    const double shktst = 0.33e0;
    const double zcut1 = 0.75e0;
    const double zcut2 = 0.85e0;
    const double dzcut = 1.0 / (zcut2 - zcut1);
    const int c = 3;
    const double qm = p[c - 1], qp = p[c + 1];
    double dp = qp - qm;
    const int ishft = (dp > 0.0) ? 1 : -1;
    double denom = vmax(small_pres, fabs(p[c + 2] - p[c - 2]));
    double zeta = fabs(dp) / denom;
    const double z = vmin(1.0, vmax(0.0, dzcut * (zeta - zcut1)));
    double tst = (un[c - 1] - un[c + 1] >= 0.0) ? 1.0 : 0.0;
    double tmp = vmin(qp, qm);
    const double chi = (fabs(dp) / tmp) > shktst ? tst : 0.0;

    const int ip = c + 1 - ishft, im = c - 1 - ishft;
    const int ipp = c + 2 - ishft, imm = c - 2 - ishft;
    dp = p[ip] - p[im];
    denom = vmax(small_pres, fabs(p[ipp] - p[imm]));
    zeta = fabs(dp) / denom;
    const double z2 = vmin(1.0, vmax(0.0, dzcut * (zeta - zcut1)));
    tst = (un[im] - un[ip] >= 0.0) ? 1.0 : 0.0;
    tmp = vmin(p[ip], p[im]);
    const double chi2 = (fabs(dp) / tmp) > shktst ? tst : 0.0;
    return 1.0 - vmax(chi2 * z2, chi * z);
}

// ---- PLM (Source/Hydro/Godunov/PLM.H) --------------------------------------
// plm_slope, PLM.H:19-79; s[0..4] = values at offsets -2..2.
CAMR_HD double mc_slope(const double *s, double flat, int order) {
    if (order == 1) return 0.;
    double dlft, drgt, dcen, dlim, dsgn, dfm = 0.0, dfp = 0.0;
    if (order == 4) {
        dlft = s[1] - s[0];
        drgt = s[2] - s[1];
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
        dfm = dsgn * vmin(dlim, fabs(dcen));
        dlft = s[3] - s[2];
        drgt = s[4] - s[3];
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
        dfp = dsgn * vmin(dlim, fabs(dcen));
    }
    dlft = s[2] - s[1];
    drgt = s[3] - s[2];
    dcen = 0.5 * (dlft + drgt);
    dsgn = copysign(1.0, dcen);
    dlim = (dlft * drgt >= 0.0) ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
    const double dtemp = 4.0 / 3.0 * dcen - 1.0 / 6.0 * (dfp + dfm);
    return flat * dsgn * vmin(dlim, fabs(dtemp));
}

// plm_pslope, PLM.H:84-165: pressure slope with the rho*g hydrostatic part
// removed.  p, r, g hold pressure, density and the normal velocity source at
// offsets -2..2.
CAMR_HD double mc_pslope(const double *p, const double *r, const double *g,
                         double dx, double flat, int order) {
    if (order == 1) return 0.;
    double dlft, drgt, dcen, dlim, dsgn, dfm = 0.0, dfp = 0.0;
    if (order == 4) {
        dlft = p[1] - p[0] - 0.25 * dx * (r[1] + r[0]) * (g[1] + g[0]);
        drgt = p[2] - p[1] - 0.25 * dx * (r[1] + r[2]) * (g[1] + g[2]);
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
        dfm = dsgn * vmin(dlim, fabs(dcen));
        dlft = p[3] - p[2] - 0.25 * dx * (r[3] + r[2]) * (g[3] + g[2]);
        drgt = p[4] - p[3] - 0.25 * dx * (r[3] + r[4]) * (g[3] + g[4]);
        dcen = 0.5 * (dlft + drgt);
        dsgn = copysign(1.0, dcen);
        dlim = (dlft * drgt >= 0.0) ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
        dfp = dsgn * vmin(dlim, fabs(dcen));
    }
    dlft = p[2] - p[1] - 0.25 * dx * (r[2] + r[1]) * (g[2] + g[1]);
    drgt = p[3] - p[2] - 0.25 * dx * (r[2] + r[3]) * (g[2] + g[3]);
    dcen = 0.5 * (dlft + drgt);
    dsgn = copysign(1.0, dcen);
    dlim = (dlft * drgt >= 0.0) ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
    const double dtemp = 4.0 / 3.0 * dcen - 1.0 / 6.0 * (dfp + dfm);
    double dp = flat * dsgn * vmin(dlim, fabs(dtemp));
    dp += r[2] * g[2] * dx;
    return dp;
}

// hydro_plm_d, PLM.H:170-323.  c[] = cell-centre primitive values in
// compact layout (SR..SY), sl[] = their slopes along d, cc = sound speed.
// Writes the right state of face i-1/2 (qp) and the left state of face
// i+1/2 (qm).
CAMR_HD void plm_trace(int d, const double *c, const double *sl, double cc,
                       double dtdx, double small_dens, double small_pres,
                       double *qp, double *qm) {
    const int l1 = vel_t1(d), l2 = vel_t2(d);
    const double cs = cc * cc;
    const double rho = c[SR], p = c[SP], rhoe = c[SE];
    const double un = c[SU + d];
    const double enth = ((rhoe + p) / rho) / cs;
    const double drho = sl[SR], dp = sl[SP], drhoe = sl[SE];
    const double dun = sl[SU + d];
    const double alpham = 0.5 * (dp / (rho * cc) - dun) * rho / cc;
    const double alphap = 0.5 * (dp / (rho * cc) + dun) * rho / cc;
    const double alpha0r = drho - dp / cs;
    const double alpha0e = drhoe - dp * enth;
    const double alpha0v = sl[l1];
    const double alpha0w = sl[l2];
    const double wv0 = un - cc, wv1 = un, wv2 = un + cc;
    {
        const double fac = 0.5 * (1.0 + dtdx * vmin(wv0, 0.0));
        const double rho_ref = rho - fac * drho;
        const double vref_n = un - fac * dun;
        const double vref_1 = c[l1] - fac * sl[l1];
        const double vref_2 = c[l2] - fac * sl[l2];
        const double p_ref = p - fac * dp;
        const double rhoe_ref = rhoe - fac * drhoe;
        const double apright = 0.25 * dtdx * (wv0 - wv2) * (1.0 - copysign(1.0, wv2)) * alphap;
        const double amright = 0.0;
        const double z = 0.25 * dtdx * (wv0 - wv1) * (1.0 - copysign(1.0, wv1));
        const double azrright = z * alpha0r;
        const double azeright = z * alpha0e;
        const double azv1rght = z * alpha0v;
        const double azw1rght = z * alpha0w;
        qp[SR] = vmax(rho_ref + apright + amright + azrright, small_dens);
        qp[SU + d] = vref_n + (apright - amright) * cc / rho;
        qp[l1] = vref_1 + azv1rght;
        qp[l2] = vref_2 + azw1rght;
        qp[SP] = vmax(p_ref + (apright + amright) * cs, small_pres);
        qp[SE] = rhoe_ref + (apright + amright) * enth * cs + azeright;
    }
    {
        const double fac = 0.5 * (1.0 - dtdx * vmax(wv2, 0.0));
        const double rho_ref = rho + fac * drho;
        const double vref_n = un + fac * dun;
        const double vref_1 = c[l1] + fac * sl[l1];
        const double vref_2 = c[l2] + fac * sl[l2];
        const double p_ref = p + fac * dp;
        const double rhoe_ref = rhoe + fac * drhoe;
        const double apleft = 0.0;
        const double amleft = 0.25 * dtdx * (wv2 - wv0) * (1.0 + copysign(1.0, wv0)) * alpham;
        const double z = 0.25 * dtdx * (wv2 - wv1) * (1.0 + copysign(1.0, wv1));
        const double azrleft = z * alpha0r;
        const double azeleft = z * alpha0e;
        const double azv1left = z * alpha0v;
        const double azw1left = z * alpha0w;
        qm[SR] = vmax(rho_ref + apleft + amleft + azrleft, small_dens);
        qm[SU + d] = vref_n + (apleft - amleft) * cc / rho;
        qm[l1] = vref_1 + azv1left;
        qm[l2] = vref_2 + azw1left;
        qm[SP] = vmax(p_ref + (apleft + amleft) * cs, small_pres);
        qm[SE] = rhoe_ref + (apleft + amleft) * enth * cs + azeleft;
    }
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) {
        const double spr = (un > 0) ? -1.0 : un * dtdx;
        qp[SY + s] = c[SY + s] + 0.5 * (-1.0 - spr) * sl[SY + s];
        const double spl = un >= 0 ? un * dtdx : 1.0;
        qm[SY + s] = c[SY + s] + 0.5 * (1.0 - spl) * sl[SY + s];
    }
}

// ---- PPM (Source/Hydro/Godunov/PPM.H:25-183) -------------------------------
CAMR_HD double vl_limited(double a, double b, double c) {
    // van Leer slope of the triple (a,b,c) as used in ppm_reconstruct
    const double dsl = 2.0 * (b - a);
    const double dsr = 2.0 * (c - b);
    double r = 0.0;
    if (dsl * dsr > 0.0) {
        const double dsc = 0.5 * (c - a);
        r = copysign(1.0, dsc) * vmin(fabs(dsc), vmin(fabs(dsl), fabs(dsr)));
    }
    return r;
}

// ppm_reconstruct: s[0..4] at offsets -2..2 -> edge values sm, sp.
CAMR_HD void ppm_edges(const double *s, double flatn, double &smo, double &spo) {
    const double dl0 = vl_limited(s[0], s[1], s[2]);
    const double dc = vl_limited(s[1], s[2], s[3]);
    const double dr0 = vl_limited(s[2], s[3], s[4]);
    double sm = 0.5 * (s[2] + s[1]) - (1.0 / 6.0) * (dc - dl0);
    sm = vmax(sm, vmin(s[2], s[1]));
    sm = vmin(sm, vmax(s[2], s[1]));
    double sp = 0.5 * (s[3] + s[2]) - (1.0 / 6.0) * (dr0 - dc);
    sp = vmax(sp, vmin(s[3], s[2]));
    sp = vmin(sp, vmax(s[3], s[2]));
    const double omf = 1.0 - flatn;
    sm = flatn * sm + omf * s[2];
    sp = flatn * sp + omf * s[2];
    if ((sp - s[2]) * (s[2] - sm) <= 0.0) {
        sp = s[2];
        sm = s[2];
    } else if (fabs(sp - s[2]) >= 2.0 * fabs(sm - s[2])) {
        sp = 3.0 * s[2] - 2.0 * sm;
    } else if (fabs(sm - s[2]) >= 2.0 * fabs(sp - s[2])) {
        sm = 3.0 * s[2] - 2.0 * sp;
    }
    smo = sm;
    spo = sp;
}

// One wave of ppm_int_profile (PPM.H:126-183): integral over the domain of
// dependence of speed `speed` from the right (ip) and left (im) edges.
CAMR_HD void ppm_integral(double sm, double sp, double s6, double speed,
                          double dtdx, double &ip, double &im) {
    const double sigma = fabs(speed) * dtdx;
    if (speed <= 0.0) {
        ip = sp;
        im = sm + 0.5 * sigma * (sp - sm + (1.0 - (2.0 / 3.0) * sigma) * s6);
    } else {
        ip = sp - 0.5 * sigma * (sp - sm - (1.0 - (2.0 / 3.0) * sigma) * s6);
        im = sm;
    }
}

CAMR_HD double ppm_s6(double sm, double sp, double sc) {
    return 6.0 * sc - 3.0 * (sm + sp);
}

// Characteristic projection of trace_ppm (PPM.cpp:204-409) for one face.
// side = +1 builds the plus state on face i (waves integrated to the left
// edge, reference = u-c wave), side = -1 the minus state on face i+1.
// I[w] arrays: rho, un, p, rhoe, gam under waves w = 0 (u-c), 1 (u), 2 (u+c).
struct PPMWaves {
    double rho[3], un[3], p[3], rhoe[3], gam[3];
};

CAMR_HD void ppm_project(int side, const PPMWaves &I, double un, double cc,
                         double small_dens, double small_pres,
                         double &rho_o, double &un_o, double &p_o,
                         double &rhoe_o) {
    const int w = (side > 0) ? 0 : 2;
    double rho_ref = I.rho[w];
    const double un_ref = I.un[w];
    double p_ref = I.p[w];
    const double rhoe_g_ref = I.rhoe[w];
    const double gam_ref = I.gam[w];
    rho_ref = vmax(rho_ref, small_dens);
    const double rho_ref_inv = 1.0 / rho_ref;
    p_ref = vmax(p_ref, small_pres);
    const double cc_ref = sqrt(gam_ref * p_ref / rho_ref);
    const double csq_ref = cc_ref * cc_ref;
    const double cc_ref_inv = 1.0 / cc_ref;
    const double h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv / csq_ref;
    const double dum = un_ref - I.un[0];
    const double dptotm = p_ref - I.p[0];
    const double drho = rho_ref - I.rho[1];
    const double dptot = p_ref - I.p[1];
    const double drhoe_g = rhoe_g_ref - I.rhoe[1];
    const double dup = un_ref - I.un[2];
    const double dptotp = p_ref - I.p[2];
    double alpham = 0.5 * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
    double alphap = 0.5 * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
    double alpha0r = drho - dptot / csq_ref;
    double alpha0e_g = drhoe_g - dptot * h_g_ref;
    const double sm_ = un - cc, sp_ = un + cc;
    if (side > 0) {
        if (sm_ > 0.) alpham = 0.; else if (sm_ < 0.) alpham *= -1.0; else alpham *= -0.5;
        if (sp_ > 0.) alphap = 0.; else if (sp_ < 0.) alphap *= -1.0; else alphap *= -0.5;
        if (un > 0.) { alpha0r = 0.; alpha0e_g = 0.; }
        else if (un < 0.) { alpha0r *= -1.0; alpha0e_g *= -1.0; }
        else { alpha0r *= -0.5; alpha0e_g *= -0.5; }
    } else {
        if (sm_ > 0.) alpham *= -1.0; else if (sm_ < 0.) alpham = 0.; else alpham *= -0.5;
        if (sp_ > 0.) alphap *= -1.0; else if (sp_ < 0.) alphap = 0.; else alphap *= -0.5;
        if (un > 0.) { alpha0r *= -1.0; alpha0e_g *= -1.0; }
        else if (un < 0.) { alpha0r = 0.; alpha0e_g = 0.; }
        else { alpha0r *= -0.5; alpha0e_g *= -0.5; }
    }
    const double apm = alphap + alpham;
    rho_o = vmax(rho_ref + alphap + alpham + alpha0r, small_dens);
    un_o = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
    // PPM.cpp:295 clamps the plus state with small_pres, PPM.cpp:398 the
    // minus state with small_dens.
    p_o = vmax(p_ref + apm * csq_ref, (side > 0) ? small_pres : small_dens);
    rhoe_o = rhoe_g_ref + apm * h_g_ref * csq_ref + alpha0e_g;
}

// ---- transverse corrections (Godunov_utils_3D.H) ---------------------------
// Conservative flux differences across a cell in the transverse direction(s)
// (already multiplied by the dt/dx factors) and the p*u work terms: the
// only inputs the transverse predictors need from the neighbouring fluxes.
struct TransDiff {
    double rho, mu, mv, mw, e, y[NSPEC];
    double dAup, pav, dAu;  // transdo: p u difference, mean p, u difference
    double pnew;            // transdd: summed pressure correction
};

// Differences of one transverse direction for hydro_transdo
// (Godunov_utils_3D.H:74-85, 112-122): fl/fr are the compact fluxes on the
// faces iv and iv+e_t, gl/gr their Godunov (normal velocity, pressure).
CAMR_HD void trans_diff_one(const double *fl, const double *fr, const double *gl,
                            const double *gr, double cdtdx, TransDiff &t) {
    t.rho = cdtdx * (fr[FR] - fl[FR]);
    t.mu = cdtdx * (fr[FMX] - fl[FMX]);
    t.mv = cdtdx * (fr[FMY] - fl[FMY]);
    t.mw = cdtdx * (fr[FMZ] - fl[FMZ]);
    t.e = cdtdx * (fr[FE] - fl[FE]);
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) t.y[s] = cdtdx * (fr[FY + s] - fl[FY + s]);
    const double pggp = gr[GP], pggm = gl[GP];
    const double ugp = gr[GUN], ugm = gl[GUN];
    t.dAup = pggp * ugp - pggm * ugm;
    t.pav = 0.5 * (pggp + pggm);
    t.dAu = ugp - ugm;
}

// hydro_transdo (Godunov_utils_3D.H:22-239), non-EB, for one side: qn is
// the normal-predictor state (qnormp(iv) for the qp update, qnormm(iv+e_d)
// for the qm update); both sides use the differences of cell iv.
CAMR_HD void transverse_one_d(const double *qn, const TransDiff &t, double cdtdx,
                              double gamc, int trd, double small_pres,
                              double *qo) {
    const double flxrho = t.rho, flxu = t.mu, flxv = t.mv, flxw = t.mw, flxe = t.e;
    const double rr = qn[SR];
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) {
        const double rrnew = rr - flxrho;
        const double compo = rr * qn[SY + s] - t.y[s];
        qo[SY + s] = compo / rrnew;
    }
    const double dAup = t.dAup, pav = t.pav, dAu = t.dAu;

    double ru = qn[SU], rv = qn[SV], rw = qn[SW];
    const double ekin = 0.5 * rr * (ru * ru + rv * rv + rw * rw);
    ru *= rr;
    rv *= rr;
    rw *= rr;
    const double re = qn[SE] + ekin;
    double rrnew = rr - flxrho;
    double runew = ru - flxu;
    double rvnew = rv - flxv;
    double rwnew = rw - flxw;
    double renew = re - flxe;
    bool reset = false;
    if (trd == 1 && rrnew < 0.) {
        rrnew = rr; runew = ru; rvnew = rv; rwnew = rw; renew = re;
        reset = true;
    }
    qo[SR] = rrnew;
    qo[SU] = runew / rrnew;
    qo[SV] = rvnew / rrnew;
    qo[SW] = rwnew / rrnew;
    const double rhoekin = 0.5 * (runew * runew + rvnew * rvnew + rwnew * rwnew) / rrnew;
    qo[SE] = renew - rhoekin;
    if (!reset) {
        const double pnew = qn[SP] - cdtdx * (dAup + pav * dAu * (gamc - 1.0));
        qo[SP] = vmax(pnew, small_pres);
    } else {
        qo[SP] = qn[SP];
    }
}

CAMR_HD void transverse_one(const double *qn, const double *fl, const double *fr,
                            const double *gl, const double *gr, double cdtdx,
                            double gamc, int trd, double small_pres,
                            double *qo) {
    TransDiff t;
    trans_diff_one(fl, fr, gl, gr, cdtdx, t);
    transverse_one_d(qn, t, cdtdx, gamc, trd, small_pres, qo);
}

// Differences of the two transverse directions a (fluxes fal/far_, GD
// gal/gar, factor ca = hdt/dx_a) and b for hydro_transdd
// (Godunov_utils_3D.H:307-316, 362-384).
CAMR_HD void trans_diff_two(const double *fal, const double *far_, const double *gal,
                            const double *gar, const double *fbl, const double *fbr,
                            const double *gbl, const double *gbr, double ca,
                            double cb, double gamc, TransDiff &t) {
#define CAMR_TF(c) (ca * (far_[c] - fal[c]) + cb * (fbr[c] - fbl[c]))
    t.rho = CAMR_TF(FR);
    t.mu = CAMR_TF(FMX);
    t.mv = CAMR_TF(FMY);
    t.mw = CAMR_TF(FMZ);
    t.e = CAMR_TF(FE);
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) t.y[s] = CAMR_TF(FY + s);
#undef CAMR_TF
    const double dAupx = gar[GP] * gar[GUN] - gal[GP] * gal[GUN];
    const double pavx = 0.5 * (gar[GP] + gal[GP]);
    const double dAux = gar[GUN] - gal[GUN];
    const double dAupy = gbr[GP] * gbr[GUN] - gbl[GP] * gbl[GUN];
    const double pavy = 0.5 * (gbr[GP] + gbl[GP]);
    const double dAuy = gbr[GUN] - gbl[GUN];
    const double pxnew = ca * (dAupx + pavx * dAux * (gamc - 1.0));
    const double pynew = cb * (dAupy + pavy * dAuy * (gamc - 1.0));
    // kept separate so that qn[SP] - pxnew - pynew keeps the reference order
    t.dAup = pxnew;
    t.pnew = pynew;
}

// hydro_transdd (Godunov_utils_3D.H:245-482), non-EB, for one side, with
// hdt * source (src in compact layout; all +0 when the sources vanish).
CAMR_HD void transverse_two_d(const double *qn, const TransDiff &t, double hdt,
                              const double *src, int trd, double small_pres,
                              double *qo) {
    const double flxrho = t.rho, flxu = t.mu, flxv = t.mv, flxw = t.mw, flxe = t.e;
    const double s_r = src[SR], s_u = src[SU], s_v = src[SV], s_w = src[SW];
    const double s_e = src[SE], s_p = src[SP];
    const double rr = qn[SR];
    double rrnew = rr - flxrho;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) {
        const double compo = rr * qn[SY + s] - t.y[s];
        qo[SY + s] = compo / rrnew + hdt * src[SY + s];
    }
    const double pxnew = t.dAup, pynew = t.pnew;

    const double ru = rr * qn[SU];
    const double rv = rr * qn[SV];
    const double rw = rr * qn[SW];
    const double ekin = 0.5 * (ru * ru + rv * rv + rw * rw) / rr;
    const double re = qn[SE] + ekin;
    double runew = ru - flxu;
    double rvnew = rv - flxv;
    double rwnew = rw - flxw;
    double renew = re - flxe;
    bool reset = false;
    if (trd == 1 && rrnew < 0.) {
        rrnew = rr; runew = ru; rvnew = rv; rwnew = rw; renew = re;
        reset = true;
    }
    qo[SR] = rrnew + hdt * s_r;
    qo[SU] = runew / rrnew + hdt * s_u;
    qo[SV] = rvnew / rrnew + hdt * s_v;
    qo[SW] = rwnew / rrnew + hdt * s_w;
    const double rhoekin = 0.5 * (runew * runew + rvnew * rvnew + rwnew * rwnew) / rrnew;
    qo[SE] = renew - rhoekin + hdt * s_e;
    double pn;
    if (!reset)
        pn = qn[SP] - pxnew - pynew + hdt * s_p;
    else
        pn = qn[SP] + hdt * s_p;
    qo[SP] = vmax(pn, small_pres);
}

CAMR_HD void transverse_two(const double *qn, const double *fal, const double *far_,
                            const double *gal, const double *gar,
                            const double *fbl, const double *fbr,
                            const double *gbl, const double *gbr, double ca,
                            double cb, double hdt, const double *src,
                            double gamc, int trd, double small_pres,
                            double *qo) {
    TransDiff t;
    trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, gamc, t);
    double z[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) z[s] = src ? src[s] : 0.0;
    transverse_two_d(qn, t, hdt, z, trd, small_pres, qo);
}

} // namespace camr
