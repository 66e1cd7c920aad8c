// adjoint_math.h -- reverse-mode derivatives of the per-point hydro
// arithmetic of hydro_math.h / state_math.h (host and device).
//
// Convention: f_adj(inputs..., output adjoints..., input adjoints...)
// recomputes the forward function internally and ADDS J^T * ybar to the
// input adjoints.  Piecewise branches (limiters, clamps, upwinding) take
// the derivative of the branch the forward evaluation takes; min/max
// follow vmin / vmax (ties go to the same operand the forward returns).
//
// Scope: the variables that carry the hydrodynamics -- rho, velocity,
// pressure, rho e (rho, rho u, rho E, rho e in conserved form).  The
// species and the temperature never enter those (species are advected
// passively, T is recomputed from e), so their adjoints are left out: for
// a loss on the hydro variables their gradients are exactly zero.
#pragma once

#include "hydro_math.h"

namespace camr {

// ---------------------------------------------------------------- helpers
CAMR_HD double sgn_of(double x) { return (x < 0.0) ? -1.0 : 1.0; }

// ybar of y = vmax(a, b) to abar / bbar.  Both accumulators are updated
// (one with +0): a branch between the two would become a select between
// their addresses and keep the callers' locals out of registers.
CAMR_HD void vmax_adj(double a, double b, double yb, double &ab, double &bb) {
    const bool to_b = a < b;
    ab += to_b ? 0.0 : yb;
    bb += to_b ? yb : 0.0;
}
// ybar of y = vmin(a, b)
CAMR_HD void vmin_adj(double a, double b, double yb, double &ab, double &bb) {
    const bool to_b = b < a;
    ab += to_b ? 0.0 : yb;
    bb += to_b ? yb : 0.0;
}

// ---------------------------------------------------------------- Riemann
// Adjoints of the riemann_two_shock outputs that later stages read.
struct RiemannBar {
    double f_rho, f_un, f_ut1, f_ut2, f_eden, f_eint, g_un, g_p;
};
// Adjoints of its varying inputs (the gammas are the EOS constant).
struct RiemannInBar {
    double rl, ul, vl, v2l, pl, rel, rr, ur, vr, v2r, pr, rer, cav, smallc;
};

CAMR_HD void riemann_two_shock_adj(double rl_in, double ul, double vl, double v2l,
                                   double pl_in, double rel, double rr_in, double ur,
                                   double vr, double v2r, double pr_in, double rer,
                                   double cav, double gamcl, double gamcr, double small,
                                   double smallc, double small_dens, double small_pres,
                                   double smallu, const RiemannBar &ob, RiemannInBar &ib) {
    // ---- forward, keeping the intermediates
    const double wsmall = smallc * small_dens;
    const double rl = vmax(rl_in, small_dens);
    const double pl = vmax(pl_in, small_pres);
    const double rr = vmax(rr_in, small_dens);
    const double pr = vmax(pr_in, small_pres);
    const double zl = gamcl * pl * rl, zr = gamcr * pr * rr;
    const double sl = sqrt(fabs(zl)), sr = sqrt(fabs(zr));
    const double wl = vmax(wsmall, sl);
    const double wr = vmax(wsmall, sr);
    const double wsum = wl + wr;
    const double pstar0 = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / wsum;
    const double pstar = vmax(pstar0, small_pres);
    const double ustar0 = ((wl * ul + wr * ur) + (pl - pr)) / wsum;
    const bool uzero = fabs(ustar0) < smallu * 0.5 * (fabs(ul) + fabs(ur));
    const double ustar = uzero ? 0.0 : ustar0;
    const bool upos = ustar > 0., uneg = ustar < 0.;
    const double ro = upos ? rl : (uneg ? rr : 0.5 * (rl + rr));
    const double uo = upos ? ul : (uneg ? ur : 0.5 * (ul + ur));
    const double po = upos ? pl : (uneg ? pr : 0.5 * (pl + pr));
    const double reo = upos ? rel : (uneg ? rer : 0.5 * (rel + rer));
    const double gamco = upos ? gamcl : (uneg ? gamcr : 0.5 * (gamcl + gamcr));
    const double ro2 = vmax(ro, small_dens);
    const double roinv = 1.0 / ro2;
    const double yco = gamco * po * roinv;
    const double co0 = sqrt(fabs(yco));
    const double co = vmax(smallc, co0);
    const double co2inv = 1.0 / (co * co);
    const double dpo = pstar - po;
    const double rstar0 = ro2 + dpo * co2inv;
    const double rstar = vmax(small_dens, rstar0);
    const double tho = (reo + po);
    const double entho = tho * roinv * co2inv;
    const double estar = reo + dpo * entho;
    const double xcs = gamco * pstar / rstar;
    const double cstar0 = sqrt(fabs(xcs));
    const double cstar = vmax(cstar0, smallc);
    const double sgnm = (ustar < 0.) ? -1.0 : 1.0;
    const double spout0 = co - sgnm * uo;
    const double spin0 = cstar - sgnm * ustar;
    const double ushock = 0.5 * (spin0 + spout0);
    const bool shock = dpo > 0.0;
    const double spin = shock ? ushock : spin0;
    const double spout = shock ? ushock : spout0;
    const double sdiff = spout - spin;
    const double scr = (sdiff == 0.0) ? small * cav : sdiff;
    const double frac0 = (1.0 + (spout + spin) / scr) * 0.5;
    const double fmin = vmin(1.0, frac0);
    const double frac = vmax(0.0, fmin);
    double t1 = (ustar > 0.0) ? vl : vr;
    double t2 = (ustar > 0.0) ? v2l : v2r;
    if (ustar == 0.0) {
        t1 = 0.5 * (vl + vr);
        t2 = 0.5 * (v2l + v2r);
    }
    const double omf = 1.0 - frac;
    int sel;  // 0 mixed, 1 outer (spout < 0), 2 star (spin >= 0)
    double rho_gd, un, pg0, re_gd;
    if (spin >= 0.0) {
        sel = 2; rho_gd = rstar; un = ustar; pg0 = pstar; re_gd = estar;
    } else if (spout < 0.0) {
        sel = 1; rho_gd = ro2; un = uo; pg0 = po; re_gd = reo;
    } else {
        sel = 0;
        rho_gd = frac * rstar + omf * ro2;
        un = frac * ustar + omf * uo;
        pg0 = frac * pstar + omf * po;
        re_gd = frac * estar + omf * reo;
    }
    const double pg = vmax(pg0, small_pres);
    const double f_rho = rho_gd * un;
    const double rhoetot = re_gd + 0.5 * rho_gd * (un * un + t1 * t1 + t2 * t2);

    // ---- reverse
    double b_un = 0.0, b_pg = 0.0, b_re_gd = 0.0, b_rho_gd = 0.0, b_t1 = 0.0, b_t2 = 0.0;
    // f_eden = un * (rhoetot + pg); f_eint = un * re_gd; g_un = un; g_p = pg
    const double b_rhoetot = ob.f_eden * un;
    b_un += ob.f_eden * (rhoetot + pg) + ob.f_eint * re_gd + ob.g_un;
    b_pg += ob.f_eden * un + ob.g_p;
    b_re_gd += ob.f_eint * un;
    // rhoetot = re_gd + 0.5 * rho_gd * (un^2 + t1^2 + t2^2)
    b_re_gd += b_rhoetot;
    b_rho_gd += b_rhoetot * 0.5 * (un * un + t1 * t1 + t2 * t2);
    b_un += b_rhoetot * rho_gd * un;
    b_t1 += b_rhoetot * rho_gd * t1;
    b_t2 += b_rhoetot * rho_gd * t2;
    // f_un = f_rho * un + pg; f_ut1 = f_rho * t1; f_ut2 = f_rho * t2
    const double b_frho = ob.f_rho + ob.f_un * un + ob.f_ut1 * t1 + ob.f_ut2 * t2;
    b_un += ob.f_un * f_rho;
    b_pg += ob.f_un;
    b_t1 += ob.f_ut1 * f_rho;
    b_t2 += ob.f_ut2 * f_rho;
    // f_rho = rho_gd * un
    b_rho_gd += b_frho * un;
    b_un += b_frho * rho_gd;
    // pg = vmax(pg0, small_pres)
    const double b_pg0 = (pg0 < small_pres) ? 0.0 : b_pg;
    double b_rstar = 0.0, b_ustar = 0.0, b_pstar = 0.0, b_estar = 0.0;
    double b_ro2 = 0.0, b_uo = 0.0, b_po = 0.0, b_reo = 0.0, b_frac = 0.0;
    if (sel == 2) {
        b_rstar += b_rho_gd; b_ustar += b_un; b_pstar += b_pg0; b_estar += b_re_gd;
    } else if (sel == 1) {
        b_ro2 += b_rho_gd; b_uo += b_un; b_po += b_pg0; b_reo += b_re_gd;
    } else {
        b_rstar += b_rho_gd * frac; b_ro2 += b_rho_gd * omf; b_frac += b_rho_gd * (rstar - ro2);
        b_ustar += b_un * frac; b_uo += b_un * omf; b_frac += b_un * (ustar - uo);
        b_pstar += b_pg0 * frac; b_po += b_pg0 * omf; b_frac += b_pg0 * (pstar - po);
        b_estar += b_re_gd * frac; b_reo += b_re_gd * omf; b_frac += b_re_gd * (estar - reo);
    }
    // upwind weights of the left / right states (1, 0 or 1/2 each: exact
    // factors); weighted updates instead of branches keep ib in registers
    const double wgl = upos ? 1.0 : (uneg ? 0.0 : 0.5);
    const double wgr = uneg ? 1.0 : (upos ? 0.0 : 0.5);
    // transverse velocities
    ib.vl += wgl * b_t1; ib.v2l += wgl * b_t2;
    ib.vr += wgr * b_t1; ib.v2r += wgr * b_t2;
    // frac = vmax(0, vmin(1, frac0))
    double b_fmin = 0.0, b_dummy = 0.0, b_frac0 = 0.0;
    vmax_adj(0.0, fmin, b_frac, b_dummy, b_fmin);
    vmin_adj(1.0, frac0, b_fmin, b_dummy, b_frac0);
    // frac0 = (1 + (spout + spin) / scr) * 0.5
    const double b_sum = b_frac0 * 0.5 / scr;
    const double b_scr = -b_frac0 * 0.5 * (spout + spin) / (scr * scr);
    double b_spout = b_sum, b_spin = b_sum;
    if (sdiff == 0.0) {
        ib.cav += b_scr * small;
    } else {
        b_spout += b_scr;
        b_spin -= b_scr;
    }
    double b_spout0, b_spin0;
    if (shock) {
        const double b_ushock = b_spin + b_spout;
        b_spin0 = 0.5 * b_ushock;
        b_spout0 = 0.5 * b_ushock;
    } else {
        b_spin0 = b_spin;
        b_spout0 = b_spout;
    }
    // spout0 = co - sgnm * uo ; spin0 = cstar - sgnm * ustar
    double b_co = b_spout0;
    b_uo -= sgnm * b_spout0;
    double b_cstar = b_spin0;
    b_ustar -= sgnm * b_spin0;
    // cstar = vmax(cstar0, smallc)
    double b_cstar0 = 0.0;
    vmax_adj(cstar0, smallc, b_cstar, b_cstar0, ib.smallc);
    if (cstar0 > 0.0) {
        const double b_xcs = b_cstar0 * 0.5 / cstar0 * sgn_of(xcs);
        b_pstar += b_xcs * gamco / rstar;
        b_rstar -= b_xcs * xcs / rstar;
    }
    // estar = reo + dpo * entho
    b_reo += b_estar;
    double b_dpo = b_estar * entho;
    const double b_entho = b_estar * dpo;
    // entho = tho * roinv * co2inv
    const double b_tho = b_entho * roinv * co2inv;
    double b_roinv = b_entho * tho * co2inv;
    double b_co2inv = b_entho * tho * roinv;
    b_reo += b_tho;
    b_po += b_tho;
    // rstar = vmax(small_dens, rstar0); rstar0 = ro2 + dpo * co2inv
    double b_rstar0 = 0.0;
    vmax_adj(small_dens, rstar0, b_rstar, b_dummy, b_rstar0);
    b_ro2 += b_rstar0;
    b_dpo += b_rstar0 * co2inv;
    b_co2inv += b_rstar0 * dpo;
    // dpo = pstar - po
    b_pstar += b_dpo;
    b_po -= b_dpo;
    // co2inv = 1 / (co * co)
    b_co += -2.0 * b_co2inv * co2inv / co;
    // co = vmax(smallc, co0)
    double b_co0 = 0.0;
    vmax_adj(smallc, co0, b_co, ib.smallc, b_co0);
    if (co0 > 0.0) {
        const double b_yco = b_co0 * 0.5 / co0 * sgn_of(yco);
        b_po += b_yco * gamco * roinv;
        b_roinv += b_yco * gamco * po;
    }
    // roinv = 1 / ro2
    b_ro2 -= b_roinv * roinv * roinv;
    // ro2 = vmax(ro, small_dens)
    double b_ro = 0.0;
    vmax_adj(ro, small_dens, b_ro2, b_ro, b_dummy);
    // outer state choice
    double b_rl = wgl * b_ro, b_pl = wgl * b_po, b_rr = wgr * b_ro, b_pr = wgr * b_po;
    ib.ul += wgl * b_uo; ib.rel += wgl * b_reo;
    ib.ur += wgr * b_uo; ib.rer += wgr * b_reo;
    // ustar (zeroed in the dead band)
    const double b_ustar0 = uzero ? 0.0 : b_ustar;
    double b_wl = 0.0, b_wr = 0.0, b_wsum = 0.0;
    {
        const double b_n = b_ustar0 / wsum;
        b_wsum -= b_ustar0 * ustar0 / wsum;
        b_wl += b_n * ul; ib.ul += b_n * wl;
        b_wr += b_n * ur; ib.ur += b_n * wr;
        b_pl += b_n; b_pr -= b_n;
    }
    // pstar = vmax(pstar0, small_pres)
    double b_pstar0 = 0.0;
    vmax_adj(pstar0, small_pres, b_pstar, b_pstar0, b_dummy);
    {
        const double b_m = b_pstar0 / wsum;
        b_wsum -= b_pstar0 * pstar0 / wsum;
        b_wr += b_m * pl; b_pl += b_m * wr;
        b_wl += b_m * pr; b_pr += b_m * wl;
        const double du = ul - ur;
        b_wl += b_m * wr * du;
        b_wr += b_m * wl * du;
        ib.ul += b_m * wl * wr;
        ib.ur -= b_m * wl * wr;
    }
    b_wl += b_wsum;
    b_wr += b_wsum;
    // wl = vmax(wsmall, sl); sl = sqrt(|gamcl pl rl|)
    double b_sl = 0.0, b_sr = 0.0, b_wsmall = 0.0;
    vmax_adj(wsmall, sl, b_wl, b_wsmall, b_sl);
    vmax_adj(wsmall, sr, b_wr, b_wsmall, b_sr);
    if (sl > 0.0) {
        const double b_z = b_sl * 0.5 / sl * sgn_of(zl);
        b_pl += b_z * gamcl * rl;
        b_rl += b_z * gamcl * pl;
    }
    if (sr > 0.0) {
        const double b_z = b_sr * 0.5 / sr * sgn_of(zr);
        b_pr += b_z * gamcr * rr;
        b_rr += b_z * gamcr * pr;
    }
    ib.smallc += b_wsmall * small_dens;
    // input clamps
    vmax_adj(rl_in, small_dens, b_rl, ib.rl, b_dummy);
    vmax_adj(pl_in, small_pres, b_pl, ib.pl, b_dummy);
    vmax_adj(rr_in, small_dens, b_rr, ib.rr, b_dummy);
    vmax_adj(pr_in, small_pres, b_pr, ib.pr, b_dummy);
}

// ---------------------------------------------------------------- face_flux
// Adjoint of face_flux for an interior face (no domain-boundary override):
// fb / gb are the adjoints of the compact fluxes and Godunov (u, p);
// qlb / qrb (compact states) and the sound-speed adjoints are accumulated.
struct FaceAuxBar {
    double c_l, c_r, csml_l, csml_r;
};

CAMR_HD void face_flux_adj(int d, const HydroConst &hc, const double *ql, const double *qr,
                           const FaceAux &a, const double *fb, const double *gb, double *qlb,
                           double *qrb, FaceAuxBar &ab) {
    const double cav = 0.5 * (a.c_r + a.c_l);
    const double smallc = vmax(a.csml_r, a.csml_l);
    const int IU = vel_n(d), IV = vel_t1(d), IV2 = vel_t2(d);
    RiemannBar ob;
    ob.f_rho = fb[FR];
    ob.f_un = fb[FMX + d];
    ob.f_ut1 = fb[FMX + (IV - SU)];
    ob.f_ut2 = fb[FMX + (IV2 - SU)];
    ob.f_eden = fb[FE];
    ob.f_eint = fb[FEI];
    ob.g_un = gb[GUN];
    ob.g_p = gb[GP];
    RiemannInBar ib{};
    riemann_two_shock_adj(ql[SR], ql[IU], ql[IV], ql[IV2], ql[SP], ql[SE], qr[SR], qr[IU],
                          qr[IV], qr[IV2], qr[SP], qr[SE], cav, a.gam_l, a.gam_r, hc.small,
                          smallc, hc.small_dens, hc.small_pres, hc.smallu, ob, ib);
    qlb[SR] += ib.rl; qlb[IU] += ib.ul; qlb[IV] += ib.vl; qlb[IV2] += ib.v2l;
    qlb[SP] += ib.pl; qlb[SE] += ib.rel;
    qrb[SR] += ib.rr; qrb[IU] += ib.ur; qrb[IV] += ib.vr; qrb[IV2] += ib.v2r;
    qrb[SP] += ib.pr; qrb[SE] += ib.rer;
    ab.c_r += 0.5 * ib.cav;
    ab.c_l += 0.5 * ib.cav;
    vmax_adj(a.csml_r, a.csml_l, ib.smallc, ab.csml_r, ab.csml_l);
}

// ------------------------------------------------------- transverse terms
// Adjoint of the differences of trans_diff_one / trans_diff_two (TransDiff
// fields: rho mu mv mw e dAup pav dAu pnew; species omitted).
struct TransDiffBar {
    double rho, mu, mv, mw, e, dAup, pav, dAu, pnew;
};

CAMR_HD void trans_diff_one_adj(const double *fl, const double *fr, const double *gl,
                                const double *gr, double cdtdx, const TransDiffBar &tb,
                                double *flb, double *frb, double *glb, double *grb) {
    frb[FR] += cdtdx * tb.rho;  flb[FR] -= cdtdx * tb.rho;
    frb[FMX] += cdtdx * tb.mu;  flb[FMX] -= cdtdx * tb.mu;
    frb[FMY] += cdtdx * tb.mv;  flb[FMY] -= cdtdx * tb.mv;
    frb[FMZ] += cdtdx * tb.mw;  flb[FMZ] -= cdtdx * tb.mw;
    frb[FE] += cdtdx * tb.e;    flb[FE] -= cdtdx * tb.e;
    const double pggp = gr[GP], pggm = gl[GP], ugp = gr[GUN], ugm = gl[GUN];
    // dAup = pggp*ugp - pggm*ugm ; pav = 0.5 (pggp + pggm) ; dAu = ugp - ugm
    grb[GP] += tb.dAup * ugp + 0.5 * tb.pav;
    grb[GUN] += tb.dAup * pggp + tb.dAu;
    glb[GP] += -tb.dAup * ugm + 0.5 * tb.pav;
    glb[GUN] += -tb.dAup * pggm - tb.dAu;
    (void)fl; (void)fr;
}

// transverse_one_d (hydro_transdo, one side); qob: adjoint of the output
// state, qnb / tb accumulate.
CAMR_HD void transverse_one_d_adj(const double *qn, const TransDiff &t, double cdtdx,
                                  double gamc, int trd, double small_pres, const double *qob,
                                  double *qnb, TransDiffBar &tb) {
    const double flxrho = t.rho, flxu = t.mu, flxv = t.mv, flxw = t.mw, flxe = t.e;
    const double rr = qn[SR];
    const double u = qn[SU], v = qn[SV], w = qn[SW];
    const double ekin = 0.5 * rr * (u * u + v * v + w * w);
    const double ru = u * rr, rv = v * rr, rw = w * rr;
    const double re = qn[SE] + ekin;
    double rrnew = rr - flxrho, runew = ru - flxu, rvnew = rv - flxv, rwnew = rw - flxw;
    double renew = re - flxe;
    bool reset = false;
    if (trd == 1 && rrnew < 0.) {
        rrnew = rr; runew = ru; rvnew = rv; rwnew = rw; renew = re;
        reset = true;
    }
    const double K = runew * runew + rvnew * rvnew + rwnew * rwnew;
    const double rhoekin = 0.5 * K / rrnew;
    // pressure
    if (!reset) {
        const double x = t.dAup + t.pav * t.dAu * (gamc - 1.0);
        const double pnew = qn[SP] - cdtdx * x;
        const double b_pnew = (pnew < small_pres) ? 0.0 : qob[SP];
        qnb[SP] += b_pnew;
        const double b_x = -cdtdx * b_pnew;
        tb.dAup += b_x;
        tb.pav += b_x * t.dAu * (gamc - 1.0);
        tb.dAu += b_x * t.pav * (gamc - 1.0);
    } else {
        qnb[SP] += qob[SP];
    }
    // rho e = renew - rhoekin
    const double b_renew = qob[SE];
    const double b_rhoekin = -qob[SE];
    double b_rrnew = qob[SR];
    const double b_K = b_rhoekin * 0.5 / rrnew;
    b_rrnew -= b_rhoekin * rhoekin / rrnew;
    double b_runew = b_K * 2.0 * runew, b_rvnew = b_K * 2.0 * rvnew, b_rwnew = b_K * 2.0 * rwnew;
    // velocities = r*new / rrnew
    b_runew += qob[SU] / rrnew;
    b_rvnew += qob[SV] / rrnew;
    b_rwnew += qob[SW] / rrnew;
    b_rrnew -= (qob[SU] * (runew / rrnew) + qob[SV] * (rvnew / rrnew) + qob[SW] * (rwnew / rrnew)) / rrnew;
    double b_rr = b_rrnew, b_ru = b_runew, b_rv = b_rvnew, b_rw = b_rwnew, b_re = b_renew;
    if (!reset) {
        tb.rho -= b_rrnew; tb.mu -= b_runew; tb.mv -= b_rvnew; tb.mw -= b_rwnew; tb.e -= b_renew;
    }
    // re = qn[SE] + ekin
    qnb[SE] += b_re;
    const double b_ekin = b_re;
    // ru = u * rr ...
    double b_u = b_ru * rr, b_v = b_rv * rr, b_w = b_rw * rr;
    b_rr += b_ru * u + b_rv * v + b_rw * w;
    // ekin = 0.5 rr (u^2 + v^2 + w^2)
    b_rr += b_ekin * 0.5 * (u * u + v * v + w * w);
    b_u += b_ekin * rr * u;
    b_v += b_ekin * rr * v;
    b_w += b_ekin * rr * w;
    qnb[SR] += b_rr;
    qnb[SU] += b_u;
    qnb[SV] += b_v;
    qnb[SW] += b_w;
}

CAMR_HD void trans_diff_two_adj(const double *gal, const double *gar, const double *gbl,
                                const double *gbr, double ca, double cb, double gamc,
                                const TransDiffBar &tb, double *falb, double *farb,
                                double *galb, double *garb, double *fblb, double *fbrb,
                                double *gblb, double *gbrb) {
    const double sa[5] = {tb.rho, tb.mu, tb.mv, tb.mw, tb.e};
    const int cmp[5] = {FR, FMX, FMY, FMZ, FE};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        farb[cmp[q]] += ca * sa[q];
        falb[cmp[q]] -= ca * sa[q];
        fbrb[cmp[q]] += cb * sa[q];
        fblb[cmp[q]] -= cb * sa[q];
    }
    // pxnew = ca * (dAupx + pavx * dAux * (gamc - 1)) -> stored in dAup
    {
        const double pavx = 0.5 * (gar[GP] + gal[GP]);
        const double dAux = gar[GUN] - gal[GUN];
        const double b_in = ca * tb.dAup;
        const double b_dAupx = b_in, b_pavx = b_in * dAux * (gamc - 1.0);
        const double b_dAux = b_in * pavx * (gamc - 1.0);
        garb[GP] += b_dAupx * gar[GUN] + 0.5 * b_pavx;
        garb[GUN] += b_dAupx * gar[GP] + b_dAux;
        galb[GP] += -b_dAupx * gal[GUN] + 0.5 * b_pavx;
        galb[GUN] += -b_dAupx * gal[GP] - b_dAux;
    }
    {
        const double pavy = 0.5 * (gbr[GP] + gbl[GP]);
        const double dAuy = gbr[GUN] - gbl[GUN];
        const double b_in = cb * tb.pnew;
        const double b_dAupy = b_in, b_pavy = b_in * dAuy * (gamc - 1.0);
        const double b_dAuy = b_in * pavy * (gamc - 1.0);
        gbrb[GP] += b_dAupy * gbr[GUN] + 0.5 * b_pavy;
        gbrb[GUN] += b_dAupy * gbr[GP] + b_dAuy;
        gblb[GP] += -b_dAupy * gbl[GUN] + 0.5 * b_pavy;
        gblb[GUN] += -b_dAupy * gbl[GP] - b_dAuy;
    }
}

// transverse_two_d (hydro_transdd, one side, zero sources).
CAMR_HD void transverse_two_d_adj(const double *qn, const TransDiff &t, int trd,
                                  double small_pres, const double *qob, double *qnb,
                                  TransDiffBar &tb) {
    const double flxrho = t.rho, flxu = t.mu, flxv = t.mv, flxw = t.mw, flxe = t.e;
    const double rr = qn[SR];
    double rrnew = rr - flxrho;
    const double ru = rr * qn[SU], rv = rr * qn[SV], rw = rr * qn[SW];
    const double E2 = ru * ru + rv * rv + rw * rw;
    const double ekin = 0.5 * E2 / rr;
    const double re = qn[SE] + ekin;
    double runew = ru - flxu, rvnew = rv - flxv, rwnew = rw - flxw, renew = re - flxe;
    bool reset = false;
    if (trd == 1 && rrnew < 0.) {
        rrnew = rr; runew = ru; rvnew = rv; rwnew = rw; renew = re;
        reset = true;
    }
    const double K = runew * runew + rvnew * rvnew + rwnew * rwnew;
    const double rhoekin = 0.5 * K / rrnew;
    const double pn = reset ? qn[SP] : qn[SP] - t.dAup - t.pnew;
    const double b_pn = (pn < small_pres) ? 0.0 : qob[SP];
    qnb[SP] += b_pn;
    if (!reset) {
        tb.dAup -= b_pn;
        tb.pnew -= b_pn;
    }
    const double b_renew = qob[SE];
    const double b_rhoekin = -qob[SE];
    double b_rrnew = qob[SR];
    const double b_K = b_rhoekin * 0.5 / rrnew;
    b_rrnew -= b_rhoekin * rhoekin / rrnew;
    double b_runew = b_K * 2.0 * runew + qob[SU] / rrnew;
    double b_rvnew = b_K * 2.0 * rvnew + qob[SV] / rrnew;
    double b_rwnew = b_K * 2.0 * rwnew + qob[SW] / rrnew;
    b_rrnew -= (qob[SU] * (runew / rrnew) + qob[SV] * (rvnew / rrnew) + qob[SW] * (rwnew / rrnew)) / rrnew;
    double b_rr = b_rrnew, b_ru = b_runew, b_rv = b_rvnew, b_rw = b_rwnew, b_re = b_renew;
    if (!reset) {
        tb.rho -= b_rrnew; tb.mu -= b_runew; tb.mv -= b_rvnew; tb.mw -= b_rwnew; tb.e -= b_renew;
    }
    qnb[SE] += b_re;
    const double b_ekin = b_re;
    // ekin = 0.5 * E2 / rr
    const double b_E2 = b_ekin * 0.5 / rr;
    b_rr -= b_ekin * ekin / rr;
    b_ru += b_E2 * 2.0 * ru;
    b_rv += b_E2 * 2.0 * rv;
    b_rw += b_E2 * 2.0 * rw;
    // ru = rr * u
    qnb[SU] += b_ru * rr;
    qnb[SV] += b_rv * rr;
    qnb[SW] += b_rw * rr;
    b_rr += b_ru * qn[SU] + b_rv * qn[SV] + b_rw * qn[SW];
    qnb[SR] += b_rr;
}

// ------------------------------------------------------------------ PPM
CAMR_HD void vl_limited_adj(double a, double b, double c, double rb, double &ab, double &bb,
                            double &cb) {
    const double dsl = 2.0 * (b - a);
    const double dsr = 2.0 * (c - b);
    if (!(dsl * dsr > 0.0)) return;
    const double dsc = 0.5 * (c - a);
    const double sg = copysign(1.0, dsc);
    const double inner = vmin(fabs(dsl), fabs(dsr));
    const double mb = rb * sg;
    double b_inner = 0.0, b_adsc = 0.0;
    vmin_adj(fabs(dsc), inner, mb, b_adsc, b_inner);
    double b_adsl = 0.0, b_adsr = 0.0;
    vmin_adj(fabs(dsl), fabs(dsr), b_inner, b_adsl, b_adsr);
    const double b_dsc = b_adsc * sgn_of(dsc);
    const double b_dsl = b_adsl * sgn_of(dsl);
    const double b_dsr = b_adsr * sgn_of(dsr);
    cb += 0.5 * b_dsc; ab -= 0.5 * b_dsc;
    bb += 2.0 * b_dsl; ab -= 2.0 * b_dsl;
    cb += 2.0 * b_dsr; bb -= 2.0 * b_dsr;
}

// ppm_edges: s[0..4], flatn -> (sm, sp); smb / spb -> sb[0..4], flatb
CAMR_HD void ppm_edges_adj(const double *s, double flatn, double smb, double spb, double *sb,
                           double &flatb) {
    const double dl0 = vl_limited(s[0], s[1], s[2]);
    const double dc = vl_limited(s[1], s[2], s[3]);
    const double dr0 = vl_limited(s[2], s[3], s[4]);
    const double sm1 = 0.5 * (s[2] + s[1]) - (1.0 / 6.0) * (dc - dl0);
    const double lo_m = vmin(s[2], s[1]), hi_m = vmax(s[2], s[1]);
    const double sm2 = vmax(sm1, lo_m);
    const double sm3 = vmin(sm2, hi_m);
    const double sp1 = 0.5 * (s[3] + s[2]) - (1.0 / 6.0) * (dr0 - dc);
    const double lo_p = vmin(s[3], s[2]), hi_p = vmax(s[3], s[2]);
    const double sp2 = vmax(sp1, lo_p);
    const double sp3 = vmin(sp2, hi_p);
    const double omf = 1.0 - flatn;
    const double sm4 = flatn * sm3 + omf * s[2];
    const double sp4 = flatn * sp3 + omf * s[2];
    double b_sm4 = 0.0, b_sp4 = 0.0;
    if ((sp4 - s[2]) * (s[2] - sm4) <= 0.0) {
        sb[2] += smb + spb;
    } else if (fabs(sp4 - s[2]) >= 2.0 * fabs(sm4 - s[2])) {
        // sp = 3 s2 - 2 sm4, sm = sm4
        sb[2] += 3.0 * spb;
        b_sm4 += -2.0 * spb + smb;
    } else if (fabs(sm4 - s[2]) >= 2.0 * fabs(sp4 - s[2])) {
        sb[2] += 3.0 * smb;
        b_sp4 += -2.0 * smb + spb;
    } else {
        b_sm4 += smb;
        b_sp4 += spb;
    }
    // sm4 = flatn sm3 + (1 - flatn) s2
    double b_sm3 = b_sm4 * flatn, b_sp3 = b_sp4 * flatn;
    sb[2] += (b_sm4 + b_sp4) * omf;
    flatb += b_sm4 * (sm3 - s[2]) + b_sp4 * (sp3 - s[2]);
    // sm3 = vmin(sm2, hi_m); sm2 = vmax(sm1, lo_m)
    double b_sm2 = 0.0, b_him = 0.0, b_sm1 = 0.0, b_lom = 0.0;
    vmin_adj(sm2, hi_m, b_sm3, b_sm2, b_him);
    vmax_adj(sm1, lo_m, b_sm2, b_sm1, b_lom);
    vmin_adj(s[2], s[1], b_lom, sb[2], sb[1]);
    vmax_adj(s[2], s[1], b_him, sb[2], sb[1]);
    double b_sp2 = 0.0, b_hip = 0.0, b_sp1 = 0.0, b_lop = 0.0;
    vmin_adj(sp2, hi_p, b_sp3, b_sp2, b_hip);
    vmax_adj(sp1, lo_p, b_sp2, b_sp1, b_lop);
    vmin_adj(s[3], s[2], b_lop, sb[3], sb[2]);
    vmax_adj(s[3], s[2], b_hip, sb[3], sb[2]);
    // sm1 = 0.5 (s2 + s1) - (dc - dl0)/6 ; sp1 = 0.5 (s3 + s2) - (dr0 - dc)/6
    sb[2] += 0.5 * b_sm1 + 0.5 * b_sp1;
    sb[1] += 0.5 * b_sm1;
    sb[3] += 0.5 * b_sp1;
    const double b_dc = -(1.0 / 6.0) * b_sm1 + (1.0 / 6.0) * b_sp1;
    const double b_dl0 = (1.0 / 6.0) * b_sm1;
    const double b_dr0 = -(1.0 / 6.0) * b_sp1;
    vl_limited_adj(s[0], s[1], s[2], b_dl0, sb[0], sb[1], sb[2]);
    vl_limited_adj(s[1], s[2], s[3], b_dc, sb[1], sb[2], sb[3]);
    vl_limited_adj(s[2], s[3], s[4], b_dr0, sb[2], sb[3], sb[4]);
}

// ppm_integral: (sm, sp, s6, speed) -> (ip, im)
CAMR_HD void ppm_integral_adj(double sm, double sp, double s6, double speed, double dtdx,
                              double ipb, double imb, double &smb, double &spb, double &s6b,
                              double &speedb) {
    const double sigma = fabs(speed) * dtdx;
    double b_sigma = 0.0;
    const double k = 1.0 - (2.0 / 3.0) * sigma;
    if (speed <= 0.0) {
        spb += ipb;
        const double X = sp - sm + k * s6;
        smb += imb;
        b_sigma += imb * 0.5 * X;
        const double b_X = imb * 0.5 * sigma;
        spb += b_X;
        smb -= b_X;
        s6b += b_X * k;
        b_sigma += b_X * (-(2.0 / 3.0)) * s6;
    } else {
        smb += imb;
        const double Y = sp - sm - k * s6;
        spb += ipb;
        b_sigma -= ipb * 0.5 * Y;
        const double b_Y = -ipb * 0.5 * sigma;
        spb += b_Y;
        smb -= b_Y;
        s6b -= b_Y * k;
        b_sigma += b_Y * (2.0 / 3.0) * s6;
    }
    speedb += b_sigma * dtdx * sgn_of(speed);
}

// Adjoint of ppm_project for the hydro outputs (rho, un, p, rhoe); Ib
// accumulates the adjoints of the wave integrals, unb / ccb those of the
// cell's normal velocity and sound speed (only through the upwinding signs,
// which are piecewise constant: no contribution).
struct PPMWavesBar {
    double rho[3], un[3], p[3], rhoe[3];
};

CAMR_HD void ppm_project_adj(int side, const PPMWaves &I, double un, double cc,
                             double small_dens, double small_pres, double rob, double uob,
                             double pob, double reob, PPMWavesBar &Ib) {
    const int w = (side > 0) ? 0 : 2;
    const double rho_ref0 = I.rho[w];
    const double un_ref = I.un[w];
    const double p_ref0 = I.p[w];
    const double rhoe_g_ref = I.rhoe[w];
    const double gam_ref = I.gam[w];
    const double rho_ref = vmax(rho_ref0, small_dens);
    const double rinv = 1.0 / rho_ref;
    const double p_ref = vmax(p_ref0, small_pres);
    const double xcc = gam_ref * p_ref / rho_ref;
    const double cc_ref = sqrt(xcc);
    const double csq = cc_ref * cc_ref;
    const double ccinv = 1.0 / cc_ref;
    const double hnum = (p_ref + rhoe_g_ref);
    const double h_g = hnum * rinv / csq;
    const double dum = un_ref - I.un[0];
    const double dptotm = p_ref - I.p[0];
    const double drho = rho_ref - I.rho[1];
    const double dptot = p_ref - I.p[1];
    const double drhoe_g = rhoe_g_ref - I.rhoe[1];
    const double dup = un_ref - I.un[2];
    const double dptotp = p_ref - I.p[2];
    const double Am = dptotm * rinv * ccinv;
    const double Ap = dptotp * rinv * ccinv;
    const double alpham0 = 0.5 * (Am - dum) * rho_ref * ccinv;
    const double alphap0 = 0.5 * (Ap + dup) * rho_ref * ccinv;
    const double alpha0r0 = drho - dptot / csq;
    const double alpha0e0 = drhoe_g - dptot * h_g;
    const double sm_ = un - cc, sp_ = un + cc;
    double fm, fp, f0;
    if (side > 0) {
        fm = (sm_ > 0.) ? 0.0 : ((sm_ < 0.) ? -1.0 : -0.5);
        fp = (sp_ > 0.) ? 0.0 : ((sp_ < 0.) ? -1.0 : -0.5);
        f0 = (un > 0.) ? 0.0 : ((un < 0.) ? -1.0 : -0.5);
    } else {
        fm = (sm_ > 0.) ? -1.0 : ((sm_ < 0.) ? 0.0 : -0.5);
        fp = (sp_ > 0.) ? -1.0 : ((sp_ < 0.) ? 0.0 : -0.5);
        f0 = (un > 0.) ? -1.0 : ((un < 0.) ? 0.0 : -0.5);
    }
    const double alpham = alpham0 * fm, alphap = alphap0 * fp;
    const double alpha0r = alpha0r0 * f0, alpha0e = alpha0e0 * f0;
    const double apm = alphap + alpham;
    const double rsum = rho_ref + alphap + alpham + alpha0r;
    const double pt = p_ref + apm * csq;
    // ---- reverse
    double b_rho_ref = 0.0, b_un_ref = 0.0, b_p_ref = 0.0, b_rhoe_ref = 0.0;
    double b_csq = 0.0, b_cc = 0.0, b_rinv = 0.0, b_h = 0.0, b_apm = 0.0;
    double b_ap = 0.0, b_am = 0.0, b_a0r = 0.0, b_a0e = 0.0;
    double dummy = 0.0;
    // rhoe_o = rhoe_g_ref + apm * h_g * csq + alpha0e
    b_rhoe_ref += reob;
    b_apm += reob * h_g * csq;
    b_h += reob * apm * csq;
    b_csq += reob * apm * h_g;
    b_a0e += reob;
    // p_o = vmax(pt, clamp)
    double b_pt = 0.0;
    vmax_adj(pt, (side > 0) ? small_pres : small_dens, pob, b_pt, dummy);
    b_p_ref += b_pt;
    b_apm += b_pt * csq;
    b_csq += b_pt * apm;
    // un_o = un_ref + (alphap - alpham) * cc_ref * rinv
    b_un_ref += uob;
    const double b_d = uob * cc_ref * rinv;
    b_ap += b_d;
    b_am -= b_d;
    b_cc += uob * (alphap - alpham) * rinv;
    b_rinv += uob * (alphap - alpham) * cc_ref;
    // rho_o = vmax(rsum, small_dens)
    double b_rsum = 0.0;
    vmax_adj(rsum, small_dens, rob, b_rsum, dummy);
    b_rho_ref += b_rsum;
    b_ap += b_rsum;
    b_am += b_rsum;
    b_a0r += b_rsum;
    b_ap += b_apm;
    b_am += b_apm;
    const double b_am0 = b_am * fm, b_ap0 = b_ap * fp;
    const double b_a0r0 = b_a0r * f0, b_a0e0 = b_a0e * f0;
    // alpha0e0 = drhoe_g - dptot * h_g
    double b_drhoe = b_a0e0;
    double b_dptot = -b_a0e0 * h_g;
    b_h -= b_a0e0 * dptot;
    // alpha0r0 = drho - dptot / csq
    double b_drho = b_a0r0;
    b_dptot -= b_a0r0 / csq;
    b_csq += b_a0r0 * dptot / (csq * csq);
    // alphap0 = 0.5 (Ap + dup) rho_ref ccinv
    double b_ccinv = 0.0;
    {
        const double b_t = b_ap0 * 0.5 * rho_ref * ccinv;
        b_rho_ref += b_ap0 * 0.5 * (Ap + dup) * ccinv;
        b_ccinv += b_ap0 * 0.5 * (Ap + dup) * rho_ref;
        const double b_Ap = b_t;
        const double b_dup = b_t;
        // Ap = dptotp * rinv * ccinv
        const double b_dptotp = b_Ap * rinv * ccinv;
        b_rinv += b_Ap * dptotp * ccinv;
        b_ccinv += b_Ap * dptotp * rinv;
        b_p_ref += b_dptotp;
        Ib.p[2] -= b_dptotp;
        b_un_ref += b_dup;
        Ib.un[2] -= b_dup;
    }
    {
        const double b_t = b_am0 * 0.5 * rho_ref * ccinv;
        b_rho_ref += b_am0 * 0.5 * (Am - dum) * ccinv;
        b_ccinv += b_am0 * 0.5 * (Am - dum) * rho_ref;
        const double b_Am = b_t;
        const double b_dum = -b_t;
        const double b_dptotm = b_Am * rinv * ccinv;
        b_rinv += b_Am * dptotm * ccinv;
        b_ccinv += b_Am * dptotm * rinv;
        b_p_ref += b_dptotm;
        Ib.p[0] -= b_dptotm;
        b_un_ref += b_dum;
        Ib.un[0] -= b_dum;
    }
    // differences
    b_rhoe_ref += b_drhoe;
    Ib.rhoe[1] -= b_drhoe;
    b_p_ref += b_dptot;
    Ib.p[1] -= b_dptot;
    b_rho_ref += b_drho;
    Ib.rho[1] -= b_drho;
    // h_g = hnum * rinv / csq
    {
        const double b_hnum = b_h * rinv / csq;
        b_rinv += b_h * hnum / csq;
        b_csq -= b_h * h_g / csq;
        b_p_ref += b_hnum;
        b_rhoe_ref += b_hnum;
    }
    // ccinv = 1 / cc_ref ; csq = cc_ref^2
    b_cc -= b_ccinv * ccinv * ccinv;
    b_cc += b_csq * 2.0 * cc_ref;
    // cc_ref = sqrt(gam p_ref / rho_ref)
    {
        const double b_x = b_cc * 0.5 / cc_ref;
        b_p_ref += b_x * gam_ref / rho_ref;
        b_rho_ref -= b_x * xcc / rho_ref;
    }
    // rinv = 1 / rho_ref
    b_rho_ref -= b_rinv * rinv * rinv;
    double b_p0 = 0.0, b_r0 = 0.0;
    vmax_adj(p_ref0, small_pres, b_p_ref, b_p0, dummy);
    vmax_adj(rho_ref0, small_dens, b_rho_ref, b_r0, dummy);
    Ib.rho[w] += b_r0;
    Ib.un[w] += b_un_ref;
    Ib.p[w] += b_p0;
    Ib.rhoe[w] += b_rhoe_ref;
    (void)un; (void)cc;
}

// ------------------------------------------------------------- flattening
// flatten_1d depends on the pressures only through the shock-strength ramps
// (the shock tests are piecewise constant): pb[0..6] accumulate.
CAMR_HD void flatten_1d_adj(const double *p, const double *un, double fb, double *pb) {
    const double small_pres = 1.e-200;
    const double shktst = 0.33e0;
    const double zcut1 = 0.75e0;
    const double zcut2 = 0.85e0;
    const double dzcut = 1.0 / (zcut2 - zcut1);
    const int c = 3;
    const double qm = p[c - 1], qp = p[c + 1];
    const double dp1 = qp - qm;
    const int ishft = (dp1 > 0.0) ? 1 : -1;
    const double den1 = vmax(small_pres, fabs(p[c + 2] - p[c - 2]));
    const double zeta1 = fabs(dp1) / den1;
    const double r1 = dzcut * (zeta1 - zcut1);
    const double z1a = vmax(0.0, r1);
    const double z = vmin(1.0, z1a);
    const double tst1 = (un[c - 1] - un[c + 1] >= 0.0) ? 1.0 : 0.0;
    const double chi = (fabs(dp1) / vmin(qp, qm)) > shktst ? tst1 : 0.0;
    const int ip = c + 1 - ishft, im = c - 1 - ishft;
    const int ipp = c + 2 - ishft, imm = c - 2 - ishft;
    const double dp2 = p[ip] - p[im];
    const double den2 = vmax(small_pres, fabs(p[ipp] - p[imm]));
    const double zeta2 = fabs(dp2) / den2;
    const double r2 = dzcut * (zeta2 - zcut1);
    const double z2a = vmax(0.0, r2);
    const double z2 = vmin(1.0, z2a);
    const double tst2 = (un[im] - un[ip] >= 0.0) ? 1.0 : 0.0;
    const double chi2 = (fabs(dp2) / vmin(p[ip], p[im])) > shktst ? tst2 : 0.0;
    // flat = 1 - vmax(chi2 z2, chi z)
    double b_a = 0.0, b_b = 0.0, dummy = 0.0;
    vmax_adj(chi2 * z2, chi * z, -fb, b_a, b_b);
    const double b_z = b_b * chi, b_z2 = b_a * chi2;
    double b_z1a = 0.0, b_r1 = 0.0;
    vmin_adj(1.0, z1a, b_z, dummy, b_z1a);
    vmax_adj(0.0, r1, b_z1a, dummy, b_r1);
    double b_z2a = 0.0, b_r2 = 0.0;
    vmin_adj(1.0, z2a, b_z2, dummy, b_z2a);
    vmax_adj(0.0, r2, b_z2a, dummy, b_r2);
    {
        const double b_zeta = b_r1 * dzcut;
        const double b_dp = b_zeta * sgn_of(dp1) / den1;
        const double b_den = -b_zeta * zeta1 / den1;
        pb[c + 1] += b_dp;
        pb[c - 1] -= b_dp;
        double b_ad = 0.0;
        vmax_adj(small_pres, fabs(p[c + 2] - p[c - 2]), b_den, dummy, b_ad);
        const double sg = sgn_of(p[c + 2] - p[c - 2]);
        pb[c + 2] += b_ad * sg;
        pb[c - 2] -= b_ad * sg;
    }
    {
        const double b_zeta = b_r2 * dzcut;
        const double b_dp = b_zeta * sgn_of(dp2) / den2;
        const double b_den = -b_zeta * zeta2 / den2;
        pb[ip] += b_dp;
        pb[im] -= b_dp;
        double b_ad = 0.0;
        vmax_adj(small_pres, fabs(p[ipp] - p[imm]), b_den, dummy, b_ad);
        const double sg = sgn_of(p[ipp] - p[imm]);
        pb[ipp] += b_ad * sg;
        pb[imm] -= b_ad * sg;
    }
}

// ------------------------------------------------------------- primitives
// hydro_ctoprim into the compact layout (prim_compact): v -> o[SR..SE, SC,
// SCS]; ob are their adjoints (species omitted), vb accumulates for
// rho, rho u/v/w, rho E, rho e.
CAMR_HD void prim_compact_adj(const double *v, const double *ob, double eta1, double small_num,
                              double *vb) {
    const double rho = v[URHO];
    const double rhoinv = 1.0 / rho;
    const double vx = v[UMX] * rhoinv, vy = v[UMY] * rhoinv, vz = v[UMZ] * rhoinv;
    const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    const double ueden = v[UEDEN], ueint = v[UEINT];
    const bool use_E = (ueden - kineng) / ueden > eta1;
    const double e = use_E ? (ueden - kineng) * rhoinv : ueint * rhoinv;
    const double p = (kGamma - 1.0) * rho * e;
    const double xc = kGamma * p / rho;
    const double cs = sqrt(xc);
    double b_cs = ob[SC];
    // SCS = vmax(small_num, small_num * cs)
    if (small_num < small_num * cs) b_cs += ob[SCS] * small_num;
    double b_rho = ob[SR], b_p = ob[SP], b_e = 0.0;
    // cs = sqrt(gamma p / rho)
    if (cs > 0.0) {
        const double b_x = b_cs * 0.5 / cs;
        b_p += b_x * kGamma / rho;
        b_rho -= b_x * xc / rho;
    }
    // rhoe output = e * rho
    b_e += ob[SE] * rho;
    b_rho += ob[SE] * e;
    // p = ((gamma - 1) rho) e
    b_e += b_p * (kGamma - 1.0) * rho;
    b_rho += b_p * (kGamma - 1.0) * e;
    double b_rhoinv = 0.0, b_kin = 0.0;
    if (use_E) {
        vb[UEDEN] += b_e * rhoinv;
        b_kin -= b_e * rhoinv;
        b_rhoinv += b_e * (ueden - kineng);
    } else {
        vb[UEINT] += b_e * rhoinv;
        b_rhoinv += b_e * ueint;
    }
    // kineng = 0.5 rho (vx^2 + vy^2 + vz^2)
    b_rho += b_kin * 0.5 * (vx * vx + vy * vy + vz * vz);
    const double b_vx = ob[SU] + b_kin * rho * vx;
    const double b_vy = ob[SV] + b_kin * rho * vy;
    const double b_vz = ob[SW] + b_kin * rho * vz;
    vb[UMX] += b_vx * rhoinv;
    vb[UMY] += b_vy * rhoinv;
    vb[UMZ] += b_vz * rhoinv;
    b_rhoinv += b_vx * v[UMX] + b_vy * v[UMY] + b_vz * v[UMZ];
    b_rho -= b_rhoinv * rhoinv * rhoinv;
    vb[URHO] += b_rho;
}

// ---------------------------------------------------------------- clean
// reset_e_and_T for allow_small_energy = 1, allow_negative_energy = 0 (the
// defaults): sb holds the adjoint of the output state on entry and of the
// input state on return (hydro components; species and T are passive).
// Returns false if the state takes a branch this adjoint does not model
// (a small-energy reset, which involves the species through the EOS).
CAMR_HD bool reset_e_adj(const double *s, const HydroConst &hc, double *sb) {
    const double rho = s[URHO];
    const double rhoInv = 1.0 / rho;
    const double Up = s[UMX] * rhoInv, Vp = s[UMY] * rhoInv, Wp = s[UMZ] * rhoInv;
    const double ke = 0.5 * (Up * Up + Vp * Vp + Wp * Wp);
    double b_ke = 0.0, b_rho = 0.0;
    bool ok = true;
    if (s[UEDEN] < 0.) {
        if (s[UEINT] < 0.) ok = false;
        // UEDEN = UEINT + rho ke
        const double b_E = sb[UEDEN];
        sb[UEDEN] = 0.0;
        sb[UEINT] += b_E;
        b_rho += b_E * ke;
        b_ke += b_E * rho;
    } else {
        const double rho_eint = s[UEDEN] - rho * ke;
        if (rho_eint > 0. && rho_eint / s[UEDEN] > hc.eta2) {
            // UEINT = UEDEN - rho ke
            const double b_e = sb[UEINT];
            sb[UEINT] = 0.0;
            sb[UEDEN] += b_e;
            b_rho -= b_e * ke;
            b_ke -= b_e * rho;
        } else if (s[UEINT] > 0 && hc.update_E_from_e == 1) {
            const double b_E = sb[UEDEN];
            sb[UEDEN] = 0.0;
            sb[UEINT] += b_E;
            b_rho += b_E * ke;
            b_ke += b_E * rho;
        } else if (s[UEINT] <= 0.) {
            ok = false;
        }
    }
    // ke = 0.5 (Up^2 + Vp^2 + Wp^2), Up = UMX / rho
    const double b_Up = b_ke * Up, b_Vp = b_ke * Vp, b_Wp = b_ke * Wp;
    sb[UMX] += b_Up * rhoInv;
    sb[UMY] += b_Vp * rhoInv;
    sb[UMZ] += b_Wp * rhoInv;
    const double b_rinv = b_Up * s[UMX] + b_Vp * s[UMY] + b_Wp * s[UMZ];
    b_rho -= b_rinv * rhoInv * rhoInv;
    sb[URHO] += b_rho;
    return ok;
}

} // namespace camr

