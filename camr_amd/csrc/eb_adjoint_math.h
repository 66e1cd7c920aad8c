// eb_adjoint_math.h -- reverse-mode derivatives of the per-point arithmetic
// of the embedded-boundary (PLM) path: the limited slopes and the
// characteristic trace of the PLM predictor, the face flux with the
// domain-boundary overrides, and the wall flux of a cut cell.  Host and
// device; same conventions as adjoint_math.h (the forward is recomputed
// inside, J^T ybar is ADDED to the input adjoints, piecewise branches take
// the derivative of the branch the forward evaluation takes).
#pragma once

#include "adjoint_math.h"

namespace camr {

// d|x|/dx with the sign convention of sgn_of (|0| counts as increasing)
CAMR_HD double dabs(double x) { return sgn_of(x); }

// ---------------------------------------------------------------- slopes
// mc_slope (plm_slope, PLM.H:19-79) for order 2 / 4: s[0..4] at offsets
// -2..2, flat the flattening coefficient.  yb: adjoint of the slope.
CAMR_HD void mc_slope_adj(const double *s, double flat, int order, double yb, double *sb,
                          double &flatb) {
    if (order == 1) return;
    double dfm = 0.0, dfp = 0.0;
    // the one-sided limited slopes of order 4, forward
    double m_dlft[2], m_drgt[2], m_dcen[2], m_dsgn[2], m_dlim[2];
    bool m_cond[2];
    if (order == 4) {
        for (int h = 0; h < 2; ++h) {
            const int o = h == 0 ? 0 : 2;  // dfm from s[0..2], dfp from s[2..4]
            m_dlft[h] = s[o + 1] - s[o];
            m_drgt[h] = s[o + 2] - s[o + 1];
            m_dcen[h] = 0.5 * (m_dlft[h] + m_drgt[h]);
            m_dsgn[h] = copysign(1.0, m_dcen[h]);
            m_cond[h] = m_dlft[h] * m_drgt[h] >= 0.0;
            m_dlim[h] = m_cond[h] ? 2.0 * vmin(fabs(m_dlft[h]), fabs(m_drgt[h])) : 0.0;
        }
        dfm = m_dsgn[0] * vmin(m_dlim[0], fabs(m_dcen[0]));
        dfp = m_dsgn[1] * vmin(m_dlim[1], fabs(m_dcen[1]));
    }
    const double dlft = s[2] - s[1], drgt = s[3] - s[2];
    const double dcen = 0.5 * (dlft + drgt);
    const double dsgn = copysign(1.0, dcen);
    const bool cond = dlft * drgt >= 0.0;
    const double dlim = cond ? 2.0 * vmin(fabs(dlft), fabs(drgt)) : 0.0;
    const double dtemp = 4.0 / 3.0 * dcen - 1.0 / 6.0 * (dfp + dfm);
    const double m = vmin(dlim, fabs(dtemp));
    // ---- reverse
    flatb += yb * dsgn * m;
    const double mb = yb * flat * dsgn;
    double dlimb = 0.0, atb = 0.0;
    vmin_adj(dlim, fabs(dtemp), mb, dlimb, atb);
    const double dtempb = atb * dabs(dtemp);
    double dcenb = 4.0 / 3.0 * dtempb;
    const double dfb = -1.0 / 6.0 * dtempb;  // dfp and dfm alike
    double dlftb = 0.5 * dcenb, drgtb = 0.5 * dcenb;
    if (cond) {
        double al = 0.0, ar = 0.0;
        vmin_adj(fabs(dlft), fabs(drgt), 2.0 * dlimb, al, ar);
        dlftb += al * dabs(dlft);
        drgtb += ar * dabs(drgt);
    }
    sb[2] += dlftb;
    sb[1] -= dlftb;
    sb[3] += drgtb;
    sb[2] -= drgtb;
    if (order == 4) {
        for (int h = 0; h < 2; ++h) {
            const int o = h == 0 ? 0 : 2;
            // df = dsgn * vmin(dlim, |dcen|)
            const double vb = dfb * m_dsgn[h];
            double lb = 0.0, cb = 0.0;
            vmin_adj(m_dlim[h], fabs(m_dcen[h]), vb, lb, cb);
            const double dcb = cb * dabs(m_dcen[h]);
            double lftb = 0.5 * dcb, rgtb = 0.5 * dcb;
            if (m_cond[h]) {
                double al = 0.0, ar = 0.0;
                vmin_adj(fabs(m_dlft[h]), fabs(m_drgt[h]), 2.0 * lb, al, ar);
                lftb += al * dabs(m_dlft[h]);
                rgtb += ar * dabs(m_drgt[h]);
            }
            sb[o + 1] += lftb;
            sb[o] -= lftb;
            sb[o + 2] += rgtb;
            sb[o + 1] -= rgtb;
        }
    }
}

// ---------------------------------------------------------------- trace
// plm_trace (hydro_plm_d, PLM.H:170-323) for the hydro components SR..SE:
// qpb / qmb the adjoints of the two face states; cb / slb (compact layout)
// and ccb accumulate.  Species are passive and left out.
CAMR_HD void plm_trace_adj(int d, const double *c, const double *sl, double cc, double dtdx,
                           double small_dens, double small_pres, const double *qpb,
                           const double *qmb, double *cb, double *slb, double &ccb) {
    const int un_i = SU + d, l1 = vel_t1(d), l2 = vel_t2(d);
    const double cs = cc * cc;
    const double rho = c[SR], p = c[SP], rhoe = c[SE];
    const double un = c[un_i];
    const double t1 = (rhoe + p) / rho;
    const double enth = t1 / cs;
    const double drho = sl[SR], dp = sl[SP], drhoe = sl[SE];
    const double dun = sl[un_i];
    const double u1 = dp / (rho * cc);
    const double inm = 0.5 * (u1 - dun), inp = 0.5 * (u1 + dun);
    const double alpham = inm * rho / cc;
    const double alphap = inp * rho / cc;
    const double alpha0r = drho - dp / cs;
    const double alpha0e = drhoe - dp * enth;
    const double alpha0v = sl[l1];
    const double alpha0w = sl[l2];
    const double wv0 = un - cc, wv1 = un, wv2 = un + cc;

    double rhob = 0.0, pb = 0.0, rhoeb = 0.0, unb = 0.0, c1b = 0.0, c2b = 0.0;
    double drhob = 0.0, dpb = 0.0, drhoeb = 0.0, dunb = 0.0, s1b = 0.0, s2b = 0.0;
    double csb = 0.0, enthb = 0.0, alphamb = 0.0, alphapb = 0.0;
    double a0rb = 0.0, a0eb = 0.0, a0vb = 0.0, a0wb = 0.0;
    double wv0b = 0.0, wv1b = 0.0, wv2b = 0.0;
    double ccb_l = 0.0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        // side 0: the plus state (qp, face i-1/2); side 1: the minus state
        const double *ob = side == 0 ? qpb : qmb;
        const double sg = side == 0 ? -1.0 : 1.0;  // rho_ref = rho + sg * fac * drho
        double fac, W, kz, wa;
        if (side == 0) {
            fac = 0.5 * (1.0 + dtdx * vmin(wv0, 0.0));
            W = 0.25 * dtdx * (wv0 - wv2) * (1.0 - copysign(1.0, wv2)) * alphap;
            kz = 0.25 * dtdx;
            wa = wv0;
        } else {
            fac = 0.5 * (1.0 - dtdx * vmax(wv2, 0.0));
            W = 0.25 * dtdx * (wv2 - wv0) * (1.0 + copysign(1.0, wv0)) * alpham;
            kz = 0.25 * dtdx;
            wa = wv2;
        }
        const double zs = side == 0 ? (1.0 - copysign(1.0, wv1)) : (1.0 + copysign(1.0, wv1));
        const double z = kz * (wa - wv1) * zs;
        const double kU = side == 0 ? 1.0 : -1.0;  // velocity: + apright / - amleft
        const double rho_ref = rho + sg * fac * drho;
        const double p_ref = p + sg * fac * dp;
        const double azr = z * alpha0r, aze = z * alpha0e;
        double Wb = 0.0, zb = 0.0, facb = 0.0;
        double rho_refb = 0.0, vnb = 0.0, v1b = 0.0, v2b = 0.0, p_refb = 0.0, rhoe_refb = 0.0;
        double azrb = 0.0, azeb = 0.0, azvb = 0.0, azwb = 0.0;
        {   // SR = vmax(rho_ref + W + azr, small_dens)
            double tb = 0.0, dm = 0.0;
            vmax_adj(rho_ref + W + 0.0 + azr, small_dens, ob[SR], tb, dm);
            rho_refb += tb;
            Wb += tb;
            azrb += tb;
        }
        {   // un = vref_n + (kU W) cc / rho
            const double b = ob[un_i];
            vnb += b;
            Wb += b * kU * cc / rho;
            ccb_l += b * kU * W / rho;
            rhob -= b * (kU * W * cc) / (rho * rho);
        }
        v1b += ob[l1];
        azvb += ob[l1];
        v2b += ob[l2];
        azwb += ob[l2];
        {   // SP = vmax(p_ref + W cs, small_pres)
            double tb = 0.0, dm = 0.0;
            vmax_adj(p_ref + W * cs, small_pres, ob[SP], tb, dm);
            p_refb += tb;
            Wb += tb * cs;
            csb += tb * W;
        }
        {   // SE = rhoe_ref + W enth cs + aze
            const double b = ob[SE];
            rhoe_refb += b;
            Wb += b * enth * cs;
            enthb += b * W * cs;
            csb += b * W * enth;
            azeb += b;
        }
        // az* = z * alpha0*
        zb += azrb * alpha0r + azeb * alpha0e + azvb * alpha0v + azwb * alpha0w;
        a0rb += azrb * z;
        a0eb += azeb * z;
        a0vb += azvb * z;
        a0wb += azwb * z;
        // z = kz (wa - wv1) zs
        const double wdb = zb * kz * zs;
        if (side == 0) wv0b += wdb; else wv2b += wdb;
        wv1b -= wdb;
        // W
        if (side == 0) {
            const double k = 0.25 * dtdx * (1.0 - copysign(1.0, wv2));
            wv0b += Wb * k * alphap;
            wv2b -= Wb * k * alphap;
            alphapb += Wb * 0.25 * dtdx * (wv0 - wv2) * (1.0 - copysign(1.0, wv2));
        } else {
            const double k = 0.25 * dtdx * (1.0 + copysign(1.0, wv0));
            wv2b += Wb * k * alpham;
            wv0b -= Wb * k * alpham;
            alphamb += Wb * 0.25 * dtdx * (wv2 - wv0) * (1.0 + copysign(1.0, wv0));
        }
        // refs: x_ref = x + sg fac dx
        rhob += rho_refb;
        drhob += sg * fac * rho_refb;
        unb += vnb;
        dunb += sg * fac * vnb;
        c1b += v1b;
        s1b += sg * fac * v1b;
        c2b += v2b;
        s2b += sg * fac * v2b;
        pb += p_refb;
        dpb += sg * fac * p_refb;
        rhoeb += rhoe_refb;
        drhoeb += sg * fac * rhoe_refb;
        facb += sg * (rho_refb * drho + vnb * dun + v1b * sl[l1] + v2b * sl[l2] +
                      p_refb * dp + rhoe_refb * drhoe);
        // fac
        if (side == 0) {
            if (!(0.0 < wv0)) wv0b += facb * 0.5 * dtdx;
        } else {
            if (!(wv2 < 0.0)) wv2b -= facb * 0.5 * dtdx;
        }
    }
    // waves
    unb += wv0b + wv1b + wv2b;
    ccb_l += wv2b - wv0b;
    s1b += a0vb;
    s2b += a0wb;
    // alpha0e = drhoe - dp enth
    drhoeb += a0eb;
    dpb -= a0eb * enth;
    enthb -= a0eb * dp;
    // alpha0r = drho - dp / cs
    drhob += a0rb;
    dpb -= a0rb / cs;
    csb += a0rb * dp / (cs * cs);
    // alpha(m|p) = in(m|p) * rho / cc, in = 0.5 (u1 -+ dun)
    {
        const double inb_m = alphamb * rho / cc, inb_p = alphapb * rho / cc;
        rhob += (alphamb * inm + alphapb * inp) / cc;
        ccb_l -= (alphamb * inm * rho + alphapb * inp * rho) / (cc * cc);
        const double u1b = 0.5 * (inb_m + inb_p);
        dunb += 0.5 * (inb_p - inb_m);
        // u1 = dp / (rho cc)
        dpb += u1b / (rho * cc);
        rhob -= u1b * u1 / rho;
        ccb_l -= u1b * u1 / cc;
    }
    // enth = ((rhoe + p) / rho) / cs
    {
        const double t1b = enthb / cs;
        csb -= enthb * enth / cs;
        rhoeb += t1b / rho;
        pb += t1b / rho;
        rhob -= t1b * t1 / rho;
    }
    ccb_l += csb * 2.0 * cc;
    cb[SR] += rhob;
    cb[SP] += pb;
    cb[SE] += rhoeb;
    cb[un_i] += unb;
    cb[l1] += c1b;
    cb[l2] += c2b;
    slb[SR] += drhob;
    slb[SP] += dpb;
    slb[SE] += drhoeb;
    slb[un_i] += dunb;
    slb[l1] += s1b;
    slb[l2] += s2b;
    ccb += ccb_l;
}

// ------------------------------------------------------------ face flux
// face_flux (hydro_cmpflx) on face idx of direction d, the domain-boundary
// overrides included: a Symmetry / SlipWall / NoSlipWall or Outflow face
// replaces one side's density, pressure, rho e and velocities by the other
// side's (the normal velocity negated at a wall).  ql / qr: the states as
// the face reads them; qlb / qrb accumulate (the replaced components of the
// replaced side get nothing); ab the sound-speed adjoints.
CAMR_HD void face_flux_bc_adj(int d, int idx, const HydroConst &hc, const double *ql,
                              const double *qr, const FaceAux &a, const double *fb,
                              const double *gb, double *qlb, double *qrb, FaceAuxBar &ab) {
    const int IU = vel_n(d), IV = vel_t1(d), IV2 = vel_t2(d);
    double el[NS], er[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) { el[s] = ql[s]; er[s] = qr[s]; }
    int lo = 0, hi = 0;  // 1: wall, 2: outflow
    if (idx == hc.domlo[d]) {
        const int bc = hc.bclo[d];
        if (bc == BC_SYMMETRY || bc == BC_SLIPWALL || bc == BC_NOSLIPWALL) lo = 1;
        else if (bc == BC_OUTFLOW) lo = 2;
    }
    if (lo) {
        el[SR] = er[SR]; el[SP] = er[SP]; el[SE] = er[SE];
        el[IU] = lo == 1 ? -er[IU] : er[IU];
        el[IV] = er[IV];
        el[IV2] = er[IV2];
    }
    if (idx == hc.domhi[d] + 1) {
        const int bc = hc.bchi[d];
        if (bc == BC_SYMMETRY || bc == BC_SLIPWALL || bc == BC_NOSLIPWALL) hi = 1;
        else if (bc == BC_OUTFLOW) hi = 2;
    }
    if (hi) {
        er[SR] = el[SR]; er[SP] = el[SP]; er[SE] = el[SE];
        er[IU] = hi == 1 ? -el[IU] : el[IU];
        er[IV] = el[IV];
        er[IV2] = el[IV2];
    }
    double elb[NS] = {}, erb[NS] = {};
    face_flux_adj(d, hc, el, er, a, fb, gb, elb, erb, ab);
    // the high side first (it read the low side's effective values)
    if (hi) {
        elb[SR] += erb[SR]; elb[SP] += erb[SP]; elb[SE] += erb[SE];
        elb[IU] += hi == 1 ? -erb[IU] : erb[IU];
        elb[IV] += erb[IV];
        elb[IV2] += erb[IV2];
        erb[SR] = erb[SP] = erb[SE] = erb[IU] = erb[IV] = erb[IV2] = 0.0;
    }
    if (lo) {
        erb[SR] += elb[SR]; erb[SP] += elb[SP]; erb[SE] += elb[SE];
        erb[IU] += lo == 1 ? -elb[IU] : elb[IU];
        erb[IV] += elb[IV];
        erb[IV2] += elb[IV2];
        elb[SR] = elb[SP] = elb[SE] = elb[IU] = elb[IV] = elb[IV2] = 0.0;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) { qlb[s] += elb[s]; qrb[s] += erb[s]; }
}

// ------------------------------------------------------------ wall flux
// compute_hyp_wallflux (Hydro_hypflux_eb_K.H:12-68) of a cut cell: the
// cell's state w[0..5] = (rho, u, v, w, p, rho e) and sound speeds c, csml
// against its mirror across the wall normal (the aperture differences
// nx, ny, nz, not normalised).  fwb: adjoints of the wall fluxes
// (URHO, UMX, UMY, UMZ, UEDEN, UEINT); wb / cb / csmlb accumulate.
CAMR_HD void wallflux_adj(const double *w, double c, double csml, double nx, double ny, double nz,
                          const HydroConst &hc, const double *fwb, double *wb, double &cb,
                          double &csmlb) {
    const double apnorm = sqrt(nx * nx + ny * ny + nz * nz);
    const double apnorminv = 1. / apnorm;
    const double un = w[1] * nx * apnorminv + w[2] * ny * apnorminv + w[3] * nz * apnorminv;
    RiemannOut o;
    riemann_two_shock(w[0], un, 0., 0., w[4], w[5], w[0], -un, 0., 0., w[4], w[5], c, kGamma,
                      kGamma, hc.small, csml, hc.small_dens, hc.small_pres, hc.smallu, o);
    RiemannBar ob{};
    ob.f_rho = fwb[URHO];
    ob.f_eden = fwb[UEDEN];
    ob.f_eint = fwb[UEINT];
    ob.f_un = fwb[UMX] * nx + fwb[UMY] * ny + fwb[UMZ] * nz;
    RiemannInBar ib{};
    riemann_two_shock_adj(w[0], un, 0., 0., w[4], w[5], w[0], -un, 0., 0., w[4], w[5], c, kGamma,
                          kGamma, hc.small, csml, hc.small_dens, hc.small_pres, hc.smallu, ob,
                          ib);
    wb[0] += ib.rl + ib.rr;
    wb[4] += ib.pl + ib.pr;
    wb[5] += ib.rel + ib.rer;
    const double unb = ib.ul - ib.ur;
    wb[1] += unb * nx * apnorminv;
    wb[2] += unb * ny * apnorminv;
    wb[3] += unb * nz * apnorminv;
    cb += ib.cav;
    csmlb += ib.smallc;
}

} // namespace camr
