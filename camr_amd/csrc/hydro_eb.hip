// hydro_eb.hip -- gfx950 kernels of hydro_umdrv_eb, the cut-cell box path
// (Source/Hydro/Hydro_umdrv_eb.cpp:14-191): Godunov_umeth_eb
// (Godunov/Godunov_3D_eb.cpp:16-567) with the aperture guards of the PLM
// trace, the transverse predictors and the Riemann solves; adjust_fluxes_eb
// (Hydro_adjust_fluxes.cpp:48-170); hydro_consup_eb (Hydro_consup.cpp:
// 29-98: the face-centroid divergence, the wall flux and p div u); and
// AMReX's single-level flux redistribution.
//
// One kernel per stage over the points of a table of boxes (grid.y = box),
// intermediates in HBM scratch per box in the compact layouts of the staged
// engine, so that a level's cut-cell boxes share each launch.  The per-point
// arithmetic is hydro_math.h's, so a cut-cell box reproduces the regular
// path's operation order wherever every face is open.  The redistribution
// scatter of the reference is a gather here: each cell adds the shares of
// its cut neighbours in the order the sequential scatter would, so the sum
// is bit-identical and needs no atomics.
#include "hydro_eb.h"
#include "state_math.h"

#include <cstring>
#include <utility>

namespace camr {

namespace {

// compact state / flux component -> QVAR / NVAR component
CAMR_HD constexpr int kQc(int s) { return s < 4 ? s : s == 4 ? QPRES : s == 5 ? QREINT : QFS + s - 6; }
CAMR_HD constexpr int kUc(int s) { return s < 6 ? s : UFS + s - 6; }

CAMR_HD int dv(int d, int e) { return d == e ? 1 : 0; }

// One direction of the cut-cell PLM trace: slopes of order 4 / 2 / 0 by the
// covered neighbours, the trace, states kept on open faces only.
template <int D>
CAMR_HD void plm_eb_dir(int i, int j, int k, const Arr &q, const EbGeom &g,
                                           const double *c, double cc, double fl, const Arr &QM,
                                           const Arr &QP, const HydroConst &hc,
                                           const double *dtp) {
    constexpr int ex = D == 0, ey = D == 1, ez = D == 2;
    const bool cm1 = g.f.covered(i - ex, j - ey, k - ez), cm2 = g.f.covered(i - 2 * ex, j - 2 * ey, k - 2 * ez);
    const bool cp1 = g.f.covered(i + ex, j + ey, k + ez), cp2 = g.f.covered(i + 2 * ex, j + 2 * ey, k + 2 * ez);
    int ord = 0;
    if (hc.plm_iorder == 4 && !cm1 && !cm2 && !cp1 && !cp2) ord = 4;
    else if (hc.plm_iorder > 1 && !cm1 && !cp1) ord = 2;
    double sl[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o) st[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, kQc(s));
        sl[s] = ord ? mc_slope(st, fl, ord) : 0.0;
    }
    double qp[NS], qm[NS];
    plm_trace(D, c, sl, cc, *dtp / hc.dx[D], hc.small_dens, hc.small_pres, qp, qm);
    const bool open_lo = g.ap[D](i, j, k, 0) > 0.0;
    const bool open_hi = g.ap[D](i + ex, j + ey, k + ez, 0) > 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        QP(i, j, k, s) = open_lo ? qp[s] : 0.0;
        QM(i + ex, j + ey, k + ez, s) = open_hi ? qm[s] : 0.0;
    }
}

// ---------------------------------------------- PLM trace (cut-cell form)
// Godunov_3D_eb.cpp:94-152: flatten_eb (flatten.H:104-134), plm_slope_eb
// (PLM_eb.H:18-54), hydro_plm_d_eb (PLM_eb.H:103-287); covered cells skip.
CAMR_HD void
b_plm_eb(int i, int j, int k, Arr q, Box qbox, Arr qaux, EbGeom g, Arr qm0, Arr qp0, Arr qm1,
         Arr qp1, Arr qm2, Arr qp2, HydroConst hc, const double *dtp) {
    if (eb_type(g.f(i, j, k)) == kEbCovered) return;
    double fl = 1.0;
    if (hc.use_flattening == 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int ex = dv(d, 0), ey = dv(d, 1), ez = dv(d, 2);
            double fd = 1.0;
            if (qbox.contains(i + 3 * ex, j + 3 * ey, k + 3 * ez) &&
                qbox.contains(i - 3 * ex, j - 3 * ey, k - 3 * ez)) {
                bool cut = eb_type(g.f(i, j, k)) == kEbSingle;
                for (int n = 1; n <= 3; ++n)
                    cut = cut || eb_type(g.f(i - n * ex, j - n * ey, k - n * ez)) == kEbSingle ||
                          eb_type(g.f(i + n * ex, j + n * ey, k + n * ez)) == kEbSingle;
                if (!cut) {
                    double p[7], un[7];
#pragma unroll
                    for (int o = -3; o <= 3; ++o) {
                        p[o + 3] = q(i + o * ex, j + o * ey, k + o * ez, QPRES);
                        un[o + 3] = q(i + o * ex, j + o * ey, k + o * ez, QU + d);
                    }
                    fd = flatten_1d(p, un);
                }
            }
            fl = vmin(fl, fd);
        }
    }
    double c[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) c[s] = q(i, j, k, kQc(s));
    const double cc = qaux(i, j, k, QC);
    plm_eb_dir<0>(i, j, k, q, g, c, cc, fl, qm0, qp0, hc, dtp);
    plm_eb_dir<1>(i, j, k, q, g, c, cc, fl, qm1, qp1, hc, dtp);
    plm_eb_dir<2>(i, j, k, q, g, c, cc, fl, qm2, qp2, hc, dtp);
}

// ------------------------------------------------ Riemann solve, open faces
// hydro_cmpflx where the aperture is > 0.  MODE 0: first solve (compact out,
// boundary-mutated states written back); 1: transverse solve (compact out);
// 2: final solve into the NVAR flux array.
template <int D, int MODE>
CAMR_HD void
b_riemann_eb(int i, int j, int k, Arr qL, Arr qR, Arr qaux, Arr fout, Arr gout, Arr ap, HydroConst hc) {
    if (!(ap(i, j, k, 0) > 0.)) return;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    double ql[NS], qr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        ql[s] = qL(i, j, k, s);
        qr[s] = qR(i, j, k, s);
    }
    FaceAux a;
    a.c_l = qaux(i - ex, j - ey, k - ez, QC);
    a.c_r = qaux(i, j, k, QC);
    a.csml_l = qaux(i - ex, j - ey, k - ez, QCSML);
    a.csml_r = qaux(i, j, k, QCSML);
    a.gam_l = qaux(i - ex, j - ey, k - ez, QGAMC);
    a.gam_r = qaux(i, j, k, QGAMC);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], gg[NG];
    int mut;
    face_flux(D, idx, hc, ql, qr, a, f, gg, &mut);
    if (MODE == 0 && mut) {
        if (mut & 1) { qL(i, j, k, SR) = ql[SR]; qL(i, j, k, SP) = ql[SP]; qL(i, j, k, SE) = ql[SE]; }
        if (mut & 2) { qR(i, j, k, SR) = qr[SR]; qR(i, j, k, SP) = qr[SP]; qR(i, j, k, SE) = qr[SE]; }
    }
    if (MODE == 2) {
#pragma unroll
        for (int s = 0; s < NF; ++s) fout(i, j, k, kUc(s)) = f[s];
    } else {
#pragma unroll
        for (int s = 0; s < NF; ++s) fout(i, j, k, s) = f[s];
    }
    gout(i, j, k, GUN) = gg[GUN];
    gout(i, j, k, GP) = gg[GP];
}

// ---------------------------------------------------- transdo, cut-cell form
// hydro_transdo with norm_area = ap[D], trans_area = ap[T]
// (Godunov_utils_3D.H:22-239): a closed transverse face zeroes the flux
// differences (the same arithmetic as zero fluxes), a closed normal face
// passes the predictor state through.
template <int D, int T>
CAMR_HD void
b_transdo_eb(int i, int j, int k, Arr qnm, Arr qnp, Arr qaux, Arr F, Arr G, Arr apD, Arr apT,
             EbFlags fl, Arr qmo, Arr qpo, const double *dtp, HydroConst hc) {
    if (fl.covered(i, j, k)) return;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    constexpr int tx = (T == 0), ty = (T == 1), tz = (T == 2);
    const double cdt = 1.0 / 3.0 * *dtp / hc.dx[T];  // Godunov_3D_eb.cpp:59-61
    const bool no_cov = apT(i + tx, j + ty, k + tz, 0) > 0.0 && apT(i, j, k, 0) > 0.0;
    const bool lo_open = apD(i, j, k, 0) > 0.0;
    const bool hi_open = apD(i + ex, j + ey, k + ez, 0) > 0.0;
    double fl_[NF], fr[NF], gl[NG], gr[NG];
#pragma unroll
    for (int s = 0; s < NF; ++s) {
        fl_[s] = no_cov ? F(i, j, k, s) : 0.0;
        fr[s] = no_cov ? F(i + tx, j + ty, k + tz, s) : 0.0;
    }
    gl[GUN] = no_cov ? G(i, j, k, GUN) : 0.0;
    gl[GP] = no_cov ? G(i, j, k, GP) : 0.0;
    gr[GUN] = no_cov ? G(i + tx, j + ty, k + tz, GUN) : 0.0;
    gr[GP] = no_cov ? G(i + tx, j + ty, k + tz, GP) : 0.0;
    const double gamc = qaux(i, j, k, QGAMC);
    double n[NS], o[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) n[s] = qnp(i, j, k, s);
    if (lo_open) transverse_one(n, fl_, fr, gl, gr, cdt, gamc, hc.trd, hc.small_pres, o);
#pragma unroll
    for (int s = 0; s < NS; ++s) qpo(i, j, k, s) = lo_open ? o[s] : n[s];
#pragma unroll
    for (int s = 0; s < NS; ++s) n[s] = qnm(i + ex, j + ey, k + ez, s);
    if (hi_open) transverse_one(n, fl_, fr, gl, gr, cdt, gamc, hc.trd, hc.small_pres, o);
#pragma unroll
    for (int s = 0; s < NS; ++s) qmo(i + ex, j + ey, k + ez, s) = hi_open ? o[s] : n[s];
}

// ---------------------------------------------------- transdd, cut-cell form
// hydro_transdd with norm_area = ap[D], trans0_area = ap[T0], trans1_area =
// ap[T1] (Godunov_utils_3D.H:245-482).
template <int D>
CAMR_HD void
b_transdd_eb(int i, int j, int k, Arr qnm, Arr qnp, Arr qaux, Arr srcq, int has_src, Arr fa, Arr ga,
             Arr fb, Arr gb, Arr apD, Arr apA, Arr apB, EbFlags fl, Arr qm, Arr qp,
             const double *dtp, HydroConst hc) {
    if (fl.covered(i, j, k)) return;
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    constexpr int ax = (T0 == 0), ay = (T0 == 1), az = (T0 == 2);
    constexpr int bx = (T1 == 0), by = (T1 == 1), bz = (T1 == 2);
    const double hdt = 0.5 * *dtp;  // Godunov_3D_eb.cpp:53-58
    const double ca = hdt / hc.dx[T0], cb = hdt / hc.dx[T1];
    const bool no_cov = apA(i + ax, j + ay, k + az, 0) > 0.0 && apB(i + bx, j + by, k + bz, 0) > 0.0 &&
                        apA(i, j, k, 0) > 0.0 && apB(i, j, k, 0) > 0.0;
    const bool lo_open = apD(i, j, k, 0) > 0.0;
    const bool hi_open = apD(i + ex, j + ey, k + ez, 0) > 0.0;
    double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
#pragma unroll
    for (int s = 0; s < NF; ++s) {
        fal[s] = no_cov ? fa(i, j, k, s) : 0.0;
        far_[s] = no_cov ? fa(i + ax, j + ay, k + az, s) : 0.0;
        fbl[s] = no_cov ? fb(i, j, k, s) : 0.0;
        fbr[s] = no_cov ? fb(i + bx, j + by, k + bz, s) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NG; ++s) {
        gal[s] = no_cov ? ga(i, j, k, s) : 0.0;
        gar[s] = no_cov ? ga(i + ax, j + ay, k + az, s) : 0.0;
        gbl[s] = no_cov ? gb(i, j, k, s) : 0.0;
        gbr[s] = no_cov ? gb(i + bx, j + by, k + bz, s) : 0.0;
    }
    const double gamc = qaux(i, j, k, QGAMC);
    double src[NS];
    if (has_src) {
#pragma unroll
        for (int s = 0; s < NS; ++s) src[s] = srcq(i, j, k, kQc(s));
    }
    double n[NS], o[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) n[s] = qnp(i, j, k, s);
    if (lo_open)
        transverse_two(n, fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, hdt,
                       has_src ? src : nullptr, gamc, hc.trd, hc.small_pres, o);
#pragma unroll
    for (int s = 0; s < NS; ++s) qp(i, j, k, s) = lo_open ? o[s] : n[s];
#pragma unroll
    for (int s = 0; s < NS; ++s) n[s] = qnm(i + ex, j + ey, k + ez, s);
    if (hi_open)
        transverse_two(n, fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, hdt,
                       has_src ? src : nullptr, gamc, hc.trd, hc.small_pres, o);
#pragma unroll
    for (int s = 0; s < NS; ++s) qm(i + ex, j + ey, k + ez, s) = hi_open ? o[s] : n[s];
}

// ------------------------------------------------------------- eb_divu
// Hydro_utils_eb_K.H:31-78 on nodes: zero unless the 8 cells around the
// node are all full.
CAMR_HD void
b_eb_divu(int i, int j, int k, Arr q, Arr vf, Arr divu, double dxi0, double dxi1, double dxi2) {
    const double sv = vf(i, j, k, 0) + vf(i - 1, j, k, 0) + vf(i, j - 1, k, 0) +
                      vf(i - 1, j - 1, k, 0) + vf(i, j, k - 1, 0) + vf(i - 1, j, k - 1, 0) +
                      vf(i, j - 1, k - 1, 0) + vf(i - 1, j - 1, k - 1, 0);
    double out = 0.0;
    if (sv >= 8.0 * (1.0 - 1.e-12)) {
#define Q(a, b, c, n) q(a, b, c, n)
        const double ux = 0.25 *
            (Q(i, j, k, QU) - Q(i - 1, j, k, QU) + Q(i, j, k - 1, QU) - Q(i - 1, j, k - 1, QU) +
             Q(i, j - 1, k, QU) - Q(i - 1, j - 1, k, QU) + Q(i, j - 1, k - 1, QU) -
             Q(i - 1, j - 1, k - 1, QU)) * dxi0;
        const double vy = 0.25 *
            (Q(i, j, k, QV) - Q(i, j - 1, k, QV) + Q(i, j, k - 1, QV) - Q(i, j - 1, k - 1, QV) +
             Q(i - 1, j, k, QV) - Q(i - 1, j - 1, k, QV) + Q(i - 1, j, k - 1, QV) -
             Q(i - 1, j - 1, k - 1, QV)) * dxi1;
        const double wz = 0.25 *
            (Q(i, j, k, QW) - Q(i, j, k - 1, QW) + Q(i - 1, j, k, QW) - Q(i - 1, j, k - 1, QW) +
             Q(i, j - 1, k, QW) - Q(i, j - 1, k - 1, QW) + Q(i - 1, j - 1, k, QW) -
             Q(i - 1, j - 1, k - 1, QW)) * dxi2;
#undef Q
        out = ux + vy + wz;
    }
    divu(i, j, k, 0) = out;
}

// ------------------------------------------------- adjust_fluxes_eb (face)
// hydro_artif_visc on every face; species normalisation and the regular
// area factor only where the aperture is open (Hydro_adjust_fluxes.cpp:
// 96-165, hydro_ext_flx_eb Hydro_utils_eb_K.H:12-27).
template <int D>
CAMR_HD void
b_adjust_eb(int i, int j, int k, Arr flx, Arr divu, Arr u, Arr ap, double area_c, HydroConst hc) {
    double f[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) f[n] = flx(i, j, k, n);
    adjust_face<D>(f, divu, u, i, j, k, area_c, hc, ap(i, j, k, 0) > 0.);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx(i, j, k, n) = f[n];
}

// compute_hyp_wallflux (Hydro_hypflux_eb_K.H:12-68): the flux through the
// embedded wall of cut cell (i,j,k), a Riemann problem of the cell state
// against its mirror across the wall normal from the apertures.
CAMR_HD void eb_wallflux(const Arr &q, const Arr &qaux, const EbGeom &g, int i,
                                            int j, int k, const HydroConst &hc, double *fw) {
    const double axm = g.ap[0](i, j, k, 0), axp = g.ap[0](i + 1, j, k, 0);
    const double aym = g.ap[1](i, j, k, 0), ayp = g.ap[1](i, j + 1, k, 0);
    const double azm = g.ap[2](i, j, k, 0), azp = g.ap[2](i, j, k + 1, 0);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) fw[n] = 0.;
    const double apnorm = sqrt((axm - axp) * (axm - axp) + (aym - ayp) * (aym - ayp) +
                               (azm - azp) * (azm - azp));
    const double apnorminv = 1. / apnorm;
    const double un = q(i, j, k, QU) * (axm - axp) * apnorminv +
                      q(i, j, k, QV) * (aym - ayp) * apnorminv +
                      q(i, j, k, QW) * (azm - azp) * apnorminv;
    const double rho = q(i, j, k, QRHO), p = q(i, j, k, QPRES), qrei = q(i, j, k, QREINT);
    const double gamc = qaux(i, j, k, QGAMC);
    RiemannOut o;
    riemann_two_shock(rho, un, 0., 0., p, qrei, rho, -un, 0., 0., p, qrei, qaux(i, j, k, QC),
                      gamc, gamc, hc.small, qaux(i, j, k, QCSML), hc.small_dens,
                      hc.small_pres, hc.smallu, o);
    fw[URHO] = o.f_rho;
    fw[UEDEN] = o.f_eden;
    fw[UEINT] = o.f_eint;
    fw[UMX] = (axm - axp) * o.f_un;
    fw[UMY] = (aym - ayp) * o.f_un;
    fw[UMZ] = (azm - azp) * o.f_un;
}

// -------------------------------------------- hydro_consup_eb (one cell)
// eb_compute_div (Hydro_utils_eb_K.H:283-575) for all NVAR components, the
// redistribution weight (eb_weights_type 2: rho), the face-centroid fluxes
// of the valid box, then eb_pdivu (:82-280).
template <bool CUT>
CAMR_HD void
b_eb_div(int i, int j, int k, Box bx, Arr ft0, Arr ft1, Arr ft2, Arr ge0, Arr ge1, Arr ge2, EbGeom g,
         Arr q, Arr qaux, Arr divc, Arr wt, Arr fo0, Arr fo1, Arr fo2, HydroConst hc,
         double dxi0, double dxi1, double dxi2, bool out = true) {
    const int32_t w = g.f(i, j, k);
    if (eb_type(w) == kEbCovered) return;
    // CUT: the cut cells only; otherwise the regular cells and the weight of
    // every cell (two launches: the cut-cell form's registers would hold
    // the regular cells' occupancy down)
    if (CUT != (eb_type(w) == kEbSingle)) return;
    // out: copy the face fluxes of the valid box into fo0..fo2
    const bool valid = out && bx.contains(i, j, k);
    const bool xh = i == bx.hi[0], yh = j == bx.hi[1], zh = k == bx.hi[2];
    const double cvi = dxi0 * dxi1 * dxi2;
    if constexpr (!CUT) {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) {
            const double dvv = cvi * (ft0(i + 1, j, k, n) - ft0(i, j, k, n)) +
                               cvi * (ft1(i, j + 1, k, n) - ft1(i, j, k, n)) +
                               cvi * (ft2(i, j, k + 1, n) - ft2(i, j, k, n));
            divc(i, j, k, n) = dvv * -1.0;
            if (valid) {
                fo0(i, j, k, n) = ft0(i, j, k, n);
                fo1(i, j, k, n) = ft1(i, j, k, n);
                fo2(i, j, k, n) = ft2(i, j, k, n);
                if (xh) fo0(i + 1, j, k, n) = ft0(i + 1, j, k, n);
                if (yh) fo1(i, j + 1, k, n) = ft1(i, j + 1, k, n);
                if (zh) fo2(i, j, k + 1, n) = ft2(i, j, k + 1, n);
            }
        }
        const double pdivu =
            0.5 * ((ge0(i + 1, j, k, GP) + ge0(i, j, k, GP)) * (ge0(i + 1, j, k, GUN) - ge0(i, j, k, GUN)) * dxi0 +
                   (ge1(i, j + 1, k, GP) + ge1(i, j, k, GP)) * (ge1(i, j + 1, k, GUN) - ge1(i, j, k, GUN)) * dxi1 +
                   (ge2(i, j, k + 1, GP) + ge2(i, j, k, GP)) * (ge2(i, j, k + 1, GUN) - ge2(i, j, k, GUN)) * dxi2);
        divc(i, j, k, UEINT) -= pdivu;
        wt(i, j, k, 0) = q(i, j, k, QRHO);
    } else {
        const double axm = g.ap[0](i, j, k, 0), axp = g.ap[0](i + 1, j, k, 0);
        const double aym = g.ap[1](i, j, k, 0), ayp = g.ap[1](i, j + 1, k, 0);
        const double azm = g.ap[2](i, j, k, 0), azp = g.ap[2](i, j, k + 1, 0);
        const double vfr = g.vf(i, j, k, 0);
        double fw[NVAR];
        eb_wallflux(q, qaux, g, i, j, k, hc, fw);
#pragma unroll 1
        for (int n = 0; n < NVAR; ++n) {
            const double fxm = centroid_value<0>(ft0, n, g, i, j, k);
            const double fxp = centroid_value<0>(ft0, n, g, i + 1, j, k);
            const double fym = centroid_value<1>(ft1, n, g, i, j, k);
            const double fyp = centroid_value<1>(ft1, n, g, i, j + 1, k);
            const double fzm = centroid_value<2>(ft2, n, g, i, j, k);
            const double fzp = centroid_value<2>(ft2, n, g, i, j, k + 1);
            if (valid) {
                fo0(i, j, k, n) = fxm;
                fo1(i, j, k, n) = fym;
                fo2(i, j, k, n) = fzm;
                if (xh) fo0(i + 1, j, k, n) = fxp;
                if (yh) fo1(i, j + 1, k, n) = fyp;
                if (zh) fo2(i, j, k + 1, n) = fzp;
            }
            double dvv = cvi / vfr * (axp * fxp - axm * fxm + ayp * fyp - aym * fym + azp * fzp - azm * fzm);
            dvv += fw[n] * dxi0 / vfr;
            divc(i, j, k, n) = dvv * -1.0;
        }
        const double fxm = axm > 0.0 ? centroid_value<0>(ge0, GUN, g, i, j, k) : 0.0;
        const double fxp = axp > 0.0 ? centroid_value<0>(ge0, GUN, g, i + 1, j, k) : 0.0;
        const double fym = aym > 0.0 ? centroid_value<1>(ge1, GUN, g, i, j, k) : 0.0;
        const double fyp = ayp > 0.0 ? centroid_value<1>(ge1, GUN, g, i, j + 1, k) : 0.0;
        const double fzm = azm > 0.0 ? centroid_value<2>(ge2, GUN, g, i, j, k) : 0.0;
        const double fzp = azp > 0.0 ? centroid_value<2>(ge2, GUN, g, i, j, k + 1) : 0.0;
        divc(i, j, k, UEINT) -= (1.0 / vfr) * q(i, j, k, QPRES) *
                                (dxi0 * (axp * fxp - axm * fxm) + dxi1 * (ayp * fyp - aym * fym) +
                                 dxi2 * (azp * fzp - azm * fzm));
        wt(i, j, k, 0) = q(i, j, k, QRHO);
    }
}

// ------------------------------------------------- flux redistribution
// Source side of AMReX's amrex_flux_redistribute (single level, no flux
// registers, use_wts_in_divnc false), over grow(bx, 1): the cut cell's
// own share optmp0 and the mass deficit delm it hands out, and 1 / (sum of
// its neighbours' weight * vfrac).
CAMR_HD void
b_redist_src(int i, int j, int k, EbGeom g, Box dbox2, Arr divc, Arr wt, Arr optmp0, Arr delm, Arr winv) {
    const int32_t w = g.f(i, j, k);
    if (eb_type(w) != kEbSingle) {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) { optmp0(i, j, k, n) = 0.0; delm(i, j, k, n) = 0.0; }
        winv(i, j, k, 0) = 0.0;
        return;
    }
    const double vfc = g.vf(i, j, k, 0);
    const double mask = dbox2.contains(i, j, k) ? 1.0 : 0.0;
    // connected neighbours (bit (ii+1) + 3 (jj+1) + 9 (kk+1), the centre
    // cleared), and those of them inside dbox2; visited in increasing bit
    // order, i.e. k, then j, then i, as the reference's loops
    uint32_t con = (uint32_t)w & ((1u << 27) - 1) & ~(1u << 13), in = 0;
    for (uint32_t m = con; m; m &= m - 1) {
        const int t = __builtin_ctz(m);
        if (dbox2.contains(i + t % 3 - 1, j + (t / 3) % 3 - 1, k + t / 9 - 1)) in |= 1u << t;
    }
    double vtot = 0.0;
    for (uint32_t m = in; m; m &= m - 1) {
        const int t = __builtin_ctz(m);
        vtot += g.vf(i + t % 3 - 1, j + (t / 3) % 3 - 1, k + t / 9 - 1, 0) * 1.0;
    }
#pragma unroll 1
    for (int n = 0; n < NVAR; ++n) {
        double divnc = 0.0;
        for (uint32_t m = in; m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const int ii = t % 3 - 1, jj = (t / 3) % 3 - 1, kk = t / 9 - 1;
            const double wf = g.vf(i + ii, j + jj, k + kk, 0) * 1.0;
            divnc += wf * divc(i + ii, j + jj, k + kk, n);
        }
        divnc /= vtot;
        const double op = (1 - vfc) * (divnc - divc(i, j, k, n)) * mask;
        optmp0(i, j, k, n) = op;
        delm(i, j, k, n) = -(vfc) * op;
    }
    double wtot = 0.0;
    for (uint32_t m = con; m; m &= m - 1) {
        const int t = __builtin_ctz(m);
        const int ii = t % 3 - 1, jj = (t / 3) % 3 - 1, kk = t / 9 - 1;
        wtot += wt(i + ii, j + jj, k + kk, 0) * g.vf(i + ii, j + jj, k + kk, 0) *
                (dbox2.contains(i + ii, j + jj, k + kk) ? 1.0 : 0.0);
    }
    winv(i, j, k, 0) = 1.0 / wtot;
}

// Receiving side over bx: the shares of the cut neighbours in the order of
// the reference's sequential scatter (source cells in increasing k, j, i),
// then dsdt = divc + optmp on non-covered cells.
CAMR_HD void
b_redist_dst(int i, int j, int k, EbGeom g, Box dbox2, Arr divc, Arr wt, Arr optmp0, Arr delm, Arr winv,
             Arr dsdt) {
    if (g.f.covered(i, j, k)) return;
    const double mask = dbox2.contains(i, j, k) ? 1.0 : 0.0;
    const double wtt = wt(i, j, k, 0);
    // the sending neighbours, bit (oi+1) + 3 (oj+1) + 9 (ok+1)
    uint32_t src = 0;
#pragma unroll
    for (int ok = -1; ok <= 1; ok++)
#pragma unroll
    for (int oj = -1; oj <= 1; oj++)
#pragma unroll
    for (int oi = -1; oi <= 1; oi++) {
        if (oi == 0 && oj == 0 && ok == 0) continue;
        const int32_t ws = g.f(i + oi, j + oj, k + ok);
        if (eb_type(ws) == kEbSingle && eb_conn(ws, -oi, -oj, -ok))
            src |= 1u << ((oi + 1) + 3 * (oj + 1) + 9 * (ok + 1));
    }
    // the cell's own share (zero unless it is a cut cell; b_redist_src
    // writes that zero when it runs over every cell, not over a cell list)
    const bool own = eb_type(g.f(i, j, k)) == kEbSingle;
#pragma unroll 1
    for (int n = 0; n < NVAR; ++n) {
        double acc = own ? optmp0(i, j, k, n) : 0.0;
        // in increasing bit order: k, then j, then i of the sender
        for (uint32_t m = src; m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const int ok = t / 9 - 1, oj = (t / 3) % 3 - 1, oi = t % 3 - 1;
            const double drho = delm(i + oi, j + oj, k + ok, n) * winv(i + oi, j + oj, k + ok, 0) * wtt * mask;
            acc += drho;
        }
        dsdt(i, j, k, n) = divc(i, j, k, n) + acc;
    }
}

CAMR_HD void
b_zero(int i, int j, int k, Arr a, int nc) {
    for (int n = 0; n < nc; ++n) a(i, j, k, n) = 0.0;
}

// ---------------------------------------------------------------- driver
// The stages of hydro_umdrv_eb in launch order: one launch site each, run
// over every box of a table (grid.y = box).  The transverse pairs D|T go in
// the order of Godunov_3D_eb.cpp:214-455, (0|1, 0|2, 1|0, 1|2, 2|0, 2|1),
// each as transdo then Riemann; the final states per direction as transdd
// then Riemann.
enum EbSite {
    ES_ZERO = 0,                    // the accumulated arrays, all in one launch
    ES_PLM,
    ES_R1,                          // + d (3)
    ES_PAIR = ES_R1 + 3,            // + 2 p (transdo), + 2 p + 1 (Riemann), p < 6
    ES_FINAL = ES_PAIR + 12,        // + 2 d (transdd), + 2 d + 1 (Riemann)
    ES_DIVU = ES_FINAL + 6,
    ES_ADJ,                         // + d (3)
    ES_DIV = ES_ADJ + 3,
    ES_DIVCUT,
    ES_RSRC,
    ES_RDST,
    ES_COUNT
};
template <int S>
__host__ __device__ inline Box site_box(const EbBox &B, const Layout &L) {
    // ES_ZERO: the nodes of grow(bx, 5), which hold ft, qe, nb and btf
    if constexpr (S == ES_ZERO) return nodes(nodes(nodes(grow(B.bx, 5), 0), 1), 2);
    else if constexpr (S == ES_PLM) return L.bxg2;
    else if constexpr (S >= ES_R1 && S < ES_R1 + 3) return L.f1[S - ES_R1];
    else if constexpr (S >= ES_PAIR && S < ES_PAIR + 12) {
        constexpr int p = (S - ES_PAIR) / 2;
        if constexpr ((S - ES_PAIR) % 2 == 0) return grow_dir(L.bxg2, pair_t(p), -1);
        else return grow_dir(nodes(L.bxg1, pair_d(p)), pair_o(p), 1);
    } else if constexpr (S >= ES_FINAL && S < ES_FINAL + 6) {
        constexpr int d = (S - ES_FINAL) / 2;
        if constexpr ((S - ES_FINAL) % 2 == 0) return L.bxg1;
        else return grow_dir(grow_dir(nodes(L.btf, d), t0_of(d), 1), t1_of(d), 1);
    } else if constexpr (S == ES_DIVU) return B.nb;
    else if constexpr (S >= ES_ADJ && S < ES_ADJ + 3) return L.ft[S - ES_ADJ];
    else if constexpr (S == ES_DIV || S == ES_DIVCUT) return L.btf;
    else if constexpr (S == ES_RSRC) return L.g1;
    else return B.bx;
}

// The stage's per-point work with box B's arrays (the arguments of the
// former per-box launches, unchanged).
template <int S>
CAMR_HD void site_body(const EbBox &B, const Layout &L, int i, int j, int k,
                                          const HydroConst &hc, const double *dtp, int has_src) {
    if constexpr (S == ES_ZERO) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (L.ft[d].contains(i, j, k)) b_zero(i, j, k, B.FT[d], (int)NVAR);
            if (L.qe[d].contains(i, j, k)) b_zero(i, j, k, B.QE[d], (int)NG);
        }
        if (L.nb.contains(i, j, k)) b_zero(i, j, k, B.divu, 1);
        // (a regular box's divergence stage sets every divc and wt)
        if (!B.regular && L.btf.contains(i, j, k)) {
            b_zero(i, j, k, B.divc, (int)NVAR);
            b_zero(i, j, k, B.wt, 1);
        }
    } else if constexpr (S == ES_PLM) {
        b_plm_eb(i, j, k, B.q, B.qbox, B.qaux, B.g, B.QM[0], B.QP[0], B.QM[1], B.QP[1], B.QM[2],
                 B.QP[2], hc, dtp);
    } else if constexpr (S >= ES_R1 && S < ES_R1 + 3) {
        constexpr int d = S - ES_R1;
        b_riemann_eb<d, 0>(i, j, k, B.QM[d], B.QP[d], B.qaux, B.F1[d], B.G1[d], B.g.ap[d], hc);
    } else if constexpr (S >= ES_PAIR && S < ES_PAIR + 12) {
        constexpr int p = (S - ES_PAIR) / 2, D = pair_d(p), T = pair_t(p);
        if constexpr ((S - ES_PAIR) % 2 == 0)
            b_transdo_eb<D, T>(i, j, k, B.QM[D], B.QP[D], B.qaux, B.F1[T], B.G1[T], B.g.ap[D],
                               B.g.ap[T], B.g.f, B.qmT[p], B.qpT[p], dtp, hc);
        else
            b_riemann_eb<D, 1>(i, j, k, B.qmT[p], B.qpT[p], B.qaux, B.F2[D][T], B.G2[D][T],
                               B.g.ap[D], hc);
    } else if constexpr (S >= ES_FINAL && S < ES_FINAL + 6) {
        constexpr int D = (S - ES_FINAL) / 2, T0 = t0_of(D), T1 = t1_of(D);
        if constexpr ((S - ES_FINAL) % 2 == 0)
            b_transdd_eb<D>(i, j, k, B.QM[D], B.QP[D], B.qaux, B.srcq, has_src, B.F2[T0][T1],
                            B.G2[T0][T1], B.F2[T1][T0], B.G2[T1][T0], B.g.ap[D], B.g.ap[T0],
                            B.g.ap[T1], B.g.f, B.qmF[D], B.qpF[D], dtp, hc);
        else
            b_riemann_eb<D, 2>(i, j, k, B.qmF[D], B.qpF[D], B.qaux, B.FT[D], B.QE[D], B.g.ap[D],
                               hc);
    } else if constexpr (S == ES_DIVU) {
        b_eb_divu(i, j, k, B.q, B.g.vf, B.divu, 1.0 / hc.dx[0], 1.0 / hc.dx[1], 1.0 / hc.dx[2]);
    } else if constexpr (S >= ES_ADJ && S < ES_ADJ + 3) {
        constexpr int D = S - ES_ADJ;
        const double area = D == 0 ? hc.dx[1] * hc.dx[2]
                          : D == 1 ? hc.dx[0] * hc.dx[2] : hc.dx[1] * hc.dx[0];
        b_adjust_eb<D>(i, j, k, B.FT[D], B.divu, B.uin, B.g.ap[D], area, hc);
    } else if constexpr (S == ES_DIV || S == ES_DIVCUT) {
        if (S == ES_DIVCUT && B.regular) return;  // no cut cell
        b_eb_div<S == ES_DIVCUT>(i, j, k, B.bx, B.FT[0], B.FT[1], B.FT[2], B.QE[0], B.QE[1], B.QE[2], B.g, B.q,
                 B.qaux, B.divc, B.wt, B.flx[0], B.flx[1], B.flx[2], hc, 1.0 / hc.dx[0],
                 1.0 / hc.dx[1], 1.0 / hc.dx[2], !B.noflx);
    } else if constexpr (S == ES_RSRC) {
        // a regular box sends nothing (the receiving side below does not
        // read optmp0 / delm / winv there)
        if (B.regular) return;
        b_redist_src(i, j, k, B.g, B.dbox2, B.divc, B.wt, B.optmp0, B.delm, B.winv);
    } else if (B.regular) {
        // b_redist_dst with optmp0 = 0 and no sending neighbour
#pragma unroll
        for (int n = 0; n < NVAR; ++n) B.dsdt(i, j, k, n) = B.divc(i, j, k, n) + 0.0;
    } else {
        b_redist_dst(i, j, k, B.g, B.dbox2, B.divc, B.wt, B.optmp0, B.delm, B.winv, B.dsdt);
    }
}

template <int S>
__global__ void __launch_bounds__(kBlock)
k_eb_site(const EbBox *tab, HydroConst hc, const double *dtp, int has_src, FlatBlocks fb) {
    int u, lb;
    flat_entry(fb, u, lb);
    const EbBox &B = tab[u];
    if constexpr (S == ES_DIVCUT || S == ES_RSRC)
        if (B.regular) return;
    const Layout L = layout_of(B.bx, B.qbox);
    const Box sb = site_box<S>(B, L);
    if (!sb.ok()) return;
    int i, j, k;
    if (!range_point(range_of(sb), (int64_t)lb * blockDim.x + threadIdx.x, i, j, k))
        return;
    site_body<S>(B, L, i, j, k, hc, dtp, has_src);
}

constexpr const char *site_name(int s) {
    return s < ES_PLM ? "eb_zero"
         : s == ES_PLM ? "eb_trace_plm"
         : s < ES_PAIR ? "eb_riemann_1"
         : s < ES_FINAL ? ((s - ES_PAIR) % 2 ? "eb_riemann_2" : "eb_transverse_1")
         : s < ES_DIVU ? ((s - ES_FINAL) % 2 ? "eb_riemann_3" : "eb_transverse_2")
         : s == ES_DIVU ? "eb_divu"
         : s < ES_DIV ? "eb_adjust"
         : s <= ES_DIVCUT ? "eb_divergence" : "eb_redistribute";
}

// The cut-cell divergence over a list of the cut cells.
__global__ void __launch_bounds__(kBlock)
k_eb_divcut_list(const EbBox *tab, const CutCell *cells, int n, HydroConst hc, const double *dtp,
                 int has_src) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const CutCell c = cells[t];
    const EbBox &B = tab[c.e];
    const Layout L = layout_of(B.bx, B.qbox);
    site_body<ES_DIVCUT>(B, L, c.i, c.j, c.k, hc, dtp, has_src);
}

// The redistribution's source side over the cut cells of grow(bx, 1) in
// the list (the other cells send nothing).
__global__ void __launch_bounds__(kBlock)
k_eb_rsrc_list(const EbBox *tab, const CutCell *cells, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const CutCell c = cells[t];
    const EbBox &B = tab[c.e];
    if (!grow(B.bx, 1).contains(c.i, c.j, c.k)) return;
    b_redist_src(c.i, c.j, c.k, B.g, B.dbox2, B.divc, B.wt, B.optmp0, B.delm, B.winv);
}

template <int S>
void launch_site(const EbBox *tab, const EbBox *htab, int n, const HydroConst &hc,
                 const double *dtp, int has_src, hipStream_t st, Prof *prof,
                 const CutCell *cells, int ncells) {
    if ((S == ES_DIVCUT || S == ES_RSRC) && cells) {
        if (ncells == 0) return;
        Region reg(prof, site_name(S), st);
        if (S == ES_DIVCUT)
            hipLaunchKernelGGL(k_eb_divcut_list, dim3(grid_for(ncells)), dim3(kBlock), 0, st, tab,
                               cells, ncells, hc, dtp, has_src);
        else
            hipLaunchKernelGGL(k_eb_rsrc_list, dim3(grid_for(ncells)), dim3(kBlock), 0, st, tab,
                               cells, ncells);
        return;
    }
    // blocks per box: none where the stage has no point or nothing to do
    auto cnt = [&](int b) -> int64_t {
        if ((S == ES_DIVCUT || S == ES_RSRC) && htab[b].regular) return 0;
        const Box sb = site_box<S>(htab[b], layout_of(htab[b].bx, htab[b].qbox));
        return sb.ok() ? (sb.npts() + kBlock - 1) / kBlock : 0;
    };
    FlatBlocks fb;
    const dim3 grid = flat_blocks(n, cnt, fb);
    if (grid.x == 0) return;
    Region reg(prof, site_name(S), st);
    hipLaunchKernelGGL(k_eb_site<S>, grid, dim3(kBlock), 0, st, tab, hc, dtp, has_src, fb);
}

template <int... S>
void launch_sites(std::integer_sequence<int, S...>, const EbBox *tab, const EbBox *htab, int n,
                  const HydroConst &hc, const double *dtp, int has_src, hipStream_t st,
                  Prof *prof, const CutCell *cells, int ncells) {
    (launch_site<S>(tab, htab, n, hc, dtp, has_src, st, prof, cells, ncells), ...);
}

// The EbBox of a one-box call, written into device memory by kernel
// arguments (no host buffer whose lifetime a captured copy would outlive).
struct Words {
    uint64_t w[240];
};
__global__ void k_put_words(uint64_t *dst, Words w, int n) {
    for (int t = (int)threadIdx.x; t < n; t += (int)blockDim.x) dst[t] = w.w[t];
}

constexpr size_t kEbHead = (sizeof(EbBox) + 255) / 256 * 32;  // doubles, 256-byte multiple

} // namespace

size_t umdrv_eb_scratch_doubles(const Box &bx, const Box &qbox, bool keep) {
    const Layout L = layout_of(bx, qbox);
    size_t n = 0;
    for (int d = 0; d < 3; ++d) {
        n += (size_t)NS * (L.qm[d].npts() + L.bxg2.npts());          // QM, QP
        n += (size_t)(NF + NG) * L.f1[d].npts();                      // F1, G1
        n += (size_t)2 * (NF + NG) * L.f2[d].npts();                  // F2, G2 pairs
        n += (size_t)NVAR * L.ft[d].npts() + (size_t)NG * L.qe[d].npts();
    }
    size_t t = 0;
    for (int d = 0; d < 3; ++d) t = std::max<size_t>(t, L.mbx[d].npts());
    n += (size_t)(keep ? 6 : 1) * 2 * NS * t;                         // qmT, qpT
    n += (size_t)(keep ? 3 : 1) * NS * (L.bxg2.npts() + L.bxg1.npts());  // final qm, qp
    n += L.nb.npts();                                                 // divu
    n += (size_t)(NVAR + 1) * L.btf.npts();                           // divc, wt
    n += (size_t)(2 * NVAR + 1) * L.g1.npts();                        // optmp0, delm, winv
    return n + 64 + kEbHead;
}

void eb_box_setup(const UmdrvEbArgs &A, double *scratch, EbBox *out, bool keep) {
    const Layout L = layout_of(A.bx, A.qbox);
    EbBox B;
    memset(&B, 0, sizeof(B));
    B.bx = A.bx;
    B.qbox = A.qbox;
    B.g = A.g;
    B.regular = A.regular;
    B.noflx = A.noflx;
    B.q = A.q;
    B.qaux = A.qaux;
    B.srcq = A.srcq;
    B.uin = A.uin;
    B.dsdt = A.dsdt;
    for (int d = 0; d < 3; ++d) B.flx[d] = A.flx[d];
    double *cur = scratch;
    auto take = [&](const Box &b, int nc) {
        Arr a = dense(cur, b);
        cur += (size_t)nc * b.npts();
        return a;
    };
    for (int d = 0; d < 3; ++d) {
        B.QM[d] = take(L.qm[d], NS);
        B.QP[d] = take(L.bxg2, NS);
        B.F1[d] = take(L.f1[d], NF);
        B.G1[d] = take(L.f1[d], NG);
        for (int t = 0; t < 3; ++t)
            if (t != d) { B.F2[d][t] = take(L.f2[d], NF); B.G2[d][t] = take(L.f2[d], NG); }
        B.FT[d] = take(L.ft[d], NVAR);
        B.QE[d] = take(L.qe[d], NG);
    }
    size_t tmax = 0;
    for (int d = 0; d < 3; ++d) tmax = std::max<size_t>(tmax, L.mbx[d].npts());
    // the pair states of pair p: the layout of mbx[D] (D the pair's normal
    // direction) in one buffer for all pairs, or one buffer each
    for (int p = 0; p < 6; ++p) {
        if (p == 0 || keep) {
            double *tq0 = cur; cur += (size_t)NS * tmax;
            double *tq1 = cur; cur += (size_t)NS * tmax;
            B.qmT[p] = dense(tq0, L.mbx[pair_d(p)]);
            B.qpT[p] = dense(tq1, L.mbx[pair_d(p)]);
        } else {
            B.qmT[p] = dense(B.qmT[0].p, L.mbx[pair_d(p)]);
            B.qpT[p] = dense(B.qpT[0].p, L.mbx[pair_d(p)]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (d == 0 || keep) {
            B.qmF[d] = take(L.bxg2, NS);
            B.qpF[d] = take(L.bxg1, NS);
        } else {
            B.qmF[d] = B.qmF[0];
            B.qpF[d] = B.qpF[0];
        }
    }
    B.divu = take(L.nb, 1);
    B.divc = take(L.btf, NVAR);
    B.wt = take(L.btf, 1);
    B.optmp0 = take(L.g1, NVAR);
    B.delm = take(L.g1, NVAR);
    B.winv = take(L.g1, 1);
    // adjust_fluxes_eb: divu on the nodes of grow(qbox, -1) that lie in the
    // domain grown by 16 in periodic directions (Hydro_adjust_fluxes.cpp:73-84)
    B.nb = L.nb;
    for (int d = 0; d < 3; ++d) {
        const int gp = A.periodic[d] ? 16 : 0;
        B.nb.lo[d] = std::max(B.nb.lo[d], A.hc.domlo[d] - gp);
        B.nb.hi[d] = std::min(B.nb.hi[d], A.hc.domhi[d] + 1 + gp);
    }
    for (int d = 0; d < 3; ++d) {
        const int gp = A.periodic[d] ? 2 : 0;
        B.dbox2.lo[d] = A.hc.domlo[d] - gp;
        B.dbox2.hi[d] = A.hc.domhi[d] + gp;
    }
    *out = B;
}

// hydro_ctoprim over qbox and the zero fill of dS/dt over bx
// (CAMR_construct_hydro_source.cpp:108-120, the hyd_src setVal), every box
// of the table in one launch.
__global__ void __launch_bounds__(kBlock)
k_eb_prim(const EbBox *tab, HydroConst hc, int *status, FlatBlocks fb) {
    int u, lb;
    flat_entry(fb, u, lb);
    const EbBox &B = tab[u];
    int i, j, k;
    if (!range_point(range_of(B.qbox), (int64_t)lb * blockDim.x + threadIdx.x, i, j, k))
        return;
    pt_ctoprim(B.uin, B.q, B.qaux, hc, status, i, j, k);
    if (B.bx.contains(i, j, k))
        for (int n = 0; n < NVAR; ++n) B.dsdt(i, j, k, n) = 0.0;
}

int eb_prim_batch(const EbBox *tab, const EbBox *htab, int n, const HydroConst &hc, int *status,
                  hipStream_t st, Prof *prof) {
    if (n <= 0) return CAMR_OK;
    FlatBlocks fb;
    const dim3 grid =
        flat_blocks(n, [&](int b) -> int64_t { return grid_for(htab[b].qbox.npts()); }, fb);
    Region reg(prof, "eb_ctoprim", st);
    hipLaunchKernelGGL(k_eb_prim, grid, dim3(kBlock), 0, st, tab, hc, status, fb);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

int umdrv_eb_batch(const EbBox *tab, const EbBox *htab, int n, const HydroConst &hc,
                   const double *dtp, int has_src, hipStream_t st, Prof *prof,
                   const CutCell *cells, int ncells) {
    if (n <= 0) return CAMR_OK;
    launch_sites(std::make_integer_sequence<int, ES_COUNT>{}, tab, htab, n, hc, dtp, has_src, st,
                 prof, cells, ncells);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

int umdrv_eb(const UmdrvEbArgs &A, double *scratch, hipStream_t st, Prof *prof) {
    EbBox hb;
    eb_box_setup(A, scratch + kEbHead, &hb);
    static_assert(sizeof(EbBox) % 8 == 0, "EbBox words");
    const int nw = (int)(sizeof(EbBox) / 8);
    const uint64_t *src = reinterpret_cast<const uint64_t *>(&hb);
    for (int w0 = 0; w0 < nw; w0 += 240) {
        Words w;
        const int m = std::min(240, nw - w0);
        memcpy(w.w, src + w0, (size_t)m * 8);
        hipLaunchKernelGGL(k_put_words, dim3(1), dim3(256), 0, st,
                           reinterpret_cast<uint64_t *>(scratch) + w0, w, m);
    }
    return umdrv_eb_batch(reinterpret_cast<const EbBox *>(scratch), &hb, 1, A.hc, A.dtp,
                          A.has_src, st, prof);
}

// The force on the embedded body from one box's cut cells: per valid cell
// the wall momentum flux times dx^2 (fluxw[UMX..UMZ] of
// compute_hyp_wallflux, Hydro_hypflux_eb_K.H:65), 0 elsewhere; the caller
// sums the cells in a fixed order.
__global__ void __launch_bounds__(kBlock)
k_eb_force(Range r, Arr q, Arr qaux, EbGeom g, HydroConst hc, Arr out, double dx2) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    double f[3] = {0.0, 0.0, 0.0};
    if (eb_type(g.f(i, j, k)) == kEbSingle) {
        double fw[NVAR];
        eb_wallflux(q, qaux, g, i, j, k, hc, fw);
        f[0] = fw[UMX] * dx2;
        f[1] = fw[UMY] * dx2;
        f[2] = fw[UMZ] * dx2;
    }
    out(i, j, k, 0) = f[0];
    out(i, j, k, 1) = f[1];
    out(i, j, k, 2) = f[2];
}

int eb_force_box(const Box &bx, Arr q, Arr qaux, const EbGeom &g, const HydroConst &hc, Arr out,
                 double dx2, hipStream_t st) {
    Range r = range_of(bx);
    hipLaunchKernelGGL(k_eb_force, dim3(grid_for(r.n)), dim3(kBlock), 0, st, r, q, qaux, g, hc,
                       out, dx2);
    return hipGetLastError() == hipSuccess ? 0 : CAMR_ERR_HIP;
}

} // namespace camr
