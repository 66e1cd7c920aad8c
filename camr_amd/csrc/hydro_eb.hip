// hydro_eb.hip -- gfx950 kernels of hydro_umdrv_eb, the cut-cell box path
// (Source/Hydro/Hydro_umdrv_eb.cpp:14-191): Godunov_umeth_eb
// (Godunov/Godunov_3D_eb.cpp:16-567) with the aperture guards of the PLM
// trace, the transverse predictors and the Riemann solves; adjust_fluxes_eb
// (Hydro_adjust_fluxes.cpp:48-170); hydro_consup_eb (Hydro_consup.cpp:
// 29-98: the face-centroid divergence, the wall flux and p div u); and
// AMReX's single-level flux redistribution.
//
// One kernel per stage over the points of one box, intermediates in HBM
// scratch in the compact layouts of the staged engine.  The per-point
// arithmetic is hydro_math.h's, so a cut-cell box reproduces the regular
// path's operation order wherever every face is open.  The redistribution
// scatter of the reference is a gather here: each cell adds the shares of
// its cut neighbours in the order the sequential scatter would, so the sum
// is bit-identical and needs no atomics.
#include "hydro_eb.h"

namespace camr {

namespace {

__device__ __constant__ int kQc[NS] = {QRHO, QU, QV, QW, QPRES, QREINT, QFS, QFS + 1, QFS + 2};
__device__ __constant__ int kUc[NF] = {URHO, UMX, UMY, UMZ, UEDEN, UEINT, UFS, UFS + 1, UFS + 2};

__device__ __forceinline__ int dv(int d, int e) { return d == e ? 1 : 0; }

// ---------------------------------------------- PLM trace (cut-cell form)
// Godunov_3D_eb.cpp:94-152: flatten_eb (flatten.H:104-134), plm_slope_eb
// (PLM_eb.H:18-54), hydro_plm_d_eb (PLM_eb.H:103-287); covered cells skip.
__global__ void __launch_bounds__(kBlock)
k_plm_eb(Range r, Arr q, Box qbox, Arr qaux, EbGeom g, Arr qm0, Arr qp0, Arr qm1,
         Arr qp1, Arr qm2, Arr qp2, HydroConst hc, const double *dtp) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    if (eb_type(g.f(i, j, k)) == kEbCovered) return;
    double fl = 1.0;
    if (hc.use_flattening == 1) {
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
    for (int s = 0; s < NS; ++s) c[s] = q(i, j, k, kQc[s]);
    const double cc = qaux(i, j, k, QC);
    Arr QM[3] = {qm0, qm1, qm2};
    Arr QP[3] = {qp0, qp1, qp2};
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        const int ex = dv(d, 0), ey = dv(d, 1), ez = dv(d, 2);
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
            for (int o = -2; o <= 2; ++o) st[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, kQc[s]);
            sl[s] = ord ? mc_slope(st, fl, ord) : 0.0;
        }
        double qp[NS], qm[NS];
        plm_trace(d, c, sl, cc, *dtp / hc.dx[d], hc.small_dens, hc.small_pres, qp, qm);
        const bool open_lo = g.ap[d](i, j, k, 0) > 0.0;
        const bool open_hi = g.ap[d](i + ex, j + ey, k + ez, 0) > 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            QP[d](i, j, k, s) = open_lo ? qp[s] : 0.0;
            QM[d](i + ex, j + ey, k + ez, s) = open_hi ? qm[s] : 0.0;
        }
    }
}

// ------------------------------------------------ Riemann solve, open faces
// hydro_cmpflx where the aperture is > 0.  MODE 0: first solve (compact out,
// boundary-mutated states written back); 1: transverse solve (compact out);
// 2: final solve into the NVAR flux array.
template <int D, int MODE>
__global__ void __launch_bounds__(kBlock)
k_riemann_eb(Range r, Arr qL, Arr qR, Arr qaux, Arr fout, Arr gout, Arr ap, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
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
        for (int s = 0; s < NF; ++s) fout(i, j, k, kUc[s]) = f[s];
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
__global__ void __launch_bounds__(kBlock)
k_transdo_eb(Range r, Arr qnm, Arr qnp, Arr qaux, Arr F, Arr G, Arr apD, Arr apT,
             EbFlags fl, Arr qmo, Arr qpo, const double *dtp, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
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
__global__ void __launch_bounds__(kBlock)
k_transdd_eb(Range r, Arr qnm, Arr qnp, Arr qaux, Arr srcq, int has_src, Arr fa, Arr ga,
             Arr fb, Arr gb, Arr apD, Arr apA, Arr apB, EbFlags fl, Arr qm, Arr qp,
             const double *dtp, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
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
        for (int s = 0; s < NS; ++s) src[s] = srcq(i, j, k, kQc[s]);
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
__global__ void __launch_bounds__(kBlock)
k_eb_divu(Range r, Arr q, Arr vf, Arr divu, double dxi0, double dxi1, double dxi2) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
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
__global__ void __launch_bounds__(kBlock)
k_adjust_eb(Range r, Arr flx, Arr divu, Arr u, Arr ap, double area_c, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    double f[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) f[n] = flx(i, j, k, n);
    adjust_face<D>(f, divu, u, i, j, k, area_c, hc, ap(i, j, k, 0) > 0.);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx(i, j, k, n) = f[n];
}

// Value of face array `u` (component n) of direction d at face (i,j,k),
// moved to the face centroid (eb_compute_div, Hydro_utils_eb_K.H:410-520).
template <int D>
__device__ __forceinline__ double centroid_value(const Arr &u, int n, const EbGeom &g, int i,
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

// compute_hyp_wallflux (Hydro_hypflux_eb_K.H:12-68): the flux through the
// embedded wall of cut cell (i,j,k), a Riemann problem of the cell state
// against its mirror across the wall normal from the apertures.
__device__ __forceinline__ void eb_wallflux(const Arr &q, const Arr &qaux, const EbGeom &g, int i,
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
__global__ void __launch_bounds__(kBlock)
k_eb_div(Range r, Box bx, Arr ft0, Arr ft1, Arr ft2, Arr ge0, Arr ge1, Arr ge2, EbGeom g,
         Arr q, Arr qaux, Arr divc, Arr wt, Arr fo0, Arr fo1, Arr fo2, HydroConst hc,
         double dxi0, double dxi1, double dxi2) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int32_t w = g.f(i, j, k);
    if (eb_type(w) == kEbCovered) return;
    const bool valid = bx.contains(i, j, k);
    const bool xh = i == bx.hi[0], yh = j == bx.hi[1], zh = k == bx.hi[2];
    const double cvi = dxi0 * dxi1 * dxi2;
    if (eb_type(w) == kEbRegular) {
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
    }
    wt(i, j, k, 0) = q(i, j, k, QRHO);
}

// ------------------------------------------------- flux redistribution
// Source side of AMReX's amrex_flux_redistribute (single level, no flux
// registers, use_wts_in_divnc false), over grow(bx, 1): the cut cell's
// own share optmp0 and the mass deficit delm it hands out, and 1 / (sum of
// its neighbours' weight * vfrac).
__global__ void __launch_bounds__(kBlock)
k_redist_src(Range r, EbGeom g, Box dbox2, Arr divc, Arr wt, Arr optmp0, Arr delm, Arr winv) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int32_t w = g.f(i, j, k);
    if (eb_type(w) != kEbSingle) {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) { optmp0(i, j, k, n) = 0.0; delm(i, j, k, n) = 0.0; }
        winv(i, j, k, 0) = 0.0;
        return;
    }
    const double vfc = g.vf(i, j, k, 0);
    const double mask = dbox2.contains(i, j, k) ? 1.0 : 0.0;
#pragma unroll 1
    for (int n = 0; n < NVAR; ++n) {
        double vtot = 0.0, divnc = 0.0;
        for (int kk = -1; kk <= 1; kk++)
        for (int jj = -1; jj <= 1; jj++)
        for (int ii = -1; ii <= 1; ii++)
            if ((ii != 0 || jj != 0 || kk != 0) && eb_conn(w, ii, jj, kk) &&
                dbox2.contains(i + ii, j + jj, k + kk)) {
                const double wf = g.vf(i + ii, j + jj, k + kk, 0) * 1.0;
                vtot += wf;
                divnc += wf * divc(i + ii, j + jj, k + kk, n);
            }
        divnc /= vtot;
        const double op = (1 - vfc) * (divnc - divc(i, j, k, n)) * mask;
        optmp0(i, j, k, n) = op;
        delm(i, j, k, n) = -(vfc) * op;
    }
    double wtot = 0.0;
    for (int kk = -1; kk <= 1; kk++)
    for (int jj = -1; jj <= 1; jj++)
    for (int ii = -1; ii <= 1; ii++)
        if ((ii != 0 || jj != 0 || kk != 0) && eb_conn(w, ii, jj, kk))
            wtot += wt(i + ii, j + jj, k + kk, 0) * g.vf(i + ii, j + jj, k + kk, 0) *
                    (dbox2.contains(i + ii, j + jj, k + kk) ? 1.0 : 0.0);
    winv(i, j, k, 0) = 1.0 / wtot;
}

// Receiving side over bx: the shares of the cut neighbours in the order of
// the reference's sequential scatter (source cells in increasing k, j, i),
// then dsdt = divc + optmp on non-covered cells.
__global__ void __launch_bounds__(kBlock)
k_redist_dst(Range r, EbGeom g, Box dbox2, Arr divc, Arr wt, Arr optmp0, Arr delm, Arr winv,
             Arr dsdt) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    if (g.f.covered(i, j, k)) return;
    const double mask = dbox2.contains(i, j, k) ? 1.0 : 0.0;
    const double wtt = wt(i, j, k, 0);
#pragma unroll 1
    for (int n = 0; n < NVAR; ++n) {
        double acc = optmp0(i, j, k, n);
        for (int ok = -1; ok <= 1; ok++)
        for (int oj = -1; oj <= 1; oj++)
        for (int oi = -1; oi <= 1; oi++) {
            if (oi == 0 && oj == 0 && ok == 0) continue;
            const int32_t ws = g.f(i + oi, j + oj, k + ok);
            if (eb_type(ws) == kEbSingle && eb_conn(ws, -oi, -oj, -ok)) {
                const double drho = delm(i + oi, j + oj, k + ok, n) * winv(i + oi, j + oj, k + ok, 0) * wtt * mask;
                acc += drho;
            }
        }
        dsdt(i, j, k, n) = divc(i, j, k, n) + acc;
    }
}

__global__ void __launch_bounds__(kBlock)
k_zero(Range r, Arr a, int nc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    for (int n = 0; n < nc; ++n) a(i, j, k, n) = 0.0;
}

Prof *g_prof = nullptr;
const char *g_name = "";

template <typename K, typename... A>
void launch(K kern, const Box &b, hipStream_t s, A... args) {
    Range r = range_of(b);
    Region reg(g_prof, g_name, s);
    hipLaunchKernelGGL(kern, dim3(grid_for(r.n)), dim3(kBlock), 0, s, r, args...);
}

struct Layout {
    Box bx, btf, bxg1, bxg2, g1;
    Box qm[3], f1[3], f2[3], ft[3], qe[3], nb, mbx[3];
};

Layout layout_of(const Box &bx, const Box &qbox) {
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
        L.nb.lo[d] = std::max(a.lo[d], b.lo[d]);
        L.nb.hi[d] = std::min(a.hi[d], b.hi[d]);
    }
    return L;
}

} // namespace

size_t umdrv_eb_scratch_doubles(const Box &bx, const Box &qbox) {
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
    n += (size_t)2 * NS * t;                                          // qmT, qpT
    n += (size_t)NS * (L.bxg2.npts() + L.bxg1.npts());               // final qm, qp
    n += L.nb.npts();                                                 // divu
    n += (size_t)(NVAR + 1) * L.btf.npts();                           // divc, wt
    n += (size_t)(2 * NVAR + 1) * L.g1.npts();                        // optmp0, delm, winv
    return n + 64;
}

int umdrv_eb(const UmdrvEbArgs &A, double *scratch, hipStream_t st, Prof *prof) {
    g_prof = prof;
    const HydroConst &hc = A.hc;
    const Layout L = layout_of(A.bx, A.qbox);
    const EbGeom &g = A.g;
    double *cur = scratch;
    auto take = [&](const Box &b, int nc) {
        Arr a = dense(cur, b);
        cur += (size_t)nc * b.npts();
        return a;
    };
    const double dxi[3] = {1.0 / hc.dx[0], 1.0 / hc.dx[1], 1.0 / hc.dx[2]};

    Arr QM[3], QP[3], F1[3], G1[3], F2[3][3], G2[3][3], FT[3], QE[3];
    for (int d = 0; d < 3; ++d) {
        QM[d] = take(L.qm[d], NS);
        QP[d] = take(L.bxg2, NS);
        F1[d] = take(L.f1[d], NF);
        G1[d] = take(L.f1[d], NG);
        for (int t = 0; t < 3; ++t)
            if (t != d) { F2[d][t] = take(L.f2[d], NF); G2[d][t] = take(L.f2[d], NG); }
        FT[d] = take(L.ft[d], NVAR);
        QE[d] = take(L.qe[d], NG);
    }
    size_t tmax = 0;
    for (int d = 0; d < 3; ++d) tmax = std::max<size_t>(tmax, L.mbx[d].npts());
    double *tq0 = cur; cur += (size_t)NS * tmax;
    double *tq1 = cur; cur += (size_t)NS * tmax;
    Arr qmF = take(L.bxg2, NS), qpF = take(L.bxg1, NS);
    Arr divu = take(L.nb, 1);
    Arr divc = take(L.btf, NVAR), wt = take(L.btf, 1);
    Arr optmp0 = take(L.g1, NVAR), delm = take(L.g1, NVAR), winv = take(L.g1, 1);

    g_name = "eb_zero";
    for (int d = 0; d < 3; ++d) {
        launch(k_zero, L.ft[d], st, FT[d], (int)NVAR);
        launch(k_zero, L.qe[d], st, QE[d], (int)NG);
    }
    launch(k_zero, L.nb, st, divu, 1);
    launch(k_zero, L.btf, st, divc, (int)NVAR);
    launch(k_zero, L.btf, st, wt, 1);

    g_name = "eb_trace_plm";
    launch(k_plm_eb, L.bxg2, st, A.q, A.qbox, A.qaux, g, QM[0], QP[0], QM[1], QP[1], QM[2],
           QP[2], hc, A.dtp);
    g_name = "eb_riemann_1";
    launch(k_riemann_eb<0, 0>, L.f1[0], st, QM[0], QP[0], A.qaux, F1[0], G1[0], g.ap[0], hc);
    launch(k_riemann_eb<1, 0>, L.f1[1], st, QM[1], QP[1], A.qaux, F1[1], G1[1], g.ap[1], hc);
    launch(k_riemann_eb<2, 0>, L.f1[2], st, QM[2], QP[2], A.qaux, F1[2], G1[2], g.ap[2], hc);

    // transverse pairs D|T over bxg2 shrunk in T; Riemann over the faces of
    // bxg1 in D grown by one in the third direction (Godunov_3D_eb.cpp:214-455)
#define CAMR_EB_PAIR(D, T, O)                                                                \
    {                                                                                        \
        Arr qmT = dense(tq0, L.mbx[D]), qpT = dense(tq1, L.mbx[D]);                          \
        g_name = "eb_transverse_1";                                                          \
        launch(k_transdo_eb<D, T>, grow_dir(L.bxg2, T, -1), st, QM[D], QP[D], A.qaux, F1[T], \
               G1[T], g.ap[D], g.ap[T], g.f, qmT, qpT, A.dtp, hc);                          \
        g_name = "eb_riemann_2";                                                             \
        launch(k_riemann_eb<D, 1>, grow_dir(nodes(L.bxg1, D), O, 1), st, qmT, qpT, A.qaux,  \
               F2[D][T], G2[D][T], g.ap[D], hc);                                             \
    }
    CAMR_EB_PAIR(0, 1, 2)
    CAMR_EB_PAIR(0, 2, 1)
    CAMR_EB_PAIR(1, 0, 2)
    CAMR_EB_PAIR(1, 2, 0)
    CAMR_EB_PAIR(2, 0, 1)
    CAMR_EB_PAIR(2, 1, 0)
#undef CAMR_EB_PAIR

    // final states on bxg1, fluxes on faces of btf grown tangentially
    // (Godunov_3D_eb.cpp:457-565)
#define CAMR_EB_FINAL(D, T0, T1)                                                              \
    {                                                                                         \
        g_name = "eb_transverse_2";                                                           \
        launch(k_transdd_eb<D>, L.bxg1, st, QM[D], QP[D], A.qaux, A.srcq, A.has_src,          \
               F2[T0][T1], G2[T0][T1], F2[T1][T0], G2[T1][T0], g.ap[D], g.ap[T0], g.ap[T1],   \
               g.f, qmF, qpF, A.dtp, hc);                                                     \
        g_name = "eb_riemann_3";                                                              \
        launch(k_riemann_eb<D, 2>, grow_dir(grow_dir(nodes(L.btf, D), T0, 1), T1, 1), st, qmF, \
               qpF, A.qaux, FT[D], QE[D], g.ap[D], hc);                                       \
    }
    CAMR_EB_FINAL(0, 1, 2)
    CAMR_EB_FINAL(1, 0, 2)
    CAMR_EB_FINAL(2, 0, 1)
#undef CAMR_EB_FINAL

    // adjust_fluxes_eb: divu on the nodes of grow(qbox, -1) that lie in the
    // domain grown by 16 in periodic directions (Hydro_adjust_fluxes.cpp:73-84)
    Box nb = L.nb;
    for (int d = 0; d < 3; ++d) {
        const int gp = A.periodic[d] ? 16 : 0;
        nb.lo[d] = std::max(nb.lo[d], hc.domlo[d] - gp);
        nb.hi[d] = std::min(nb.hi[d], hc.domhi[d] + 1 + gp);
    }
    g_name = "eb_divu";
    if (nb.ok()) launch(k_eb_divu, nb, st, A.q, g.vf, divu, dxi[0], dxi[1], dxi[2]);
    g_name = "eb_adjust";
    launch(k_adjust_eb<0>, L.ft[0], st, FT[0], divu, A.uin, g.ap[0], hc.dx[1] * hc.dx[2], hc);
    launch(k_adjust_eb<1>, L.ft[1], st, FT[1], divu, A.uin, g.ap[1], hc.dx[0] * hc.dx[2], hc);
    launch(k_adjust_eb<2>, L.ft[2], st, FT[2], divu, A.uin, g.ap[2], hc.dx[1] * hc.dx[0], hc);

    g_name = "eb_divergence";
    launch(k_eb_div, L.btf, st, A.bx, FT[0], FT[1], FT[2], QE[0], QE[1], QE[2], g, A.q, A.qaux,
           divc, wt, A.flx[0], A.flx[1], A.flx[2], hc, dxi[0], dxi[1], dxi[2]);

    Box dbox2;
    for (int d = 0; d < 3; ++d) {
        const int gp = A.periodic[d] ? 2 : 0;
        dbox2.lo[d] = hc.domlo[d] - gp;
        dbox2.hi[d] = hc.domhi[d] + gp;
    }
    g_name = "eb_redistribute";
    launch(k_redist_src, L.g1, st, g, dbox2, divc, wt, optmp0, delm, winv);
    launch(k_redist_dst, A.bx, st, g, dbox2, divc, wt, optmp0, delm, winv, A.dsdt);
    g_prof = nullptr;
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
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
