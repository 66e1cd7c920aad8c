// hydro_staged.hip -- per-stage gfx950 kernels of the CTU Godunov advance.
//
// One kernel per stage of Godunov_umeth / hydro_umdrv
// (Source/Hydro/Godunov/Godunov_3D.cpp:12-461, Hydro_umdrv.cpp:11-93),
// with compact intermediates (9 primitive / 9 flux / 2 Godunov components)
// in HBM scratch.  This is the reference-shaped path; the fused engine in
// hydro_fused.hip is the fast path.  Both reproduce the reference bit for
// bit (-ffp-contract=off, identical operation order).
#include "hydro_kernels.h"

namespace camr {

namespace {

// q component of each compact slot (SR..SY): QRHO QU QV QW QPRES QREINT QFS..
__device__ __constant__ int kQcomp[NS] = {QRHO, QU, QV, QW, QPRES, QREINT,
                                          QFS, QFS + 1, QFS + 2};
// U component of each compact flux slot (FR..FY).
__device__ __constant__ int kUcomp[NF] = {URHO, UMX, UMY, UMZ, UEDEN, UEINT,
                                          UFS, UFS + 1, UFS + 2};

__device__ __forceinline__ int dv(int d, int e) { return d == e ? 1 : 0; }

// ---------------------------------------------------------------- flatten
__global__ void __launch_bounds__(kBlock)
k_flat(Range r, Arr q, Box qbox, Arr flat, int use_flattening) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    double f = 1.0;
    if (use_flattening == 1) {
        for (int d = 0; d < 3; ++d) {
            const int ex = dv(d, 0), ey = dv(d, 1), ez = dv(d, 2);
            double fd = 1.0;
            if (qbox.contains(i + 3 * ex, j + 3 * ey, k + 3 * ez) &&
                qbox.contains(i - 3 * ex, j - 3 * ey, k - 3 * ez)) {
                double p[7], un[7];
#pragma unroll
                for (int o = -3; o <= 3; ++o) {
                    p[o + 3] = q(i + o * ex, j + o * ey, k + o * ez, QPRES);
                    un[o + 3] = q(i + o * ex, j + o * ey, k + o * ez, QU + d);
                }
                fd = flatten_1d(p, un);
            }
            f = vmin(f, fd);
        }
    }
    flat(i, j, k, 0) = f;
}

// --------------------------------------------------------------- PLM trace
// hydro_plm_d for all three directions (Godunov_3D.cpp:104-152).
__global__ void __launch_bounds__(kBlock)
k_plm(Range r, Arr q, Arr qaux, Arr srcq, int has_src, Arr flat, Arr qm0,
      Arr qp0, Arr qm1, Arr qp1, Arr qm2, Arr qp2, HydroConst hc, const double *dtp) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const double fl = flat(i, j, k, 0);
    double c[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) c[s] = q(i, j, k, kQcomp[s]);
    const double cc = qaux(i, j, k, QC);
    Arr QM[3] = {qm0, qm1, qm2};
    Arr QP[3] = {qp0, qp1, qp2};
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        const int ex = dv(d, 0), ey = dv(d, 1), ez = dv(d, 2);
        double sl[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double st[5];
#pragma unroll
            for (int o = -2; o <= 2; ++o)
                st[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, kQcomp[s]);
            if (s == SP && hc.use_pslope) {
                double rr[5], gg[5];
#pragma unroll
                for (int o = -2; o <= 2; ++o) {
                    rr[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, QRHO);
                    gg[o + 2] = has_src ? srcq(i + o * ex, j + o * ey, k + o * ez, QU + d) : 0.0;
                }
                sl[s] = mc_pslope(st, rr, gg, hc.dx[d], fl, hc.plm_iorder);
            } else {
                sl[s] = mc_slope(st, fl, hc.plm_iorder);
            }
        }
        double qp[NS], qm[NS];
        plm_trace(d, c, sl, cc, *dtp / hc.dx[d], hc.small_dens, hc.small_pres, qp, qm);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            QP[d](i, j, k, s) = qp[s];
            QM[d](i + ex, j + ey, k + ez, s) = qm[s];
        }
    }
}

// --------------------------------------------------------------- PPM trace
// trace_ppm for one direction (PPM.cpp:11-411).
template <int D>
__global__ void __launch_bounds__(kBlock)
k_ppm(Range r, Arr q, Arr qaux, Arr flat, Arr qm, Arr qp, HydroConst hc, const double *dtp) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    const double dtdx = *dtp / hc.dx[D];
    const double cc = sqrt(qaux(i, j, k, QGAMC) * q(i, j, k, QPRES) / q(i, j, k, QRHO));
    const double un = q(i, j, k, QU + D);
    const double fl = flat(i, j, k, 0);
    double Ip[NS][3], Im[NS][3];
    double gIp[3], gIm[3];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o)
            st[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, kQcomp[s]);
        double sm, sp;
        ppm_edges(st, fl, sm, sp);
        const double s6 = ppm_s6(sm, sp, st[2]);
        ppm_integral(sm, sp, s6, un - cc, dtdx, Ip[s][0], Im[s][0]);
        ppm_integral(sm, sp, s6, un, dtdx, Ip[s][1], Im[s][1]);
        ppm_integral(sm, sp, s6, un + cc, dtdx, Ip[s][2], Im[s][2]);
    }
    {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o)
            st[o + 2] = qaux(i + o * ex, j + o * ey, k + o * ez, QGAMC);
        double sm, sp;
        ppm_edges(st, fl, sm, sp);
        const double s6 = ppm_s6(sm, sp, st[2]);
        ppm_integral(sm, sp, s6, un - cc, dtdx, gIp[0], gIm[0]);
        ppm_integral(sm, sp, s6, un, dtdx, gIp[1], gIm[1]);
        ppm_integral(sm, sp, s6, un + cc, dtdx, gIp[2], gIm[2]);
    }
    constexpr int SN = SU + D;
    constexpr int T1 = (D == 0) ? SV : SU;
    constexpr int T2 = (D == 2) ? SV : SW;
    PPMWaves W;
    // plus state on face i
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Im[SR][w]; W.un[w] = Im[SN][w]; W.p[w] = Im[SP][w];
        W.rhoe[w] = Im[SE][w]; W.gam[w] = gIm[w];
    }
    {
        double ro, uo, po, eo;
        ppm_project(+1, W, un, cc, hc.small_dens, hc.small_pres, ro, uo, po, eo);
        qp(i, j, k, SR) = ro; qp(i, j, k, SN) = uo; qp(i, j, k, SP) = po;
        qp(i, j, k, SE) = eo;
        qp(i, j, k, T1) = Im[T1][1]; qp(i, j, k, T2) = Im[T2][1];
#pragma unroll
        for (int s = 0; s < NSPEC; ++s) qp(i, j, k, SY + s) = Im[SY + s][1];
    }
    // minus state on face i+1
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Ip[SR][w]; W.un[w] = Ip[SN][w]; W.p[w] = Ip[SP][w];
        W.rhoe[w] = Ip[SE][w]; W.gam[w] = gIp[w];
    }
    {
        double ro, uo, po, eo;
        ppm_project(-1, W, un, cc, hc.small_dens, hc.small_pres, ro, uo, po, eo);
        const int a = i + ex, b = j + ey, c = k + ez;
        qm(a, b, c, SR) = ro; qm(a, b, c, SN) = uo; qm(a, b, c, SP) = po;
        qm(a, b, c, SE) = eo;
        qm(a, b, c, T1) = Ip[T1][1]; qm(a, b, c, T2) = Ip[T2][1];
#pragma unroll
        for (int s = 0; s < NSPEC; ++s) qm(a, b, c, SY + s) = Ip[SY + s][1];
    }
}

// ------------------------------------------------------------ Riemann solve
// hydro_cmpflx over faces of direction D.  mode 0: compact flux + GD out,
// writing the boundary-mutated states back (first solve); mode 1: compact
// out without write-back (temporaries); mode 2: final solve into the
// caller's NVAR flux array + GD out, with the face's adjust_fluxes
// (artificial viscosity from the divu nodes, species normalisation, area)
// applied in registers.
struct AdjustIn {
    Arr divu, u, area;
    int has_area;
    double area_c;
};

template <int D, int MODE>
__global__ void __launch_bounds__(kBlock)
k_riemann(Range r, Arr qL, Arr qR, Arr qaux, Arr fout, Arr gout, HydroConst hc, AdjustIn adj) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
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
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, ql, qr, a, f, g, &mut);
    if (MODE == 0 && mut) {
        if (mut & 1) {
            qL(i, j, k, SR) = ql[SR]; qL(i, j, k, SP) = ql[SP]; qL(i, j, k, SE) = ql[SE];
        }
        if (mut & 2) {
            qR(i, j, k, SR) = qr[SR]; qR(i, j, k, SP) = qr[SP]; qR(i, j, k, SE) = qr[SE];
        }
    }
    if (MODE == 2) {
        double fa[NVAR];
        fa[UTEMP] = 0.0;
#pragma unroll
        for (int s = 0; s < NF; ++s) fa[kUcomp[s]] = f[s];
        adjust_face<D>(fa, adj.divu, adj.u, i, j, k,
                       adj.has_area ? adj.area(i, j, k, 0) : adj.area_c, hc);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) fout(i, j, k, n) = fa[n];
    } else {
#pragma unroll
        for (int s = 0; s < NF; ++s) fout(i, j, k, s) = f[s];
    }
    gout(i, j, k, GUN) = g[GUN];
    gout(i, j, k, GP) = g[GP];
}

// ------------------------------------------------------ transverse (transdo)
// Both pairs D|T0 and D|T1 (Godunov_3D.cpp:218-241, 271-293, 325-348).
template <int D>
__global__ void __launch_bounds__(kBlock)
k_transdo(Range r, Arr qnm, Arr qnp, Arr qaux, Arr fa, Arr ga, Arr fb, Arr gb,
          Arr qmA, Arr qpA, Arr qmB, Arr qpB, const double *dtp, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    const double gamc = qaux(i, j, k, QGAMC);
    double np[NS], nm[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        np[s] = qnp(i, j, k, s);
        nm[s] = qnm(i + ex, j + ey, k + ez, s);
    }
#pragma unroll
    for (int pair = 0; pair < 2; ++pair) {
        const int t = pair ? T1 : T0;
        const Arr &F = pair ? fb : fa;
        const Arr &G = pair ? gb : ga;
        // cdtdx = dt/(3 dx_t), Godunov_3D.cpp:50-56
        const double cdt = 1.0 / 3.0 * *dtp / hc.dx[t];
        const int tx = (t == 0), ty = (t == 1), tz = (t == 2);
        double fl[NF], fr[NF], gl[NG], gr[NG];
#pragma unroll
        for (int s = 0; s < NF; ++s) {
            fl[s] = F(i, j, k, s);
            fr[s] = F(i + tx, j + ty, k + tz, s);
        }
        gl[GUN] = G(i, j, k, GUN); gl[GP] = G(i, j, k, GP);
        gr[GUN] = G(i + tx, j + ty, k + tz, GUN); gr[GP] = G(i + tx, j + ty, k + tz, GP);
        double op[NS], om[NS];
        transverse_one(np, fl, fr, gl, gr, cdt, gamc, hc.trd, hc.small_pres, op);
        transverse_one(nm, fl, fr, gl, gr, cdt, gamc, hc.trd, hc.small_pres, om);
        const Arr &QP = pair ? qpB : qpA;
        const Arr &QM = pair ? qmB : qmA;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            QP(i, j, k, s) = op[s];
            QM(i + ex, j + ey, k + ez, s) = om[s];
        }
    }
}

// ------------------------------------------------------ transverse (transdd)
// Godunov_3D.cpp:389-396, 411-418, 433-440.  fa: flux of direction T0
// corrected by T1 (faces of T0), fb: direction T1 corrected by T0.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_transdd(Range r, Arr qnm, Arr qnp, Arr qaux, Arr srcq, int has_src, Arr fa,
          Arr ga, Arr fb, Arr gb, Arr qm, Arr qp, const double *dtp, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    // hdt = dt/2 and hdt/dx of the two transverse directions (Godunov_3D.cpp:50-56)
    const double hdt = 0.5 * *dtp;
    const double ca = 0.5 * *dtp / hc.dx[(D == 0) ? 1 : 0];
    const double cb = 0.5 * *dtp / hc.dx[(D == 2) ? 1 : 2];
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    constexpr int ax = (T0 == 0), ay = (T0 == 1), az = (T0 == 2);
    constexpr int bx = (T1 == 0), by = (T1 == 1), bz = (T1 == 2);
    const double gamc = qaux(i, j, k, QGAMC);
    double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
#pragma unroll
    for (int s = 0; s < NF; ++s) {
        fal[s] = fa(i, j, k, s);
        far_[s] = fa(i + ax, j + ay, k + az, s);
        fbl[s] = fb(i, j, k, s);
        fbr[s] = fb(i + bx, j + by, k + bz, s);
    }
    gal[GUN] = ga(i, j, k, GUN); gal[GP] = ga(i, j, k, GP);
    gar[GUN] = ga(i + ax, j + ay, k + az, GUN); gar[GP] = ga(i + ax, j + ay, k + az, GP);
    gbl[GUN] = gb(i, j, k, GUN); gbl[GP] = gb(i, j, k, GP);
    gbr[GUN] = gb(i + bx, j + by, k + bz, GUN); gbr[GP] = gb(i + bx, j + by, k + bz, GP);
    double src[NS];
    if (has_src) {
#pragma unroll
        for (int s = 0; s < NS; ++s) src[s] = srcq(i, j, k, kQcomp[s]);
    }
    double n[NS], o[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) n[s] = qnp(i, j, k, s);
    transverse_two(n, fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, hdt,
                   has_src ? src : nullptr, gamc, hc.trd, hc.small_pres, o);
#pragma unroll
    for (int s = 0; s < NS; ++s) qp(i, j, k, s) = o[s];
#pragma unroll
    for (int s = 0; s < NS; ++s) n[s] = qnm(i + ex, j + ey, k + ez, s);
    transverse_two(n, fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, hdt,
                   has_src ? src : nullptr, gamc, hc.trd, hc.small_pres, o);
#pragma unroll
    for (int s = 0; s < NS; ++s) qm(i + ex, j + ey, k + ez, s) = o[s];
}

// ---------------------------------------------------- method of lines face
// MOL_umeth's per-direction pair (MOL_umeth.cpp:74-139): the slopes of the
// two cells beside the face (mol_slope_*), their face states
// (mol_riemann_*) and hydro_cmpflx, one face per thread, followed in
// registers by the face's adjust_fluxes (artificial viscosity from the
// divu nodes, species normalisation, area).  The slope of a cell is formed
// by both faces that read it, with the same operations, so no slope array
// is kept.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_mol_face(Range r, Arr q, Arr qaux, Arr fout, Arr gout, HydroConst hc, Arr divu, Arr u,
           Arr area, int has_area, double area_c) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    // compact states of the cells i-2 .. i+1 along D
    double c[4][NS];
#pragma unroll
    for (int o = 0; o < 4; ++o)
#pragma unroll
        for (int s = 0; s < NS; ++s)
            c[o][s] = q(i + (o - 2) * ex, j + (o - 2) * ey, k + (o - 2) * ez, kQcomp[s]);
    const double cl = qaux(i - ex, j - ey, k - ez, QC), cr = qaux(i, j, k, QC);
    double dql[NMS], dqr[NMS];
    mol_slopes(D, c[0], c[1], c[2], cl, hc.plm_iorder, hc.small_dens, dql);
    mol_slopes(D, c[1], c[2], c[3], cr, hc.plm_iorder, hc.small_dens, dqr);
    double ql[NS], qr[NS];
    mol_face_state(D, 1.0, c[1], dql, cl, &c[2][SY], hc.small_dens, hc.small_pres, ql);
    mol_face_state(D, -1.0, c[2], dqr, cr, &c[2][SY], hc.small_dens, hc.small_pres, qr);
    FaceAux a;
    a.c_l = cl;
    a.c_r = cr;
    a.csml_l = qaux(i - ex, j - ey, k - ez, QCSML);
    a.csml_r = qaux(i, j, k, QCSML);
    a.gam_l = qaux(i - ex, j - ey, k - ez, QGAMC);
    a.gam_r = qaux(i, j, k, QGAMC);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, ql, qr, a, f, g, &mut);
    double fa[NVAR];
    fa[UTEMP] = 0.0;
#pragma unroll
    for (int s = 0; s < NF; ++s) fa[kUcomp[s]] = f[s];
    adjust_face<D>(fa, divu, u, i, j, k, has_area ? area(i, j, k, 0) : area_c, hc);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) fout(i, j, k, n) = fa[n];
    gout(i, j, k, GUN) = g[GUN];
    gout(i, j, k, GP) = g[GP];
}

} // namespace

// ---------------------------------------------------------------- divu
// hydro_divu (Hydro_utils_K.H:12-169) on nodes.
__global__ void __launch_bounds__(kBlock)
k_divu(Range r, Arr q, Arr divu, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    divu(i, j, k, 0) = divu_node(q, i, j, k, hc);
}

// ------------------------------------------------------ pdivu + update
// hydro_pdivu (Hydro_utils_K.H:175-202) + hydro_update (:239-257).
__global__ void __launch_bounds__(kBlock)
k_update(Range r, Arr dsdt, Arr f0, Arr f1, Arr f2, Arr g0, Arr g1, Arr g2,
         Arr a0, Arr a1, Arr a2, int has_area, double ac0, double ac1,
         double ac2, Arr vol, int has_vol, double volc, int from_zero) {
    int i, j, k;
    if (!range_point(r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const double A0l = has_area ? a0(i, j, k, 0) : ac0;
    const double A0r = has_area ? a0(i + 1, j, k, 0) : ac0;
    const double A1l = has_area ? a1(i, j, k, 0) : ac1;
    const double A1r = has_area ? a1(i, j + 1, k, 0) : ac1;
    const double A2l = has_area ? a2(i, j, k, 0) : ac2;
    const double A2r = has_area ? a2(i, j, k + 1, 0) : ac2;
    const double v = has_vol ? vol(i, j, k, 0) : volc;
    const double pdivu =
        0.5 *
        ((g0(i + 1, j, k, GP) + g0(i, j, k, GP)) * (g0(i + 1, j, k, GUN) * A0r - g0(i, j, k, GUN) * A0l) +
         (g1(i, j + 1, k, GP) + g1(i, j, k, GP)) * (g1(i, j + 1, k, GUN) * A1r - g1(i, j, k, GUN) * A1l) +
         (g2(i, j, k + 1, GP) + g2(i, j, k, GP)) * (g2(i, j, k + 1, GUN) * A2r - g2(i, j, k, GUN) * A2l)) /
        v;
    const double volinv = 1.0 / v;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        const double base = from_zero ? 0.0 : dsdt(i, j, k, n);
        dsdt(i, j, k, n) = base + (f0(i, j, k, n) - f0(i + 1, j, k, n) + f1(i, j, k, n) -
                                   f1(i, j + 1, k, n) + f2(i, j, k, n) - f2(i, j, k + 1, n)) *
                                      volinv;
    }
    dsdt(i, j, k, UEINT) -= pdivu;
}

// ---------------------------------------------------------------- driver
namespace {
Prof *g_prof = nullptr;  // set for the duration of one umdrv_staged call
const char *g_name = "";

template <typename K, typename... A>
void launch(K kern, const Box &b, hipStream_t s, A... args) {
    Range r = range_of(b);
    Region reg(g_prof, g_name, s);
    hipLaunchKernelGGL(kern, dim3(grid_for(r.n)), dim3(kBlock), 0, s, r, args...);
}
#define NAMED(n) g_name = n
} // namespace

size_t umdrv_staged_scratch_doubles(const Box &bx) {
    const Box bxg1 = grow(bx, 1), bxg2 = grow(bx, 2);
    size_t n = 0;
    n += bxg2.npts();                                            // flat
    for (int d = 0; d < 3; ++d) {
        n += (size_t)NS * grow_hi(bxg2, d, 1).npts();            // qdm
        n += (size_t)NS * bxg2.npts();                           // qdp
        Box fb = nodes(grow_dir(bxg2, d, -1), d);
        n += (size_t)(NF + NG) * fb.npts();                      // F1, G1
        Box f2 = nodes(bxg1, d);
        n += (size_t)2 * (NF + NG) * f2.npts();                  // F2 pairs
        n += (size_t)NG * nodes(bx, d).npts();                   // G3
    }
    size_t tq = 0, fq = 0;
    for (int d = 0; d < 3; ++d) {
        Box tb = grow_dir(bxg1, d, 1);
        size_t t = (size_t)2 * NS * (grow_hi(tb, d, 1).npts() + tb.npts());
        tq = t > tq ? t : tq;
        Box fbx = grow_dir(bx, d, 1);
        size_t f = (size_t)NS * (grow_hi(fbx, d, 1).npts() + fbx.npts());
        fq = f > fq ? f : fq;
    }
    n += tq + fq;
    n += nodes(nodes(nodes(bx, 0), 1), 2).npts();               // divu
    return n + 64;
}

int umdrv_staged(const UmdrvArgs &A, double *scratch, hipStream_t st, Prof *prof) {
    g_prof = prof;
    const HydroConst &hc = A.hc;
    const Box bx = A.bx, bxg1 = grow(bx, 1), bxg2 = grow(bx, 2);
    double *cur = scratch;
    auto take = [&](const Box &b, int nc) {
        Arr a = dense(cur, b);
        cur += (size_t)nc * b.npts();
        return a;
    };

    Arr flat = take(bxg2, 1);
    NAMED("flatten");
    launch(k_flat, bxg2, st, A.q, A.qbox, flat, hc.use_flattening);

    Arr qdm[3], qdp[3];
    for (int d = 0; d < 3; ++d) {
        qdm[d] = take(grow_hi(bxg2, d, 1), NS);
        qdp[d] = take(bxg2, NS);
    }
    if (hc.ppm_type == 0) {
        NAMED("trace_plm");
        launch(k_plm, bxg2, st, A.q, A.qaux, A.srcq, A.has_src, flat, qdm[0],
               qdp[0], qdm[1], qdp[1], qdm[2], qdp[2], hc, A.dtp);
    } else {
        NAMED("trace_ppm");
        launch(k_ppm<0>, bxg2, st, A.q, A.qaux, flat, qdm[0], qdp[0], hc, A.dtp);
        launch(k_ppm<1>, bxg2, st, A.q, A.qaux, flat, qdm[1], qdp[1], hc, A.dtp);
        launch(k_ppm<2>, bxg2, st, A.q, A.qaux, flat, qdm[2], qdp[2], hc, A.dtp);
    }

    // first fluxes
    Arr F1[3], G1[3];
    for (int d = 0; d < 3; ++d) {
        Box fb = nodes(grow_dir(bxg2, d, -1), d);
        F1[d] = take(fb, NF);
        G1[d] = take(fb, NG);
    }
    NAMED("riemann_normal");
    launch(k_riemann<0, 0>, nodes(grow_dir(bxg2, 0, -1), 0), st, qdm[0], qdp[0], A.qaux, F1[0], G1[0], hc, AdjustIn{});
    launch(k_riemann<1, 0>, nodes(grow_dir(bxg2, 1, -1), 1), st, qdm[1], qdp[1], A.qaux, F1[1], G1[1], hc, AdjustIn{});
    launch(k_riemann<2, 0>, nodes(grow_dir(bxg2, 2, -1), 2), st, qdm[2], qdp[2], A.qaux, F1[2], G1[2], hc, AdjustIn{});

    // transverse corrections + Riemann solves of the 6 pairs
    Arr F2[3][3], G2[3][3];
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < 3; ++t)
            if (t != d) {
                F2[d][t] = take(nodes(bxg1, d), NF);
                G2[d][t] = take(nodes(bxg1, d), NG);
            }
    double *tq_base = cur;
    NAMED("transverse_1");
    for (int d = 0; d < 3; ++d) {
        cur = tq_base;
        const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2;
        const Box tb = grow_dir(bxg1, d, 1);
        Arr qmA = take(grow_hi(tb, d, 1), NS), qpA = take(tb, NS);
        Arr qmB = take(grow_hi(tb, d, 1), NS), qpB = take(tb, NS);
        const Box fb = nodes(bxg1, d);
        if (d == 0) {
            launch(k_transdo<0>, tb, st, qdm[0], qdp[0], A.qaux, F1[t0], G1[t0], F1[t1], G1[t1], qmA, qpA, qmB, qpB, A.dtp, hc);
            launch(k_riemann<0, 1>, fb, st, qmA, qpA, A.qaux, F2[0][t0], G2[0][t0], hc, AdjustIn{});
            launch(k_riemann<0, 1>, fb, st, qmB, qpB, A.qaux, F2[0][t1], G2[0][t1], hc, AdjustIn{});
        } else if (d == 1) {
            launch(k_transdo<1>, tb, st, qdm[1], qdp[1], A.qaux, F1[t0], G1[t0], F1[t1], G1[t1], qmA, qpA, qmB, qpB, A.dtp, hc);
            launch(k_riemann<1, 1>, fb, st, qmA, qpA, A.qaux, F2[1][t0], G2[1][t0], hc, AdjustIn{});
            launch(k_riemann<1, 1>, fb, st, qmB, qpB, A.qaux, F2[1][t1], G2[1][t1], hc, AdjustIn{});
        } else {
            launch(k_transdo<2>, tb, st, qdm[2], qdp[2], A.qaux, F1[t0], G1[t0], F1[t1], G1[t1], qmA, qpA, qmB, qpB, A.dtp, hc);
            launch(k_riemann<2, 1>, fb, st, qmA, qpA, A.qaux, F2[2][t0], G2[2][t0], hc, AdjustIn{});
            launch(k_riemann<2, 1>, fb, st, qmB, qpB, A.qaux, F2[2][t1], G2[2][t1], hc, AdjustIn{});
        }
    }

    // final states and fluxes
    Arr G3[3];
    double *fq_base = cur;
    // G3 and divu stay live until the update: allocate before the reused region
    cur = fq_base;
    for (int d = 0; d < 3; ++d) G3[d] = take(nodes(bx, d), NG);
    Arr divu = take(nodes(nodes(nodes(bx, 0), 1), 2), 1);
    double *qf_base = cur;
    // divu first: the final solves apply adjust_fluxes in registers
    NAMED("divu");
    launch(k_divu, nodes(nodes(nodes(bx, 0), 1), 2), st, A.q, divu, hc);
    AdjustIn adj[3];
    for (int d = 0; d < 3; ++d) adj[d] = AdjustIn{divu, A.uin, A.area[d], A.has_area, A.area_c[d]};
    NAMED("transverse_2");
    for (int d = 0; d < 3; ++d) {
        cur = qf_base;
        const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2;
        const Box tb = grow_dir(bx, d, 1);
        Arr qm = take(grow_hi(tb, d, 1), NS), qp = take(tb, NS);
        const Box fb = nodes(bx, d);
        if (d == 0) {
            launch(k_transdd<0>, tb, st, qdm[0], qdp[0], A.qaux, A.srcq, A.has_src, F2[t0][t1], G2[t0][t1], F2[t1][t0], G2[t1][t0], qm, qp, A.dtp, hc);
            launch(k_riemann<0, 2>, fb, st, qm, qp, A.qaux, A.flx[0], G3[0], hc, adj[0]);
        } else if (d == 1) {
            launch(k_transdd<1>, tb, st, qdm[1], qdp[1], A.qaux, A.srcq, A.has_src, F2[t0][t1], G2[t0][t1], F2[t1][t0], G2[t1][t0], qm, qp, A.dtp, hc);
            launch(k_riemann<1, 2>, fb, st, qm, qp, A.qaux, A.flx[1], G3[1], hc, adj[1]);
        } else {
            launch(k_transdd<2>, tb, st, qdm[2], qdp[2], A.qaux, A.srcq, A.has_src, F2[t0][t1], G2[t0][t1], F2[t1][t0], G2[t1][t0], qm, qp, A.dtp, hc);
            launch(k_riemann<2, 2>, fb, st, qm, qp, A.qaux, A.flx[2], G3[2], hc, adj[2]);
        }
    }

    NAMED("consup");
    launch(k_update, bx, st, A.dsdt, A.flx[0], A.flx[1], A.flx[2], G3[0], G3[1], G3[2],
           A.area[0], A.area[1], A.area[2], A.has_area, A.area_c[0], A.area_c[1],
           A.area_c[2], A.vol, A.has_vol, A.vol_c, A.dsdt_zero);
    g_prof = nullptr;
    return hipGetLastError() == hipSuccess ? 0 : CAMR_ERR_HIP;
}

// hydro_umdrv with do_mol (Hydro_umdrv.cpp:57-63): divu, then MOL_umeth's
// fluxes and Godunov face states with adjust_fluxes applied in the same
// kernel, then the same consup as the Godunov path.  The face states reach two cells past bx along the face
// direction; divu one cell past it.
size_t umdrv_mol_scratch_doubles(const Box &bx) {
    size_t n = 0;
    for (int d = 0; d < 3; ++d) n += (size_t)NG * nodes(bx, d).npts();
    n += nodes(nodes(nodes(bx, 0), 1), 2).npts();
    return n + 64;
}

int umdrv_mol(const UmdrvArgs &A, double *scratch, hipStream_t st, Prof *prof) {
    g_prof = prof;
    const HydroConst &hc = A.hc;
    const Box bx = A.bx;
    double *cur = scratch;
    auto take = [&](const Box &b, int nc) {
        Arr a = dense(cur, b);
        cur += (size_t)nc * b.npts();
        return a;
    };
    Arr G[3];
    for (int d = 0; d < 3; ++d) G[d] = take(nodes(bx, d), NG);
    Arr divu = take(nodes(nodes(nodes(bx, 0), 1), 2), 1);
    NAMED("divu");
    launch(k_divu, nodes(nodes(nodes(bx, 0), 1), 2), st, A.q, divu, hc);
    NAMED("mol_faces");
    launch(k_mol_face<0>, nodes(bx, 0), st, A.q, A.qaux, A.flx[0], G[0], hc, divu, A.uin,
           A.area[0], A.has_area, A.area_c[0]);
    launch(k_mol_face<1>, nodes(bx, 1), st, A.q, A.qaux, A.flx[1], G[1], hc, divu, A.uin,
           A.area[1], A.has_area, A.area_c[1]);
    launch(k_mol_face<2>, nodes(bx, 2), st, A.q, A.qaux, A.flx[2], G[2], hc, divu, A.uin,
           A.area[2], A.has_area, A.area_c[2]);
    NAMED("consup");
    launch(k_update, bx, st, A.dsdt, A.flx[0], A.flx[1], A.flx[2], G[0], G[1], G[2],
           A.area[0], A.area[1], A.area[2], A.has_area, A.area_c[0], A.area_c[1],
           A.area_c[2], A.vol, A.has_vol, A.vol_c, A.dsdt_zero);
    g_prof = nullptr;
    return hipGetLastError() == hipSuccess ? 0 : CAMR_ERR_HIP;
}

} // namespace camr
