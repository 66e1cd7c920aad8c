// hydro_fused.hip -- the fast gfx950 path of the CTU Godunov advance.
//
// Same arithmetic as the reference (Godunov_3D.cpp:12-461,
// Hydro_umdrv.cpp:11-93), re-organised around faces so that no transverse
// or final interface state ever goes to HBM:
//   k_prim_c    U -> compact primitives (rho u v w p rhoe Y0..2 c csml)
//   k_flat_c    shock flattening coefficient
//   k_ppm_c / k_plm_c   normal predictor states, face-indexed
//   k_riemann1  first fluxes (with the reference's boundary-face mutation)
//   k_trans1<D> per D-face: the two transdo states of each side are formed
//               in registers and solved at once (both pairs D|t)
//   k_trans2<D> per D-face: transdd states, final Riemann solve, artificial
//               viscosity, species normalisation and area weighting
//   k_update_c  flux divergence, p div u, S_new += dt dS/dt, floor check
// Every face state is computed exactly once, as in the reference; the gamma
// of the gamma-law EOS is the constant 1.4, which the PPM reconstruction of
// QGAMC reproduces exactly (a constant profile integrates to itself).
#include "hydro_fused.h"

namespace camr {

namespace {

constexpr int ex_(int D, int d) { return D == d ? 1 : 0; }

// occupancy request for the register-heavy face kernels (waves per SIMD)
#ifdef CAMR_TRANS_MINW
#define TRANS_BOUNDS __launch_bounds__(kBlock, CAMR_TRANS_MINW)
#else
#define TRANS_BOUNDS __launch_bounds__(kBlock)
#endif

// ------------------------------------------------------------ primitives
// hydro_ctoprim (Hydro_ctoprim.H:11-101) into the compact layout, with
// hydro_srctoprim (:106-157) when sources are non-zero.
__global__ void __launch_bounds__(kBlock)
k_prim_c(Range r, Arr u, Arr q, Arr src, Arr srcq, int has_src, HydroConst hc, int *status) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    double v[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) v[n] = u(i, j, k, n);
    const double rho = v[URHO];
    if (!(rho > 0.)) return;
    const double e = prim_compact_point(v, q, i, j, k, hc, status);
    if (has_src) {
        const double rhoinv = 1.0 / rho;
        const double vx = v[UMX] * rhoinv, vy = v[UMY] * rhoinv, vz = v[UMZ] * rhoinv;
        const double srcrho = src(i, j, k, URHO);
        srcq(i, j, k, SR) = srcrho;
        srcq(i, j, k, SU) = (src(i, j, k, UMX) - vx * srcrho) * rhoinv;
        srcq(i, j, k, SV) = (src(i, j, k, UMY) - vy * srcrho) * rhoinv;
        srcq(i, j, k, SW) = (src(i, j, k, UMZ) - vz * srcrho) * rhoinv;
        const double sre = src(i, j, k, UEINT);
        srcq(i, j, k, SE) = sre;
#pragma unroll
        for (int s = 0; s < NSPEC; ++s)
            srcq(i, j, k, SY + s) = (src(i, j, k, UFS + s) - (v[UFS + s] * rhoinv) * srcrho) * rhoinv;
        const double dpde = (kGamma - 1.0) * rho, dpdr = (kGamma - 1.0) * e;
        srcq(i, j, k, SP) = dpde * (sre - (e * rho) * srcrho * rhoinv) * rhoinv + dpdr * srcrho;
    }
}

// ------------------------------------------------------------ flattening
// min over directions of flatten (flatten.H:19-99); 1 when the +-3 stencil
// leaves q's box.
__global__ void __launch_bounds__(kBlock)
k_flat_c(Range r, Arr q, Box qbox, Arr flat, int use_flattening) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    double f = 1.0;
    if (use_flattening == 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
            double fd = 1.0;
            if (qbox.contains(i + 3 * ex, j + 3 * ey, k + 3 * ez) &&
                qbox.contains(i - 3 * ex, j - 3 * ey, k - 3 * ez)) {
                double p[7], un[7];
#pragma unroll
                for (int o = -3; o <= 3; ++o) {
                    p[o + 3] = q(i + o * ex, j + o * ey, k + o * ez, SP);
                    un[o + 3] = q(i + o * ex, j + o * ey, k + o * ez, SU + d);
                }
                fd = flatten_1d(p, un);
            }
            f = vmin(f, fd);
        }
    }
    flat(i, j, k, 0) = f;
}

// ------------------------------------------------------------ PPM trace
// trace_ppm (PPM.cpp:11-411) for direction D; qp(i) is the plus state of
// face i, qm(i+1) the minus state of face i+1.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_ppm_c(Range r, Arr q, Arr flat, Arr qm, Arr qp, HydroConst hc, double dtdx) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    constexpr int SN = SU + D;
    constexpr int T1 = (D == 0) ? SV : SU;
    constexpr int T2 = (D == 2) ? SV : SW;
    const double cc = q(i, j, k, SC);
    const double un = q(i, j, k, SN);
    const double fl = flat(i, j, k, 0);
    double Ip[NS][3], Im[NS][3];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o) st[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, s);
        double sm, sp;
        ppm_edges(st, fl, sm, sp);
        const double s6 = ppm_s6(sm, sp, st[2]);
        ppm_integral(sm, sp, s6, un - cc, dtdx, Ip[s][0], Im[s][0]);
        ppm_integral(sm, sp, s6, un, dtdx, Ip[s][1], Im[s][1]);
        ppm_integral(sm, sp, s6, un + cc, dtdx, Ip[s][2], Im[s][2]);
    }
    PPMWaves W;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Im[SR][w]; W.un[w] = Im[SN][w]; W.p[w] = Im[SP][w];
        W.rhoe[w] = Im[SE][w]; W.gam[w] = kGamma;
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
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Ip[SR][w]; W.un[w] = Ip[SN][w]; W.p[w] = Ip[SP][w];
        W.rhoe[w] = Ip[SE][w];
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

// ------------------------------------------------------------ PLM trace
// hydro_plm_d (PLM.H:170-323) for direction D.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_plm_c(Range r, Arr q, Arr srcq, int has_src, Arr flat, Arr qm, Arr qp,
        HydroConst hc, double dtdx) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    const double fl = flat(i, j, k, 0);
    double c[NS], sl[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o) st[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, s);
        c[s] = st[2];
        if (s == SP && hc.use_pslope) {
            double rr[5], gg[5];
#pragma unroll
            for (int o = -2; o <= 2; ++o) {
                rr[o + 2] = q(i + o * ex, j + o * ey, k + o * ez, SR);
                gg[o + 2] = has_src ? srcq(i + o * ex, j + o * ey, k + o * ez, SU + D) : 0.0;
            }
            sl[s] = mc_pslope(st, rr, gg, hc.dx[D], fl, hc.plm_iorder);
        } else {
            sl[s] = mc_slope(st, fl, hc.plm_iorder);
        }
    }
    double op[NS], om[NS];
    plm_trace(D, c, sl, q(i, j, k, SC), dtdx, hc.small_dens, hc.small_pres, op, om);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        qp(i, j, k, s) = op[s];
        qm(i + ex, j + ey, k + ez, s) = om[s];
    }
}

__device__ __forceinline__ FaceAux face_aux(const Arr &q, int il, int jl, int kl,
                                            int i, int j, int k) {
    FaceAux a;
    a.c_l = q(il, jl, kl, SC);
    a.c_r = q(i, j, k, SC);
    a.csml_l = q(il, jl, kl, SCS);
    a.csml_r = q(i, j, k, SCS);
    a.gam_l = kGamma;
    a.gam_r = kGamma;
    return a;
}

// compact transverse flux layout: FR FMX FMY FMZ FE Y0 Y1 Y2 (no rho e)
enum { TR = 0, TMX = 1, TMY = 2, TMZ = 3, TE = 4, TY = 5, NT = 8 };

__device__ __forceinline__ void load_tflux(const Arr &F, const Arr &G, int i, int j,
                                           int k, double *f, double *g) {
    f[FR] = F(i, j, k, TR);
    f[FMX] = F(i, j, k, TMX);
    f[FMY] = F(i, j, k, TMY);
    f[FMZ] = F(i, j, k, TMZ);
    f[FE] = F(i, j, k, TE);
    f[FEI] = 0.0;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) f[FY + s] = F(i, j, k, TY + s);
    g[GUN] = G(i, j, k, GUN);
    g[GP] = G(i, j, k, GP);
}

__device__ __forceinline__ void store_tflux(const Arr &F, const Arr &G, int i, int j,
                                            int k, const double *f, const double *g) {
    F(i, j, k, TR) = f[FR];
    F(i, j, k, TMX) = f[FMX];
    F(i, j, k, TMY) = f[FMY];
    F(i, j, k, TMZ) = f[FMZ];
    F(i, j, k, TE) = f[FE];
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) F(i, j, k, TY + s) = f[FY + s];
    G(i, j, k, GUN) = g[GUN];
    G(i, j, k, GP) = g[GP];
}

// ------------------------------------------------------------ first fluxes
// hydro_cmpflx on the normal predictor states (Godunov_3D.cpp:180-216);
// the boundary overrides of Hydro_cmpflx.H:239-282 are written back to the
// state arrays because the transverse stages read them.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_riemann1(Range r, Arr qm, Arr qp, Arr q, Arr F, Arr G, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    double ql[NS], qr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        ql[s] = qm(i, j, k, s);
        qr[s] = qp(i, j, k, s);
    }
    const FaceAux a = face_aux(q, i - ex, j - ey, k - ez, i, j, k);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, ql, qr, a, f, g, &mut);
    if (mut & 1) {
        qm(i, j, k, SR) = ql[SR]; qm(i, j, k, SP) = ql[SP]; qm(i, j, k, SE) = ql[SE];
    }
    if (mut & 2) {
        qp(i, j, k, SR) = qr[SR]; qp(i, j, k, SP) = qr[SP]; qp(i, j, k, SE) = qr[SE];
    }
    store_tflux(F, G, i, j, k, f, g);
}

// ------------------------------------------------ transdo + Riemann (pairs)
// For D-face f: the plus state comes from transdo at cell f, the minus state
// from transdo at cell f-e_D (Godunov_utils_3D.H:22-239 writes qp(iv) and
// qm(iv+e_D)); both pairs D|t.
template <int D, int T>
__device__ __forceinline__ void trans1_pair(int i, int j, int k, const double *np,
                                            const double *nm, const FaceAux &a, int idx,
                                            const Arr &F, const Arr &G, const Arr &O,
                                            const Arr &OG, double cdt, const HydroConst &hc) {
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    constexpr int tx = ex_(T, 0), ty = ex_(T, 1), tz = ex_(T, 2);
    double sp[NS], sm[NS];
    {
        double fl[NF], fr[NF], gl[NG], gr[NG];
        TransDiff t;
        load_tflux(F, G, i, j, k, fl, gl);
        load_tflux(F, G, i + tx, j + ty, k + tz, fr, gr);
        trans_diff_one(fl, fr, gl, gr, cdt, t);
        transverse_one_d(np, t, cdt, kGamma, hc.trd, hc.small_pres, sp);
    }
    {
        double fl[NF], fr[NF], gl[NG], gr[NG];
        TransDiff t;
        load_tflux(F, G, i - ex, j - ey, k - ez, fl, gl);
        load_tflux(F, G, i - ex + tx, j - ey + ty, k - ez + tz, fr, gr);
        trans_diff_one(fl, fr, gl, gr, cdt, t);
        transverse_one_d(nm, t, cdt, kGamma, hc.trd, hc.small_pres, sm);
    }
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, sm, sp, a, f, g, &mut);
    store_tflux(O, OG, i, j, k, f, g);
}

template <int D>
__global__ void TRANS_BOUNDS
k_trans1(Range r, Arr qm, Arr qp, Arr q, Arr Fa, Arr Ga, Arr Fb, Arr Gb,
         Arr Oa, Arr OGa, Arr Ob, Arr OGb, double cda, double cdb, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    double np[NS], nm[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        np[s] = qp(i, j, k, s);
        nm[s] = qm(i, j, k, s);
    }
    const FaceAux a = face_aux(q, i - ex, j - ey, k - ez, i, j, k);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    trans1_pair<D, T0>(i, j, k, np, nm, a, idx, Fa, Ga, Oa, OGa, cda, hc);
    trans1_pair<D, T1>(i, j, k, np, nm, a, idx, Fb, Gb, Ob, OGb, cdb, hc);
}

// transdd differences of cell (i,j,k) for direction D: Fa is the flux of
// direction T0 corrected by T1 (faces of T0), Fb the flux of T1 corrected
// by T0.
template <int D>
__device__ __forceinline__ void trans2_diff(int i, int j, int k, const Arr &Fa,
                                            const Arr &Ga, const Arr &Fb, const Arr &Gb,
                                            double ca, double cb, TransDiff &t) {
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    constexpr int ax = ex_(T0, 0), ay = ex_(T0, 1), az = ex_(T0, 2);
    constexpr int bx = ex_(T1, 0), by = ex_(T1, 1), bz = ex_(T1, 2);
    double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
    load_tflux(Fa, Ga, i, j, k, fal, gal);
    load_tflux(Fa, Ga, i + ax, j + ay, k + az, far_, gar);
    load_tflux(Fb, Gb, i, j, k, fbl, gbl);
    load_tflux(Fb, Gb, i + bx, j + by, k + bz, fbr, gbr);
    trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, kGamma, t);
}

// ---------------------------------- transdd + final Riemann + adjust_fluxes
// D-face f: plus state from transdd at cell f, minus state from transdd at
// cell f-e_D (Godunov_utils_3D.H:245-482); then hydro_cmpflx, then
// hydro_artif_visc / hydro_norm_spec_flx / hydro_ext_flx
// (Hydro_adjust_fluxes.cpp:13-44).
template <int D>
__global__ void TRANS_BOUNDS
k_trans2(Range r, Arr qm, Arr qp, Arr q, Arr srcq, int has_src, Arr Fa, Arr Ga,
         Arr Fb, Arr Gb, Arr divu, Arr u, Arr flx, Arr G3, double ca, double cb,
         double hdt, double area, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    double sp[NS], sm[NS];
    {
        double n[NS], src[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            n[s] = qp(i, j, k, s);
            src[s] = has_src ? srcq(i, j, k, s) : 0.0;
        }
        TransDiff t;
        trans2_diff<D>(i, j, k, Fa, Ga, Fb, Gb, ca, cb, t);
        transverse_two_d(n, t, hdt, src, hc.trd, hc.small_pres, sp);
    }
    {
        const int i0 = i - ex, j0 = j - ey, k0 = k - ez;
        double n[NS], src[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            n[s] = qm(i, j, k, s);
            src[s] = has_src ? srcq(i0, j0, k0, s) : 0.0;
        }
        TransDiff t;
        trans2_diff<D>(i0, j0, k0, Fa, Ga, Fb, Gb, ca, cb, t);
        transverse_two_d(n, t, hdt, src, hc.trd, hc.small_pres, sm);
    }
    const FaceAux a = face_aux(q, i - ex, j - ey, k - ez, i, j, k);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, sm, sp, a, f, g, &mut);
    double fu[NVAR];
    fu[URHO] = f[FR]; fu[UMX] = f[FMX]; fu[UMY] = f[FMY]; fu[UMZ] = f[FMZ];
    fu[UEDEN] = f[FE]; fu[UEINT] = f[FEI]; fu[UTEMP] = 0.0;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) fu[UFS + s] = f[FY + s];
    adjust_face<D>(fu, divu, u, i, j, k, area, hc);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx(i, j, k, n) = fu[n];
    G3(i, j, k, GUN) = g[GUN];
    G3(i, j, k, GP) = g[GP];
}

__global__ void __launch_bounds__(kBlock)
k_divu_c(Range r, Arr q, Arr divu, HydroConst hc) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    // compact q has u, v, w in the same slots as QU, QV, QW
    divu(i, j, k, 0) = divu_node(q, i, j, k, hc);
}

// ------------------------------------------------------------ update
// hydro_pdivu (Hydro_utils_K.H:175-202), hydro_update (:239-257) on a zeroed
// dS/dt (CAMR_construct_hydro_source.cpp:14), then the Saxpy of
// CAMR_advance.cpp:141 and the floor check of the following clean_state.
__global__ void __launch_bounds__(kBlock)
k_update_c(Range r, Arr snew, Arr f0, Arr f1, Arr f2, Arr g0, Arr g1, Arr g2,
           double a0, double a1, double a2, double vol, double dt,
           double small_dens, int *flags) {
    int i, j, k;
    if (!range_point(r, logical_thread(), i, j, k)) return;
    const double pdivu =
        0.5 *
        ((g0(i + 1, j, k, GP) + g0(i, j, k, GP)) * (g0(i + 1, j, k, GUN) * a0 - g0(i, j, k, GUN) * a0) +
         (g1(i, j + 1, k, GP) + g1(i, j, k, GP)) * (g1(i, j + 1, k, GUN) * a1 - g1(i, j, k, GUN) * a1) +
         (g2(i, j, k + 1, GP) + g2(i, j, k, GP)) * (g2(i, j, k + 1, GUN) * a2 - g2(i, j, k, GUN) * a2)) /
        vol;
    const double volinv = 1.0 / vol;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        double ds = 0.0;
        ds += (f0(i, j, k, n) - f0(i + 1, j, k, n) + f1(i, j, k, n) - f1(i, j + 1, k, n) +
               f2(i, j, k, n) - f2(i, j, k + 1, n)) *
              volinv;
        if (n == UEINT) ds -= pdivu;
        snew(i, j, k, n) += dt * ds;
    }
    const double rho = snew(i, j, k, URHO);
    if (rho == 0.) atomicOr(flags, 1 << CAMR_ERR_ZERO_DENSITY);
    else if (rho < small_dens) atomicOr(flags, kFloorFlag);
}

template <typename K, typename... A>
void go(const char *name, Prof *prof, K kern, const Box &b, hipStream_t s, A... args) {
    Range r = range_of(b);
    Region reg(prof, name, s);
    hipLaunchKernelGGL(kern, dim3(grid_for(r.n)), dim3(kBlock), 0, s, r, args...);
}

} // namespace

// --------------------------------------------------------------- engine
size_t FusedEngine::scratch_doubles(const Box &bx) {
    const Box bxg1 = grow(bx, 1), bxg2 = grow(bx, 2), g4 = grow(bx, CAMR_NGROW);
    size_t n = 0;
    n += (size_t)NQC * g4.npts();                             // compact q
    n += (size_t)NS * g4.npts();                              // compact srcq
    n += bxg2.npts();                                         // flat
    for (int d = 0; d < 3; ++d) {
        n += (size_t)NS * grow_hi(bxg2, d, 1).npts();         // qm
        n += (size_t)NS * bxg2.npts();                        // qp
        n += (size_t)(NT + NG) * nodes(grow_dir(bxg2, d, -1), d).npts();  // F1
        n += (size_t)2 * (NT + NG) * nodes(bxg1, d).npts();   // F2 pairs
        n += (size_t)(NVAR + NG) * nodes(bx, d).npts();       // flx, G3
    }
    n += nodes(nodes(nodes(bx, 0), 1), 2).npts();             // divu
    return n + 32 * 64;
}

int FusedEngine::reserve(const std::vector<Box> &boxes) {
    size_t n = 0;
    for (const Box &b : boxes) n = std::max(n, scratch_doubles(b));
    return scratch_.reserve(n * sizeof(double)) ? CAMR_OK : CAMR_ERR_ALLOC;
}

int FusedEngine::advance_box(const Box &bx, const Arr &sborder, const Arr &srchyd,
                             int has_src, const Arr &snew, const HydroConst &hc,
                             double dt, hipStream_t st, int *status, Prof *prof,
                             bool have_prim) {
    const Box bxg1 = grow(bx, 1), bxg2 = grow(bx, 2), g4 = grow(bx, CAMR_NGROW);
    scratch_.reset();
    auto take = [&](const Box &b, int nc) { return dense(scratch_.take((size_t)nc * b.npts()), b); };
    Arr q = take(g4, NQC);
    Arr sq = has_src ? take(g4, NS) : q;
    Arr flat = take(bxg2, 1);
    Arr qm[3], qp[3], F1[3], G1[3], F2[3][3], G2[3][3], flx[3], G3[3];
    for (int d = 0; d < 3; ++d) {
        qm[d] = take(grow_hi(bxg2, d, 1), NS);
        qp[d] = take(bxg2, NS);
        const Box fb = nodes(grow_dir(bxg2, d, -1), d);
        F1[d] = take(fb, NT);
        G1[d] = take(fb, NG);
        for (int t = 0; t < 3; ++t)
            if (t != d) {
                F2[d][t] = take(nodes(bxg1, d), NT);
                G2[d][t] = take(nodes(bxg1, d), NG);
            }
        flx[d] = take(nodes(bx, d), NVAR);
        G3[d] = take(nodes(bx, d), NG);
    }
    Arr divu = take(nodes(nodes(nodes(bx, 0), 1), 2), 1);
    if (!divu.p) {
        set_error("fused engine scratch too small");
        return CAMR_ERR_ALLOC;
    }

    if (!have_prim || has_src)
        go("prim", prof, k_prim_c, g4, st, sborder, q, srchyd, sq, has_src, hc, status);
    go("flatten", prof, k_flat_c, bxg2, st, q, g4, flat, hc.use_flattening);
    const double dtdx[3] = {dt / hc.dx[0], dt / hc.dx[1], dt / hc.dx[2]};
    if (hc.ppm_type == 1) {
        go("trace_ppm", prof, k_ppm_c<0>, bxg2, st, q, flat, qm[0], qp[0], hc, dtdx[0]);
        go("trace_ppm", prof, k_ppm_c<1>, bxg2, st, q, flat, qm[1], qp[1], hc, dtdx[1]);
        go("trace_ppm", prof, k_ppm_c<2>, bxg2, st, q, flat, qm[2], qp[2], hc, dtdx[2]);
    } else {
        go("trace_plm", prof, k_plm_c<0>, bxg2, st, q, sq, has_src, flat, qm[0], qp[0], hc, dtdx[0]);
        go("trace_plm", prof, k_plm_c<1>, bxg2, st, q, sq, has_src, flat, qm[1], qp[1], hc, dtdx[1]);
        go("trace_plm", prof, k_plm_c<2>, bxg2, st, q, sq, has_src, flat, qm[2], qp[2], hc, dtdx[2]);
    }
    go("riemann_normal", prof, k_riemann1<0>, nodes(grow_dir(bxg2, 0, -1), 0), st, qm[0], qp[0], q, F1[0], G1[0], hc);
    go("riemann_normal", prof, k_riemann1<1>, nodes(grow_dir(bxg2, 1, -1), 1), st, qm[1], qp[1], q, F1[1], G1[1], hc);
    go("riemann_normal", prof, k_riemann1<2>, nodes(grow_dir(bxg2, 2, -1), 2), st, qm[2], qp[2], q, F1[2], G1[2], hc);

    const double cdt[3] = {1.0 / 3.0 * dt / hc.dx[0], 1.0 / 3.0 * dt / hc.dx[1],
                           1.0 / 3.0 * dt / hc.dx[2]};
    go("transverse_1", prof, k_trans1<0>, nodes(bxg1, 0), st, qm[0], qp[0], q, F1[1], G1[1], F1[2], G1[2],
       F2[0][1], G2[0][1], F2[0][2], G2[0][2], cdt[1], cdt[2], hc);
    go("transverse_1", prof, k_trans1<1>, nodes(bxg1, 1), st, qm[1], qp[1], q, F1[0], G1[0], F1[2], G1[2],
       F2[1][0], G2[1][0], F2[1][2], G2[1][2], cdt[0], cdt[2], hc);
    go("transverse_1", prof, k_trans1<2>, nodes(bxg1, 2), st, qm[2], qp[2], q, F1[0], G1[0], F1[1], G1[1],
       F2[2][0], G2[2][0], F2[2][1], G2[2][1], cdt[0], cdt[1], hc);

    go("divu", prof, k_divu_c, nodes(nodes(nodes(bx, 0), 1), 2), st, q, divu, hc);
    const double hdt = 0.5 * dt;
    const double hdtd[3] = {0.5 * dt / hc.dx[0], 0.5 * dt / hc.dx[1], 0.5 * dt / hc.dx[2]};
    const double area[3] = {hc.dx[1] * hc.dx[2], hc.dx[0] * hc.dx[2], hc.dx[0] * hc.dx[1]};
    go("transverse_2", prof, k_trans2<0>, nodes(bx, 0), st, qm[0], qp[0], q, sq, has_src,
       F2[1][2], G2[1][2], F2[2][1], G2[2][1], divu, sborder, flx[0], G3[0], hdtd[1], hdtd[2], hdt, area[0], hc);
    go("transverse_2", prof, k_trans2<1>, nodes(bx, 1), st, qm[1], qp[1], q, sq, has_src,
       F2[0][2], G2[0][2], F2[2][0], G2[2][0], divu, sborder, flx[1], G3[1], hdtd[0], hdtd[2], hdt, area[1], hc);
    go("transverse_2", prof, k_trans2<2>, nodes(bx, 2), st, qm[2], qp[2], q, sq, has_src,
       F2[0][1], G2[0][1], F2[1][0], G2[1][0], divu, sborder, flx[2], G3[2], hdtd[0], hdtd[1], hdt, area[2], hc);

    const double vol = hc.dx[0] * hc.dx[1] * hc.dx[2];
    go("consup", prof, k_update_c, bx, st, snew, flx[0], flx[1], flx[2], G3[0], G3[1], G3[2],
       area[0], area[1], area[2], vol, dt, hc.small_dens, status);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

} // namespace camr
