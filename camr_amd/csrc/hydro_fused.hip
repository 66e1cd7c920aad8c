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
//
// All arrays of one box share the canonical layout of canon.h: a thread
// computes one 32-bit byte offset for its cell and reaches components and
// neighbours through uniform offsets.
#include "canon.h"
#include "hydro_fused.h"
#include "hydro_kernels.h"

#include <cstring>

namespace camr {

namespace {

constexpr int ex_(int D, int d) { return D == d ? 1 : 0; }

// occupancy request for the register-heavy face kernels (waves per SIMD)
#ifdef CAMR_TRANS_MINW
#define TRANS_BOUNDS __launch_bounds__(kBlock, CAMR_TRANS_MINW)
#else
#define TRANS_BOUNDS __launch_bounds__(kBlock)
#endif

// byte shift of the unit vector e_D in the canonical layout
template <int D>
__device__ __forceinline__ unsigned shift(const Canon &L) {
    return D == 0 ? 8u : (D == 1 ? L.sy() : L.sz());
}


// ------------------------------------------------------------ primitives
// hydro_ctoprim (Hydro_ctoprim.H:11-101) into the compact layout, with
// hydro_srctoprim (:106-157) when sources are non-zero; also copies Sborder
// into the canonical layout (read back by the artificial viscosity).
__global__ void __launch_bounds__(kBlock)
k_prim_c(Batch B, const DevFab *sb_tab, const DevFab *src_tab, int has_src, HydroConst hc,
         int *status) {
    BATCH_POINT(ST_G4)
    const Arr u = sb_tab[E.id].a;
    const Arr src = has_src ? src_tab[E.id].a : u;
    double *srcq_ = slot_ptr<S_SRCQ>(E);
    const CView uc = cview(L, slot_ptr<S_U>(E), NVAR), q = cview(L, slot_ptr<S_Q>(E), NQC);
    const unsigned v0 = L.off(i, j, k);
    double v[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        v[n] = u(i, j, k, n);
        uc.st(v0, n, v[n]);
    }
    const double rho = v[URHO];
    double o[NQC];
    double e;
    if (!prim_compact(v, o, e, hc, status)) {
        if (hc.eb) {
            // covered cell of an EB level: q = 0 (CAMR_construct_hydro_source.cpp:
            // 116-118), and hydro_srctoprim's zeros for such cells
#pragma unroll
            for (int n = 0; n < NQC; ++n) q.st(v0, n, 0.0);
            if (has_src) {
                const CView sq = cview(L, srcq_, NS);
                sq.st(v0, SR, 0.0);
                sq.st(v0, SU, 0.0);
                sq.st(v0, SV, 0.0);
                sq.st(v0, SW, 0.0);
                sq.st(v0, SP, 0.0);
                sq.st(v0, SE, 0.0);
            }
        }
        return;
    }
#pragma unroll
    for (int n = 0; n < NQC; ++n) q.st(v0, n, o[n]);
    if (has_src) {
        const CView sq = cview(L, srcq_, NS);
        const double rhoinv = 1.0 / rho;
        const double vx = v[UMX] * rhoinv, vy = v[UMY] * rhoinv, vz = v[UMZ] * rhoinv;
        const double srcrho = src(i, j, k, URHO);
        sq.st(v0, SR, srcrho);
        sq.st(v0, SU, (src(i, j, k, UMX) - vx * srcrho) * rhoinv);
        sq.st(v0, SV, (src(i, j, k, UMY) - vy * srcrho) * rhoinv);
        sq.st(v0, SW, (src(i, j, k, UMZ) - vz * srcrho) * rhoinv);
        const double sre = src(i, j, k, UEINT);
        sq.st(v0, SE, sre);
#pragma unroll
        for (int s = 0; s < NSPEC; ++s)
            sq.st(v0, SY + s, (src(i, j, k, UFS + s) - (v[UFS + s] * rhoinv) * srcrho) * rhoinv);
        const double dpde = (kGamma - 1.0) * rho, dpdr = (kGamma - 1.0) * e;
        sq.st(v0, SP, dpde * (sre - (e * rho) * srcrho * rhoinv) * rhoinv + dpdr * srcrho);
    }
}

// ------------------------------------------------------------ flattening
// min over directions of flatten (flatten.H:19-99); 1 when the +-3 stencil
// leaves q's box grow(bx, qgrow).  The canonical primitives cover
// grow(bx, 4); with qgrow > 4 (the 6-ghost q of an EB level) the pressure
// and normal velocity of a stencil point beyond them come from hydro_ctoprim
// of Sborder (u), covered cells (rho <= 0) giving 0 as in k_prim_c.
template <int D, bool WIDE>
__device__ __forceinline__ double flat_dir(const CView &q, const Canon &L, unsigned v0,
                                           int i, int j, int k, const Box &bx, int qgrow,
                                           const Arr &u, const HydroConst &hc, int *status) {
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    const Box qbox = grow(bx, WIDE ? qgrow : CAMR_NGROW), cbox = grow(bx, CAMR_NGROW);
    if (!(qbox.contains(i + 3 * ex, j + 3 * ey, k + 3 * ez) &&
          qbox.contains(i - 3 * ex, j - 3 * ey, k - 3 * ez)))
        return 1.0;
    const unsigned sd = shift<D>(L);
    double p[7], un[7];
#pragma unroll
    for (int o = -3; o <= 3; ++o) {
        const int ii = i + o * ex, jj = j + o * ey, kk = k + o * ez;
        if (!WIDE || cbox.contains(ii, jj, kk)) {
            const unsigned vo = v0 + (unsigned)o * sd;
            p[o + 3] = q.ld(vo, SP);
            un[o + 3] = q.ld(vo, SU + D);
        } else {
            double v[NVAR], w[NQC], e;
#pragma unroll
            for (int n = 0; n < NVAR; ++n) v[n] = u(ii, jj, kk, n);
            if (prim_compact(v, w, e, hc, status)) {
                p[o + 3] = w[SP];
                un[o + 3] = w[SU + D];
            } else {
                p[o + 3] = 0.0;
                un[o + 3] = 0.0;
            }
        }
    }
    return flatten_1d(p, un);
}

// WIDE: q's box is grow(bx, qgrow) with qgrow > 4 (points beyond the
// canonical primitives from sb_tab); else grow(bx, 4)
template <bool WIDE>
__global__ void __launch_bounds__(kBlock)
k_flat_c(Batch B, int use_flattening, int qgrow, const DevFab *sb_tab, HydroConst hc,
         int *status) {
    BATCH_POINT(ST_G2)
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), flat = cview(L, slot_ptr<S_FLAT>(E), 1);
    const Arr u = WIDE ? sb_tab[E.id].a : Arr{};
    const unsigned v0 = L.off(i, j, k);
    double f = 1.0;
    if (use_flattening == 1) {
        f = vmin(f, flat_dir<0, WIDE>(q, L, v0, i, j, k, E.bx, qgrow, u, hc, status));
        f = vmin(f, flat_dir<1, WIDE>(q, L, v0, i, j, k, E.bx, qgrow, u, hc, status));
        f = vmin(f, flat_dir<2, WIDE>(q, L, v0, i, j, k, E.bx, qgrow, u, hc, status));
    }
    flat.st(v0, 0, f);
}

// ------------------------------------------------------------ PPM trace
// trace_ppm (PPM.cpp:11-411) for direction D; qp(i) is the plus state of
// face i, qm(i+1) the minus state of face i+1.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_ppm_c(Batch B, HydroConst hc, const double *dtp) {
    BATCH_POINT(ST_G2)
    const double dtdx = *dtp / hc.dx[D];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), flat = cview(L, slot_ptr<S_FLAT>(E), 1);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    constexpr int SN = SU + D;
    constexpr int T1 = (D == 0) ? SV : SU;
    constexpr int T2 = (D == 2) ? SV : SW;
    const unsigned v0 = L.off(i, j, k), sd = shift<D>(L);
    const double cc = q.ld(v0, SC);
    const double un = q.ld(v0, SN);
    const double fl = flat.ld(v0, 0);
    double Ip[NS][3], Im[NS][3];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o) st[o + 2] = q.ld(v0 + (unsigned)o * sd, s);
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
        qp.st(v0, SR, ro); qp.st(v0, SN, uo); qp.st(v0, SP, po); qp.st(v0, SE, eo);
        qp.st(v0, T1, Im[T1][1]); qp.st(v0, T2, Im[T2][1]);
#pragma unroll
        for (int s = 0; s < NSPEC; ++s) qp.st(v0, SY + s, Im[SY + s][1]);
    }
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Ip[SR][w]; W.un[w] = Ip[SN][w]; W.p[w] = Ip[SP][w];
        W.rhoe[w] = Ip[SE][w];
    }
    {
        double ro, uo, po, eo;
        ppm_project(-1, W, un, cc, hc.small_dens, hc.small_pres, ro, uo, po, eo);
        const unsigned v1 = v0 + sd;
        qm.st(v1, SR, ro); qm.st(v1, SN, uo); qm.st(v1, SP, po); qm.st(v1, SE, eo);
        qm.st(v1, T1, Ip[T1][1]); qm.st(v1, T2, Ip[T2][1]);
#pragma unroll
        for (int s = 0; s < NSPEC; ++s) qm.st(v1, SY + s, Ip[SY + s][1]);
    }
}

// ------------------------------------------------------------ PLM trace
// hydro_plm_d (PLM.H:170-323) for direction D.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_plm_c(Batch B, int has_src, HydroConst hc, const double *dtp) {
    BATCH_POINT(ST_G2)
    const double dtdx = *dtp / hc.dx[D];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), flat = cview(L, slot_ptr<S_FLAT>(E), 1);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const CView sq = cview(L, slot_ptr<S_SRCQ>(E), NS);
    const unsigned v0 = L.off(i, j, k), sd = shift<D>(L);
    const double fl = flat.ld(v0, 0);
    double c[NS], sl[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double st[5];
#pragma unroll
        for (int o = -2; o <= 2; ++o) st[o + 2] = q.ld(v0 + (unsigned)o * sd, s);
        c[s] = st[2];
        if (s == SP && hc.use_pslope) {
            double rr[5], gg[5];
#pragma unroll
            for (int o = -2; o <= 2; ++o) {
                rr[o + 2] = q.ld(v0 + (unsigned)o * sd, SR);
                gg[o + 2] = has_src ? sq.ld(v0 + (unsigned)o * sd, SU + D) : 0.0;
            }
            sl[s] = mc_pslope(st, rr, gg, hc.dx[D], fl, hc.plm_iorder);
        } else {
            sl[s] = mc_slope(st, fl, hc.plm_iorder);
        }
    }
    double op[NS], om[NS];
    plm_trace(D, c, sl, q.ld(v0, SC), dtdx, hc.small_dens, hc.small_pres, op, om);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        qp.st(v0, s, op[s]);
        qm.st(v0 + sd, s, om[s]);
    }
}

// sound speeds of the cells left (offset vl) and right (vr) of a face
__device__ __forceinline__ FaceAux face_aux(const CView &q, unsigned vl, unsigned vr) {
    FaceAux a;
    a.c_l = q.ld(vl, SC);
    a.c_r = q.ld(vr, SC);
    a.csml_l = q.ld(vl, SCS);
    a.csml_r = q.ld(vr, SCS);
    a.gam_l = kGamma;
    a.gam_r = kGamma;
    return a;
}

__device__ __forceinline__ void load_tflux(const CView &F, unsigned v, double *f, double *g) {
    f[FR] = F.ld(v, TR);
    f[FMX] = F.ld(v, TMX);
    f[FMY] = F.ld(v, TMY);
    f[FMZ] = F.ld(v, TMZ);
    f[FE] = F.ld(v, TE);
    f[FEI] = 0.0;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) f[FY + s] = F.ld(v, TY + s);
    g[GUN] = F.ld(v, TGU);
    g[GP] = F.ld(v, TGP);
}

__device__ __forceinline__ void store_tflux(const CView &F, unsigned v, const double *f,
                                            const double *g) {
    F.st(v, TR, f[FR]);
    F.st(v, TMX, f[FMX]);
    F.st(v, TMY, f[FMY]);
    F.st(v, TMZ, f[FMZ]);
    F.st(v, TE, f[FE]);
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) F.st(v, TY + s, f[FY + s]);
    F.st(v, TGU, g[GUN]);
    F.st(v, TGP, g[GP]);
}

// ------------------------------------------------------------ first fluxes
// hydro_cmpflx on the normal predictor states (Godunov_3D.cpp:180-216);
// the boundary overrides of Hydro_cmpflx.H:239-282 are written back to the
// state arrays because the transverse stages read them.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_riemann1(Batch B, HydroConst hc) {
    BATCH_POINT(ST_R1 + D)
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), F = cview(L, slot_ptr<S_F1 + D>(E), NTF);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const unsigned v0 = L.off(i, j, k), sd = shift<D>(L);
    double ql[NS], qr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        ql[s] = qm.ld(v0, s);
        qr[s] = qp.ld(v0, s);
    }
    const FaceAux a = face_aux(q, v0 - sd, v0);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, ql, qr, a, f, g, &mut);
    if (mut & 1) {
        qm.st(v0, SR, ql[SR]); qm.st(v0, SP, ql[SP]); qm.st(v0, SE, ql[SE]);
    }
    if (mut & 2) {
        qp.st(v0, SR, qr[SR]); qp.st(v0, SP, qr[SP]); qp.st(v0, SE, qr[SE]);
    }
    store_tflux(F, v0, f, g);
}

// ------------------------------------------------ transdo + Riemann (pairs)
// For D-face f: the plus state comes from transdo at cell f, the minus state
// from transdo at cell f-e_D (Godunov_utils_3D.H:22-239 writes qp(iv) and
// qm(iv+e_D)); both pairs D|t.
template <int D, int T>
__device__ __forceinline__ void trans1_pair(unsigned v0, const Canon &L, const double *np,
                                            const double *nm, const FaceAux &a, int idx,
                                            const CView &F, const CView &O, double cdt,
                                            const HydroConst &hc) {
    const unsigned sd = shift<D>(L), st_ = shift<T>(L);
    double sp[NS], sm[NS];
    {
        double fl[NF], fr[NF], gl[NG], gr[NG];
        TransDiff t;
        load_tflux(F, v0, fl, gl);
        load_tflux(F, v0 + st_, fr, gr);
        trans_diff_one(fl, fr, gl, gr, cdt, t);
        transverse_one_d(np, t, cdt, kGamma, hc.trd, hc.small_pres, sp);
    }
    {
        double fl[NF], fr[NF], gl[NG], gr[NG];
        TransDiff t;
        load_tflux(F, v0 - sd, fl, gl);
        load_tflux(F, v0 - sd + st_, fr, gr);
        trans_diff_one(fl, fr, gl, gr, cdt, t);
        transverse_one_d(nm, t, cdt, kGamma, hc.trd, hc.small_pres, sm);
    }
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, sm, sp, a, f, g, &mut);
    store_tflux(O, v0, f, g);
}

template <int D>
__global__ void TRANS_BOUNDS
k_trans1(Batch B, HydroConst hc, const double *dtp) {
    BATCH_POINT(ST_T1 + D)
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    const double cda = 1.0 / 3.0 * *dtp / hc.dx[T0];
    const double cdb = 1.0 / 3.0 * *dtp / hc.dx[T1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const unsigned v0 = L.off(i, j, k), sd = shift<D>(L);
    double np[NS], nm[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        np[s] = qp.ld(v0, s);
        nm[s] = qm.ld(v0, s);
    }
    const FaceAux a = face_aux(q, v0 - sd, v0);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    trans1_pair<D, T0>(v0, L, np, nm, a, idx, cview(L, slot_ptr<S_F1 + T0>(E), NTF),
                       cview(L, slot_ptr<f2_slot(D, T0)>(E), NTF), cda, hc);
    trans1_pair<D, T1>(v0, L, np, nm, a, idx, cview(L, slot_ptr<S_F1 + T1>(E), NTF),
                       cview(L, slot_ptr<f2_slot(D, T1)>(E), NTF), cdb, hc);
}

// transdd differences of the cell at offset v for direction D: Fa is the
// flux of direction T0 corrected by T1 (faces of T0), Fb the flux of T1
// corrected by T0.
template <int D>
__device__ __forceinline__ void trans2_diff(unsigned v, const Canon &L, const CView &Fa,
                                            const CView &Fb, double ca, double cb,
                                            TransDiff &t) {
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
    load_tflux(Fa, v, fal, gal);
    load_tflux(Fa, v + shift<T0>(L), far_, gar);
    load_tflux(Fb, v, fbl, gbl);
    load_tflux(Fb, v + shift<T1>(L), fbr, gbr);
    trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, kGamma, t);
}

// ---------------------------------- transdd + final Riemann + adjust_fluxes
// D-face f: plus state from transdd at cell f, minus state from transdd at
// cell f-e_D (Godunov_utils_3D.H:245-482); then hydro_cmpflx, then
// hydro_artif_visc / hydro_norm_spec_flx / hydro_ext_flx
// (Hydro_adjust_fluxes.cpp:13-44).
template <int D, int HAS_SRC>
__global__ void TRANS_BOUNDS
k_trans2(Batch B, HydroConst hc, const double *dtp) {
    BATCH_POINT(ST_T2 + D)
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    const double dt = *dtp;
    const double hdt = 0.5 * dt;
    const double ca = 0.5 * dt / hc.dx[T0], cb = 0.5 * dt / hc.dx[T1];
    const double area = (D == 0) ? hc.dx[1] * hc.dx[2]
                      : (D == 1) ? hc.dx[0] * hc.dx[2] : hc.dx[0] * hc.dx[1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), sq = cview(L, slot_ptr<S_SRCQ>(E), NS);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const CView Fa = cview(L, slot_ptr<f2_slot(T0, T1)>(E), NTF);
    const CView Fb = cview(L, slot_ptr<f2_slot(T1, T0)>(E), NTF);
    const unsigned v0 = L.off(i, j, k), sd = shift<D>(L);
    double sp[NS], sm[NS];
    {
        double n[NS], src[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            n[s] = qp.ld(v0, s);
            src[s] = HAS_SRC ? sq.ld(v0, s) : 0.0;
        }
        TransDiff t;
        trans2_diff<D>(v0, L, Fa, Fb, ca, cb, t);
        transverse_two_d(n, t, hdt, src, hc.trd, hc.small_pres, sp);
    }
    {
        const unsigned vm = v0 - sd;
        double n[NS], src[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            n[s] = qm.ld(v0, s);
            src[s] = HAS_SRC ? sq.ld(vm, s) : 0.0;
        }
        TransDiff t;
        trans2_diff<D>(vm, L, Fa, Fb, ca, cb, t);
        transverse_two_d(n, t, hdt, src, hc.trd, hc.small_pres, sm);
    }
    const FaceAux a = face_aux(q, v0 - sd, v0);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, sm, sp, a, f, g, &mut);
    double fu[NVAR];
    fu[URHO] = f[FR]; fu[UMX] = f[FMX]; fu[UMY] = f[FMY]; fu[UMZ] = f[FMZ];
    fu[UEDEN] = f[FE]; fu[UEINT] = f[FEI]; fu[UTEMP] = 0.0;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) fu[UFS + s] = f[FY + s];
    adjust_face<D>(fu, cview(L, slot_ptr<S_DIVU>(E), 1), cview(L, slot_ptr<S_U>(E), NVAR), i, j,
                   k, area, hc);
    const CView flx = cview(L, slot_ptr<S_FLX + D>(E), NXF);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx.st(v0, n, fu[n]);
    flx.st(v0, XGU, g[GUN]);
    flx.st(v0, XGP, g[GP]);
}

// ------------------------------------------------ z-marching z-face kernels
// The z-face transverse stages with each thread walking a column of faces
// k0..k1: the minus side of face k belongs to cell k-1, the plus side of
// face k-1, so its transverse differences, sound speeds and the values the
// flux adjustment reads come from the previous iteration's registers
// instead of a second load of the plane below.  Same per-point arithmetic
// as k_trans1<2> / k_trans2<2, S>, hence the same results.

// adjust_face<D> (hydro_artif_visc, hydro_norm_spec_flx, hydro_ext_flx) on
// values: d4 the face's four nodal divergences in adjust_face's order (the
// face's node, +e_T0, +e_T1, +e_T0+e_T1); uc/um the conserved state of the
// cells above and below the face.
template <int D>
__device__ __forceinline__ void adjust_face_vals(double *f, const double *d4,
                                                 const double *uc, const double *um, int idx,
                                                 double area, const HydroConst &hc) {
    const double div = hc.difmag * vmin(0.0, 0.25 * (d4[0] + d4[1] + d4[2] + d4[3]));
    const bool at_bndry = (idx == hc.domlo[D] && hc.bclo[D] == BC_NOSLIPWALL) ||
                          (idx == hc.domhi[D] + 1 && hc.bchi[D] == BC_NOSLIPWALL);
    const double dxd = hc.dx[D];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        if (n == UTEMP) continue;
        const bool tang = (n == UMX || n == UMY || n == UMZ) && (n != UMX + D);
        if (!(at_bndry && tang)) f[n] += dxd * div * (uc[n] - um[n]);
    }
    f[UTEMP] = 0.0;
    double summ = 0.0;
    summ += f[UFS];
    summ += f[UFS + 1];
    summ += f[UFS + 2];
    const double fac = (summ != 0.0) ? f[URHO] / summ : 1.0;
    f[UFS] *= fac;
    f[UFS + 1] *= fac;
    f[UFS + 2] *= fac;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) f[n] *= area;
}

// transdo differences of the cell at offset v in direction T (faces v and
// v + e_T of F)
template <int T>
__device__ __forceinline__ void trans1_diff(const CView &F, unsigned v, const Canon &L,
                                            double cdt, TransDiff &t) {
    double fl[NF], fr[NF], gl[NG], gr[NG];
    load_tflux(F, v, fl, gl);
    load_tflux(F, v + shift<T>(L), fr, gr);
    trans_diff_one(fl, fr, gl, gr, cdt, t);
}

// waves per SIMD requested for the z-marched transdo kernel: 2 keeps every
// value in registers; 3 trades a few spilled bytes per lane for occupancy
#ifndef CAMR_T1ZM_MINW
#define CAMR_T1ZM_MINW 2
#endif
__global__ void __launch_bounds__(kBlock, CAMR_T1ZM_MINW)
k_trans1_zm(Batch B, HydroConst hc, const double *dtp) {
    BATCH_COLUMN(ST_T1 + 2)
    const double cda = 1.0 / 3.0 * *dtp / hc.dx[0];
    const double cdb = 1.0 / 3.0 * *dtp / hc.dx[1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + 2>(E), NS), qp = cview(L, slot_ptr<S_QP0 + 2>(E), NS);
    const CView Fx = cview(L, slot_ptr<S_F1 + 0>(E), NTF), Fy = cview(L, slot_ptr<S_F1 + 1>(E), NTF);
    const CView Ox = cview(L, slot_ptr<f2_slot(2, 0)>(E), NTF);
    const CView Oy = cview(L, slot_ptr<f2_slot(2, 1)>(E), NTF);
    // the cell below the first face
    TransDiff tmx, tmy;
    unsigned v = L.off(i, j, k0 - 1);
    trans1_diff<0>(Fx, v, L, cda, tmx);
    trans1_diff<1>(Fy, v, L, cdb, tmy);
    double cm = q.ld(v, SC), csm = q.ld(v, SCS);
    for (int k = k0; k <= k1; ++k) {
        v = L.off(i, j, k);
        double np[NS], nm[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            np[s] = qp.ld(v, s);
            nm[s] = qm.ld(v, s);
        }
        FaceAux a;
        a.c_l = cm;
        a.csml_l = csm;
        a.c_r = q.ld(v, SC);
        a.csml_r = q.ld(v, SCS);
        a.gam_l = kGamma;
        a.gam_r = kGamma;
        {
            double sp[NS], sm[NS], f[NF], g[NG];
            int mut;
            TransDiff tp;
            trans1_diff<0>(Fx, v, L, cda, tp);
            transverse_one_d(np, tp, cda, kGamma, hc.trd, hc.small_pres, sp);
            transverse_one_d(nm, tmx, cda, kGamma, hc.trd, hc.small_pres, sm);
            tmx = tp;
            face_flux(2, k, hc, sm, sp, a, f, g, &mut);
            store_tflux(Ox, v, f, g);
        }
        {
            double sp[NS], sm[NS], f[NF], g[NG];
            int mut;
            TransDiff tp;
            trans1_diff<1>(Fy, v, L, cdb, tp);
            transverse_one_d(np, tp, cdb, kGamma, hc.trd, hc.small_pres, sp);
            transverse_one_d(nm, tmy, cdb, kGamma, hc.trd, hc.small_pres, sm);
            tmy = tp;
            face_flux(2, k, hc, sm, sp, a, f, g, &mut);
            store_tflux(Oy, v, f, g);
        }
        cm = a.c_r;
        csm = a.csml_r;
    }
}

template <int HAS_SRC>
__global__ void TRANS_BOUNDS
k_trans2_zm(Batch B, HydroConst hc, const double *dtp) {
    BATCH_COLUMN(ST_T2 + 2)
    const double dt = *dtp;
    const double hdt = 0.5 * dt;
    const double ca = 0.5 * dt / hc.dx[0], cb = 0.5 * dt / hc.dx[1];
    const double area = hc.dx[0] * hc.dx[1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), sq = cview(L, slot_ptr<S_SRCQ>(E), NS);
    const CView qm = cview(L, slot_ptr<S_QM0 + 2>(E), NS), qp = cview(L, slot_ptr<S_QP0 + 2>(E), NS);
    const CView Fa = cview(L, slot_ptr<f2_slot(0, 1)>(E), NTF);
    const CView Fb = cview(L, slot_ptr<f2_slot(1, 0)>(E), NTF);
    const CView dv = cview(L, slot_ptr<S_DIVU>(E), 1), uc = cview(L, slot_ptr<S_U>(E), NVAR);
    const CView flx = cview(L, slot_ptr<S_FLX + 2>(E), NXF);
    const unsigned sx = 8u, sy = L.sy();
    // the cell below the first face
    TransDiff tm;
    double srcm[NS], um[NVAR];
    unsigned v = L.off(i, j, k0 - 1);
    trans2_diff<2>(v, L, Fa, Fb, ca, cb, tm);
#pragma unroll
    for (int s = 0; s < NS; ++s) srcm[s] = HAS_SRC ? sq.ld(v, s) : 0.0;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) um[n] = uc.ld(v, n);
    double cm = q.ld(v, SC), csm = q.ld(v, SCS);
    for (int k = k0; k <= k1; ++k) {
        v = L.off(i, j, k);
        TransDiff tp;
        trans2_diff<2>(v, L, Fa, Fb, ca, cb, tp);
        double srcp[NS], sp[NS], sm[NS];
        {
            double n[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                n[s] = qp.ld(v, s);
                srcp[s] = HAS_SRC ? sq.ld(v, s) : 0.0;
            }
            transverse_two_d(n, tp, hdt, srcp, hc.trd, hc.small_pres, sp);
        }
        {
            double n[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) n[s] = qm.ld(v, s);
            transverse_two_d(n, tm, hdt, srcm, hc.trd, hc.small_pres, sm);
        }
        FaceAux a;
        a.c_l = cm;
        a.csml_l = csm;
        a.c_r = q.ld(v, SC);
        a.csml_r = q.ld(v, SCS);
        a.gam_l = kGamma;
        a.gam_r = kGamma;
        double f[NF], g[NG];
        int mut;
        face_flux(2, k, hc, sm, sp, a, f, g, &mut);
        double fu[NVAR];
        fu[URHO] = f[FR]; fu[UMX] = f[FMX]; fu[UMY] = f[FMY]; fu[UMZ] = f[FMZ];
        fu[UEDEN] = f[FE]; fu[UEINT] = f[FEI]; fu[UTEMP] = 0.0;
#pragma unroll
        for (int s = 0; s < NSPEC; ++s) fu[UFS + s] = f[FY + s];
        double up[NVAR], d4[4];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) up[n] = uc.ld(v, n);
        d4[0] = dv.ld(v, 0);
        d4[1] = dv.ld(v + sx, 0);
        d4[2] = dv.ld(v + sy, 0);
        d4[3] = dv.ld(v + sx + sy, 0);
        adjust_face_vals<2>(fu, d4, up, um, k, area, hc);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) flx.st(v, n, fu[n]);
        flx.st(v, XGU, g[GUN]);
        flx.st(v, XGP, g[GP]);
        tm = tp;
#pragma unroll
        for (int s = 0; s < NS; ++s) srcm[s] = srcp[s];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) um[n] = up[n];
        cm = a.c_r;
        csm = a.csml_r;
    }
}

// ------------------------------------------------ x-face kernels, lane-shared
// A wavefront covers consecutive x points of a row, so the cell left of
// lane l's x-face is the cell right of lane l-1's: its transverse
// differences, sound speeds, sources and state come from lane l-1 through a
// cross-lane move instead of a second load.  The first lane of a wavefront
// and the first face of a row load them.  Same per-point arithmetic as
// k_trans1<0> / k_trans2<0, S>, hence the same results.

__device__ __forceinline__ double lane_up(double x) { return __shfl_up(x, 1, 64); }

// True when lane l-1 holds the point (i-1, j, k): the blocks take
// consecutive x points (linear blocks, or the rows of a tile of width
// 2^lx >= 8), except at a wavefront's first lane, a row's first point and
// a tile row's first point.
__device__ __forceinline__ bool left_in_lane(const Batch &B, int i, int lo0) {
    const int t = (int)threadIdx.x;
    if ((t & 63) == 0 || i == lo0) return false;
    return B.lx < 0 || (t & ((1 << B.lx) - 1)) != 0;
}

__device__ __forceinline__ void lane_up_td1(const TransDiff &a, TransDiff &b) {
    b.rho = lane_up(a.rho);
    b.mu = lane_up(a.mu);
    b.mv = lane_up(a.mv);
    b.mw = lane_up(a.mw);
    b.e = lane_up(a.e);
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) b.y[s] = lane_up(a.y[s]);
    b.dAup = lane_up(a.dAup);
    b.pav = lane_up(a.pav);
    b.dAu = lane_up(a.dAu);
}
// the fields transverse_two_d reads: the flux differences and the two
// pressure corrections (dAup holds the first, pnew the second)
__device__ __forceinline__ void lane_up_td2(const TransDiff &a, TransDiff &b) {
    b.rho = lane_up(a.rho);
    b.mu = lane_up(a.mu);
    b.mv = lane_up(a.mv);
    b.mw = lane_up(a.mw);
    b.e = lane_up(a.e);
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) b.y[s] = lane_up(a.y[s]);
    b.dAup = lane_up(a.dAup);
    b.pnew = lane_up(a.pnew);
}

__global__ void TRANS_BOUNDS
k_trans1_x(Batch B, HydroConst hc, const double *dtp) {
    BATCH_POINT(ST_T1 + 0)
    const double cda = 1.0 / 3.0 * *dtp / hc.dx[1];
    const double cdb = 1.0 / 3.0 * *dtp / hc.dx[2];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + 0>(E), NS), qp = cview(L, slot_ptr<S_QP0 + 0>(E), NS);
    const CView Fy = cview(L, slot_ptr<S_F1 + 1>(E), NTF), Fz = cview(L, slot_ptr<S_F1 + 2>(E), NTF);
    const CView Oy = cview(L, slot_ptr<f2_slot(0, 1)>(E), NTF);
    const CView Oz = cview(L, slot_ptr<f2_slot(0, 2)>(E), NTF);
    const unsigned v0 = L.off(i, j, k), vm = v0 - 8u;
    const bool own = !left_in_lane(B, i, stage_box(ST_T1 + 0, E.bx).lo[0]);
    double np[NS], nm[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        np[s] = qp.ld(v0, s);
        nm[s] = qm.ld(v0, s);
    }
    FaceAux a;
    a.c_r = q.ld(v0, SC);
    a.csml_r = q.ld(v0, SCS);
    a.c_l = lane_up(a.c_r);
    a.csml_l = lane_up(a.csml_r);
    if (own) {
        a.c_l = q.ld(vm, SC);
        a.csml_l = q.ld(vm, SCS);
    }
    a.gam_l = kGamma;
    a.gam_r = kGamma;
    {
        double sp[NS], sm[NS], f[NF], g[NG];
        int mut;
        TransDiff tp, tm;
        trans1_diff<1>(Fy, v0, L, cda, tp);
        lane_up_td1(tp, tm);
        if (own) trans1_diff<1>(Fy, vm, L, cda, tm);
        transverse_one_d(np, tp, cda, kGamma, hc.trd, hc.small_pres, sp);
        transverse_one_d(nm, tm, cda, kGamma, hc.trd, hc.small_pres, sm);
        face_flux(0, i, hc, sm, sp, a, f, g, &mut);
        store_tflux(Oy, v0, f, g);
    }
    {
        double sp[NS], sm[NS], f[NF], g[NG];
        int mut;
        TransDiff tp, tm;
        trans1_diff<2>(Fz, v0, L, cdb, tp);
        lane_up_td1(tp, tm);
        if (own) trans1_diff<2>(Fz, vm, L, cdb, tm);
        transverse_one_d(np, tp, cdb, kGamma, hc.trd, hc.small_pres, sp);
        transverse_one_d(nm, tm, cdb, kGamma, hc.trd, hc.small_pres, sm);
        face_flux(0, i, hc, sm, sp, a, f, g, &mut);
        store_tflux(Oz, v0, f, g);
    }
}

template <int HAS_SRC>
__global__ void TRANS_BOUNDS
k_trans2_x(Batch B, HydroConst hc, const double *dtp) {
    BATCH_POINT(ST_T2 + 0)
    const double dt = *dtp;
    const double hdt = 0.5 * dt;
    const double ca = 0.5 * dt / hc.dx[1], cb = 0.5 * dt / hc.dx[2];
    const double area = hc.dx[1] * hc.dx[2];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), sq = cview(L, slot_ptr<S_SRCQ>(E), NS);
    const CView qm = cview(L, slot_ptr<S_QM0 + 0>(E), NS), qp = cview(L, slot_ptr<S_QP0 + 0>(E), NS);
    const CView Fa = cview(L, slot_ptr<f2_slot(1, 2)>(E), NTF);
    const CView Fb = cview(L, slot_ptr<f2_slot(2, 1)>(E), NTF);
    const CView dv = cview(L, slot_ptr<S_DIVU>(E), 1), uc = cview(L, slot_ptr<S_U>(E), NVAR);
    const unsigned v0 = L.off(i, j, k), vm = v0 - 8u, sy = L.sy(), sz = L.sz();
    const bool own = !left_in_lane(B, i, stage_box(ST_T2 + 0, E.bx).lo[0]);
    double sp[NS], sm[NS];
    {
        double n[NS], srcp[NS], srcm[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            srcp[s] = HAS_SRC ? sq.ld(v0, s) : 0.0;
            srcm[s] = HAS_SRC ? lane_up(srcp[s]) : 0.0;
        }
        if (HAS_SRC && own) {
#pragma unroll
            for (int s = 0; s < NS; ++s) srcm[s] = sq.ld(vm, s);
        }
        TransDiff tp, tm;
        trans2_diff<0>(v0, L, Fa, Fb, ca, cb, tp);
        lane_up_td2(tp, tm);
        if (own) trans2_diff<0>(vm, L, Fa, Fb, ca, cb, tm);
#pragma unroll
        for (int s = 0; s < NS; ++s) n[s] = qp.ld(v0, s);
        transverse_two_d(n, tp, hdt, srcp, hc.trd, hc.small_pres, sp);
#pragma unroll
        for (int s = 0; s < NS; ++s) n[s] = qm.ld(v0, s);
        transverse_two_d(n, tm, hdt, srcm, hc.trd, hc.small_pres, sm);
    }
    FaceAux a;
    a.c_r = q.ld(v0, SC);
    a.csml_r = q.ld(v0, SCS);
    a.c_l = lane_up(a.c_r);
    a.csml_l = lane_up(a.csml_r);
    if (own) {
        a.c_l = q.ld(vm, SC);
        a.csml_l = q.ld(vm, SCS);
    }
    a.gam_l = kGamma;
    a.gam_r = kGamma;
    double f[NF], g[NG];
    int mut;
    face_flux(0, i, hc, sm, sp, a, f, g, &mut);
    double fu[NVAR];
    fu[URHO] = f[FR]; fu[UMX] = f[FMX]; fu[UMY] = f[FMY]; fu[UMZ] = f[FMZ];
    fu[UEDEN] = f[FE]; fu[UEINT] = f[FEI]; fu[UTEMP] = 0.0;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) fu[UFS + s] = f[FY + s];
    double up[NVAR], um[NVAR], d4[4];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        up[n] = uc.ld(v0, n);
        um[n] = lane_up(up[n]);
    }
    if (own) {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) um[n] = uc.ld(vm, n);
    }
    d4[0] = dv.ld(v0, 0);
    d4[1] = dv.ld(v0 + sy, 0);
    d4[2] = dv.ld(v0 + sz, 0);
    d4[3] = dv.ld(v0 + sy + sz, 0);
    adjust_face_vals<0>(fu, d4, up, um, i, area, hc);
    const CView flx = cview(L, slot_ptr<S_FLX + 0>(E), NXF);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx.st(v0, n, fu[n]);
    flx.st(v0, XGU, g[GUN]);
    flx.st(v0, XGP, g[GP]);
}

// ------------------------------------------------ LDS-staged CTU face kernels
// k_ctu_x / k_ctu_y / k_ctu_z: the transdo pairs, transdd and the final
// Riemann solve of one face direction D in one kernel, so that the
// transdo-corrected fluxes (the S_F2 slots) never go to memory.  A block
// owns a tile of kCtuX x kCtuY points of the x-y plane and walks B.zm
// planes in z.  Per plane:
//   1. each thread solves the transdo pairs on the low faces of its cell:
//      the in-plane transverse direction into LDS, the z direction in
//      registers (face k+1 of one plane is face k of the next); the faces
//      on the tile's far edge and the halo cells' faces go to a few
//      designated threads;
//   2. after a barrier, each thread forms the transdd differences of its
//      cell from the pair fluxes on the cell's four transverse faces;
//   3. after a barrier, the D-face between the thread's cell and the cell
//      below it in D (LDS for x and y, the previous plane for z) is finished
//      as in k_trans2: transdd states, Riemann solve, adjust_fluxes, store.
// Every value is the same per-point arithmetic as k_trans1 + k_trans2 on
// the same inputs, hence the same results.

// transdo pair D|T at the D-face of offset v (index idx along D): the
// arithmetic of trans1_pair without the store; f/g the compact flux
__device__ __forceinline__ void rec_put(double *s, int n, int p, const double *f,
                                        const double *g) {
    s[TR * n + p] = f[FR];
    s[TMX * n + p] = f[FMX];
    s[TMY * n + p] = f[FMY];
    s[TMZ * n + p] = f[FMZ];
    s[TE * n + p] = f[FE];
#pragma unroll
    for (int c = 0; c < NSPEC; ++c) s[(TY + c) * n + p] = f[FY + c];
    s[TGU * n + p] = g[GUN];
    s[TGP * n + p] = g[GP];
}
__device__ __forceinline__ void rec_get(const double *s, int n, int p, double *f, double *g) {
    f[FR] = s[TR * n + p];
    f[FMX] = s[TMX * n + p];
    f[FMY] = s[TMY * n + p];
    f[FMZ] = s[TMZ * n + p];
    f[FE] = s[TE * n + p];
    f[FEI] = 0.0;
#pragma unroll
    for (int c = 0; c < NSPEC; ++c) f[FY + c] = s[(TY + c) * n + p];
    g[GUN] = s[TGU * n + p];
    g[GP] = s[TGP * n + p];
}
// the transdd fields of a TransDiff (what transverse_two_d reads)
enum { NTD = 10 };
__device__ __forceinline__ void td_put(double *s, int n, int p, const TransDiff &t) {
    s[0 * n + p] = t.rho;
    s[1 * n + p] = t.mu;
    s[2 * n + p] = t.mv;
    s[3 * n + p] = t.mw;
    s[4 * n + p] = t.e;
#pragma unroll
    for (int c = 0; c < NSPEC; ++c) s[(5 + c) * n + p] = t.y[c];
    s[8 * n + p] = t.dAup;
    s[9 * n + p] = t.pnew;
}
__device__ __forceinline__ void td_get(const double *s, int n, int p, TransDiff &t) {
    t.rho = s[0 * n + p];
    t.mu = s[1 * n + p];
    t.mv = s[2 * n + p];
    t.mw = s[3 * n + p];
    t.e = s[4 * n + p];
#pragma unroll
    for (int c = 0; c < NSPEC; ++c) t.y[c] = s[(5 + c) * n + p];
    t.dAup = s[8 * n + p];
    t.pnew = s[9 * n + p];
}

template <int D, int T>
__device__ __forceinline__ void ctu_pair(const BoxEnt &E, const Canon &L, unsigned v, int idx,
                                         double cdt, const HydroConst &hc, double *f,
                                         double *g) {
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const CView F = cview(L, slot_ptr<S_F1 + T>(E), NTF);
    const unsigned sd = shift<D>(L);
    double sp[NS], sm[NS];
    {
        double n[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) n[s] = qp.ld(v, s);
        TransDiff t;
        trans1_diff<T>(F, v, L, cdt, t);
        transverse_one_d(n, t, cdt, kGamma, hc.trd, hc.small_pres, sp);
    }
    {
        double n[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) n[s] = qm.ld(v, s);
        TransDiff t;
        trans1_diff<T>(F, v - sd, L, cdt, t);
        transverse_one_d(n, t, cdt, kGamma, hc.trd, hc.small_pres, sm);
    }
    const FaceAux a = face_aux(q, v - sd, v);
    int mut;
    face_flux(D, idx, hc, sm, sp, a, f, g, &mut);
}

// transdd states of the D-face at offset v of point (i,j,k) from the
// differences tp of the cell above the face and tm of the cell below it,
// then the Riemann solve, adjust_fluxes and the store of the final flux
// (the tail of k_trans2)
template <int D, int HAS_SRC>
__device__ __forceinline__ void ctu_face(const BoxEnt &E, const Canon &L, unsigned v, int i,
                                         int j, int k, const TransDiff &tp,
                                         const TransDiff &tm, double hdt, double area,
                                         const HydroConst &hc) {
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), sq = cview(L, slot_ptr<S_SRCQ>(E), NS);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const unsigned sd = shift<D>(L);
    double sp[NS], sm[NS];
    {
        double n[NS], src[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            n[s] = qp.ld(v, s);
            src[s] = HAS_SRC ? sq.ld(v, s) : 0.0;
        }
        transverse_two_d(n, tp, hdt, src, hc.trd, hc.small_pres, sp);
    }
    {
        double n[NS], src[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            n[s] = qm.ld(v, s);
            src[s] = HAS_SRC ? sq.ld(v - sd, s) : 0.0;
        }
        transverse_two_d(n, tm, hdt, src, hc.trd, hc.small_pres, sm);
    }
    const FaceAux a = face_aux(q, v - sd, v);
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    double f[NF], g[NG];
    int mut;
    face_flux(D, idx, hc, sm, sp, a, f, g, &mut);
    double fu[NVAR];
    fu[URHO] = f[FR]; fu[UMX] = f[FMX]; fu[UMY] = f[FMY]; fu[UMZ] = f[FMZ];
    fu[UEDEN] = f[FE]; fu[UEINT] = f[FEI]; fu[UTEMP] = 0.0;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) fu[UFS + s] = f[FY + s];
    adjust_face<D>(fu, cview(L, slot_ptr<S_DIVU>(E), 1), cview(L, slot_ptr<S_U>(E), NVAR), i, j,
                   k, area, hc);
    const CView flx = cview(L, slot_ptr<S_FLX + D>(E), NXF);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx.st(v, n, fu[n]);
    flx.st(v, XGU, g[GUN]);
    flx.st(v, XGP, g[GP]);
}

#ifndef CAMR_CTU_MINW
#define CAMR_CTU_MINW 3
#endif
#define CTU_BOUNDS __launch_bounds__(kCtuBlock, CAMR_CTU_MINW)

// First statements of the three kernels: the block's box, stage range sb
// and tile T; (tx, ty) the thread's point in the tile.
#define CTU_TILE(D)                                                            \
    if (B.halt && *B.halt) return;                                             \
    int lblk_;                                                                 \
    const BoxEnt &E = B.tab[batch_box(B, lblk_)];                              \
    const Canon L = E.L;                                                       \
    const Box sb = stage_box(ST_T2 + (D), E.bx);                               \
    const CtuTile T = ctu_tile(sb, lblk_, B.zm, (D) == 0);                     \
    const int t = (int)threadIdx.x, tx = t % kCtuX, ty = t / kCtuX;            \
    const int i = T.i0 + tx, j = T.j0 + ty;                                    \
    const int hx = sb.hi[0], hy = sb.hi[1];

// z-faces.  Pair records: x|y on x-faces (columns 0..kCtuX of each row),
// y|x on y-faces (rows 0..kCtuY of each column); the transdd differences
// of the cell below the face come from the previous plane.
template <int HAS_SRC>
__global__ void CTU_BOUNDS k_ctu_z(Batch B, HydroConst hc, const double *dtp) {
    CTU_TILE(2)
    constexpr int NA = kCtuY * (kCtuX + 1), NB = (kCtuY + 1) * kCtuX;
    __shared__ double sA[NTF * NA], sB[NTF * NB];
    const double dt = *dtp;
    const double cdx = 1.0 / 3.0 * dt / hc.dx[0], cdy = 1.0 / 3.0 * dt / hc.dx[1];
    const double ca = 0.5 * dt / hc.dx[0], cb = 0.5 * dt / hc.dx[1];
    const double hdt = 0.5 * dt, area = hc.dx[0] * hc.dx[1];
    const bool cell = i <= hx && j <= hy;
    const bool ownA = i <= hx + 1 && j <= hy;  // x|y on the cell's low x-face
    const bool ownB = i <= hx && j <= hy + 1;  // y|x on the cell's low y-face
    // far edges: y|x on row kCtuY (wavefront 0), x|y on column kCtuX (wavefront 1)
    const int eB = t, eA = t - 64;
    const bool extB = eB < kCtuX && T.i0 + eB <= hx && T.j0 + kCtuY <= hy + 1;
    const bool extA = eA >= 0 && eA < kCtuY && T.i0 + kCtuX <= hx + 1 && T.j0 + eA <= hy;
    TransDiff tm = {};
    for (int k = T.k0 - 1; k <= T.k1; ++k) {
        if (ownA) {
            double f[NF], g[NG];
            ctu_pair<0, 1>(E, L, L.off(i, j, k), i, cdy, hc, f, g);
            rec_put(sA, NA, ty * (kCtuX + 1) + tx, f, g);
        }
        if (ownB) {
            double f[NF], g[NG];
            ctu_pair<1, 0>(E, L, L.off(i, j, k), j, cdx, hc, f, g);
            rec_put(sB, NB, ty * kCtuX + tx, f, g);
        }
        if (extB) {
            double f[NF], g[NG];
            const int jb = T.j0 + kCtuY;
            ctu_pair<1, 0>(E, L, L.off(T.i0 + eB, jb, k), jb, cdx, hc, f, g);
            rec_put(sB, NB, kCtuY * kCtuX + eB, f, g);
        }
        if (extA) {
            double f[NF], g[NG];
            const int ia = T.i0 + kCtuX;
            ctu_pair<0, 1>(E, L, L.off(ia, T.j0 + eA, k), ia, cdy, hc, f, g);
            rec_put(sA, NA, eA * (kCtuX + 1) + kCtuX, f, g);
        }
        __syncthreads();
        TransDiff tp = {};
        if (cell) {
            double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
            rec_get(sA, NA, ty * (kCtuX + 1) + tx, fal, gal);
            rec_get(sA, NA, ty * (kCtuX + 1) + tx + 1, far_, gar);
            rec_get(sB, NB, ty * kCtuX + tx, fbl, gbl);
            rec_get(sB, NB, (ty + 1) * kCtuX + tx, fbr, gbr);
            trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, kGamma, tp);
        }
        __syncthreads();
        if (cell && k >= T.k0)
            ctu_face<2, HAS_SRC>(E, L, L.off(i, j, k), i, j, k, tp, tm, hdt, area, hc);
        tm = tp;
    }
}

// x-faces.  The thread at column tx owns face i = i0 + tx and the cell
// right of it; the cell left of the tile's first face (column 0 of the
// LDS arrays, the cells sitting at columns 1..kCtuX) is a halo cell, and
// a merged last tile (T.wx = kCtuX + 1) has a second one right of its
// last cell (column kCtuX + 1) with one more face.  y|z pairs on y-faces
// in LDS (rows 0..kCtuY), z|y pairs on z-faces in registers (halo cells:
// in LDS, two planes), transdd differences in LDS.
template <int HAS_SRC>
__global__ void CTU_BOUNDS k_ctu_x(Batch B, HydroConst hc, const double *dtp) {
    CTU_TILE(0)
    constexpr int NC = kCtuX + 2;
    constexpr int NY = (kCtuY + 1) * NC, NT = kCtuY * NC, NH = 2 * kCtuY;
    __shared__ double sY[NTF * NY], sT[NTD * NT], sH[2 * NTF * NH];
    const double dt = *dtp;
    const double cdy = 1.0 / 3.0 * dt / hc.dx[1], cdz = 1.0 / 3.0 * dt / hc.dx[2];
    const double ca = 0.5 * dt / hc.dx[1], cb = 0.5 * dt / hc.dx[2];
    const double hdt = 0.5 * dt, area = hc.dx[1] * hc.dx[2];
    const int wx = T.wx < kCtuX ? T.wx : kCtuX;  // own columns of this tile
    const bool face = tx < wx && j <= hy;          // face (i, j) and the cell right of it
    const bool ownY = tx < wx && j <= hy + 1;      // y|z on the cell's low y-face
    // y|z on far or halo faces (wavefront 0): row kCtuY of the own columns,
    // then the left halo column (rows 0..kCtuY), then the right one
    int ye_i = 0, ye_j = 0, ye_p = 0;
    bool extY = false;
    if (t < kCtuX) {
        ye_i = T.i0 + t; ye_j = T.j0 + kCtuY; ye_p = kCtuY * NC + t + 1;
        extY = t < wx && ye_j <= hy + 1;
    } else if (t < kCtuX + kCtuY + 1) {
        const int r = t - kCtuX;
        ye_i = T.i0 - 1; ye_j = T.j0 + r; ye_p = r * NC;
        extY = ye_j <= hy + 1;
    } else if (t < kCtuX + 2 * (kCtuY + 1)) {
        const int r = t - kCtuX - (kCtuY + 1);
        ye_i = T.i0 + kCtuX; ye_j = T.j0 + r; ye_p = r * NC + kCtuX + 1;
        extY = T.wx == kCtuX + 1 && ye_j <= hy + 1;
    }
    // halo cells (wavefront 1): left column rows 0..kCtuY-1, then the right one
    const int h = t - 64;
    const bool halo = h >= 0 && h < 2 * kCtuY &&
                      (h < kCtuY ? T.j0 + h <= hy : (T.wx == kCtuX + 1 && T.j0 + h - kCtuY <= hy));
    const int hc_i = h < kCtuY ? T.i0 - 1 : T.i0 + kCtuX;
    const int hc_j = T.j0 + (h < kCtuY ? h : h - kCtuY);
    const int hc_col = h < kCtuY ? 0 : kCtuX + 1;
    double zlo[NF], zglo[NG];
    if (face) ctu_pair<2, 1>(E, L, L.off(i, j, T.k0), T.k0, cdy, hc, zlo, zglo);
    if (halo) {
        double f[NF], g[NG];
        ctu_pair<2, 1>(E, L, L.off(hc_i, hc_j, T.k0), T.k0, cdy, hc, f, g);
        rec_put(sH + (T.k0 & 1) * NTF * NH, NH, h, f, g);
    }
    for (int k = T.k0; k <= T.k1; ++k) {
        double zhi[NF], zghi[NG];
        if (face) ctu_pair<2, 1>(E, L, L.off(i, j, k + 1), k + 1, cdy, hc, zhi, zghi);
        if (ownY) {
            double f[NF], g[NG];
            ctu_pair<1, 2>(E, L, L.off(i, j, k), j, cdz, hc, f, g);
            rec_put(sY, NY, ty * NC + tx + 1, f, g);
        }
        if (extY) {
            double f[NF], g[NG];
            ctu_pair<1, 2>(E, L, L.off(ye_i, ye_j, k), ye_j, cdz, hc, f, g);
            rec_put(sY, NY, ye_p, f, g);
        }
        if (halo) {
            double f[NF], g[NG];
            ctu_pair<2, 1>(E, L, L.off(hc_i, hc_j, k + 1), k + 1, cdy, hc, f, g);
            rec_put(sH + ((k + 1) & 1) * NTF * NH, NH, h, f, g);
        }
        __syncthreads();
        TransDiff tp = {};
        if (face) {
            double fal[NF], far_[NF], gal[NG], gar[NG];
            rec_get(sY, NY, ty * NC + tx + 1, fal, gal);
            rec_get(sY, NY, (ty + 1) * NC + tx + 1, far_, gar);
            trans_diff_two(fal, far_, gal, gar, zlo, zhi, zglo, zghi, ca, cb, kGamma, tp);
            td_put(sT, NT, ty * NC + tx + 1, tp);
        }
        if (halo) {
            const int r = hc_j - T.j0;
            double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
            rec_get(sY, NY, r * NC + hc_col, fal, gal);
            rec_get(sY, NY, (r + 1) * NC + hc_col, far_, gar);
            rec_get(sH + (k & 1) * NTF * NH, NH, h, fbl, gbl);
            rec_get(sH + ((k + 1) & 1) * NTF * NH, NH, h, fbr, gbr);
            TransDiff th;
            trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, kGamma, th);
            td_put(sT, NT, r * NC + hc_col, th);
        }
        __syncthreads();
        if (face) {
            TransDiff tm;
            td_get(sT, NT, ty * NC + tx, tm);
            ctu_face<0, HAS_SRC>(E, L, L.off(i, j, k), i, j, k, tp, tm, hdt, area, hc);
        }
        if (halo && h >= kCtuY) {
            // the merged tile's last face, between its last cell and the halo cell
            const int r = hc_j - T.j0;
            TransDiff tl, tr;
            td_get(sT, NT, r * NC + kCtuX, tl);
            td_get(sT, NT, r * NC + kCtuX + 1, tr);
            ctu_face<0, HAS_SRC>(E, L, L.off(hc_i, hc_j, k), hc_i, hc_j, k, tr, tl, hdt, area, hc);
        }
#pragma unroll
        for (int n = 0; n < NF; ++n) zlo[n] = zhi[n];
        zglo[GUN] = zghi[GUN];
        zglo[GP] = zghi[GP];
    }
}

// y-faces.  The thread at row ty owns face j = j0 + ty and the cell above
// it; row 0 of the LDS arrays is the halo row of cells below the tile's
// first face (rows 1..kCtuY the own cells).  x|z pairs on x-faces in LDS
// (columns 0..kCtuX), z|x pairs in registers (halo cells: in LDS), transdd
// differences in LDS.
template <int HAS_SRC>
__global__ void CTU_BOUNDS k_ctu_y(Batch B, HydroConst hc, const double *dtp) {
    CTU_TILE(1)
    constexpr int NC = kCtuX + 1;
    constexpr int NX = (kCtuY + 1) * NC, NT = (kCtuY + 1) * kCtuX, NH = kCtuX;
    __shared__ double sX[NTF * NX], sT[NTD * NT], sH[2 * NTF * NH];
    const double dt = *dtp;
    const double cdx = 1.0 / 3.0 * dt / hc.dx[0], cdz = 1.0 / 3.0 * dt / hc.dx[2];
    const double ca = 0.5 * dt / hc.dx[0], cb = 0.5 * dt / hc.dx[2];
    const double hdt = 0.5 * dt, area = hc.dx[0] * hc.dx[2];
    const bool face = i <= hx && j <= hy;        // face (i, j) and the cell above it
    const bool ownX = i <= hx + 1 && j <= hy;    // x|z on the cell's low x-face
    // x|z on far or halo faces (wavefront 0): column kCtuX of the own rows,
    // then the halo row (columns 0..kCtuX)
    int xe_i = 0, xe_j = 0, xe_p = 0;
    bool extX = false;
    if (t < kCtuY) {
        xe_i = T.i0 + kCtuX; xe_j = T.j0 + t; xe_p = (t + 1) * NC + kCtuX;
        extX = xe_i <= hx + 1 && xe_j <= hy;
    } else if (t < kCtuY + kCtuX + 1) {
        const int c = t - kCtuY;
        xe_i = T.i0 + c; xe_j = T.j0 - 1; xe_p = c;
        extX = xe_i <= hx + 1;
    }
    // halo cells (wavefront 1): the row below the tile
    const int h = t - 64;
    const bool halo = h >= 0 && h < kCtuX && T.i0 + h <= hx;
    const int hc_i = T.i0 + h, hc_j = T.j0 - 1;
    double zlo[NF], zglo[NG];
    if (face) ctu_pair<2, 0>(E, L, L.off(i, j, T.k0), T.k0, cdx, hc, zlo, zglo);
    if (halo) {
        double f[NF], g[NG];
        ctu_pair<2, 0>(E, L, L.off(hc_i, hc_j, T.k0), T.k0, cdx, hc, f, g);
        rec_put(sH + (T.k0 & 1) * NTF * NH, NH, h, f, g);
    }
    for (int k = T.k0; k <= T.k1; ++k) {
        double zhi[NF], zghi[NG];
        if (face) ctu_pair<2, 0>(E, L, L.off(i, j, k + 1), k + 1, cdx, hc, zhi, zghi);
        if (ownX) {
            double f[NF], g[NG];
            ctu_pair<0, 2>(E, L, L.off(i, j, k), i, cdz, hc, f, g);
            rec_put(sX, NX, (ty + 1) * NC + tx, f, g);
        }
        if (extX) {
            double f[NF], g[NG];
            ctu_pair<0, 2>(E, L, L.off(xe_i, xe_j, k), xe_i, cdz, hc, f, g);
            rec_put(sX, NX, xe_p, f, g);
        }
        if (halo) {
            double f[NF], g[NG];
            ctu_pair<2, 0>(E, L, L.off(hc_i, hc_j, k + 1), k + 1, cdx, hc, f, g);
            rec_put(sH + ((k + 1) & 1) * NTF * NH, NH, h, f, g);
        }
        __syncthreads();
        TransDiff tp = {};
        if (face) {
            double fal[NF], far_[NF], gal[NG], gar[NG];
            rec_get(sX, NX, (ty + 1) * NC + tx, fal, gal);
            rec_get(sX, NX, (ty + 1) * NC + tx + 1, far_, gar);
            trans_diff_two(fal, far_, gal, gar, zlo, zhi, zglo, zghi, ca, cb, kGamma, tp);
            td_put(sT, NT, (ty + 1) * kCtuX + tx, tp);
        }
        if (halo) {
            double fal[NF], far_[NF], fbl[NF], fbr[NF], gal[NG], gar[NG], gbl[NG], gbr[NG];
            rec_get(sX, NX, h, fal, gal);
            rec_get(sX, NX, h + 1, far_, gar);
            rec_get(sH + (k & 1) * NTF * NH, NH, h, fbl, gbl);
            rec_get(sH + ((k + 1) & 1) * NTF * NH, NH, h, fbr, gbr);
            TransDiff th;
            trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, kGamma, th);
            td_put(sT, NT, h, th);
        }
        __syncthreads();
        if (face) {
            TransDiff tm;
            td_get(sT, NT, ty * kCtuX + tx, tm);
            ctu_face<1, HAS_SRC>(E, L, L.off(i, j, k), i, j, k, tp, tm, hdt, area, hc);
        }
#pragma unroll
        for (int n = 0; n < NF; ++n) zlo[n] = zhi[n];
        zglo[GUN] = zghi[GUN];
        zglo[GP] = zghi[GP];
    }
}

// ------------------------------------------------ one-box route (per-box ABI)
// The table of a one-box launch: its entry and a block prefix of {0}.
struct OneBox {
    BoxEnt e;
    int blk0;
};
__global__ void k_put_onebox(OneBox *dst, OneBox v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

// The caller's primitives (QVAR / NQAUX layout), primitive sources and
// conserved state into the canonical compact arrays.
__global__ void __launch_bounds__(kBlock)
k_box_in(Batch B, Arr q, Arr qa, Arr srcq, int has_src, Arr uin, Box ubox) {
    BATCH_POINT(ST_G4)
    const unsigned v0 = L.off(i, j, k);
    const CView qc = cview(L, slot_ptr<S_Q>(E), NQC);
    qc.st(v0, SR, q(i, j, k, QRHO));
    qc.st(v0, SU, q(i, j, k, QU));
    qc.st(v0, SV, q(i, j, k, QV));
    qc.st(v0, SW, q(i, j, k, QW));
    qc.st(v0, SP, q(i, j, k, QPRES));
    qc.st(v0, SE, q(i, j, k, QREINT));
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) qc.st(v0, SY + s, q(i, j, k, QFS + s));
    qc.st(v0, SC, qa(i, j, k, QC));
    qc.st(v0, SCS, qa(i, j, k, QCSML));
    if (has_src) {
        const CView sq = cview(L, slot_ptr<S_SRCQ>(E), NS);
        sq.st(v0, SR, srcq(i, j, k, QRHO));
        sq.st(v0, SU, srcq(i, j, k, QU));
        sq.st(v0, SV, srcq(i, j, k, QV));
        sq.st(v0, SW, srcq(i, j, k, QW));
        sq.st(v0, SP, srcq(i, j, k, QPRES));
        sq.st(v0, SE, srcq(i, j, k, QREINT));
#pragma unroll
        for (int s = 0; s < NSPEC; ++s) sq.st(v0, SY + s, srcq(i, j, k, QFS + s));
    }
    if (ubox.contains(i, j, k)) {
        const CView uc = cview(L, slot_ptr<S_U>(E), NVAR);
#pragma unroll
        for (int n = 0; n < NVAR; ++n) uc.st(v0, n, uin(i, j, k, n));
    }
}

// The final d-fluxes of the canonical array into the caller's flx[d].
template <int D>
__global__ void __launch_bounds__(kBlock)
k_box_flux_out(Batch B, Arr flx) {
    BATCH_POINT(ST_T2 + D)
    const CView f = cview(L, slot_ptr<S_FLX + D>(E), NXF);
    const unsigned v0 = L.off(i, j, k);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) flx(i, j, k, n) = f.ld(v0, n);
}

// hydro_pdivu (Hydro_utils_K.H:175-202) and hydro_update (:239-257) into
// dsdt (+=), in the operation order of the staged k_update with the
// uniform-grid area and volume.
__global__ void __launch_bounds__(kBlock)
k_box_update(Batch B, Arr dsdt, double ac0, double ac1, double ac2, double volc) {
    BATCH_POINT(ST_VALID)
    const CView f0 = cview(L, slot_ptr<S_FLX + 0>(E), NXF), f1 = cview(L, slot_ptr<S_FLX + 1>(E), NXF),
                f2 = cview(L, slot_ptr<S_FLX + 2>(E), NXF);
    const unsigned v0 = L.off(i, j, k), sx = 8u, sy = L.sy(), sz = L.sz();
    const double pdivu =
        0.5 *
        ((f0.ld(v0 + sx, XGP) + f0.ld(v0, XGP)) * (f0.ld(v0 + sx, XGU) * ac0 - f0.ld(v0, XGU) * ac0) +
         (f1.ld(v0 + sy, XGP) + f1.ld(v0, XGP)) * (f1.ld(v0 + sy, XGU) * ac1 - f1.ld(v0, XGU) * ac1) +
         (f2.ld(v0 + sz, XGP) + f2.ld(v0, XGP)) * (f2.ld(v0 + sz, XGU) * ac2 - f2.ld(v0, XGU) * ac2)) /
        volc;
    const double volinv = 1.0 / volc;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        dsdt(i, j, k, n) += (f0.ld(v0, n) - f0.ld(v0 + sx, n) + f1.ld(v0, n) - f1.ld(v0 + sy, n) +
                             f2.ld(v0, n) - f2.ld(v0 + sz, n)) *
                            volinv;
    }
    dsdt(i, j, k, UEINT) -= pdivu;
}

// The same over a box batch into dsdt_tab[E.id], overwriting: 0 + the
// divergence, then - pdivu, as hydro_update's += on the zeroed dS/dt.
__global__ void __launch_bounds__(kBlock)
k_dsdt_c(Batch B, const DevFab *dsdt_tab, double ac0, double ac1, double ac2, double volc) {
    BATCH_POINT(ST_VALID)
    const Arr dsdt = dsdt_tab[E.id].a;
    const CView f0 = cview(L, slot_ptr<S_FLX + 0>(E), NXF), f1 = cview(L, slot_ptr<S_FLX + 1>(E), NXF),
                f2 = cview(L, slot_ptr<S_FLX + 2>(E), NXF);
    const unsigned v0 = L.off(i, j, k), sx = 8u, sy = L.sy(), sz = L.sz();
    const double pdivu =
        0.5 *
        ((f0.ld(v0 + sx, XGP) + f0.ld(v0, XGP)) * (f0.ld(v0 + sx, XGU) * ac0 - f0.ld(v0, XGU) * ac0) +
         (f1.ld(v0 + sy, XGP) + f1.ld(v0, XGP)) * (f1.ld(v0 + sy, XGU) * ac1 - f1.ld(v0, XGU) * ac1) +
         (f2.ld(v0 + sz, XGP) + f2.ld(v0, XGP)) * (f2.ld(v0 + sz, XGU) * ac2 - f2.ld(v0, XGU) * ac2)) /
        volc;
    const double volinv = 1.0 / volc;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        double d = 0.0;
        d += (f0.ld(v0, n) - f0.ld(v0 + sx, n) + f1.ld(v0, n) - f1.ld(v0 + sy, n) +
              f2.ld(v0, n) - f2.ld(v0 + sz, n)) *
             volinv;
        if (n == UEINT) d -= pdivu;
        dsdt(i, j, k, n) = d;
    }
}

__global__ void __launch_bounds__(kBlock)
k_divu_c(Batch B, HydroConst hc) {
    BATCH_POINT(ST_DIVU)
    // compact q has u, v, w in the same slots as QU, QV, QW
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), divu = cview(L, slot_ptr<S_DIVU>(E), 1);
    divu.st(L.off(i, j, k), 0, divu_node(q, i, j, k, hc));
}

// ------------------------------------------------------------ update
// hydro_pdivu (Hydro_utils_K.H:175-202), hydro_update (:239-257) on a zeroed
// dS/dt (CAMR_construct_hydro_source.cpp:14), then the Saxpy of
// CAMR_advance.cpp:141 and the floor check of the following clean_state.
__global__ void __launch_bounds__(kBlock)
k_update_c(Batch B, HydroConst hc, const double *dtp, int *flags) {
    BATCH_POINT(ST_VALID)
    (void)L;
    const Arr snew = E.snew;
    double v[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) v[n] = snew(i, j, k, n);
    update_values(E, i, j, k, hc, *dtp, v);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) snew(i, j, k, n) = v[n];
    const double rho = v[URHO];
    if (rho == 0.) atomicOr(flags, 1 << CAMR_ERR_ZERO_DENSITY);
    else if (rho < hc.small_dens) atomicOr(flags, kFloorFlag);
}

} // namespace

// --------------------------------------------------------------- engine
// block prefix rows: the stage kinds as tuned, the same without z-march
// (LaunchCfg linear), then the CTU face kernels of the three directions
constexpr int kCtuRow0 = 2 * ST_COUNT, kBlkRows = 2 * ST_COUNT + 3;

Tuning default_tuning() {
    Tuning t;
    for (int k = 0; k < ST_COUNT; ++k) {
        t.block[k] = kBlock;
        t.tile[k][0] = -1;
        t.tile[k][1] = t.tile[k][2] = 0;
        t.xcd[k] = 16;
        t.zm[k] = 0;
    }
    t.zm[ST_T1 + 2] = t.zm[ST_T2 + 2] = 2;
    t.order = 0;
    t.ctu = 0;
    t.ctu_planes = 8;
    return t;
}

size_t FusedEngine::scratch_doubles(const Box &bx) {
    return (size_t)slot_off(S_COUNT) * (size_t)canon_box(bx).npts();
}
bool FusedEngine::fits(const Box &bx) { return canon_fits(bx); }

int FusedEngine::setup(const std::vector<Box> &boxes, const std::vector<Arr> &bufA,
                       const std::vector<Arr> &bufB, const Tuning &t,
                       const std::vector<int> &ids) {
    boxes_ = boxes;
    ids_ = ids;
    if (!ids_.empty() && ids_.size() != boxes_.size()) {
        set_error("fused engine: one level index per box");
        return CAMR_ERR_PARAM;
    }
    bufA_ = bufA;
    bufB_ = bufB;
    const int n = (int)boxes_.size();
    off_.assign(n, 0);
    size_t total = 0;
    for (int b = 0; b < n; ++b) {
        if (!canon_fits(boxes_[b])) {
            set_error("box %d (%lld cells) is too large for the fused engine "
                      "(32-bit buffer offsets)", b, (long long)boxes_[b].npts());
            return CAMR_ERR_SHAPE;
        }
        off_[b] = total;
        total += (scratch_doubles(boxes_[b]) + 31) & ~size_t(31);  // 256-byte aligned
    }
    if (!scratch_.reserve(total * sizeof(double))) {
        set_error("fused-engine scratch allocation of %zu bytes failed", total * sizeof(double));
        return CAMR_ERR_ALLOC;
    }
    total_ = total;
    for (int p = 0; p < 2; ++p) {
        if (tab_[p]) (void)hipFree(tab_[p]);
        tab_[p] = nullptr;
        if (hipMalloc(&tab_[p], std::max(n, 1) * sizeof(BoxEnt)) != hipSuccess) {
            set_error("box table allocation failed");
            return CAMR_ERR_ALLOC;
        }
    }
    if (blk_dev_) (void)hipFree(blk_dev_);
    blk_dev_ = nullptr;
    if (hipMalloc(&blk_dev_, (size_t)kBlkRows * (n + 1) * sizeof(int)) != hipSuccess) {
        set_error("block table allocation failed");
        return CAMR_ERR_ALLOC;
    }
    return set_tuning(t);
}

FusedEngine::~FusedEngine() {
    for (int p = 0; p < 2; ++p)
        if (tab_[p]) (void)hipFree(tab_[p]);
    if (blk_dev_) (void)hipFree(blk_dev_);
}

int FusedEngine::set_tuning(const Tuning &t) {
    for (int k = 0; k < ST_COUNT; ++k)
        if (t.block[k] != 64 && t.block[k] != 128 && t.block[k] != 256) {
            set_error("block size must be 64, 128 or 256");
            return CAMR_ERR_PARAM;
        }
    if (t.order != 0 && t.order != 1) {
        set_error("box order must be 0 (level order) or 1 (largest first)");
        return CAMR_ERR_PARAM;
    }
    if ((t.ctu != 0 && t.ctu != 1) || t.ctu_planes < 1 || t.ctu_planes > 1024) {
        set_error("CTU face kernels: switch 0 or 1, 1..1024 planes per block");
        return CAMR_ERR_PARAM;
    }
    for (int k = 0; k < ST_COUNT; ++k) {
        if (t.xcd[k] < 0 || t.xcd[k] > 4096) {
            set_error("stage %d: XCD chunk must be 0..4096", k);
            return CAMR_ERR_PARAM;
        }
        if (t.zm[k] < 0 || t.zm[k] > 1024 ||
            (t.zm[k] > 0 && ((k != ST_T1 + 2 && k != ST_T2 + 2) || t.tile[k][0] >= 0))) {
            set_error("stage %d: z-march (1..1024 planes) exists for the linear blocks of "
                      "the z-face transverse stages only", k);
            return CAMR_ERR_PARAM;
        }
        const int *l = t.tile[k];
        if (l[0] < 0) continue;
        if (l[0] > 8 || l[1] < 0 || l[1] > 8 || l[2] < 0 || l[2] > 8 ||
            (1 << (l[0] + l[1] + l[2])) != t.block[k] || l[0] < 3) {
            set_error("stage %d: a tile is 2^lx x 2^ly x 2^lz points with lx >= 3 and "
                      "2^(lx+ly+lz) equal to its block size", k);
            return CAMR_ERR_PARAM;
        }
    }
    tune_ = t;
    return build_tables();
}

int FusedEngine::build_tables() {
    const int n = (int)boxes_.size();
    order_.resize(n);
    for (int b = 0; b < n; ++b) order_[b] = b;
    if (tune_.order == 1)
        std::stable_sort(order_.begin(), order_.end(), [&](int a, int b) {
            return boxes_[a].npts() > boxes_[b].npts();
        });
    pos_.assign(n, 0);
    for (int p = 0; p < n; ++p) pos_[order_[p]] = p;
    std::vector<BoxEnt> t0(n), t1(n);
    for (int p = 0; p < n; ++p) {
        const int b = order_[p];
        BoxEnt e;
        e.id = ids_.empty() ? b : ids_[b];
        e.bx = boxes_[b];
        e.L = canon_of(boxes_[b]);
        e.base = scratch_.base() + off_[b];
        e.snew = bufA_[b];
        e.sold = bufB_[b];
        e.abase = ascratch_.base() ? ascratch_.base() + off_[b] : nullptr;
        t0[p] = e;
        e.snew = bufB_[b];
        e.sold = bufA_[b];
        t1[p] = e;
    }
    blk_host_.assign((size_t)kBlkRows * (n + 1), 0);
    for (int k = 0; k < kBlkRows; ++k) {
        int *row = &blk_host_[(size_t)k * (n + 1)];
        const bool ctu = k >= kCtuRow0;
        const int kind = ctu ? ST_T2 + (k - kCtuRow0) : k % ST_COUNT;
        const int zm = k < ST_COUNT ? tune_.zm[kind] : 0;
        for (int p = 0; p < n; ++p) {
            const int *l = tune_.tile[kind];
            const Box sb = stage_box(kind, boxes_[order_[p]]);
            const int64_t nb =
                ctu ? ctu_blocks(sb, tune_.ctu_planes, kind == ST_T2)
                    : batch_blocks(sb, tune_.block[kind], l[0], l[1], l[2], zm);
            if ((int64_t)row[p] + nb > 0x7fffffff) {
                set_error("stage grid exceeds 2^31 blocks");
                return CAMR_ERR_SHAPE;
            }
            row[p + 1] = row[p] + (int)nb;
        }
    }
    if (n > 0) {
        if (hipMemcpy(tab_[0], t0.data(), n * sizeof(BoxEnt), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(tab_[1], t1.data(), n * sizeof(BoxEnt), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(blk_dev_, blk_host_.data(), blk_host_.size() * sizeof(int),
                      hipMemcpyHostToDevice) != hipSuccess) {
            set_error("box table upload failed");
            return CAMR_ERR_HIP;
        }
    }
    return CAMR_OK;
}

int FusedEngine::enable_adjoint() {
    if (ascratch_.base()) return CAMR_OK;
    if (!ascratch_.reserve(total_ * sizeof(double))) {
        set_error("adjoint scratch allocation of %zu bytes failed", total_ * sizeof(double));
        return CAMR_ERR_ALLOC;
    }
    return build_tables();
}

LaunchCfg FusedEngine::cfg(int kind, int b, int parity, bool linear) const {
    const int n = (int)boxes_.size();
    const int rk = kind + (linear ? ST_COUNT : 0);
    const int *row = &blk_host_[(size_t)rk * (n + 1)];
    LaunchCfg c;
    c.B.tab = tab_[parity];
    c.B.blk = blk_dev_ + (size_t)rk * (n + 1);
    c.B.nbox = n;
    if (b < 0) {
        c.B.first = 0;
        c.grid = (unsigned)row[n];
    } else {
        const int p = pos_[b];
        c.B.first = row[p];
        c.grid = (unsigned)(row[p + 1] - row[p]);
    }
    c.block = tune_.block[kind];
    c.B.halt = halt_;
    c.B.lx = tune_.tile[kind][0];
    c.B.ly = tune_.tile[kind][1];
    c.B.lz = tune_.tile[kind][2];
    c.B.xk = tune_.xcd[kind];
    c.B.zm = linear ? 0 : tune_.zm[kind];
    return c;
}

template <typename K, typename... A>
void FusedEngine::launch(const char *name, Prof *prof, K kern, int kind, int b, int parity,
                         hipStream_t st, A... args) {
    const LaunchCfg c = cfg(kind, b, parity);
    if (c.grid == 0) return;
    Region reg(prof, name, st);
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.block), 0, st, c.B, args...);
}

LaunchCfg FusedEngine::cfg_ctu(int d, int b, int parity) const {
    const int n = (int)boxes_.size();
    const int rk = kCtuRow0 + d;
    const int *row = &blk_host_[(size_t)rk * (n + 1)];
    LaunchCfg c;
    c.B.tab = tab_[parity];
    c.B.blk = blk_dev_ + (size_t)rk * (n + 1);
    c.B.nbox = n;
    if (b < 0) {
        c.B.first = 0;
        c.grid = (unsigned)row[n];
    } else {
        const int p = pos_[b];
        c.B.first = row[p];
        c.grid = (unsigned)(row[p + 1] - row[p]);
    }
    c.block = kCtuBlock;
    c.B.halt = halt_;
    c.B.lx = -1;
    c.B.ly = c.B.lz = 0;
    c.B.xk = tune_.xcd[ST_T2 + d];
    c.B.zm = tune_.ctu_planes;
    return c;
}

template <typename K, typename... A>
void FusedEngine::launch_ctu(const char *name, Prof *prof, K kern, int d, int b, int parity,
                             hipStream_t st, A... args) {
    const LaunchCfg c = cfg_ctu(d, b, parity);
    if (c.grid == 0) return;
    Region reg(prof, name, st);
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(kCtuBlock), 0, st, c.B, args...);
}

int FusedEngine::hydro(int b, int parity, bool have_prim, const DevFab *sb_tab,
                       const DevFab *src_tab, int has_src, const HydroConst &hc,
                       const double *dtp, hipStream_t st, int *status, Prof *prof,
                       const PostArgs *post, bool no_update) {
    if (!have_prim || has_src)
        launch("prim", prof, k_prim_c, ST_G4, b, parity, st, sb_tab, src_tab, has_src, hc, status);
    // q beyond the canonical grow(bx, 4) comes from sb_tab (given with !have_prim)
    const int qgrow = (!have_prim && sb_tab) ? qgrow_ : CAMR_NGROW;
    if (qgrow > CAMR_NGROW)
        launch("flatten", prof, k_flat_c<true>, ST_G2, b, parity, st, hc.use_flattening, qgrow,
               sb_tab, hc, status);
    else
        launch("flatten", prof, k_flat_c<false>, ST_G2, b, parity, st, hc.use_flattening, qgrow,
               sb_tab, hc, status);
    if (hc.ppm_type == 1) {
        launch("trace_ppm", prof, k_ppm_c<0>, ST_G2, b, parity, st, hc, dtp);
        launch("trace_ppm", prof, k_ppm_c<1>, ST_G2, b, parity, st, hc, dtp);
        launch("trace_ppm", prof, k_ppm_c<2>, ST_G2, b, parity, st, hc, dtp);
    } else {
        launch("trace_plm", prof, k_plm_c<0>, ST_G2, b, parity, st, has_src, hc, dtp);
        launch("trace_plm", prof, k_plm_c<1>, ST_G2, b, parity, st, has_src, hc, dtp);
        launch("trace_plm", prof, k_plm_c<2>, ST_G2, b, parity, st, has_src, hc, dtp);
    }
    launch("riemann_normal", prof, k_riemann1<0>, ST_R1 + 0, b, parity, st, hc);
    launch("riemann_normal", prof, k_riemann1<1>, ST_R1 + 1, b, parity, st, hc);
    launch("riemann_normal", prof, k_riemann1<2>, ST_R1 + 2, b, parity, st, hc);
    // the adjoint's recompute (no_update) reads the transdo fluxes
    const bool ctu = tune_.ctu && !no_update;
    if (ctu) {
        launch("divu", prof, k_divu_c, ST_DIVU, b, parity, st, hc);
        if (has_src) {
            launch_ctu("transverse_ctu", prof, k_ctu_x<1>, 0, b, parity, st, hc, dtp);
            launch_ctu("transverse_ctu", prof, k_ctu_y<1>, 1, b, parity, st, hc, dtp);
            launch_ctu("transverse_ctu", prof, k_ctu_z<1>, 2, b, parity, st, hc, dtp);
        } else {
            launch_ctu("transverse_ctu", prof, k_ctu_x<0>, 0, b, parity, st, hc, dtp);
            launch_ctu("transverse_ctu", prof, k_ctu_y<0>, 1, b, parity, st, hc, dtp);
            launch_ctu("transverse_ctu", prof, k_ctu_z<0>, 2, b, parity, st, hc, dtp);
        }
    } else {
        launch("transverse_1", prof, k_trans1_x, ST_T1 + 0, b, parity, st, hc, dtp);
        launch("transverse_1", prof, k_trans1<1>, ST_T1 + 1, b, parity, st, hc, dtp);
        if (tune_.zm[ST_T1 + 2] > 0)
            launch("transverse_1", prof, k_trans1_zm, ST_T1 + 2, b, parity, st, hc, dtp);
        else
            launch("transverse_1", prof, k_trans1<2>, ST_T1 + 2, b, parity, st, hc, dtp);
        launch("divu", prof, k_divu_c, ST_DIVU, b, parity, st, hc);
        const bool zm2 = tune_.zm[ST_T2 + 2] > 0;
        if (has_src) {
            launch("transverse_2", prof, k_trans2_x<1>, ST_T2 + 0, b, parity, st, hc, dtp);
            launch("transverse_2", prof, k_trans2<1, 1>, ST_T2 + 1, b, parity, st, hc, dtp);
            if (zm2)
                launch("transverse_2", prof, k_trans2_zm<1>, ST_T2 + 2, b, parity, st, hc, dtp);
            else
                launch("transverse_2", prof, k_trans2<2, 1>, ST_T2 + 2, b, parity, st, hc, dtp);
        } else {
            launch("transverse_2", prof, k_trans2_x<0>, ST_T2 + 0, b, parity, st, hc, dtp);
            launch("transverse_2", prof, k_trans2<1, 0>, ST_T2 + 1, b, parity, st, hc, dtp);
            if (zm2)
                launch("transverse_2", prof, k_trans2_zm<0>, ST_T2 + 2, b, parity, st, hc, dtp);
            else
                launch("transverse_2", prof, k_trans2<2, 0>, ST_T2 + 2, b, parity, st, hc, dtp);
        }
    }
    if (no_update) {
        // the caller applies the update itself
    } else if (post) {
        const LaunchCfg c = cfg(ST_VALID, b, parity);
        Region reg(prof, "consup_clean_estdt", st);
        launch_update_post_fused(c, hc, dtp, *post, st, status);
    } else {
        launch("consup", prof, k_update_c, ST_VALID, b, parity, st, hc, dtp, status);
    }
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

int FusedEngine::fill(int b, int parity, const FillArgs &a, hipStream_t st, int *status,
                      Prof *prof) {
    const LaunchCfg c = cfg(ST_G4, b, parity);
    Region reg(prof, "fill_clean_prim", st);
    launch_fill_fused(c, a, st, status);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

int FusedEngine::dsdt(int b, int parity, const DevFab *dsdt_tab, const HydroConst &hc,
                      hipStream_t st, Prof *prof) {
    launch("consup", prof, k_dsdt_c, ST_VALID, b, parity, st, dsdt_tab, hc.dx[1] * hc.dx[2],
           hc.dx[0] * hc.dx[2], hc.dx[0] * hc.dx[1], hc.dx[0] * hc.dx[1] * hc.dx[2]);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

int FusedEngine::post(int b, int parity, const PostArgs &a, hipStream_t st, int *status,
                      Prof *prof) {
    const LaunchCfg c = cfg(ST_VALID, b, parity);
    Region reg(prof, "clean_estdt", st);
    launch_post_fused(c, a, st, status);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

// ------------------------------------------------------- one-box route
namespace {
constexpr size_t kOneBoxHeader = 64;  // doubles ahead of the canonical arrays (512 bytes)
static_assert(sizeof(OneBox) <= kOneBoxHeader * sizeof(double), "one-box header");

template <typename K, typename... A>
void launch_one(K kern, int kind, const Box &bx, const Batch &B0, hipStream_t st, A... args) {
    const int64_t nb = batch_blocks(stage_box(kind, bx), kBlock, -1, 0, 0);
    if (nb <= 0) return;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kBlock), 0, st, B0, args...);
}
} // namespace

bool fused_umdrv_applies(const UmdrvArgs &A, int do_mol) {
    const Box g = grow(A.bx, CAMR_NGROW);
    bool same = true;
    for (int d = 0; d < 3; ++d) same = same && A.qbox.lo[d] == g.lo[d] && A.qbox.hi[d] == g.hi[d];
    return !do_mol && !A.has_area && !A.has_vol && same && canon_fits(A.bx);
}

size_t fused_umdrv_scratch_doubles(const Box &bx) {
    return kOneBoxHeader + FusedEngine::scratch_doubles(bx);
}

int fused_umdrv(const UmdrvArgs &A, double *scratch, hipStream_t st) {
    const Box bx = A.bx;
    const HydroConst &hc = A.hc;
    OneBox ob;
    memset(&ob, 0, sizeof(ob));
    ob.e.id = 0;
    ob.e.bx = bx;
    ob.e.L = canon_of(bx);
    ob.e.base = scratch + kOneBoxHeader;
    ob.e.abase = nullptr;
    ob.blk0 = 0;
    OneBox *dev = (OneBox *)scratch;
    hipLaunchKernelGGL(k_put_onebox, dim3(1), dim3(64), 0, st, dev, ob);
    Batch B;
    memset(&B, 0, sizeof(B));
    B.tab = &dev->e;
    B.blk = &dev->blk0;
    B.nbox = 1;
    B.first = 0;
    B.halt = nullptr;
    B.lx = -1;
    B.xk = 0;
    B.zm = 0;
    const Box ubox = grow(bx, 1);
    launch_one(k_box_in, ST_G4, bx, B, st, A.q, A.qaux, A.srcq, A.has_src, A.uin, ubox);
    const double *dtp = A.dtp;
    const int has_src = A.has_src;
    launch_one(k_flat_c<false>, ST_G2, bx, B, st, hc.use_flattening, (int)CAMR_NGROW,
               (const DevFab *)nullptr, hc, (int *)nullptr);
    if (hc.ppm_type == 1) {
        launch_one(k_ppm_c<0>, ST_G2, bx, B, st, hc, dtp);
        launch_one(k_ppm_c<1>, ST_G2, bx, B, st, hc, dtp);
        launch_one(k_ppm_c<2>, ST_G2, bx, B, st, hc, dtp);
    } else {
        launch_one(k_plm_c<0>, ST_G2, bx, B, st, has_src, hc, dtp);
        launch_one(k_plm_c<1>, ST_G2, bx, B, st, has_src, hc, dtp);
        launch_one(k_plm_c<2>, ST_G2, bx, B, st, has_src, hc, dtp);
    }
    launch_one(k_riemann1<0>, ST_R1 + 0, bx, B, st, hc);
    launch_one(k_riemann1<1>, ST_R1 + 1, bx, B, st, hc);
    launch_one(k_riemann1<2>, ST_R1 + 2, bx, B, st, hc);
    launch_one(k_trans1_x, ST_T1 + 0, bx, B, st, hc, dtp);
    launch_one(k_trans1<1>, ST_T1 + 1, bx, B, st, hc, dtp);
    launch_one(k_trans1<2>, ST_T1 + 2, bx, B, st, hc, dtp);
    launch_one(k_divu_c, ST_DIVU, bx, B, st, hc);
    if (has_src) {
        launch_one(k_trans2_x<1>, ST_T2 + 0, bx, B, st, hc, dtp);
        launch_one(k_trans2<1, 1>, ST_T2 + 1, bx, B, st, hc, dtp);
        launch_one(k_trans2<2, 1>, ST_T2 + 2, bx, B, st, hc, dtp);
    } else {
        launch_one(k_trans2_x<0>, ST_T2 + 0, bx, B, st, hc, dtp);
        launch_one(k_trans2<1, 0>, ST_T2 + 1, bx, B, st, hc, dtp);
        launch_one(k_trans2<2, 0>, ST_T2 + 2, bx, B, st, hc, dtp);
    }
    launch_one(k_box_flux_out<0>, ST_T2 + 0, bx, B, st, A.flx[0]);
    launch_one(k_box_flux_out<1>, ST_T2 + 1, bx, B, st, A.flx[1]);
    launch_one(k_box_flux_out<2>, ST_T2 + 2, bx, B, st, A.flx[2]);
    launch_one(k_box_update, ST_VALID, bx, B, st, A.dsdt, A.area_c[0], A.area_c[1], A.area_c[2],
               A.vol_c);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}


} // namespace camr
