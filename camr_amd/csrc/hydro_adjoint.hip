// hydro_adjoint.hip -- reverse-mode adjoint of one fused level step.
//
// The step S_new = post(update(hydro(fill(S_old)))) of the fused engine is
// differentiated stage by stage in reverse order.  The primal intermediates
// (Sborder copy, primitives, flattening, face states, first and corrected
// fluxes, final fluxes, divu) are those a recompute of the step leaves in
// the engine's canonical scratch; their adjoints live in a second scratch
// set of the same layout (BoxEnt::abase).  Each kernel recomputes its own
// per-point forward values and applies the per-point adjoints of
// adjoint_math.h.  Contributions to values shared between threads (stencil
// neighbours, fluxes read by several faces, the ghost images of a cell)
// are accumulated with FP64 atomics; values owned by one thread of a
// kernel are updated directly.
//
// Scope (checked by the host): periodic (Interior) domains, PPM, zero
// gravity, the default energy branches; the hydrodynamic variables (rho,
// rho u, rho E, rho e) carry the adjoint, species and T are passive.
#include "adjoint_math.h"
#include "hydro_adjoint.h"
#include "state_math.h"

namespace camr {

namespace {

constexpr int ex_(int D, int d) { return D == d ? 1 : 0; }
constexpr int NH = 6;  // hydro components: URHO..UEINT == SR..SE

template <int S>
__device__ __forceinline__ double *aslot(const BoxEnt &E) {
    constexpr size_t off = (size_t)slot_off(S);
    return E.abase + off * E.L.cs;
}
__device__ __forceinline__ size_t eoff(const Canon &L, int i, int j, int k) {
    return (size_t)((i - L.lo0) + (j - L.lo1) * L.js + (k - L.lo2) * L.ks);
}
__device__ __forceinline__ void aadd(double *p, double v) {
    if (v != 0.0) unsafeAtomicAdd(p, v);
}
template <int D>
__device__ __forceinline__ unsigned shift_b(const Canon &L) {
    return D == 0 ? 8u : (D == 1 ? L.sy() : L.sz());
}
template <int D>
__device__ __forceinline__ size_t shift_e(const Canon &L) {
    return D == 0 ? 1 : (D == 1 ? (size_t)L.js : (size_t)L.ks);
}

__device__ __forceinline__ FaceAux face_aux_c(const CView &q, unsigned vl, unsigned vr) {
    FaceAux a;
    a.c_l = q.ld(vl, SC);
    a.c_r = q.ld(vr, SC);
    a.csml_l = q.ld(vl, SCS);
    a.csml_r = q.ld(vr, SCS);
    a.gam_l = kGamma;
    a.gam_r = kGamma;
    return a;
}
__device__ __forceinline__ void aux_adj(double *Qb, size_t cs, size_t el, size_t er,
                                        const FaceAuxBar &ab) {
    aadd(Qb + SC * cs + el, ab.c_l);
    aadd(Qb + SCS * cs + el, ab.csml_l);
    aadd(Qb + SC * cs + er, ab.c_r);
    aadd(Qb + SCS * cs + er, ab.csml_r);
}

__device__ __forceinline__ void load_tf(const CView &F, unsigned v, double *f, double *g) {
    f[FR] = F.ld(v, TR); f[FMX] = F.ld(v, TMX); f[FMY] = F.ld(v, TMY); f[FMZ] = F.ld(v, TMZ);
    f[FE] = F.ld(v, TE); f[FEI] = 0.0;
    for (int s = 0; s < NSPEC; ++s) f[FY + s] = F.ld(v, TY + s);
    g[GUN] = F.ld(v, TGU);
    g[GP] = F.ld(v, TGP);
}
// accumulate compact-flux / Godunov adjoints into a transverse-flux adjoint
__device__ __forceinline__ void tf_adj(double *Fb, size_t cs, size_t e, const double *fb,
                                       const double *gb) {
    aadd(Fb + TR * cs + e, fb[FR]);
    aadd(Fb + TMX * cs + e, fb[FMX]);
    aadd(Fb + TMY * cs + e, fb[FMY]);
    aadd(Fb + TMZ * cs + e, fb[FMZ]);
    aadd(Fb + TE * cs + e, fb[FE]);
    aadd(Fb + TGU * cs + e, gb[GUN]);
    aadd(Fb + TGP * cs + e, gb[GP]);
}
__device__ __forceinline__ void read_tf_bar(const double *Fb, size_t cs, size_t e, double *fb,
                                            double *gb) {
    for (int n = 0; n < NF; ++n) fb[n] = 0.0;
    fb[FR] = Fb[TR * cs + e]; fb[FMX] = Fb[TMX * cs + e]; fb[FMY] = Fb[TMY * cs + e];
    fb[FMZ] = Fb[TMZ * cs + e]; fb[FE] = Fb[TE * cs + e];
    gb[GUN] = Fb[TGU * cs + e];
    gb[GP] = Fb[TGP * cs + e];
}

// ------------------------------------------------------------ post pass
// clean_state x2 (x3 with the zero-gravity new-source Saxpy, an identity)
// after the update: sbar_in <- adjoint of S_new before the cleaning.
__global__ void __launch_bounds__(kBlock)
k_post_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_VALID)
    (void)L;
    const Arr snew = E.snew;
    const Arr bin = a.sbar_in[E.id].a;
    double x[3][NVAR];
    double v[NVAR];
    for (int n = 0; n < NVAR; ++n) v[n] = snew(i, j, k, n);
    const int ncl = a.add_grav ? 3 : 2;
    for (int c = 0; c < ncl; ++c) {
        normalize_species(v);
        for (int n = 0; n < NVAR; ++n) x[c][n] = v[n];  // input of reset c
        reset_e_and_T(v, a.hc);
    }
    double b[NVAR] = {};
    for (int n = 0; n < NH; ++n) b[n] = bin(i, j, k, n);
    bool ok = true;
    for (int c = ncl - 1; c >= 0; --c) ok &= reset_e_adj(x[c], a.hc, b);
    if (!ok) atomicOr(a.status, 1 << CAMR_ERR_ADJOINT);
    for (int n = 0; n < NVAR; ++n) bin(i, j, k, n) = (n < NH) ? b[n] : 0.0;
}

// ------------------------------------------------------------ update
// hydro_update + p div u + S_new Saxpy: face-centred gather of the flux
// and Godunov adjoints of the D-faces of bx.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_update_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_T2 + D)
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    const HydroConst &hc = a.hc;
    const double dt = *a.dtp;
    const double vol = hc.dx[0] * hc.dx[1] * hc.dx[2];
    const double volinv = 1.0 / vol;
    const double ar = (D == 0) ? hc.dx[1] * hc.dx[2] : (D == 1) ? hc.dx[0] * hc.dx[2] : hc.dx[0] * hc.dx[1];
    const Arr bin = a.sbar_in[E.id].a;
    const bool in1 = E.bx.contains(i, j, k);
    const bool in0 = E.bx.contains(i - ex, j - ey, k - ez);
    const CView G = cview(L, slot_ptr<S_FLX + D>(E), NXF);
    const unsigned v0 = L.off(i, j, k), sd = shift_b<D>(L);
    const size_t cs = L.cs, e0 = eoff(L, i, j, k);
    double *Fb = aslot<S_FLX + D>(E);
    const double coef = dt * volinv;
    for (int n = 0; n < NH; ++n) {
        const double b1 = in1 ? bin(i, j, k, n) : 0.0;
        const double b0 = in0 ? bin(i - ex, j - ey, k - ez, n) : 0.0;
        Fb[n * cs + e0] = coef * b1 - coef * b0;
    }
    double bgp = 0.0, bgu = 0.0;
    const double P0 = G.ld(v0, XGP), U0 = G.ld(v0, XGU);
    if (in1) {
        const double tb = -dt * bin(i, j, k, UEINT) * 0.5 / vol;
        const double P1 = G.ld(v0 + sd, XGP), U1 = G.ld(v0 + sd, XGU);
        bgp += tb * (U1 * ar - U0 * ar);
        bgu += -tb * (P1 + P0) * ar;
    }
    if (in0) {
        const double tb = -dt * bin(i - ex, j - ey, k - ez, UEINT) * 0.5 / vol;
        const double Pm = G.ld(v0 - sd, XGP), Um = G.ld(v0 - sd, XGU);
        bgp += tb * (U0 * ar - Um * ar);
        bgu += tb * (P0 + Pm) * ar;
    }
    Fb[XGU * cs + e0] = bgu;
    Fb[XGP * cs + e0] = bgp;
}

// transdd differences of the cell at byte offset v (as k_trans2)
template <int D>
__device__ __forceinline__ void t2_diff(unsigned v, const Canon &L, const CView &Fa,
                                        const CView &Fb, double ca, double cb, TransDiff &t,
                                        double *gal, double *gar, double *gbl, double *gbr) {
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    double fal[NF], far_[NF], fbl[NF], fbr[NF];
    load_tf(Fa, v, fal, gal);
    load_tf(Fa, v + shift_b<T0>(L), far_, gar);
    load_tf(Fb, v, fbl, gbl);
    load_tf(Fb, v + shift_b<T1>(L), fbr, gbr);
    trans_diff_two(fal, far_, gal, gar, fbl, fbr, gbl, gbr, ca, cb, kGamma, t);
}

// ---------------------------------- transdd + final Riemann + adjust_fluxes
template <int D>
__global__ void __launch_bounds__(kBlock)
k_trans2_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_T2 + D)
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    const HydroConst &hc = a.hc;
    const double dt = *a.dtp;
    const double ca = 0.5 * dt / hc.dx[T0], cb = 0.5 * dt / hc.dx[T1];
    const double area = (D == 0) ? hc.dx[1] * hc.dx[2] : (D == 1) ? hc.dx[0] * hc.dx[2] : hc.dx[0] * hc.dx[1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const CView Fa = cview(L, slot_ptr<f2_slot(T0, T1)>(E), NTF);
    const CView Fb = cview(L, slot_ptr<f2_slot(T1, T0)>(E), NTF);
    const CView u = cview(L, slot_ptr<S_U>(E), NVAR);
    const CView dv = cview(L, slot_ptr<S_DIVU>(E), 1);
    const unsigned v0 = L.off(i, j, k), sd = shift_b<D>(L), vm = v0 - sd;
    const size_t cs = L.cs, e0 = eoff(L, i, j, k), esd = shift_e<D>(L), em = e0 - esd;
    // ---- primal
    double np[NS], nm[NS], sp[NS], sm[NS], z[NS] = {};
    TransDiff tp, tm;
    double gpl[NG], gpr[NG], gpbl[NG], gpbr[NG], gml[NG], gmr[NG], gmbl[NG], gmbr[NG];
    for (int s = 0; s < NS; ++s) { np[s] = qp.ld(v0, s); nm[s] = qm.ld(v0, s); }
    t2_diff<D>(v0, L, Fa, Fb, ca, cb, tp, gpl, gpr, gpbl, gpbr);
    t2_diff<D>(vm, L, Fa, Fb, ca, cb, tm, gml, gmr, gmbl, gmbr);
    transverse_two_d(np, tp, 0.5 * dt, z, hc.trd, hc.small_pres, sp);
    transverse_two_d(nm, tm, 0.5 * dt, z, hc.trd, hc.small_pres, sm);
    const FaceAux ax = face_aux_c(q, vm, v0);
    // ---- adjoint of area weighting and artificial viscosity
    const double *Xb = aslot<S_FLX + D>(E);
    double fb[NF] = {}, gb[NG];
    double fub[NH];
    for (int n = 0; n < NH; ++n) fub[n] = Xb[n * cs + e0] * area;
    gb[GUN] = Xb[XGU * cs + e0];
    gb[GP] = Xb[XGP * cs + e0];
    constexpr int A1 = (D == 0) ? 1 : 0;
    constexpr int A2 = (D == 2) ? 1 : 2;
    const unsigned sa = shift_b<A1>(L), sb2 = shift_b<A2>(L);
    const double xs = 0.25 * (dv.ld(v0, 0) + dv.ld(v0 + sa, 0) + dv.ld(v0 + sb2, 0) +
                              dv.ld(v0 + sa + sb2, 0));
    const double div = hc.difmag * vmin(0.0, xs);
    const double dxd = hc.dx[D];
    double *Ub = aslot<S_U>(E);
    double divb = 0.0;
    for (int n = 0; n < NH; ++n) {
        const double du = u.ld(v0, n) - u.ld(vm, n);
        aadd(Ub + n * cs + e0, fub[n] * dxd * div);
        aadd(Ub + n * cs + em, -fub[n] * dxd * div);
        divb += fub[n] * dxd * du;
    }
    if (xs < 0.0) {
        const double b = 0.25 * divb * hc.difmag;
        double *Db = aslot<S_DIVU>(E);
        const size_t ea = shift_e<A1>(L), eb = shift_e<A2>(L);
        aadd(Db + e0, b);
        aadd(Db + e0 + ea, b);
        aadd(Db + e0 + eb, b);
        aadd(Db + e0 + ea + eb, b);
    }
    fb[FR] = fub[URHO]; fb[FMX] = fub[UMX]; fb[FMY] = fub[UMY]; fb[FMZ] = fub[UMZ];
    fb[FE] = fub[UEDEN]; fb[FEI] = fub[UEINT];
    // ---- final Riemann solve
    double smb[NS] = {}, spb[NS] = {};
    FaceAuxBar ab{};
    face_flux_adj(D, hc, sm, sp, ax, fb, gb, smb, spb, ab);
    aux_adj(aslot<S_Q>(E), cs, em, e0, ab);
    // ---- transdd, both sides
    double *QPb = aslot<S_QP0 + D>(E), *QMb = aslot<S_QM0 + D>(E);
    double *FAb = aslot<f2_slot(T0, T1)>(E), *FBb = aslot<f2_slot(T1, T0)>(E);
    const size_t et0 = shift_e<T0>(L), et1 = shift_e<T1>(L);
    {
        double nb[NS] = {};
        TransDiffBar tb{};
        transverse_two_d_adj(np, tp, hc.trd, hc.small_pres, spb, nb, tb);
        for (int s = 0; s < NH; ++s) QPb[s * cs + e0] += nb[s];
        double falb[NF] = {}, farb[NF] = {}, fblb[NF] = {}, fbrb[NF] = {};
        double galb[NG] = {}, garb[NG] = {}, gblb[NG] = {}, gbrb[NG] = {};
        trans_diff_two_adj(gpl, gpr, gpbl, gpbr, ca, cb, kGamma, tb, falb, farb, galb, garb,
                           fblb, fbrb, gblb, gbrb);
        tf_adj(FAb, cs, e0, falb, galb);
        tf_adj(FAb, cs, e0 + et0, farb, garb);
        tf_adj(FBb, cs, e0, fblb, gblb);
        tf_adj(FBb, cs, e0 + et1, fbrb, gbrb);
    }
    {
        double nb[NS] = {};
        TransDiffBar tb{};
        transverse_two_d_adj(nm, tm, hc.trd, hc.small_pres, smb, nb, tb);
        for (int s = 0; s < NH; ++s) QMb[s * cs + e0] += nb[s];
        double falb[NF] = {}, farb[NF] = {}, fblb[NF] = {}, fbrb[NF] = {};
        double galb[NG] = {}, garb[NG] = {}, gblb[NG] = {}, gbrb[NG] = {};
        trans_diff_two_adj(gml, gmr, gmbl, gmbr, ca, cb, kGamma, tb, falb, farb, galb, garb,
                           fblb, fbrb, gblb, gbrb);
        tf_adj(FAb, cs, em, falb, galb);
        tf_adj(FAb, cs, em + et0, farb, garb);
        tf_adj(FBb, cs, em, fblb, gblb);
        tf_adj(FBb, cs, em + et1, fbrb, gbrb);
    }
}

// ------------------------------------------------------------ divu
__global__ void __launch_bounds__(kBlock)
k_divu_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_DIVU)
    const double db = aslot<S_DIVU>(E)[eoff(L, i, j, k)];
    if (db == 0.0) return;
    double *Qb = aslot<S_Q>(E);
    const size_t cs = L.cs, e = eoff(L, i, j, k);
    const size_t sx = 1, sy = L.js, sz = L.ks;
    const double cx = db * 0.25 / a.hc.dx[0], cy = db * 0.25 / a.hc.dx[1], cz = db * 0.25 / a.hc.dx[2];
    // ux: +i / -(i-1) over (j, j-1) x (k, k-1); likewise vy, wz
    for (int jj = 0; jj < 2; ++jj)
        for (int kk = 0; kk < 2; ++kk) {
            const size_t c = e - jj * sy - kk * sz;
            aadd(Qb + SU * cs + c, cx);
            aadd(Qb + SU * cs + c - sx, -cx);
        }
    for (int ii = 0; ii < 2; ++ii)
        for (int kk = 0; kk < 2; ++kk) {
            const size_t c = e - ii * sx - kk * sz;
            aadd(Qb + SV * cs + c, cy);
            aadd(Qb + SV * cs + c - sy, -cy);
        }
    for (int ii = 0; ii < 2; ++ii)
        for (int jj = 0; jj < 2; ++jj) {
            const size_t c = e - ii * sx - jj * sy;
            aadd(Qb + SW * cs + c, cz);
            aadd(Qb + SW * cs + c - sz, -cz);
        }
}

// ------------------------------------------------ transdo + Riemann (pairs)
template <int D, int T>
__device__ __forceinline__ void trans1_pair_adj(const BoxEnt &E, const Canon &L, unsigned v0,
                                                size_t e0, const double *np, const double *nm,
                                                const FaceAux &ax, double cdt,
                                                const HydroConst &hc) {
    const unsigned sd = shift_b<D>(L), st_ = shift_b<T>(L);
    const size_t esd = shift_e<D>(L), est = shift_e<T>(L), cs = L.cs;
    const CView F = cview(L, slot_ptr<S_F1 + T>(E), NTF);
    double flp[NF], frp[NF], glp[NG], grp[NG], flm[NF], frm[NF], glm[NG], grm[NG];
    load_tf(F, v0, flp, glp);
    load_tf(F, v0 + st_, frp, grp);
    load_tf(F, v0 - sd, flm, glm);
    load_tf(F, v0 - sd + st_, frm, grm);
    TransDiff tp, tm;
    trans_diff_one(flp, frp, glp, grp, cdt, tp);
    trans_diff_one(flm, frm, glm, grm, cdt, tm);
    double sp[NS], sm[NS];
    transverse_one_d(np, tp, cdt, kGamma, hc.trd, hc.small_pres, sp);
    transverse_one_d(nm, tm, cdt, kGamma, hc.trd, hc.small_pres, sm);
    double fb[NF], gb[NG];
    read_tf_bar(aslot<f2_slot(D, T)>(E), cs, e0, fb, gb);
    double smb[NS] = {}, spb[NS] = {};
    FaceAuxBar ab{};
    face_flux_adj(D, hc, sm, sp, ax, fb, gb, smb, spb, ab);
    aux_adj(aslot<S_Q>(E), cs, e0 - esd, e0, ab);
    double *QPb = aslot<S_QP0 + D>(E), *QMb = aslot<S_QM0 + D>(E), *F1b = aslot<S_F1 + T>(E);
    {
        double nb[NS] = {};
        TransDiffBar tb{};
        transverse_one_d_adj(np, tp, cdt, kGamma, hc.trd, hc.small_pres, spb, nb, tb);
        for (int s = 0; s < NH; ++s) QPb[s * cs + e0] += nb[s];
        double flb[NF] = {}, frb[NF] = {}, glb[NG] = {}, grb[NG] = {};
        trans_diff_one_adj(flp, frp, glp, grp, cdt, tb, flb, frb, glb, grb);
        tf_adj(F1b, cs, e0, flb, glb);
        tf_adj(F1b, cs, e0 + est, frb, grb);
    }
    {
        double nb[NS] = {};
        TransDiffBar tb{};
        transverse_one_d_adj(nm, tm, cdt, kGamma, hc.trd, hc.small_pres, smb, nb, tb);
        for (int s = 0; s < NH; ++s) QMb[s * cs + e0] += nb[s];
        double flb[NF] = {}, frb[NF] = {}, glb[NG] = {}, grb[NG] = {};
        trans_diff_one_adj(flm, frm, glm, grm, cdt, tb, flb, frb, glb, grb);
        tf_adj(F1b, cs, e0 - esd, flb, glb);
        tf_adj(F1b, cs, e0 - esd + est, frb, grb);
    }
}

template <int D>
__global__ void __launch_bounds__(kBlock)
k_trans1_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_T1 + D)
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    const double dt = *a.dtp;
    const double cda = 1.0 / 3.0 * dt / a.hc.dx[T0];
    const double cdb = 1.0 / 3.0 * dt / a.hc.dx[T1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const unsigned v0 = L.off(i, j, k), sd = shift_b<D>(L);
    double np[NS], nm[NS];
    for (int s = 0; s < NS; ++s) { np[s] = qp.ld(v0, s); nm[s] = qm.ld(v0, s); }
    const FaceAux ax = face_aux_c(q, v0 - sd, v0);
    const size_t e0 = eoff(L, i, j, k);
    trans1_pair_adj<D, T0>(E, L, v0, e0, np, nm, ax, cda, a.hc);
    trans1_pair_adj<D, T1>(E, L, v0, e0, np, nm, ax, cdb, a.hc);
}

// ------------------------------------------------------------ first fluxes
template <int D>
__global__ void __launch_bounds__(kBlock)
k_riemann1_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_R1 + D)
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const unsigned v0 = L.off(i, j, k), sd = shift_b<D>(L);
    const size_t cs = L.cs, e0 = eoff(L, i, j, k), esd = shift_e<D>(L);
    double ql[NS], qr[NS];
    for (int s = 0; s < NS; ++s) { ql[s] = qm.ld(v0, s); qr[s] = qp.ld(v0, s); }
    const FaceAux ax = face_aux_c(q, v0 - sd, v0);
    double fb[NF], gb[NG];
    read_tf_bar(aslot<S_F1 + D>(E), cs, e0, fb, gb);
    double qlb[NS] = {}, qrb[NS] = {};
    FaceAuxBar ab{};
    face_flux_adj(D, a.hc, ql, qr, ax, fb, gb, qlb, qrb, ab);
    aux_adj(aslot<S_Q>(E), cs, e0 - esd, e0, ab);
    double *QPb = aslot<S_QP0 + D>(E), *QMb = aslot<S_QM0 + D>(E);
    for (int s = 0; s < NH; ++s) {
        QMb[s * cs + e0] += qlb[s];
        QPb[s * cs + e0] += qrb[s];
    }
}

// ------------------------------------------------------------ PPM trace
template <int D>
__global__ void __launch_bounds__(kBlock)
k_ppm_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_G2)
    constexpr int SN = SU + D;
    constexpr int T1 = (D == 0) ? SV : SU;
    constexpr int T2 = (D == 2) ? SV : SW;
    const HydroConst &hc = a.hc;
    const double dtdx = *a.dtp / hc.dx[D];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC), flat = cview(L, slot_ptr<S_FLAT>(E), 1);
    const unsigned v0 = L.off(i, j, k), sd = shift_b<D>(L);
    const size_t cs = L.cs, e0 = eoff(L, i, j, k), esd = shift_e<D>(L);
    const double cc = q.ld(v0, SC);
    const double un = q.ld(v0, SN);
    const double fl = flat.ld(v0, 0);
    // forward: edges, s6 and the wave integrals of the hydro components
    double st[NH][5], smv[NH], spv[NH], s6v[NH], Ip[NH][3], Im[NH][3];
    for (int s = 0; s < NH; ++s) {
        for (int o = -2; o <= 2; ++o) st[s][o + 2] = q.ld(v0 + (unsigned)o * sd, s);
        ppm_edges(st[s], fl, smv[s], spv[s]);
        s6v[s] = ppm_s6(smv[s], spv[s], st[s][2]);
        ppm_integral(smv[s], spv[s], s6v[s], un - cc, dtdx, Ip[s][0], Im[s][0]);
        ppm_integral(smv[s], spv[s], s6v[s], un, dtdx, Ip[s][1], Im[s][1]);
        ppm_integral(smv[s], spv[s], s6v[s], un + cc, dtdx, Ip[s][2], Im[s][2]);
    }
    const double *QPb = aslot<S_QP0 + D>(E), *QMb = aslot<S_QM0 + D>(E);
    double Ipb[NH][3] = {}, Imb[NH][3] = {};
    PPMWaves W;
    // plus state of face i (from Im)
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Im[SR][w]; W.un[w] = Im[SN][w]; W.p[w] = Im[SP][w];
        W.rhoe[w] = Im[SE][w]; W.gam[w] = kGamma;
    }
    {
        PPMWavesBar Wb{};
        ppm_project_adj(+1, W, un, cc, hc.small_dens, hc.small_pres, QPb[SR * cs + e0],
                        QPb[SN * cs + e0], QPb[SP * cs + e0], QPb[SE * cs + e0], Wb);
        for (int w = 0; w < 3; ++w) {
            Imb[SR][w] += Wb.rho[w]; Imb[SN][w] += Wb.un[w];
            Imb[SP][w] += Wb.p[w]; Imb[SE][w] += Wb.rhoe[w];
        }
        Imb[T1][1] += QPb[T1 * cs + e0];
        Imb[T2][1] += QPb[T2 * cs + e0];
    }
    // minus state of face i+1 (from Ip)
    for (int w = 0; w < 3; ++w) {
        W.rho[w] = Ip[SR][w]; W.un[w] = Ip[SN][w]; W.p[w] = Ip[SP][w];
        W.rhoe[w] = Ip[SE][w];
    }
    {
        const size_t e1 = e0 + esd;
        PPMWavesBar Wb{};
        ppm_project_adj(-1, W, un, cc, hc.small_dens, hc.small_pres, QMb[SR * cs + e1],
                        QMb[SN * cs + e1], QMb[SP * cs + e1], QMb[SE * cs + e1], Wb);
        for (int w = 0; w < 3; ++w) {
            Ipb[SR][w] += Wb.rho[w]; Ipb[SN][w] += Wb.un[w];
            Ipb[SP][w] += Wb.p[w]; Ipb[SE][w] += Wb.rhoe[w];
        }
        Ipb[T1][1] += QMb[T1 * cs + e1];
        Ipb[T2][1] += QMb[T2 * cs + e1];
    }
    double unb = 0.0, ccb = 0.0, flb = 0.0;
    double *Qb = aslot<S_Q>(E);
    for (int s = 0; s < NH; ++s) {
        double smb = 0.0, spb = 0.0, s6b = 0.0;
        double vb0 = 0.0, vb1 = 0.0, vb2 = 0.0;
        ppm_integral_adj(smv[s], spv[s], s6v[s], un - cc, dtdx, Ipb[s][0], Imb[s][0], smb, spb, s6b, vb0);
        ppm_integral_adj(smv[s], spv[s], s6v[s], un, dtdx, Ipb[s][1], Imb[s][1], smb, spb, s6b, vb1);
        ppm_integral_adj(smv[s], spv[s], s6v[s], un + cc, dtdx, Ipb[s][2], Imb[s][2], smb, spb, s6b, vb2);
        unb += vb0 + vb1 + vb2;
        ccb += -vb0 + vb2;
        double sb[5] = {};
        sb[2] += 6.0 * s6b;
        smb -= 3.0 * s6b;
        spb -= 3.0 * s6b;
        ppm_edges_adj(st[s], fl, smb, spb, sb, flb);
        for (int o = -2; o <= 2; ++o) aadd(Qb + s * cs + e0 + (ptrdiff_t)o * (ptrdiff_t)esd, sb[o + 2]);
    }
    aadd(Qb + SN * cs + e0, unb);
    aadd(Qb + SC * cs + e0, ccb);
    aslot<S_FLAT>(E)[e0] += flb;
}

// ------------------------------------------------------------ flattening
template <int D>
__device__ __forceinline__ double flat_dir_v(const CView &q, const Canon &L, unsigned v0, int i,
                                             int j, int k, const Box &qbox, double *p, double *u) {
    constexpr int ex = ex_(D, 0), ey = ex_(D, 1), ez = ex_(D, 2);
    if (!(qbox.contains(i + 3 * ex, j + 3 * ey, k + 3 * ez) &&
          qbox.contains(i - 3 * ex, j - 3 * ey, k - 3 * ez)))
        return 1.0;
    const unsigned sd = shift_b<D>(L);
    for (int o = -3; o <= 3; ++o) {
        const unsigned vo = v0 + (unsigned)o * sd;
        p[o + 3] = q.ld(vo, SP);
        u[o + 3] = q.ld(vo, SU + D);
    }
    return flatten_1d(p, u);
}

__global__ void __launch_bounds__(kBlock)
k_flat_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_G2)
    if (a.hc.use_flattening != 1) return;
    const size_t cs = L.cs, e0 = eoff(L, i, j, k);
    const double fb = aslot<S_FLAT>(E)[e0];
    if (fb == 0.0) return;
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const unsigned v0 = L.off(i, j, k);
    const Box qbox = grow(E.bx, CAMR_NGROW);
    double p[3][7], u[3][7];
    const double f0 = flat_dir_v<0>(q, L, v0, i, j, k, qbox, p[0], u[0]);
    const double f1 = flat_dir_v<1>(q, L, v0, i, j, k, qbox, p[1], u[1]);
    const double f2 = flat_dir_v<2>(q, L, v0, i, j, k, qbox, p[2], u[2]);
    const double m1 = vmin(1.0, f0);
    const double m2 = vmin(m1, f1);
    double m2b = 0.0, m1b = 0.0, dummy = 0.0, fdb[3] = {};
    vmin_adj(m2, f2, fb, m2b, fdb[2]);
    vmin_adj(m1, f1, m2b, m1b, fdb[1]);
    vmin_adj(1.0, f0, m1b, dummy, fdb[0]);
    double *Qb = aslot<S_Q>(E);
    const size_t es[3] = {1, (size_t)L.js, (size_t)L.ks};
    for (int d = 0; d < 3; ++d) {
        if (fdb[d] == 0.0) continue;
        const int ddx = d == 0, ddy = d == 1, ddz = d == 2;
        if (!(qbox.contains(i + 3 * ddx, j + 3 * ddy, k + 3 * ddz) &&
              qbox.contains(i - 3 * ddx, j - 3 * ddy, k - 3 * ddz)))
            continue;
        double pb[7] = {};
        flatten_1d_adj(p[d], u[d], fdb[d], pb);
        for (int o = -3; o <= 3; ++o)
            aadd(Qb + SP * cs + e0 + (ptrdiff_t)o * (ptrdiff_t)es[d], pb[o + 3]);
    }
}

// ------------------------------------------------------------ ghost fill
// clean_state(S_old) + FillPatch + clean_state(Sborder) + ctoprim and the
// S_new copy: adjoints of Sborder, primitives and S_new (valid cells) to
// the adjoint of the raw S_old of the owning box.
__global__ void __launch_bounds__(kBlock)
k_fill_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_G4)
    int m[3];
    unsigned odd;
    const bool phys = map_cell(a.dom, i, j, k, m, &odd);
    const int ob = owner_of(a.dom, a.owner, m);
    if (ob < 0) return;  // exterior data: not differentiated
    const Arr s = a.sold_tab[ob].a;
    double x1[NVAR], x2[NVAR], v[NVAR];
    for (int n = 0; n < NVAR; ++n) v[n] = s(m[0], m[1], m[2], n);
    normalize_species(v);
    for (int n = 0; n < NVAR; ++n) x1[n] = v[n];
    reset_e_and_T(v, a.hc);
    for (int n = 0; n < NVAR; ++n)
        if ((odd >> n) & 1u) v[n] = -v[n];
    if (!phys) normalize_species(v);
    for (int n = 0; n < NVAR; ++n) x2[n] = v[n];
    reset_e_and_T(v, a.hc);
    const size_t cs = L.cs, e0 = eoff(L, i, j, k);
    const double *Ub = aslot<S_U>(E), *Qb = aslot<S_Q>(E);
    double b[NVAR] = {};
    for (int n = 0; n < NH; ++n) b[n] = Ub[n * cs + e0];
    if (E.bx.contains(i, j, k)) {
        const Arr bin = a.sbar_in[E.id].a;
        for (int n = 0; n < NH; ++n) b[n] += bin(i, j, k, n);
    }
    double qb[NQC] = {};
    for (int n = 0; n < NH; ++n) qb[n] = Qb[n * cs + e0];
    qb[SC] = Qb[SC * cs + e0];
    qb[SCS] = Qb[SCS * cs + e0];
    if (v[URHO] > 0.) prim_compact_adj(v, qb, a.hc.eta1, a.hc.small_num, b);
    bool ok = reset_e_adj(x2, a.hc, b);
    for (int n = 0; n < NH; ++n)
        if ((odd >> n) & 1u) b[n] = -b[n];
    ok &= reset_e_adj(x1, a.hc, b);
    if (!ok) atomicOr(a.status, 1 << CAMR_ERR_ADJOINT);
    const Arr out = a.sbar_out[ob].a;
    for (int n = 0; n < NH; ++n) aadd(&out(m[0], m[1], m[2], n), b[n]);
}

__global__ void k_zero(double *p, size_t n) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

template <typename K>
void go(const FusedEngine &eng, const char *name, Prof *prof, K kern, int kind, int parity,
        hipStream_t st, const AdjArgs &a) {
    const LaunchCfg c = eng.cfg(kind, -1, parity);
    if (c.grid == 0) return;
    Region reg(prof, name, st);
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.block), 0, st, c.B, a);
}

} // namespace

void zero_doubles(double *p, size_t n, hipStream_t st) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(kBlock), 0, st, p, n);
}

int adjoint_step_kernels(const FusedEngine &eng, int parity, const AdjArgs &a, hipStream_t st,
                         Prof *prof) {
    go(eng, "adj_post", prof, k_post_adj, ST_VALID, parity, st, a);
    go(eng, "adj_update", prof, k_update_adj<0>, ST_T2 + 0, parity, st, a);
    go(eng, "adj_update", prof, k_update_adj<1>, ST_T2 + 1, parity, st, a);
    go(eng, "adj_update", prof, k_update_adj<2>, ST_T2 + 2, parity, st, a);
    go(eng, "adj_transverse_2", prof, k_trans2_adj<0>, ST_T2 + 0, parity, st, a);
    go(eng, "adj_transverse_2", prof, k_trans2_adj<1>, ST_T2 + 1, parity, st, a);
    go(eng, "adj_transverse_2", prof, k_trans2_adj<2>, ST_T2 + 2, parity, st, a);
    go(eng, "adj_divu", prof, k_divu_adj, ST_DIVU, parity, st, a);
    go(eng, "adj_transverse_1", prof, k_trans1_adj<0>, ST_T1 + 0, parity, st, a);
    go(eng, "adj_transverse_1", prof, k_trans1_adj<1>, ST_T1 + 1, parity, st, a);
    go(eng, "adj_transverse_1", prof, k_trans1_adj<2>, ST_T1 + 2, parity, st, a);
    go(eng, "adj_riemann_normal", prof, k_riemann1_adj<0>, ST_R1 + 0, parity, st, a);
    go(eng, "adj_riemann_normal", prof, k_riemann1_adj<1>, ST_R1 + 1, parity, st, a);
    go(eng, "adj_riemann_normal", prof, k_riemann1_adj<2>, ST_R1 + 2, parity, st, a);
    go(eng, "adj_trace_ppm", prof, k_ppm_adj<0>, ST_G2, parity, st, a);
    go(eng, "adj_trace_ppm", prof, k_ppm_adj<1>, ST_G2, parity, st, a);
    go(eng, "adj_trace_ppm", prof, k_ppm_adj<2>, ST_G2, parity, st, a);
    go(eng, "adj_flatten", prof, k_flat_adj, ST_G2, parity, st, a);
    go(eng, "adj_fill", prof, k_fill_adj, ST_G4, parity, st, a);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

} // namespace camr
