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
// Reads the adjoint of the transverse flux at e and clears it: every
// adjoint array is cleared by the one kernel that consumes it, so the
// adjoint scratch is all zero again after each step.
__device__ __forceinline__ void take_tf_bar(double *Fb, size_t cs, size_t e, double *fb,
                                            double *gb) {
#pragma unroll
    for (int n = 0; n < NF; ++n) fb[n] = 0.0;
    fb[FR] = Fb[TR * cs + e]; fb[FMX] = Fb[TMX * cs + e]; fb[FMY] = Fb[TMY * cs + e];
    fb[FMZ] = Fb[TMZ * cs + e]; fb[FE] = Fb[TE * cs + e];
    gb[GUN] = Fb[TGU * cs + e];
    gb[GP] = Fb[TGP * cs + e];
    Fb[TR * cs + e] = 0.0; Fb[TMX * cs + e] = 0.0; Fb[TMY * cs + e] = 0.0;
    Fb[TMZ * cs + e] = 0.0; Fb[TE * cs + e] = 0.0;
    Fb[TGU * cs + e] = 0.0; Fb[TGP * cs + e] = 0.0;
}

// ------------------------------------------------------------ post pass
// clean_state x2 (x3 with the zero-gravity new-source Saxpy, an identity)
// after the update: sbar_in <- adjoint of S_new before the cleaning.  The
// update of the cell (and its floor check) is recomputed here.
__global__ void __launch_bounds__(kBlock)
k_post_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_VALID)
    (void)L;
    const Arr snew = E.snew;
    const Arr bin = a.sbar_in[E.id].a;
    // inputs of the two (three) energy resets: S_new after the update, which
    // the recompute leaves to this kernel (S_new holds the filled state)
    double x0[NVAR], x1[NVAR], x2[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) x0[n] = snew(i, j, k, n);
    update_values(E, i, j, k, a.hc, *a.dtp, x0);
    if (x0[URHO] == 0.) atomicOr(a.status, 1 << CAMR_ERR_ZERO_DENSITY);
    else if (x0[URHO] < a.hc.small_dens) atomicOr(a.status, kFloorFlag);
    normalize_species(x0);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) x1[n] = x0[n];
    reset_e_and_T(x1, a.hc);
    normalize_species(x1);
    if (a.add_grav) {
#pragma unroll
        for (int n = 0; n < NVAR; ++n) x2[n] = x1[n];
        reset_e_and_T(x2, a.hc);
        normalize_species(x2);
    }
    double b[NVAR] = {};
#pragma unroll
    for (int n = 0; n < NH; ++n) b[n] = bin(i, j, k, n);
    bool ok = true;
    if (a.add_grav) ok &= reset_e_adj(x2, a.hc, b);
    ok &= reset_e_adj(x1, a.hc, b);
    ok &= reset_e_adj(x0, a.hc, b);
    if (!ok) atomicOr(a.status, 1 << CAMR_ERR_ADJOINT);
#pragma unroll
    for (int n = 0; n < NVAR; ++n) bin(i, j, k, n) = (n < NH) ? b[n] : 0.0;
    // the output adjoint (same boxes and cells) starts this step at zero
    const Arr bout = a.sbar_out[E.id].a;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) bout(i, j, k, n) = 0.0;
}

// ------------------------------------------------------------ update
// hydro_update + p div u + S_new Saxpy, for one D-face: the adjoints of
// its NVAR flux (area-weighted, hydro components) and of its Godunov (u, p)
// from the adjoints of the updated states of the two cells it separates.
template <int D>
__device__ __forceinline__ void update_face_adj(const BoxEnt &E, const Canon &L, int i, int j,
                                                int k, const AdjArgs &a, double *fub,
                                                double *gb) {
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
    const double coef = dt * volinv;
#pragma unroll
    for (int n = 0; n < NH; ++n) {
        const double b1 = in1 ? bin(i, j, k, n) : 0.0;
        const double b0 = in0 ? bin(i - ex, j - ey, k - ez, n) : 0.0;
        fub[n] = (coef * b1 - coef * b0) * ar;
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
    gb[GUN] = bgu;
    gb[GP] = bgp;
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
// One side of the D-face: the normal state nq (qp or qm of the face) with
// the transdd differences of cell vc.  Recomputed where needed instead of
// kept, to keep the register footprint of the kernel small.
template <int D>
__device__ __forceinline__ void t2_state(const CView &qs, unsigned v0, unsigned vc,
                                         const Canon &L, const CView &Fa, const CView &Fb,
                                         double ca, double cb, double hdt, const HydroConst &hc,
                                         double *out) {
    double nq[NS], z[NS] = {};
#pragma unroll
    for (int s = 0; s < NS; ++s) nq[s] = qs.ld(v0, s);
    TransDiff t;
    double g0[NG], g1[NG], g2[NG], g3[NG];
    t2_diff<D>(vc, L, Fa, Fb, ca, cb, t, g0, g1, g2, g3);
    transverse_two_d(nq, t, hdt, z, hc.trd, hc.small_pres, out);
}
// Adjoint of that side from the adjoint `ob` of the corrected state: into
// the normal state (owned by this face) and the four corrected fluxes of
// cell vc (shared: atomics).
template <int D>
__device__ __forceinline__ void t2_side_adj(const CView &qs, double *QSb, unsigned v0, size_t e0,
                                            unsigned vc, size_t ec, const Canon &L,
                                            const CView &Fa, const CView &Fb, double *FAb,
                                            double *FBb, double ca, double cb,
                                            const HydroConst &hc, const double *ob) {
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    const size_t cs = L.cs;
    double nq[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) nq[s] = qs.ld(v0, s);
    TransDiff t;
    double gal[NG], gar[NG], gbl[NG], gbr[NG];
    t2_diff<D>(vc, L, Fa, Fb, ca, cb, t, gal, gar, gbl, gbr);
    double nb[NS] = {};
    TransDiffBar tb{};
    transverse_two_d_adj(nq, t, hc.trd, hc.small_pres, ob, nb, tb);
#pragma unroll
    for (int s = 0; s < NH; ++s) QSb[s * cs + e0] += nb[s];
    double falb[NF] = {}, farb[NF] = {}, fblb[NF] = {}, fbrb[NF] = {};
    double galb[NG] = {}, garb[NG] = {}, gblb[NG] = {}, gbrb[NG] = {};
    trans_diff_two_adj(gal, gar, gbl, gbr, ca, cb, kGamma, tb, falb, farb, galb, garb, fblb,
                       fbrb, gblb, gbrb);
    tf_adj(FAb, cs, ec, falb, galb);
    tf_adj(FAb, cs, ec + shift_e<T0>(L), farb, garb);
    tf_adj(FBb, cs, ec, fblb, gblb);
    tf_adj(FBb, cs, ec + shift_e<T1>(L), fbrb, gbrb);
}

template <int D>
__global__ void __launch_bounds__(kBlock)
k_trans2_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_T2 + D)
    constexpr int T0 = (D == 0) ? 1 : 0;
    constexpr int T1 = (D == 2) ? 1 : 2;
    const HydroConst &hc = a.hc;
    const double dt = *a.dtp;
    const double ca = 0.5 * dt / hc.dx[T0], cb = 0.5 * dt / hc.dx[T1];
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const CView qm = cview(L, slot_ptr<S_QM0 + D>(E), NS), qp = cview(L, slot_ptr<S_QP0 + D>(E), NS);
    const CView Fa = cview(L, slot_ptr<f2_slot(T0, T1)>(E), NTF);
    const CView Fb = cview(L, slot_ptr<f2_slot(T1, T0)>(E), NTF);
    const unsigned v0 = L.off(i, j, k), sd = shift_b<D>(L), vm = v0 - sd;
    const size_t cs = L.cs, e0 = eoff(L, i, j, k), esd = shift_e<D>(L), em = e0 - esd;
    // ---- adjoint of the update: the area-weighted flux and Godunov (u, p)
    double fb[NF] = {}, gb[NG];
    double fub[NH];
    update_face_adj<D>(E, L, i, j, k, a, fub, gb);
    // ---- adjoint of the artificial viscosity
    {
        const CView u = cview(L, slot_ptr<S_U>(E), NVAR);
        const CView dv = cview(L, slot_ptr<S_DIVU>(E), 1);
        constexpr int A1 = (D == 0) ? 1 : 0;
        constexpr int A2 = (D == 2) ? 1 : 2;
        const unsigned sa = shift_b<A1>(L), sb2 = shift_b<A2>(L);
        const double xs = 0.25 * (dv.ld(v0, 0) + dv.ld(v0 + sa, 0) + dv.ld(v0 + sb2, 0) +
                                  dv.ld(v0 + sa + sb2, 0));
        const double div = hc.difmag * vmin(0.0, xs);
        const double dxd = hc.dx[D];
        double *Ub = aslot<S_U>(E);
        double divb = 0.0;
#pragma unroll
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
    }
    fb[FR] = fub[URHO]; fb[FMX] = fub[UMX]; fb[FMY] = fub[UMY]; fb[FMZ] = fub[UMZ];
    fb[FE] = fub[UEDEN]; fb[FEI] = fub[UEINT];
    // ---- final Riemann solve on the transdd-corrected states
    double smb[NS] = {}, spb[NS] = {};
    {
        double sp[NS], sm[NS];
        t2_state<D>(qp, v0, v0, L, Fa, Fb, ca, cb, 0.5 * dt, hc, sp);
        t2_state<D>(qm, v0, vm, L, Fa, Fb, ca, cb, 0.5 * dt, hc, sm);
        const FaceAux ax = face_aux_c(q, vm, v0);
        FaceAuxBar ab{};
        face_flux_adj(D, hc, sm, sp, ax, fb, gb, smb, spb, ab);
        aux_adj(aslot<S_Q>(E), cs, em, e0, ab);
    }
    // ---- transdd, both sides
    double *FAb = aslot<f2_slot(T0, T1)>(E), *FBb = aslot<f2_slot(T1, T0)>(E);
    t2_side_adj<D>(qp, aslot<S_QP0 + D>(E), v0, e0, v0, e0, L, Fa, Fb, FAb, FBb, ca, cb, hc, spb);
    t2_side_adj<D>(qm, aslot<S_QM0 + D>(E), v0, e0, vm, em, L, Fa, Fb, FAb, FBb, ca, cb, hc, smb);
}

// ------------------------------------------------------------ divu
// The nodal divergence adjoint is gathered by the cells in k_fill_adj
// (divu_gather); this clears it for the next step once every cell has read
// its eight nodes.
__global__ void __launch_bounds__(kBlock)
k_divu_clear(Batch B, AdjArgs a) {
    BATCH_POINT(ST_DIVU)
    (void)a;
    aslot<S_DIVU>(E)[eoff(L, i, j, k)] = 0.0;
}

// Adjoint of hydro_divu at the 8 nodes of the upper corner of cell e:
// node (i,j,k) has ux = 1/(4 dx) sum over (j, j-1) x (k, k-1) of
// u(i) - u(i-1), so cell (a,b,c) receives +db of nodes (a, b..b+1,
// c..c+1) and -db of nodes (a+1, ...); likewise v, w.
__device__ __forceinline__ void divu_gather(const double *Db, size_t e, const Canon &L,
                                            const HydroConst &hc, double *qb) {
    double d[2][2][2];
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
                d[ii][jj][kk] = Db[e + ii + (size_t)jj * L.js + (size_t)kk * L.ks];
    double su = 0.0, sv = 0.0, sw = 0.0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            su += d[0][p][r] - d[1][p][r];
            sv += d[p][0][r] - d[p][1][r];
            sw += d[p][r][0] - d[p][r][1];
        }
    qb[SU] += su * (0.25 / hc.dx[0]);
    qb[SV] += sv * (0.25 / hc.dx[1]);
    qb[SW] += sw * (0.25 / hc.dx[2]);
}

// ------------------------------------------------ transdo + Riemann (pairs)
// transdo of one side of the D-face: the normal state (qp or qm of the
// face) with the T-flux differences of cell vc.
template <int T>
__device__ __forceinline__ void t1_state(const CView &qs, unsigned v0, unsigned vc,
                                         const Canon &L, const CView &F, double cdt,
                                         const HydroConst &hc, double *out) {
    double nq[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) nq[s] = qs.ld(v0, s);
    double fl[NF], fr[NF], gl[NG], gr[NG];
    load_tf(F, vc, fl, gl);
    load_tf(F, vc + shift_b<T>(L), fr, gr);
    TransDiff t;
    trans_diff_one(fl, fr, gl, gr, cdt, t);
    transverse_one_d(nq, t, cdt, kGamma, hc.trd, hc.small_pres, out);
}
template <int T>
__device__ __forceinline__ void t1_side_adj(const CView &qs, double *QSb, unsigned v0, size_t e0,
                                            unsigned vc, size_t ec, const Canon &L,
                                            const CView &F, double *F1b, double cdt,
                                            const HydroConst &hc, const double *ob) {
    const size_t cs = L.cs;
    double nq[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) nq[s] = qs.ld(v0, s);
    double fl[NF], fr[NF], gl[NG], gr[NG];
    load_tf(F, vc, fl, gl);
    load_tf(F, vc + shift_b<T>(L), fr, gr);
    TransDiff t;
    trans_diff_one(fl, fr, gl, gr, cdt, t);
    double nb[NS] = {};
    TransDiffBar tb{};
    transverse_one_d_adj(nq, t, cdt, kGamma, hc.trd, hc.small_pres, ob, nb, tb);
#pragma unroll
    for (int s = 0; s < NH; ++s) QSb[s * cs + e0] += nb[s];
    double flb[NF] = {}, frb[NF] = {}, glb[NG] = {}, grb[NG] = {};
    trans_diff_one_adj(fl, fr, gl, gr, cdt, tb, flb, frb, glb, grb);
    tf_adj(F1b, cs, ec, flb, glb);
    tf_adj(F1b, cs, ec + shift_e<T>(L), frb, grb);
}

template <int D, int T>
__device__ __forceinline__ void trans1_pair_adj(const BoxEnt &E, const Canon &L, const CView &q,
                                                const CView &qp, const CView &qm, unsigned v0,
                                                size_t e0, double cdt, const HydroConst &hc) {
    const unsigned sd = shift_b<D>(L);
    const size_t esd = shift_e<D>(L), cs = L.cs;
    const CView F = cview(L, slot_ptr<S_F1 + T>(E), NTF);
    double fb[NF], gb[NG];
    take_tf_bar(aslot<f2_slot(D, T)>(E), cs, e0, fb, gb);
    double smb[NS] = {}, spb[NS] = {};
    {
        double sp[NS], sm[NS];
        t1_state<T>(qp, v0, v0, L, F, cdt, hc, sp);
        t1_state<T>(qm, v0, v0 - sd, L, F, cdt, hc, sm);
        const FaceAux ax = face_aux_c(q, v0 - sd, v0);
        FaceAuxBar ab{};
        face_flux_adj(D, hc, sm, sp, ax, fb, gb, smb, spb, ab);
        aux_adj(aslot<S_Q>(E), cs, e0 - esd, e0, ab);
    }
    double *F1b = aslot<S_F1 + T>(E);
    t1_side_adj<T>(qp, aslot<S_QP0 + D>(E), v0, e0, v0, e0, L, F, F1b, cdt, hc, spb);
    t1_side_adj<T>(qm, aslot<S_QM0 + D>(E), v0, e0, v0 - sd, e0 - esd, L, F, F1b, cdt, hc, smb);
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
    const unsigned v0 = L.off(i, j, k);
    const size_t e0 = eoff(L, i, j, k);
    trans1_pair_adj<D, T0>(E, L, q, qp, qm, v0, e0, cda, a.hc);
    trans1_pair_adj<D, T1>(E, L, q, qp, qm, v0, e0, cdb, a.hc);
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
    take_tf_bar(aslot<S_F1 + D>(E), cs, e0, fb, gb);
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
// One component s of the trace: recompute its edges, s6 and the integrals'
// inputs, then push the adjoints of its three wave integrals (ipb / imb,
// zero where a wave is not used) back to the 5-point stencil, the speeds
// and the flattening coefficient.  The stencil is re-read rather than kept
// from the forward pass so that only one component is live at a time.
template <bool ALL_WAVES>
__device__ __forceinline__ void ppm_comp_adj(const CView &q, unsigned v0, unsigned sd, int s,
                                             double *Qs, size_t esd, double fl, double un,
                                             double cc, double dtdx, const double *ipb,
                                             const double *imb, double &unb, double &ccb,
                                             double &flb) {
    double st[5];
#pragma unroll
    for (int o = -2; o <= 2; ++o) st[o + 2] = q.ld(v0 + (unsigned)o * sd, s);
    double sm, sp;
    ppm_edges(st, fl, sm, sp);
    const double s6 = ppm_s6(sm, sp, st[2]);
    double smb = 0.0, spb = 0.0, s6b = 0.0, vb1 = 0.0;
    if (ALL_WAVES) {
        double vb0 = 0.0, vb2 = 0.0;
        ppm_integral_adj(sm, sp, s6, un - cc, dtdx, ipb[0], imb[0], smb, spb, s6b, vb0);
        ppm_integral_adj(sm, sp, s6, un + cc, dtdx, ipb[2], imb[2], smb, spb, s6b, vb2);
        unb += vb0 + vb2;
        ccb += -vb0 + vb2;
    }
    ppm_integral_adj(sm, sp, s6, un, dtdx, ipb[1], imb[1], smb, spb, s6b, vb1);
    unb += vb1;
    double sb[5] = {};
    sb[2] += 6.0 * s6b;
    smb -= 3.0 * s6b;
    spb -= 3.0 * s6b;
    ppm_edges_adj(st, fl, smb, spb, sb, flb);
#pragma unroll
    for (int o = -2; o <= 2; ++o) aadd(Qs + (ptrdiff_t)o * (ptrdiff_t)esd, sb[o + 2]);
}

// the three wave integrals (Ip, Im) of component s
__device__ __forceinline__ void ppm_waves(const CView &q, unsigned v0, unsigned sd, int s,
                                          double fl, double un, double cc, double dtdx,
                                          double *ip, double *im) {
    double st[5];
#pragma unroll
    for (int o = -2; o <= 2; ++o) st[o + 2] = q.ld(v0 + (unsigned)o * sd, s);
    double sm, sp;
    ppm_edges(st, fl, sm, sp);
    const double s6 = ppm_s6(sm, sp, st[2]);
    ppm_integral(sm, sp, s6, un - cc, dtdx, ip[0], im[0]);
    ppm_integral(sm, sp, s6, un, dtdx, ip[1], im[1]);
    ppm_integral(sm, sp, s6, un + cc, dtdx, ip[2], im[2]);
}

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
    const size_t cs = L.cs, e0 = eoff(L, i, j, k), esd = shift_e<D>(L), e1 = e0 + esd;
    const double cc = q.ld(v0, SC);
    const double un = q.ld(v0, SN);
    const double fl = flat.ld(v0, 0);
    // the adjoints of the two face states this cell traces (read and cleared)
    double pb[NH], mb[NH];
    {
        double *QPb = aslot<S_QP0 + D>(E) + e0, *QMb = aslot<S_QM0 + D>(E) + e1;
#pragma unroll
        for (int s = 0; s < NH; ++s) {
            pb[s] = QPb[s * cs];
            mb[s] = QMb[s * cs];
            QPb[s * cs] = 0.0;
            QMb[s * cs] = 0.0;
        }
    }
    // forward: the wave integrals of the projected components
    PPMWaves Wm, Wp;  // Im (plus state of face i), Ip (minus state of face i+1)
    {
        double ip[3], im[3];
        ppm_waves(q, v0, sd, SR, fl, un, cc, dtdx, ip, im);
#pragma unroll
        for (int w = 0; w < 3; ++w) { Wm.rho[w] = im[w]; Wp.rho[w] = ip[w]; }
        ppm_waves(q, v0, sd, SN, fl, un, cc, dtdx, ip, im);
#pragma unroll
        for (int w = 0; w < 3; ++w) { Wm.un[w] = im[w]; Wp.un[w] = ip[w]; }
        ppm_waves(q, v0, sd, SP, fl, un, cc, dtdx, ip, im);
#pragma unroll
        for (int w = 0; w < 3; ++w) { Wm.p[w] = im[w]; Wp.p[w] = ip[w]; }
        ppm_waves(q, v0, sd, SE, fl, un, cc, dtdx, ip, im);
#pragma unroll
        for (int w = 0; w < 3; ++w) {
            Wm.rhoe[w] = im[w]; Wp.rhoe[w] = ip[w];
            Wm.gam[w] = kGamma; Wp.gam[w] = kGamma;
        }
    }
    // adjoints of the characteristic projections
    PPMWavesBar Bm{}, Bp{};
    ppm_project_adj(+1, Wm, un, cc, hc.small_dens, hc.small_pres, pb[SR], pb[SN], pb[SP],
                    pb[SE], Bm);
    ppm_project_adj(-1, Wp, un, cc, hc.small_dens, hc.small_pres, mb[SR], mb[SN], mb[SP],
                    mb[SE], Bp);
    // back through the integrals and the edge reconstruction, per component
    double unb = 0.0, ccb = 0.0, flb = 0.0;
    double *Qb = aslot<S_Q>(E) + e0;
    ppm_comp_adj<true>(q, v0, sd, SR, Qb + SR * cs, esd, fl, un, cc, dtdx, Bp.rho, Bm.rho, unb, ccb, flb);
    ppm_comp_adj<true>(q, v0, sd, SN, Qb + SN * cs, esd, fl, un, cc, dtdx, Bp.un, Bm.un, unb, ccb, flb);
    ppm_comp_adj<true>(q, v0, sd, SP, Qb + SP * cs, esd, fl, un, cc, dtdx, Bp.p, Bm.p, unb, ccb, flb);
    ppm_comp_adj<true>(q, v0, sd, SE, Qb + SE * cs, esd, fl, un, cc, dtdx, Bp.rhoe, Bm.rhoe, unb, ccb, flb);
    {
        // transverse velocities: the u wave only
        const double ipb[3] = {0.0, mb[T1], 0.0}, imb[3] = {0.0, pb[T1], 0.0};
        ppm_comp_adj<false>(q, v0, sd, T1, Qb + T1 * cs, esd, fl, un, cc, dtdx, ipb, imb, unb, ccb, flb);
    }
    {
        const double ipb[3] = {0.0, mb[T2], 0.0}, imb[3] = {0.0, pb[T2], 0.0};
        ppm_comp_adj<false>(q, v0, sd, T2, Qb + T2 * cs, esd, fl, un, cc, dtdx, ipb, imb, unb, ccb, flb);
    }
    aadd(Qb + SN * cs, unb);
    aadd(Qb + SC * cs, ccb);
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

// adjoint of one direction's flattening coefficient to the pressures of
// its 7-point stencil (the velocities enter only through sign tests)
template <int D>
__device__ __forceinline__ void flat_dir_adj(const CView &q, const Canon &L, unsigned v0,
                                             size_t e0, int i, int j, int k, const Box &qbox,
                                             double fdb, double *Pb) {
    if (fdb == 0.0) return;
    double p[7], u[7];
    if (flat_dir_v<D>(q, L, v0, i, j, k, qbox, p, u) == 1.0 &&
        !(qbox.contains(i + 3 * ex_(D, 0), j + 3 * ex_(D, 1), k + 3 * ex_(D, 2)) &&
          qbox.contains(i - 3 * ex_(D, 0), j - 3 * ex_(D, 1), k - 3 * ex_(D, 2))))
        return;  // stencil outside q's box: the coefficient is the constant 1
    double pb[7] = {};
    flatten_1d_adj(p, u, fdb, pb);
    const ptrdiff_t es = (ptrdiff_t)shift_e<D>(L);
#pragma unroll
    for (int o = -3; o <= 3; ++o) aadd(Pb + e0 + o * es, pb[o + 3]);
}

__global__ void __launch_bounds__(kBlock)
k_flat_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_G2)
    const size_t cs = L.cs, e0 = eoff(L, i, j, k);
    double *Fbp = aslot<S_FLAT>(E) + e0;
    const double fb = *Fbp;
    if (fb == 0.0) return;
    *Fbp = 0.0;
    if (a.hc.use_flattening != 1) return;
    const CView q = cview(L, slot_ptr<S_Q>(E), NQC);
    const unsigned v0 = L.off(i, j, k);
    const Box qbox = grow(E.bx, CAMR_NGROW);
    double f[3];
    {
        double p[7], u[7];
        f[0] = flat_dir_v<0>(q, L, v0, i, j, k, qbox, p, u);
        f[1] = flat_dir_v<1>(q, L, v0, i, j, k, qbox, p, u);
        f[2] = flat_dir_v<2>(q, L, v0, i, j, k, qbox, p, u);
    }
    const double m1 = vmin(1.0, f[0]);
    const double m2 = vmin(m1, f[1]);
    double m2b = 0.0, m1b = 0.0, dummy = 0.0, fdb[3] = {};
    vmin_adj(m2, f[2], fb, m2b, fdb[2]);
    vmin_adj(m1, f[1], m2b, m1b, fdb[1]);
    vmin_adj(1.0, f[0], m1b, dummy, fdb[0]);
    flat_dir_adj<0>(q, L, v0, e0, i, j, k, qbox, fdb[0], aslot<S_Q>(E) + SP * cs);
    flat_dir_adj<1>(q, L, v0, e0, i, j, k, qbox, fdb[1], aslot<S_Q>(E) + SP * cs);
    flat_dir_adj<2>(q, L, v0, e0, i, j, k, qbox, fdb[2], aslot<S_Q>(E) + SP * cs);
}

// ------------------------------------------------------------ ghost fill
// clean_state(S_old) + FillPatch + clean_state(Sborder) + ctoprim and the
// S_new copy: adjoints of Sborder, primitives and S_new (valid cells) to
// the adjoint of the raw S_old of the owning box.
__global__ void __launch_bounds__(kBlock)
k_fill_adj(Batch B, AdjArgs a) {
    BATCH_POINT(ST_G4)
    const size_t cs = L.cs, e0 = eoff(L, i, j, k);
    // the adjoints of the Sborder copy and the primitives here (read and cleared)
    double b[NVAR] = {}, qb[NQC] = {};
    {
        double *Ub = aslot<S_U>(E) + e0, *Qb = aslot<S_Q>(E) + e0;
#pragma unroll
        for (int n = 0; n < NH; ++n) {
            b[n] = Ub[n * cs];
            qb[n] = Qb[n * cs];
            Ub[n * cs] = 0.0;
            Qb[n * cs] = 0.0;
        }
        qb[SC] = Qb[SC * cs];
        qb[SCS] = Qb[SCS * cs];
        Qb[SC * cs] = 0.0;
        Qb[SCS * cs] = 0.0;
    }
    divu_gather(aslot<S_DIVU>(E), e0, L, a.hc, qb);
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
    if (E.bx.contains(i, j, k)) {
        const Arr bin = a.sbar_in[E.id].a;
        for (int n = 0; n < NH; ++n) b[n] += bin(i, j, k, n);
    }
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
    const LaunchCfg c = eng.cfg(kind, -1, parity, /*linear=*/true);
    if (c.grid == 0) return;
    Region reg(prof, name, st);
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.block), 0, st, c.B, a);
}

} // namespace

// Zero ranges in one launch: block b clears kZeroChunk doubles of range u,
// where off[u] <= b < off[u + 1].
constexpr int kZeroChunk = 4096;
constexpr int kZeroRanges = 64;
struct ZeroList {
    int n;
    int off[kZeroRanges + 1];
    size_t start[kZeroRanges], len[kZeroRanges];
};

__global__ void __launch_bounds__(kBlock) k_zero_list(double *p, ZeroList z) {
    const int b = blockIdx.x;
    int lo = 0, hi = z.n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (z.off[mid] <= b) lo = mid;
        else hi = mid;
    }
    const size_t beg = (size_t)(b - z.off[lo]) * kZeroChunk;
    const size_t end = beg + kZeroChunk < z.len[lo] ? beg + kZeroChunk : z.len[lo];
    double *q = p + z.start[lo];
    for (size_t t = beg + threadIdx.x; t < end; t += blockDim.x) q[t] = 0.0;
}

void zero_ranges(double *p, const std::vector<size_t> &start, const std::vector<size_t> &len,
                 hipStream_t st) {
    const size_t n = start.size();
    if (n > (size_t)kZeroRanges) {
        for (size_t r = 0; r < n; ++r) zero_doubles(p + start[r], len[r], st);
        return;
    }
    ZeroList z;
    z.n = (int)n;
    z.off[0] = 0;
    for (size_t r = 0; r < n; ++r) {
        z.start[r] = start[r];
        z.len[r] = len[r];
        z.off[r + 1] = z.off[r] + (int)((len[r] + kZeroChunk - 1) / kZeroChunk);
    }
    if (z.off[n] == 0) return;
    hipLaunchKernelGGL(k_zero_list, dim3((unsigned)z.off[n]), dim3(kBlock), 0, st, p, z);
}

void zero_doubles(double *p, size_t n, hipStream_t st) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(kBlock), 0, st, p, n);
}

int adjoint_step_kernels(const FusedEngine &eng, int parity, const AdjArgs &a, hipStream_t st,
                         Prof *prof) {
    go(eng, "adj_post", prof, k_post_adj, ST_VALID, parity, st, a);
    go(eng, "adj_transverse_2", prof, k_trans2_adj<0>, ST_T2 + 0, parity, st, a);
    go(eng, "adj_transverse_2", prof, k_trans2_adj<1>, ST_T2 + 1, parity, st, a);
    go(eng, "adj_transverse_2", prof, k_trans2_adj<2>, ST_T2 + 2, parity, st, a);
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
    go(eng, "adj_divu_clear", prof, k_divu_clear, ST_DIVU, parity, st, a);
    return hipGetLastError() == hipSuccess ? CAMR_OK : CAMR_ERR_HIP;
}

} // namespace camr
