// Host-side check of the embedded-boundary hydro adjoint
// (camr_amd/csrc/hydro_eb_adjoint.hip) against central finite differences
// of the cut-cell hydro stages it differentiates (camr_amd/csrc/hydro_eb.hip).
// Built and run by tests/test_eb_adjoint_host.py.
//
// Both files' per-point stage bodies are host/device functions; this
// program runs the stages of one box point by point on the host, in launch
// order, as the batched kernels do on the device (grid.y = box, one thread
// per point).  L = sum W . dS/dt (valid cells, hydro components) for fixed
// random weights W; the adjoint gives dL/dSborder, which is compared with
// (L(S + h e) - L(S - h e)) / 2h at sampled Sborder entries.  Samples whose
// difference straddles a branch switch (one-sided slopes disagree) are
// skipped, as in adjoint_unit.cpp.  Geometries: all-regular, and a plane
// cutting the box (cut cells with apertures and centroids, covered cells
// behind it), with interior, wall and outflow domain faces.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../camr_amd/csrc/hydro_eb.hip"
#include "../camr_amd/csrc/hydro_eb_adjoint.hip"

namespace camr {

static std::mt19937_64 rng(20261205);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

struct HostGeom {
    Box fb;                       // flags / geometry box (bx grown by 7)
    std::vector<int32_t> flags;
    std::vector<double> vf, ap[3], fc[3];
    EbGeom g;
};

// Plane a.x + b = 0 through the box (fluid where a.x + b > 0): volume
// fractions, apertures and face centroids by sub-sampling, connectivity of
// every non-covered neighbour pair.  plane == false: all regular.
static void make_geom(const Box &bx, bool plane, HostGeom &G) {
    G.fb = grow(bx, 7);
    const Box &fb = G.fb;
    const int64_t n = fb.npts();
    const double nx = 0.6, ny = 0.5, nz = 0.62, off = -(nx * 3.3 + ny * 4.1 + nz * 3.7);
    auto phi = [&](double x, double y, double z) { return nx * x + ny * y + nz * z + off; };
    const int M = 8;
    auto cellfrac = [&](int i, int j, int k) {
        if (!plane) return 1.0;
        int c = 0;
        for (int a = 0; a < M; ++a)
            for (int b = 0; b < M; ++b)
                for (int d = 0; d < M; ++d)
                    c += phi(i + (a + 0.5) / M, j + (b + 0.5) / M, k + (d + 0.5) / M) > 0.0;
        return (double)c / (M * M * M);
    };
    G.vf.assign(n, 0.0);
    G.flags.assign(n, 0);
    auto idx = [&](int i, int j, int k) {
        return (int64_t)(i - fb.lo[0]) + (int64_t)(j - fb.lo[1]) * fb.len(0) +
               (int64_t)(k - fb.lo[2]) * fb.len(0) * fb.len(1);
    };
    for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
        for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
            for (int i = fb.lo[0]; i <= fb.hi[0]; ++i) G.vf[idx(i, j, k)] = cellfrac(i, j, k);
    auto type_of = [&](double v) { return v >= 1.0 ? kEbRegular : (v <= 0.0 ? kEbCovered : kEbSingle); };
    for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
        for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
            for (int i = fb.lo[0]; i <= fb.hi[0]; ++i) {
                const double v = G.vf[idx(i, j, k)];
                const int t = type_of(v);
                int32_t w = t << 27;
                if (t != kEbCovered)
                    for (int kk = -1; kk <= 1; ++kk)
                        for (int jj = -1; jj <= 1; ++jj)
                            for (int ii = -1; ii <= 1; ++ii) {
                                const int a = i + ii, b = j + jj, c = k + kk;
                                if (!fb.contains(a, b, c)) continue;
                                if (type_of(G.vf[idx(a, b, c)]) != kEbCovered)
                                    w |= 1 << ((ii + 1) + 3 * (jj + 1) + 9 * (kk + 1));
                            }
                G.flags[idx(i, j, k)] = w;
            }
    for (int d = 0; d < 3; ++d) {
        const Box nb = nodes(fb, d);
        G.ap[d].assign(nb.npts(), 0.0);
        G.fc[d].assign(2 * nb.npts(), 0.0);
        const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
        for (int k = nb.lo[2]; k <= nb.hi[2]; ++k)
            for (int j = nb.lo[1]; j <= nb.hi[1]; ++j)
                for (int i = nb.lo[0]; i <= nb.hi[0]; ++i) {
                    const int64_t e = (int64_t)(i - nb.lo[0]) + (int64_t)(j - nb.lo[1]) * nb.len(0) +
                                      (int64_t)(k - nb.lo[2]) * nb.len(0) * nb.len(1);
                    int c = 0;
                    double s0 = 0.0, s1 = 0.0;
                    for (int a = 0; a < M; ++a)
                        for (int b = 0; b < M; ++b) {
                            double x[3] = {(double)i, (double)j, (double)k};
                            const double u = (a + 0.5) / M, v = (b + 0.5) / M;
                            x[t0] += u;
                            x[t1] += v;
                            if (!plane || phi(x[0], x[1], x[2]) > 0.0) {
                                ++c;
                                s0 += u - 0.5;
                                s1 += v - 0.5;
                            }
                        }
                    G.ap[d][e] = (double)c / (M * M);
                    if (c > 0 && c < M * M) {
                        G.fc[d][e] = s0 / c;
                        G.fc[d][nb.npts() + e] = s1 / c;
                    }
                }
    }
    G.g.f.p = G.flags.data();
    G.g.f.lo0 = fb.lo[0]; G.g.f.lo1 = fb.lo[1]; G.g.f.lo2 = fb.lo[2];
    G.g.f.js = fb.len(0);
    G.g.f.ks = (int64_t)fb.len(0) * fb.len(1);
    G.g.vf = dense(G.vf.data(), fb);
    for (int d = 0; d < 3; ++d) {
        G.g.ap[d] = dense(G.ap[d].data(), nodes(fb, d));
        G.g.fc[d] = dense(G.fc[d].data(), nodes(fb, d));
    }
}

// hydro_ctoprim of every cell of qbox with rho > 0 (k_ctoprim); covered
// cells zero.
static void ctoprim_host(const Box &qbox, const Arr &u, const Arr &q, const Arr &qa,
                         const HydroConst &hc) {
    for (int k = qbox.lo[2]; k <= qbox.hi[2]; ++k)
        for (int j = qbox.lo[1]; j <= qbox.hi[1]; ++j)
            for (int i = qbox.lo[0]; i <= qbox.hi[0]; ++i) {
                const double rho = u(i, j, k, URHO);
                if (!(rho > 0.)) {
                    for (int n = 0; n < QVAR; ++n) q(i, j, k, n) = 0.;
                    for (int n = 0; n < NQAUX; ++n) qa(i, j, k, n) = 0.;
                    continue;
                }
                const double rhoinv = 1.0 / rho;
                const double vx = u(i, j, k, UMX) * rhoinv, vy = u(i, j, k, UMY) * rhoinv;
                const double vz = u(i, j, k, UMZ) * rhoinv;
                const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
                q(i, j, k, QRHO) = rho;
                q(i, j, k, QU) = vx;
                q(i, j, k, QV) = vy;
                q(i, j, k, QW) = vz;
                const double ueden = u(i, j, k, UEDEN), ueint = u(i, j, k, UEINT);
                const double e = ((ueden - kineng) / ueden > hc.eta1) ? (ueden - kineng) * rhoinv
                                                                     : ueint * rhoinv;
                for (int s = 0; s < NSPEC; ++s) q(i, j, k, QFS + s) = u(i, j, k, UFS + s) * rhoinv;
                const double p = (kGamma - 1.0) * rho * e;
                q(i, j, k, QTEMP) = eos_T_from_e(e);
                q(i, j, k, QREINT) = e * rho;
                q(i, j, k, QPRES) = p;
                q(i, j, k, QGAME) = p / (e * rho) + 1.0;
                qa(i, j, k, QDPDR) = (kGamma - 1.0) * e;
                qa(i, j, k, QDPDE) = (kGamma - 1.0) * rho;
                qa(i, j, k, QGAMC) = kGamma;
                const double cs = sqrt(kGamma * p / rho);
                qa(i, j, k, QC) = cs;
                qa(i, j, k, QCSML) = vmax(hc.small_num, hc.small_num * cs);
            }
}

template <class F>
static void each(const Box &b, F f) {
    if (!b.ok()) return;
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) f(i, j, k);
}

template <int S>
static void fwd_site(const EbBox &B, const HydroConst &hc, const double *dtp) {
    const Layout L = layout_of(B.bx, B.qbox);
    each(site_box<S>(B, L), [&](int i, int j, int k) { site_body<S>(B, L, i, j, k, hc, dtp, 0); });
}
template <int... S>
static void fwd_all(std::integer_sequence<int, S...>, const EbBox &B, const HydroConst &hc,
                    const double *dtp) {
    (fwd_site<S>(B, hc, dtp), ...);
}
template <int S>
static void adj_site(const EbBox &F, const EbBox &A, const HydroConst &hc, double dt) {
    const Layout L = layout_of(F.bx, F.qbox);
    each(adj_box<S>(F, L), [&](int i, int j, int k) { adj_body<S>(F, A, i, j, k, hc, dt); });
}
template <int... S>
static void adj_all(std::integer_sequence<int, S...>, const EbBox &F, const EbBox &A,
                    const HydroConst &hc, double dt) {
    (adj_site<S>(F, A, hc, dt), ...);
}

struct Case {
    const char *name;
    bool plane;
    int bc;        // BC of the domain faces the box touches (-1: none)
    int iorder;
};

static int run_case(const Case &cs, int samples) {
    rng.seed(20261205);
    const Box bx{{0, 0, 0}, {7, 7, 7}};
    const Box qbox = grow(bx, 6);
    HydroConst hc{};
    hc.ppm_type = 0; hc.plm_iorder = cs.iorder; hc.use_pslope = 0; hc.use_flattening = 1;
    hc.trd = 1; hc.allow_negative_energy = 0; hc.allow_small_energy = 1; hc.update_E_from_e = 1;
    hc.small = 1e-8; hc.small_dens = 1e-7; hc.small_pres = 1e-200; hc.small_temp = 1e2;
    hc.smallu = 1e-12; hc.small_num = 1e-8; hc.difmag = 0.1; hc.eta1 = 1.0; hc.eta2 = 1e-4;
    for (int d = 0; d < 3; ++d) {
        hc.dx[d] = 1.0 / 16;
        hc.domlo[d] = cs.bc < 0 ? -100 : 0;
        hc.domhi[d] = cs.bc < 0 ? 100 : 7;
        hc.bclo[d] = cs.bc < 0 ? BC_INTERIOR : cs.bc;
        hc.bchi[d] = cs.bc < 0 ? BC_INTERIOR : cs.bc;
    }
    hc.eb = 1;
    const double dt = 0.004;
    HostGeom G;
    make_geom(bx, cs.plane, G);
    // Sborder: a smooth flow, covered cells zero
    const size_t nq = qbox.npts();
    std::vector<double> S(NVAR * nq);
    const Arr Sv = dense(S.data(), qbox);
    each(qbox, [&](int i, int j, int k) {
        const double x = (i + 0.5) / 16, y = (j + 0.5) / 16, z = (k + 0.5) / 16;
        const double tp = 2 * M_PI;
        // every field varies along every axis: a field constant along a
        // direction puts its slopes there exactly on the limiter's kink
        const double rho = 1.0 + 0.2 * sin(tp * (0.5 * x + 0.3 * y + 0.2 * z) + 0.4) +
                           0.1 * cos(tp * (0.35 * x + z - y) + 0.2);
        const double u = 0.4 + 0.1 * sin(tp * (0.25 * x + y + z) + 0.3);
        const double v = 0.1 * sin(tp * (0.5 * x + 0.4 * y + z) + 0.7);
        const double w = 0.1 * cos(tp * (0.5 * x - y + 0.3 * z) + 0.1);
        const double p = 1.0 + 0.2 * sin(tp * (0.5 * x + y + 0.5 * z) + 1.1);
        // covered cells of the domain hold zeros; ghost cells beyond a
        // domain face hold a state whatever their flags (FillPatch fills
        // them by reflection or extrapolation)
        const bool in_dom = i >= hc.domlo[0] && i <= hc.domhi[0] && j >= hc.domlo[1] &&
                            j <= hc.domhi[1] && k >= hc.domlo[2] && k <= hc.domhi[2];
        const bool cov = in_dom && G.g.f.covered(i, j, k);
        Sv(i, j, k, URHO) = cov ? 0.0 : rho;
        Sv(i, j, k, UMX) = cov ? 0.0 : rho * u;
        Sv(i, j, k, UMY) = cov ? 0.0 : rho * v;
        Sv(i, j, k, UMZ) = cov ? 0.0 : rho * w;
        Sv(i, j, k, UEINT) = cov ? 0.0 : p / (kGamma - 1.0);
        Sv(i, j, k, UEDEN) = cov ? 0.0 : p / (kGamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
        Sv(i, j, k, UFS) = cov ? 0.0 : rho;
    });
    // forward arrays
    std::vector<double> qv(QVAR * nq), qav(NQAUX * nq), dsdt(NVAR * bx.npts());
    std::vector<double> flx[3];
    for (int d = 0; d < 3; ++d) flx[d].assign(NVAR * nodes(bx, d).npts(), 0.0);
    std::vector<double> scr(umdrv_eb_scratch_doubles(bx, qbox, true), 0.0);
    UmdrvEbArgs A;
    A.bx = bx;
    A.qbox = qbox;
    A.hc = hc;
    for (int d = 0; d < 3; ++d) A.periodic[d] = 0;
    A.dtp = &dt;
    A.uin = Sv;
    A.dsdt = dense(dsdt.data(), bx);
    A.q = dense(qv.data(), qbox);
    A.qaux = dense(qav.data(), qbox);
    A.srcq = A.q;
    A.has_src = 0;
    for (int d = 0; d < 3; ++d) A.flx[d] = dense(flx[d].data(), nodes(bx, d));
    A.g = G.g;
    A.regular = cs.plane ? 0 : 1;  // the shortcut of all-regular geometry
    EbBox F;
    eb_box_setup(A, scr.data(), &F, true);
    std::vector<double> W(NVAR * bx.npts(), 0.0);
    const Arr Wv = dense(W.data(), bx);
    each(bx, [&](int i, int j, int k) {
        if (G.g.f.covered(i, j, k)) return;
        for (int n = 0; n < 6; ++n) Wv(i, j, k, n) = U(-1, 1);
    });
    auto loss = [&]() {
        std::fill(dsdt.begin(), dsdt.end(), 0.0);
        ctoprim_host(qbox, Sv, F.q, F.qaux, hc);
        fwd_all(std::make_integer_sequence<int, ES_COUNT>{}, F, hc, &dt);
        // the hydro components only (the species are passive)
        double l = 0.0;
        for (size_t e = 0; e < W.size(); ++e)
            if (W[e] != 0.0) l += W[e] * dsdt[e];
        return l;
    };
    const double l0 = loss();
    if (!std::isfinite(l0)) {
        printf("%-28s loss not finite\n", cs.name);
        const Arr D = dense(dsdt.data(), bx);
        int shown = 0;
        each(bx, [&](int i, int j, int k) {
            for (int n = 0; n < 6; ++n)
                if (!std::isfinite(D(i, j, k, n)) && shown++ < 6)
                    printf("  dsdt not finite at (%d,%d,%d) comp %d, flag type %d vf %g divc %g\n", i, j, k, n,
                           eb_type(G.g.f(i, j, k)), G.g.vf(i, j, k, 0), F.divc(i, j, k, n));
        });
        return 1;
    }
    // adjoint
    std::vector<double> qb(QVAR * nq, 0.0), qab(NQAUX * nq, 0.0), sbb(NVAR * nq, 0.0);
    std::vector<double> ub(NVAR * bx.npts());
    for (size_t e = 0; e < ub.size(); ++e) ub[e] = W[e] / dt;
    // the adjoint side shares the pairs' and the final solves' face-state
    // buffers, as the level does
    std::vector<double> ascr(umdrv_eb_scratch_doubles(bx, qbox, false), 0.0);
    UmdrvEbArgs Ab = A;
    Ab.q = dense(qb.data(), qbox);
    Ab.qaux = dense(qab.data(), qbox);
    Ab.srcq = Ab.q;
    Ab.uin = dense(sbb.data(), qbox);
    Ab.dsdt = dense(ub.data(), bx);
    EbBox Ad;
    eb_box_setup(Ab, ascr.data(), &Ad, false);
    adj_all(std::make_integer_sequence<int, AS_COUNT>{}, F, Ad, hc, dt);
    const Arr Gv = dense(sbb.data(), qbox);
    // finite differences at sampled Sborder entries
    int checked = 0, skipped = 0, bad = 0;
    double worst = 0.0;
    for (int t = 0; t < samples; ++t) {
        const int i = (int)std::floor(U(-5, 13)), j = (int)std::floor(U(-5, 13));
        const int k = (int)std::floor(U(-5, 13)), n = (int)std::floor(U(0, 6));
        if (G.g.f.covered(i, j, k)) continue;
        double &x = Sv(i, j, k, n);
        const double x0 = x, h = 1e-6 * std::max(1e-3, std::fabs(x0));
        x = x0 + h;
        const double lp = loss();
        x = x0 - h;
        const double lm = loss();
        x = x0;
        const double fd = (lp - lm) / (2 * h), fr = (lp - l0) / h, fl = (l0 - lm) / h;
        const double scale = std::max(1.0, std::fabs(fd));
        if (std::fabs(fr - fl) > 1e-4 * std::max(std::fabs(fr), std::fabs(fl)) + 1e-7 * scale) {
            ++skipped;
            continue;
        }
        double err = std::fabs(Gv(i, j, k, n) - fd) / scale;
        if (!std::isfinite(err)) err = 1e300;
        ++checked;
        worst = std::max(worst, err);
        if (err > 1e-5 && bad++ < 12)
            printf("MISMATCH %s cell (%d,%d,%d) comp %d: adjoint %.10e fd %.10e\n", cs.name, i, j,
                   k, n, Gv(i, j, k, n), fd);
    }
    loss();  // leave the forward consistent
    printf("%-28s checked %5d skipped %4d worst rel err %.3e\n", cs.name, checked, skipped, worst);
    return (worst > 1e-5 || checked < samples / 4) ? 1 : 0;
}

} // namespace camr

int main(int argc, char **argv) {
    using namespace camr;
    const int samples = argc > 1 ? atoi(argv[1]) : 300;
    const Case cases[] = {
        {"regular, interior, plm 1", false, -1, 1},
        {"regular, interior, plm 2", false, -1, 2},
        {"regular, walls, plm 2", false, BC_SLIPWALL, 2},
        {"regular, outflow, plm 4", false, BC_OUTFLOW, 4},
        {"plane cut, interior, plm 2", true, -1, 2},
        {"plane cut, walls, plm 2", true, BC_SLIPWALL, 2},
    };
    int fail = 0;
    for (const Case &c : cases) fail |= run_case(c, samples);
    printf(fail ? "EB ADJOINT CHECK FAILED\n" : "EB ADJOINT CHECK OK\n");
    return fail;
}
