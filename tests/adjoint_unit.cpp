// Host-side check of the per-point adjoints (camr_amd/csrc/adjoint_math.h)
// against central finite differences of the forward functions they
// differentiate.  Built and run by tests/test_adjoint_unit.py.
//
// For each function: random inputs in a physical range, random output
// weights w; the adjoint gives g = J^T w; each g_i is compared with
// (w.f(x + h e_i) - w.f(x - h e_i)) / 2h.  Samples whose finite difference
// straddles a branch switch (a limiter or upwinding kink) are detected by
// comparing the two one-sided differences and skipped.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

#include "adjoint_math.h"
#include "eb_adjoint_math.h"
#include "state_math.h"

using namespace camr;

static std::mt19937_64 rng(20261204);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

using Fwd = std::function<void(const std::vector<double> &, std::vector<double> &)>;
using Adj = std::function<void(const std::vector<double> &, const std::vector<double> &,
                               std::vector<double> &)>;

struct Stat {
    int checked = 0, skipped = 0;
    double worst = 0.0;
};

static void check(const char *name, int nin, int nout, std::function<std::vector<double>()> gen,
                  Fwd f, Adj a, int samples, Stat &st) {
    for (int s = 0; s < samples; ++s) {
        std::vector<double> x = gen(), w(nout), g(nin, 0.0), y(nout);
        for (auto &v : w) v = U(-1, 1);
        f(x, y);
        a(x, w, g);
        for (int i = 0; i < nin; ++i) {
            const double h = 1e-6 * std::max(1e-3, std::fabs(x[i]));
            std::vector<double> xp = x, xm = x, yp(nout), ym(nout);
            xp[i] += h;
            xm[i] -= h;
            f(xp, yp);
            f(xm, ym);
            double lp = 0, lm = 0, l0 = 0;
            for (int o = 0; o < nout; ++o) {
                lp += w[o] * yp[o];
                lm += w[o] * ym[o];
                l0 += w[o] * y[o];
            }
            const double fd = (lp - lm) / (2 * h);
            const double fr = (lp - l0) / h, fl = (l0 - lm) / h;
            const double scale = std::max(1.0, std::fabs(fd));
            if (std::fabs(fr - fl) > 1e-4 * std::max(std::fabs(fr), std::fabs(fl)) + 1e-7 * scale) {
                st.skipped++;  // kink inside [x - h, x + h]
                continue;
            }
            const double err = std::fabs(g[i] - fd) / scale;
            st.checked++;
            if (err > st.worst) st.worst = err;
            if (err > 1e-5) {
                printf("MISMATCH %s sample %d input %d: adjoint %.12e fd %.12e\n", name, s, i,
                       g[i], fd);
            }
        }
    }
    printf("%-22s checked %6d skipped %5d worst rel err %.3e\n", name, st.checked, st.skipped,
           st.worst);
}

static HydroConst consts() {
    HydroConst hc{};
    hc.small = 1e-8;
    hc.small_dens = 1e-7;
    hc.small_pres = 1e-200;
    hc.small_temp = 1e2;
    hc.smallu = 1e-12;
    hc.small_num = 1e-8;
    hc.difmag = 0.1;
    hc.eta1 = 1.0;
    hc.eta2 = 1e-4;
    hc.trd = 1;
    hc.allow_small_energy = 1;
    hc.allow_negative_energy = 0;
    hc.update_E_from_e = 1;
    return hc;
}

int main() {
    const HydroConst hc = consts();
    bool fail = false;
    auto done = [&](const Stat &s) {
        if (s.worst > 1e-5 || s.checked < 10) fail = true;
    };

    {   // Riemann solver
        Stat st;
        auto gen = [] {
            std::vector<double> x(14);
            x[0] = U(0.2, 2); x[1] = U(-1, 1); x[2] = U(-1, 1); x[3] = U(-1, 1);
            x[4] = U(0.2, 2); x[5] = x[4] / 0.4 * U(0.9, 1.1);
            x[6] = U(0.2, 2); x[7] = U(-1, 1); x[8] = U(-1, 1); x[9] = U(-1, 1);
            x[10] = U(0.2, 2); x[11] = x[10] / 0.4 * U(0.9, 1.1);
            x[12] = U(0.5, 2); x[13] = 1e-8 * x[12];
            return x;
        };
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            RiemannOut o;
            riemann_two_shock(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10],
                              x[11], x[12], kGamma, kGamma, hc.small, x[13], hc.small_dens,
                              hc.small_pres, hc.smallu, o);
            const double v[8] = {o.f_rho, o.f_un, o.f_ut1, o.f_ut2, o.f_eden, o.f_eint, o.g_un, o.g_p};
            for (int i = 0; i < 8; ++i) y[i] = v[i];
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            RiemannBar ob{w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7]};
            RiemannInBar ib{};
            riemann_two_shock_adj(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9],
                                  x[10], x[11], x[12], kGamma, kGamma, hc.small, x[13],
                                  hc.small_dens, hc.small_pres, hc.smallu, ob, ib);
            const double v[14] = {ib.rl, ib.ul, ib.vl, ib.v2l, ib.pl, ib.rel, ib.rr,
                                  ib.ur, ib.vr, ib.v2r, ib.pr, ib.rer, ib.cav, ib.smallc};
            for (int i = 0; i < 14; ++i) g[i] = v[i];
        };
        check("riemann_two_shock", 14, 8, gen, f, a, 400, st);
        done(st);
    }

    // hydro state (SR..SE) generator
    auto state6 = [] {
        std::vector<double> q(6);
        q[SR] = U(0.3, 2); q[SU] = U(-1, 1); q[SV] = U(-1, 1); q[SW] = U(-1, 1);
        q[SP] = U(0.3, 2); q[SE] = q[SP] / 0.4 * U(0.95, 1.05);
        return q;
    };

    for (int d = 0; d < 3; ++d) {   // face_flux (interior face)
        Stat st;
        auto gen = [&] {
            std::vector<double> x = state6(), r = state6();
            x.insert(x.end(), r.begin(), r.end());
            x.push_back(U(0.5, 2)); x.push_back(U(0.5, 2));
            x.push_back(U(0.5, 2) * 1e-8); x.push_back(U(0.5, 2) * 1e-8);
            return x;
        };
        auto unpack = [](const std::vector<double> &x, double *ql, double *qr, FaceAux &a) {
            for (int i = 0; i < NS; ++i) { ql[i] = 0.0; qr[i] = 0.0; }
            for (int i = 0; i < 6; ++i) { ql[i] = x[i]; qr[i] = x[6 + i]; }
            a.c_l = x[12]; a.c_r = x[13]; a.csml_l = x[14]; a.csml_r = x[15];
            a.gam_l = kGamma; a.gam_r = kGamma;
        };
        HydroConst h2 = hc;
        for (int k = 0; k < 3; ++k) { h2.domlo[k] = -100; h2.domhi[k] = 100; }
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            double ql[NS], qr[NS], fl[NF], g[NG];
            FaceAux a;
            unpack(x, ql, qr, a);
            int mut;
            face_flux(d, 0, h2, ql, qr, a, fl, g, &mut);
            for (int i = 0; i < 6; ++i) y[i] = fl[i];
            y[6] = g[GUN]; y[7] = g[GP];
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &gr) {
            double ql[NS], qr[NS], qlb[NS] = {}, qrb[NS] = {}, fb[NF] = {}, gb[NG];
            FaceAux a;
            unpack(x, ql, qr, a);
            for (int i = 0; i < 6; ++i) fb[i] = w[i];
            gb[GUN] = w[6]; gb[GP] = w[7];
            FaceAuxBar ab{};
            face_flux_adj(d, h2, ql, qr, a, fb, gb, qlb, qrb, ab);
            for (int i = 0; i < 6; ++i) { gr[i] = qlb[i]; gr[6 + i] = qrb[i]; }
            gr[12] = ab.c_l; gr[13] = ab.c_r; gr[14] = ab.csml_l; gr[15] = ab.csml_r;
        };
        char nm[32];
        snprintf(nm, sizeof nm, "face_flux d=%d", d);
        check(nm, 16, 8, gen, f, a, 150, st);
        done(st);
    }

    {   // transdo: trans_diff_one + transverse_one_d
        Stat st;
        auto gen = [&] {
            std::vector<double> x = state6();
            for (int i = 0; i < 10; ++i) x.push_back(U(-0.3, 0.3));  // fl, fr (5 each)
            x.push_back(U(-1, 1)); x.push_back(U(0.3, 2));              // gl (un, p)
            x.push_back(U(-1, 1)); x.push_back(U(0.3, 2));              // gr
            return x;
        };
        const double cdt = 0.13;
        auto unpack = [](const std::vector<double> &x, double *qn, double *fl, double *fr,
                         double *gl, double *gr) {
            for (int i = 0; i < NS; ++i) qn[i] = 0.0;
            for (int i = 0; i < 6; ++i) qn[i] = x[i];
            for (int i = 0; i < NF; ++i) { fl[i] = 0.0; fr[i] = 0.0; }
            const int c[5] = {FR, FMX, FMY, FMZ, FE};
            for (int i = 0; i < 5; ++i) { fl[c[i]] = x[6 + i]; fr[c[i]] = x[11 + i]; }
            gl[GUN] = x[16]; gl[GP] = x[17]; gr[GUN] = x[18]; gr[GP] = x[19];
        };
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            double qn[NS], fl[NF], fr[NF], gl[NG], gr[NG], qo[NS];
            unpack(x, qn, fl, fr, gl, gr);
            TransDiff t;
            trans_diff_one(fl, fr, gl, gr, cdt, t);
            transverse_one_d(qn, t, cdt, kGamma, 1, hc.small_pres, qo);
            for (int i = 0; i < 6; ++i) y[i] = qo[i];
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            double qn[NS], fl[NF], fr[NF], gl[NG], gr[NG];
            unpack(x, qn, fl, fr, gl, gr);
            TransDiff t;
            trans_diff_one(fl, fr, gl, gr, cdt, t);
            double qob[NS] = {}, qnb[NS] = {};
            for (int i = 0; i < 6; ++i) qob[i] = w[i];
            TransDiffBar tb{};
            transverse_one_d_adj(qn, t, cdt, kGamma, 1, hc.small_pres, qob, qnb, tb);
            double flb[NF] = {}, frb[NF] = {}, glb[NG] = {}, grb[NG] = {};
            trans_diff_one_adj(fl, fr, gl, gr, cdt, tb, flb, frb, glb, grb);
            for (int i = 0; i < 6; ++i) g[i] = qnb[i];
            const int c[5] = {FR, FMX, FMY, FMZ, FE};
            for (int i = 0; i < 5; ++i) { g[6 + i] = flb[c[i]]; g[11 + i] = frb[c[i]]; }
            g[16] = glb[GUN]; g[17] = glb[GP]; g[18] = grb[GUN]; g[19] = grb[GP];
        };
        check("transdo", 20, 6, gen, f, a, 300, st);
        done(st);
    }

    {   // transdd: trans_diff_two + transverse_two_d
        Stat st;
        auto gen = [&] {
            std::vector<double> x = state6();
            for (int i = 0; i < 20; ++i) x.push_back(U(-0.3, 0.3));
            for (int i = 0; i < 4; ++i) { x.push_back(U(-1, 1)); x.push_back(U(0.3, 2)); }
            return x;
        };
        const double ca = 0.07, cb = 0.11;
        auto unpack = [](const std::vector<double> &x, double *qn, double F[4][NF], double G[4][NG]) {
            for (int i = 0; i < NS; ++i) qn[i] = 0.0;
            for (int i = 0; i < 6; ++i) qn[i] = x[i];
            const int c[5] = {FR, FMX, FMY, FMZ, FE};
            for (int k = 0; k < 4; ++k) {
                for (int i = 0; i < NF; ++i) F[k][i] = 0.0;
                for (int i = 0; i < 5; ++i) F[k][c[i]] = x[6 + 5 * k + i];
                G[k][GUN] = x[26 + 2 * k];
                G[k][GP] = x[27 + 2 * k];
            }
        };
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            double qn[NS], F[4][NF], G[4][NG], qo[NS], z[NS] = {};
            unpack(x, qn, F, G);
            TransDiff t;
            trans_diff_two(F[0], F[1], G[0], G[1], F[2], F[3], G[2], G[3], ca, cb, kGamma, t);
            transverse_two_d(qn, t, 0.05, z, 1, hc.small_pres, qo);
            for (int i = 0; i < 6; ++i) y[i] = qo[i];
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            double qn[NS], F[4][NF], G[4][NG];
            unpack(x, qn, F, G);
            TransDiff t;
            trans_diff_two(F[0], F[1], G[0], G[1], F[2], F[3], G[2], G[3], ca, cb, kGamma, t);
            double qob[NS] = {}, qnb[NS] = {};
            for (int i = 0; i < 6; ++i) qob[i] = w[i];
            TransDiffBar tb{};
            transverse_two_d_adj(qn, t, 1, hc.small_pres, qob, qnb, tb);
            double Fb[4][NF] = {}, Gb[4][NG] = {};
            trans_diff_two_adj(G[0], G[1], G[2], G[3], ca, cb, kGamma, tb, Fb[0], Fb[1], Gb[0],
                               Gb[1], Fb[2], Fb[3], Gb[2], Gb[3]);
            for (int i = 0; i < 6; ++i) g[i] = qnb[i];
            const int c[5] = {FR, FMX, FMY, FMZ, FE};
            for (int k = 0; k < 4; ++k) {
                for (int i = 0; i < 5; ++i) g[6 + 5 * k + i] = Fb[k][c[i]];
                g[26 + 2 * k] = Gb[k][GUN];
                g[27 + 2 * k] = Gb[k][GP];
            }
        };
        check("transdd", 34, 6, gen, f, a, 300, st);
        done(st);
    }

    {   // PPM edges
        Stat st;
        auto gen = [] {
            std::vector<double> x(6);
            for (int i = 0; i < 5; ++i) x[i] = U(0.5, 1.5);
            x[5] = U(0.3, 1.0);
            return x;
        };
        auto f = [](const std::vector<double> &x, std::vector<double> &y) {
            double sm, sp;
            ppm_edges(x.data(), x[5], sm, sp);
            y[0] = sm; y[1] = sp;
        };
        auto a = [](const std::vector<double> &x, const std::vector<double> &w,
                    std::vector<double> &g) {
            double sb[5] = {}, fb = 0.0;
            ppm_edges_adj(x.data(), x[5], w[0], w[1], sb, fb);
            for (int i = 0; i < 5; ++i) g[i] = sb[i];
            g[5] = fb;
        };
        check("ppm_edges", 6, 2, gen, f, a, 3000, st);
        done(st);
    }

    {   // PPM integrals
        Stat st;
        auto gen = [] {
            std::vector<double> x(4);
            x[0] = U(0.5, 1.5); x[1] = U(0.5, 1.5); x[2] = U(-0.5, 0.5); x[3] = U(-2, 2);
            return x;
        };
        auto f = [](const std::vector<double> &x, std::vector<double> &y) {
            double ip, im;
            ppm_integral(x[0], x[1], x[2], x[3], 0.3, ip, im);
            y[0] = ip; y[1] = im;
        };
        auto a = [](const std::vector<double> &x, const std::vector<double> &w,
                    std::vector<double> &g) {
            double smb = 0, spb = 0, s6b = 0, vb = 0;
            ppm_integral_adj(x[0], x[1], x[2], x[3], 0.3, w[0], w[1], smb, spb, s6b, vb);
            g[0] = smb; g[1] = spb; g[2] = s6b; g[3] = vb;
        };
        check("ppm_integral", 4, 2, gen, f, a, 500, st);
        done(st);
    }

    for (int side = -1; side <= 1; side += 2) {   // PPM projection
        Stat st;
        auto gen = [] {
            std::vector<double> x(14);
            const double r = U(0.5, 1.5), u = U(-1, 1), p = U(0.5, 1.5);
            for (int w = 0; w < 3; ++w) {
                x[w] = r * U(0.9, 1.1); x[3 + w] = u + U(-0.1, 0.1);
                x[6 + w] = p * U(0.9, 1.1); x[9 + w] = x[6 + w] / 0.4 * U(0.95, 1.05);
            }
            x[12] = u; x[13] = std::sqrt(1.4 * p / r);
            return x;
        };
        auto mk = [](const std::vector<double> &x) {
            PPMWaves W;
            for (int w = 0; w < 3; ++w) {
                W.rho[w] = x[w]; W.un[w] = x[3 + w]; W.p[w] = x[6 + w];
                W.rhoe[w] = x[9 + w]; W.gam[w] = kGamma;
            }
            return W;
        };
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            double ro, uo, po, eo;
            ppm_project(side, mk(x), x[12], x[13], hc.small_dens, hc.small_pres, ro, uo, po, eo);
            y[0] = ro; y[1] = uo; y[2] = po; y[3] = eo;
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            PPMWavesBar Ib{};
            ppm_project_adj(side, mk(x), x[12], x[13], hc.small_dens, hc.small_pres, w[0], w[1],
                            w[2], w[3], Ib);
            for (int k = 0; k < 3; ++k) {
                g[k] = Ib.rho[k]; g[3 + k] = Ib.un[k]; g[6 + k] = Ib.p[k]; g[9 + k] = Ib.rhoe[k];
            }
            g[12] = 0.0; g[13] = 0.0;
        };
        check(side > 0 ? "ppm_project +" : "ppm_project -", 14, 4, gen, f, a, 300, st);
        done(st);
    }

    {   // flattening (pressures; shock-like profiles so the ramps are active)
        Stat st;
        auto gen = [] {
            std::vector<double> x(7);
            const double j = U(0.5, 3.0);
            for (int i = 0; i < 7; ++i) x[i] = (i < 3 ? 1.0 : 1.0 + j) * U(0.9, 1.1);
            x[3] = 1.0 + 0.5 * j * U(0.8, 1.2);
            return x;
        };
        const double un[7] = {1.0, 0.8, 0.6, 0.4, 0.2, 0.0, -0.2};
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            y[0] = flatten_1d(x.data(), un);
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            double pb[7] = {};
            flatten_1d_adj(x.data(), un, w[0], pb);
            for (int i = 0; i < 7; ++i) g[i] = pb[i];
        };
        check("flatten_1d", 7, 1, gen, f, a, 2000, st);
        done(st);
    }

    {   // primitives
        Stat st;
        auto gen = [] {
            std::vector<double> v(10, 0.0);
            v[URHO] = U(0.3, 2); v[UMX] = U(-1, 1); v[UMY] = U(-1, 1); v[UMZ] = U(-1, 1);
            v[UEINT] = U(0.5, 3);
            v[UEDEN] = v[UEINT] + 0.5 * (v[UMX] * v[UMX] + v[UMY] * v[UMY] + v[UMZ] * v[UMZ]) / v[URHO];
            v[UFS] = v[URHO];
            return v;
        };
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            // the forward of prim_compact (hydro_fused.h), restated for the host
            const double rho = x[URHO], rhoinv = 1.0 / rho;
            const double vx = x[UMX] * rhoinv, vy = x[UMY] * rhoinv, vz = x[UMZ] * rhoinv;
            const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
            const double e = ((x[UEDEN] - kineng) / x[UEDEN] > hc.eta1) ? (x[UEDEN] - kineng) * rhoinv
                                                                      : x[UEINT] * rhoinv;
            const double p = (kGamma - 1.0) * rho * e;
            const double cs = sqrt(kGamma * p / rho);
            y[0] = rho; y[1] = vx; y[2] = vy; y[3] = vz; y[4] = p; y[5] = e * rho;
            y[6] = cs; y[7] = vmax(hc.small_num, hc.small_num * cs);
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            double ob[NQC] = {};
            for (int i = 0; i < 6; ++i) ob[i] = w[i];
            ob[SC] = w[6]; ob[SCS] = w[7];
            double vb[10] = {};
            prim_compact_adj(x.data(), ob, hc.eta1, hc.small_num, vb);
            for (int i = 0; i < 10; ++i) g[i] = vb[i];
        };
        check("prim_compact", 10, 8, gen, f, a, 300, st);
        done(st);
    }

    {   // energy reset (clean_state)
        Stat st;
        auto gen = [] {
            std::vector<double> v(10, 0.0);
            v[URHO] = U(0.3, 2); v[UMX] = U(-1, 1); v[UMY] = U(-1, 1); v[UMZ] = U(-1, 1);
            v[UEINT] = U(0.5, 3);
            v[UEDEN] = U(0.6, 1.4) * (v[UEINT] + 0.5 * (v[UMX] * v[UMX] + v[UMY] * v[UMY] + v[UMZ] * v[UMZ]) / v[URHO]);
            v[UFS] = v[URHO];
            v[UTEMP] = 300.0;
            return v;
        };
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            double s[10];
            for (int i = 0; i < 10; ++i) s[i] = x[i];
            reset_e_and_T(s, hc);
            for (int i = 0; i < 6; ++i) y[i] = s[i];
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            double sb[10] = {};
            for (int i = 0; i < 6; ++i) sb[i] = w[i];
            reset_e_adj(x.data(), hc, sb);
            for (int i = 0; i < 10; ++i) g[i] = (i < 6) ? sb[i] : 0.0;
        };
        check("reset_e", 10, 6, gen, f, a, 300, st);
        done(st);
    }

    for (int order : {2, 4}) {   // limited PLM slope (mc_slope)
        Stat st;
        auto gen = [] {
            std::vector<double> x(6);
            for (int i = 0; i < 5; ++i) x[i] = U(-1, 1);
            x[5] = U(0.2, 1.0);  // flattening coefficient
            return x;
        };
        auto f = [order](const std::vector<double> &x, std::vector<double> &y) {
            y[0] = mc_slope(x.data(), x[5], order);
        };
        auto a = [order](const std::vector<double> &x, const std::vector<double> &w,
                         std::vector<double> &g) {
            double sb[5] = {}, flb = 0.0;
            mc_slope_adj(x.data(), x[5], order, w[0], sb, flb);
            for (int i = 0; i < 5; ++i) g[i] = sb[i];
            g[5] = flb;
        };
        check(order == 2 ? "mc_slope order 2" : "mc_slope order 4", 6, 1, gen, f, a, 3000, st);
        done(st);
    }

    for (int d = 0; d < 3; ++d) {   // PLM characteristic trace (plm_trace)
        Stat st;
        const double dtdx = 0.37;
        auto gen = [&] {
            std::vector<double> x = state6();
            for (int i = 0; i < 6; ++i) x.push_back(U(-0.2, 0.2));  // slopes
            x.push_back(U(0.8, 2.0));                                // sound speed
            return x;
        };
        auto unpack = [](const std::vector<double> &x, double *c, double *sl, double &cc) {
            for (int i = 0; i < NS; ++i) { c[i] = 0.0; sl[i] = 0.0; }
            for (int i = 0; i < 6; ++i) { c[i] = x[i]; sl[i] = x[6 + i]; }
            cc = x[12];
        };
        auto f = [&, d](const std::vector<double> &x, std::vector<double> &y) {
            double c[NS], sl[NS], cc, qp[NS], qm[NS];
            unpack(x, c, sl, cc);
            plm_trace(d, c, sl, cc, dtdx, hc.small_dens, hc.small_pres, qp, qm);
            for (int i = 0; i < 6; ++i) { y[i] = qp[i]; y[6 + i] = qm[i]; }
        };
        auto a = [&, d](const std::vector<double> &x, const std::vector<double> &w,
                        std::vector<double> &g) {
            double c[NS], sl[NS], cc, qpb[NS] = {}, qmb[NS] = {}, cb[NS] = {}, slb[NS] = {};
            double ccb = 0.0;
            unpack(x, c, sl, cc);
            for (int i = 0; i < 6; ++i) { qpb[i] = w[i]; qmb[i] = w[6 + i]; }
            plm_trace_adj(d, c, sl, cc, dtdx, hc.small_dens, hc.small_pres, qpb, qmb, cb, slb, ccb);
            for (int i = 0; i < 6; ++i) { g[i] = cb[i]; g[6 + i] = slb[i]; }
            g[12] = ccb;
        };
        char nm[32];
        snprintf(nm, sizeof nm, "plm_trace d=%d", d);
        check(nm, 13, 12, gen, f, a, 400, st);
        done(st);
    }

    // face_flux at domain faces: wall (reflecting) and outflow overrides on
    // either side
    for (int bc : {BC_SLIPWALL, BC_OUTFLOW})
        for (int side = 0; side < 2; ++side) {
            Stat st;
            const int d = 1;
            HydroConst h2 = hc;
            for (int k = 0; k < 3; ++k) {
                h2.domlo[k] = 0; h2.domhi[k] = 9;
                h2.bclo[k] = bc; h2.bchi[k] = bc;
            }
            const int idx = side == 0 ? 0 : 10;
            auto gen = [&] {
                std::vector<double> x = state6(), r = state6();
                x.insert(x.end(), r.begin(), r.end());
                x.push_back(U(0.5, 2)); x.push_back(U(0.5, 2));
                x.push_back(U(0.5, 2) * 1e-8); x.push_back(U(0.5, 2) * 1e-8);
                return x;
            };
            auto unpack = [](const std::vector<double> &x, double *ql, double *qr, FaceAux &a) {
                for (int i = 0; i < NS; ++i) { ql[i] = 0.0; qr[i] = 0.0; }
                for (int i = 0; i < 6; ++i) { ql[i] = x[i]; qr[i] = x[6 + i]; }
                a.c_l = x[12]; a.c_r = x[13]; a.csml_l = x[14]; a.csml_r = x[15];
                a.gam_l = kGamma; a.gam_r = kGamma;
            };
            auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
                double ql[NS], qr[NS], fl[NF], g[NG];
                FaceAux a;
                unpack(x, ql, qr, a);
                int mut;
                face_flux(d, idx, h2, ql, qr, a, fl, g, &mut);
                for (int i = 0; i < 6; ++i) y[i] = fl[i];
                y[6] = g[GUN]; y[7] = g[GP];
            };
            auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                         std::vector<double> &gr) {
                double ql[NS], qr[NS], qlb[NS] = {}, qrb[NS] = {}, fb[NF] = {}, gb[NG];
                FaceAux a;
                unpack(x, ql, qr, a);
                for (int i = 0; i < 6; ++i) fb[i] = w[i];
                gb[GUN] = w[6]; gb[GP] = w[7];
                FaceAuxBar ab{};
                face_flux_bc_adj(d, idx, h2, ql, qr, a, fb, gb, qlb, qrb, ab);
                for (int i = 0; i < 6; ++i) { gr[i] = qlb[i]; gr[6 + i] = qrb[i]; }
                gr[12] = ab.c_l; gr[13] = ab.c_r; gr[14] = ab.csml_l; gr[15] = ab.csml_r;
            };
            char nm[40];
            snprintf(nm, sizeof nm, "face_flux %s %s", bc == BC_OUTFLOW ? "outflow" : "wall",
                     side ? "hi" : "lo");
            check(nm, 16, 8, gen, f, a, 150, st);
            done(st);
        }

    {   // wall flux of a cut cell (compute_hyp_wallflux)
        Stat st;
        auto gen = [&] {
            std::vector<double> x = state6();
            x.push_back(U(0.8, 2.0));        // c
            x.push_back(U(0.5, 2) * 1e-8);   // csml
            return x;
        };
        const double n0 = 0.3, n1 = -0.5, n2 = 0.2;
        auto f = [&](const std::vector<double> &x, std::vector<double> &y) {
            const double apnorm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
            const double apnorminv = 1. / apnorm;
            const double un = x[1] * n0 * apnorminv + x[2] * n1 * apnorminv + x[3] * n2 * apnorminv;
            RiemannOut o;
            riemann_two_shock(x[0], un, 0., 0., x[4], x[5], x[0], -un, 0., 0., x[4], x[5], x[6],
                              kGamma, kGamma, hc.small, x[7], hc.small_dens, hc.small_pres,
                              hc.smallu, o);
            y[0] = o.f_rho; y[1] = n0 * o.f_un; y[2] = n1 * o.f_un; y[3] = n2 * o.f_un;
            y[4] = o.f_eden; y[5] = o.f_eint;
        };
        auto a = [&](const std::vector<double> &x, const std::vector<double> &w,
                     std::vector<double> &g) {
            double fwb[NVAR] = {};
            for (int i = 0; i < 6; ++i) fwb[i] = w[i];
            double wb[6] = {}, cb = 0.0, csb = 0.0;
            wallflux_adj(x.data(), x[6], x[7], n0, n1, n2, hc, fwb, wb, cb, csb);
            for (int i = 0; i < 6; ++i) g[i] = wb[i];
            g[6] = cb; g[7] = csb;
        };
        check("wallflux", 8, 6, gen, f, a, 400, st);
        done(st);
    }

    printf(fail ? "ADJOINT CHECK FAILED\n" : "ADJOINT CHECK OK\n");
    return fail ? 1 : 0;
}
