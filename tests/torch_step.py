"""Float64 PyTorch restatement of one level step on a single periodic box,
for reference gradients by autograd (test infrastructure only).

The step is CAMR_advance on the default path for a periodic domain, PPM,
zero gravity (Source/CAMR_advance.cpp:38-164): clean_state of S_old, the
periodic ghost fill of Sborder (4 ghosts), clean_state of Sborder, ctoprim
(Hydro_ctoprim.H:11-101), flattening (flatten.H:19-99), trace_ppm
(PPM.cpp:11-411, PPM.H:25-183), the first Riemann solves (hydro_cmpflx /
riemannus, Hydro_cmpflx.H:153-310, Hydro_riemann.H:8-173), hydro_transdo
and the corrected Riemann solves (Godunov_utils_3D.H:22-239), hydro_divu
(Hydro_utils_K.H:12-169), hydro_transdd and the final Riemann solves
(Godunov_utils_3D.H:245-482), adjust_fluxes (Hydro_adjust_fluxes.cpp:13-44),
hydro_pdivu / hydro_update (Hydro_utils_K.H:175-257), the S_new Saxpy and
the two clean_state passes after it.  Only the hydro variables (rho, rho u,
rho E, rho e) are carried: T and the species never feed back into them on
this path, so their outputs are not computed.

Every stage is evaluated on the whole padded box with periodic shifts of
the array; values near the array faces are wrong, but no valid cell reads
them (the CTU dependency radius is the ghost width).  The one place where
the box edge changes a value the valid cells see, flatten's "1 when the
+-3 stencil leaves q's box" (flatten.H:41-44), is reproduced explicitly.
Branches are torch.where selections, with vmin/vmax ties resolved as
std::min/std::max do, so autograd differentiates the branch the forward
takes.
"""
from __future__ import annotations

import torch

GAMMA = 1.4
NG = 4  # ghost cells of Sborder
URHO, UMX, UMY, UMZ, UEDEN, UEINT = range(6)


def vmin(a, b):
    return torch.where(b < a, b, a)


def vmax(a, b):
    return torch.where(a < b, b, a)


def sh(x, d, o):
    """x at offset +o along direction d (0 = x, the last axis)."""
    return torch.roll(x, shifts=-o, dims=x.dim() - 1 - d)


def _c(x, v):
    return torch.full_like(x, v)


# ------------------------------------------------------------- state
def reset_e(U, eta2):
    """CAMR_reset_internal_e with allow_small_energy = 1, allow_negative_energy
    = 0 (Utils/CAMR_reset_internal_e.H:94-120): rho e <- rho E - rho ke where
    that is positive and above eta2 of rho E."""
    rho = U[URHO]
    rinv = 1.0 / rho
    up, vp, wp = U[UMX] * rinv, U[UMY] * rinv, U[UMZ] * rinv
    ke = 0.5 * (up * up + vp * vp + wp * wp)
    assert bool((U[UEDEN] >= 0.0).all())
    rho_eint = U[UEDEN] - rho * ke
    take = (rho_eint > 0.0) & (rho_eint / U[UEDEN] > eta2)
    assert bool((take | (U[UEINT] > 0.0)).all())
    ueint = torch.where(take, rho_eint, U[UEINT])
    ueden = torch.where(take, U[UEDEN], ueint + rho * ke)
    return torch.stack([U[URHO], U[UMX], U[UMY], U[UMZ], ueden, ueint])


def prim(U, small_num):
    """compact primitives rho, u, v, w, p, rho e, c, c_small (eta1 = 1:
    e = rho e / rho)."""
    rho = U[URHO]
    rinv = 1.0 / rho
    vx, vy, vz = U[UMX] * rinv, U[UMY] * rinv, U[UMZ] * rinv
    e = U[UEINT] * rinv
    p = (GAMMA - 1.0) * rho * e
    c = torch.sqrt(GAMMA * p / rho)
    return {"r": rho, "u": [vx, vy, vz], "p": p, "re": e * rho, "c": c,
            "cs": vmax(_c(c, small_num), small_num * c)}


# ------------------------------------------------------------- flattening
def flatten_dir(p, un, d):
    """flatten.H:19-99 along direction d, for every cell."""
    P = {o: sh(p, d, o) for o in range(-3, 4)}
    Un = {o: sh(un, d, o) for o in range(-3, 4)}
    small_pres, shktst, zcut1, zcut2 = 1.e-200, 0.33e0, 0.75e0, 0.85e0
    dzcut = 1.0 / (zcut2 - zcut1)
    qm, qp = P[-1], P[1]
    dp = qp - qm
    right = dp > 0.0  # ishft = 1
    denom = vmax(_c(dp, small_pres), torch.abs(P[2] - P[-2]))
    zeta = torch.abs(dp) / denom
    z = vmin(_c(dp, 1.0), vmax(_c(dp, 0.0), dzcut * (zeta - zcut1)))
    tst = torch.where(Un[-1] - Un[1] >= 0.0, _c(dp, 1.0), _c(dp, 0.0))
    chi = torch.where(torch.abs(dp) / vmin(qp, qm) > shktst, tst, _c(dp, 0.0))

    def pick(o):  # value at offset o - ishft
        return torch.where(right, P[o - 1], P[o + 1])

    def pick_u(o):
        return torch.where(right, Un[o - 1], Un[o + 1])

    pip, pim, pipp, pimm = pick(1), pick(-1), pick(2), pick(-2)
    dp2 = pip - pim
    denom2 = vmax(_c(dp, small_pres), torch.abs(pipp - pimm))
    zeta2 = torch.abs(dp2) / denom2
    z2 = vmin(_c(dp, 1.0), vmax(_c(dp, 0.0), dzcut * (zeta2 - zcut1)))
    tst2 = torch.where(pick_u(-1) - pick_u(1) >= 0.0, _c(dp, 1.0), _c(dp, 0.0))
    chi2 = torch.where(torch.abs(dp2) / vmin(pip, pim) > shktst, tst2, _c(dp, 0.0))
    return 1.0 - vmax(chi2 * z2, chi * z)


def flattening(q, N):
    f = torch.ones_like(q["p"])
    idx = torch.arange(N)
    for d in range(3):
        fd = flatten_dir(q["p"], q["u"][d], d)
        inside = (idx >= 3) & (idx <= N - 4)  # i +- 3 within the padded box
        shape = [1, 1, 1]
        shape[2 - d] = N
        fd = torch.where(inside.view(shape), fd, torch.ones_like(fd))
        f = vmin(f, fd)
    return f


# ------------------------------------------------------------- PPM
def vl_limited(a, b, c):
    dsl = 2.0 * (b - a)
    dsr = 2.0 * (c - b)
    dsc = 0.5 * (c - a)
    sg = torch.where(dsc < 0.0, -torch.ones_like(dsc), torch.ones_like(dsc))
    r = sg * vmin(torch.abs(dsc), vmin(torch.abs(dsl), torch.abs(dsr)))
    return torch.where(dsl * dsr > 0.0, r, torch.zeros_like(r))


def ppm_edges(s, flat):
    """PPM.H:25-110 on the 5-point stencil s[0..4] (offsets -2..2)."""
    dl0 = vl_limited(s[0], s[1], s[2])
    dc = vl_limited(s[1], s[2], s[3])
    dr0 = vl_limited(s[2], s[3], s[4])
    sm = 0.5 * (s[2] + s[1]) - (1.0 / 6.0) * (dc - dl0)
    sm = vmax(sm, vmin(s[2], s[1]))
    sm = vmin(sm, vmax(s[2], s[1]))
    sp = 0.5 * (s[3] + s[2]) - (1.0 / 6.0) * (dr0 - dc)
    sp = vmax(sp, vmin(s[3], s[2]))
    sp = vmin(sp, vmax(s[3], s[2]))
    omf = 1.0 - flat
    sm = flat * sm + omf * s[2]
    sp = flat * sp + omf * s[2]
    ext = (sp - s[2]) * (s[2] - sm) <= 0.0
    hi = torch.abs(sp - s[2]) >= 2.0 * torch.abs(sm - s[2])
    lo = torch.abs(sm - s[2]) >= 2.0 * torch.abs(sp - s[2])
    sp2 = torch.where(ext, s[2], torch.where(hi, 3.0 * s[2] - 2.0 * sm, sp))
    sm2 = torch.where(ext, s[2], torch.where(hi, sm, torch.where(lo, 3.0 * s[2] - 2.0 * sp, sm)))
    return sm2, sp2


def ppm_integral(sm, sp, s6, speed, dtdx):
    sigma = torch.abs(speed) * dtdx
    k = 1.0 - (2.0 / 3.0) * sigma
    neg = speed <= 0.0
    ip = torch.where(neg, sp, sp - 0.5 * sigma * (sp - sm - k * s6))
    im = torch.where(neg, sm + 0.5 * sigma * (sp - sm + k * s6), sm)
    return ip, im


def ppm_project(side, W, un, cc, small_dens, small_pres):
    """The characteristic projection of trace_ppm (PPM.cpp:204-409) for one
    face state; W maps 'r', 'u', 'p', 're' to the three wave integrals."""
    w = 0 if side > 0 else 2
    rho_ref = vmax(W["r"][w], _c(un, small_dens))
    un_ref = W["u"][w]
    p_ref = vmax(W["p"][w], _c(un, small_pres))
    re_ref = W["re"][w]
    rinv = 1.0 / rho_ref
    cc_ref = torch.sqrt(GAMMA * p_ref / rho_ref)
    csq = cc_ref * cc_ref
    ccinv = 1.0 / cc_ref
    h_g = (p_ref + re_ref) * rinv / csq
    dum = un_ref - W["u"][0]
    dptotm = p_ref - W["p"][0]
    drho = rho_ref - W["r"][1]
    dptot = p_ref - W["p"][1]
    drhoe = re_ref - W["re"][1]
    dup = un_ref - W["u"][2]
    dptotp = p_ref - W["p"][2]
    alpham = 0.5 * (dptotm * rinv * ccinv - dum) * rho_ref * ccinv
    alphap = 0.5 * (dptotp * rinv * ccinv + dup) * rho_ref * ccinv
    alpha0r = drho - dptot / csq
    alpha0e = drhoe - dptot * h_g
    smv, spv = un - cc, un + cc

    def fac(speed, keep_pos):
        # side +1: 0 for a positive speed, -1 for negative, -1/2 at zero;
        # side -1 the other way round
        one = torch.ones_like(speed)
        pos, neg = speed > 0.0, speed < 0.0
        if keep_pos:
            return torch.where(pos, 0.0 * one, torch.where(neg, -one, -0.5 * one))
        return torch.where(pos, -one, torch.where(neg, 0.0 * one, -0.5 * one))

    keep = side > 0
    alpham = alpham * fac(smv, keep)
    alphap = alphap * fac(spv, keep)
    f0 = fac(un, keep)
    alpha0r = alpha0r * f0
    alpha0e = alpha0e * f0
    apm = alphap + alpham
    rho_o = vmax(rho_ref + alphap + alpham + alpha0r, _c(un, small_dens))
    un_o = un_ref + (alphap - alpham) * cc_ref * rinv
    p_o = vmax(p_ref + apm * csq, _c(un, small_pres if side > 0 else small_dens))
    re_o = re_ref + apm * h_g * csq + alpha0e
    return rho_o, un_o, p_o, re_o


def trace_ppm(q, flat, d, dtdx, small_dens, small_pres):
    """Face states of direction d: qp at face i (plus state, from cell i) and
    qm at face i (minus state, from cell i-1)."""
    comps = {"r": q["r"], "u0": q["u"][0], "u1": q["u"][1], "u2": q["u"][2],
             "p": q["p"], "re": q["re"]}
    un, cc = q["u"][d], q["c"]
    Ip, Im = {}, {}
    for name, s in comps.items():
        st = [sh(s, d, o) for o in range(-2, 3)]
        sm, sp = ppm_edges(st, flat)
        s6 = 6.0 * st[2] - 3.0 * (sm + sp)
        Ip[name], Im[name] = zip(*[ppm_integral(sm, sp, s6, v, dtdx)
                                   for v in (un - cc, un, un + cc)])
    n = "u%d" % d
    t1, t2 = [("u%d" % t) for t in range(3) if t != d]
    Wm = {"r": Im["r"], "u": Im[n], "p": Im["p"], "re": Im["re"]}
    Wp = {"r": Ip["r"], "u": Ip[n], "p": Ip["p"], "re": Ip["re"]}
    r, u, p, re = ppm_project(+1, Wm, un, cc, small_dens, small_pres)
    qp = {"r": r, n: u, "p": p, "re": re, t1: Im[t1][1], t2: Im[t2][1]}
    r, u, p, re = ppm_project(-1, Wp, un, cc, small_dens, small_pres)
    qm_cell = {"r": r, n: u, "p": p, "re": re, t1: Ip[t1][1], t2: Ip[t2][1]}
    qm = {k: sh(v, d, -1) for k, v in qm_cell.items()}  # minus state of face i
    return qp, qm


# ------------------------------------------------------------- Riemann
def riemann(ql, qr, d, cav, smallc, hc):
    """riemannus (Hydro_riemann.H:8-173) through hydro_cmpflx's component
    choice (Hydro_cmpflx.H:181-223, 286-297) on an interior face."""
    small, small_dens, small_pres, smallu = hc["small"], hc["small_dens"], hc["small_pres"], hc["smallu"]
    n = "u%d" % d
    t1, t2 = [("u%d" % t) for t in range(3) if t != d]
    z = torch.zeros_like(cav)
    wsmall = smallc * small_dens
    rl = vmax(ql["r"], z + small_dens)
    pl = vmax(ql["p"], z + small_pres)
    rr = vmax(qr["r"], z + small_dens)
    pr = vmax(qr["p"], z + small_pres)
    ul, ur = ql[n], qr[n]
    wl = vmax(wsmall, torch.sqrt(torch.abs(GAMMA * pl * rl)))
    wr = vmax(wsmall, torch.sqrt(torch.abs(GAMMA * pr * rr)))
    wsum = wl + wr
    pstar = vmax(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / wsum, z + small_pres)
    ustar = ((wl * ul + wr * ur) + (pl - pr)) / wsum
    ustar = torch.where(torch.abs(ustar) < smallu * 0.5 * (torch.abs(ul) + torch.abs(ur)), z, ustar)
    pos, neg = ustar > 0.0, ustar < 0.0

    def choose(a, b):
        return torch.where(pos, a, torch.where(neg, b, 0.5 * (a + b)))

    ro = vmax(choose(rl, rr), z + small_dens)
    uo = choose(ul, ur)
    po = choose(pl, pr)
    reo = choose(ql["re"], qr["re"])
    gamco = choose(z + GAMMA, z + GAMMA)
    roinv = 1.0 / ro
    co = vmax(smallc, torch.sqrt(torch.abs(gamco * po * roinv)))
    co2inv = 1.0 / (co * co)
    dpo = pstar - po
    rstar = vmax(z + small_dens, ro + dpo * co2inv)
    entho = (reo + po) * roinv * co2inv
    estar = reo + dpo * entho
    cstar = vmax(torch.sqrt(torch.abs(gamco * pstar / rstar)), smallc)
    sgnm = torch.where(neg, z - 1.0, z + 1.0)
    spout = co - sgnm * uo
    spin = cstar - sgnm * ustar
    ushock = 0.5 * (spin + spout)
    shock = dpo > 0.0
    spin = torch.where(shock, ushock, spin)
    spout = torch.where(shock, ushock, spout)
    sdiff = spout - spin
    scr = torch.where(sdiff == 0.0, small * cav, sdiff)
    frac = vmax(z, vmin(z + 1.0, (1.0 + (spout + spin) / scr) * 0.5))
    v1 = torch.where(pos, ql[t1], torch.where(neg, qr[t1], 0.5 * (ql[t1] + qr[t1])))
    v2 = torch.where(pos, ql[t2], torch.where(neg, qr[t2], 0.5 * (ql[t2] + qr[t2])))
    omf = 1.0 - frac
    rho_gd = frac * rstar + omf * ro
    un = frac * ustar + omf * uo
    pg = frac * pstar + omf * po
    re_gd = frac * estar + omf * reo
    out = spout < 0.0
    rho_gd, un, pg, re_gd = (torch.where(out, ro, rho_gd), torch.where(out, uo, un),
                             torch.where(out, po, pg), torch.where(out, reo, re_gd))
    star = spin >= 0.0
    rho_gd, un, pg, re_gd = (torch.where(star, rstar, rho_gd), torch.where(star, ustar, un),
                             torch.where(star, pstar, pg), torch.where(star, estar, re_gd))
    pg = vmax(pg, z + small_pres)
    f_rho = rho_gd * un
    f = {"r": f_rho, n: f_rho * un + pg, t1: f_rho * v1, t2: f_rho * v2}
    rhoetot = re_gd + 0.5 * rho_gd * (un * un + v1 * v1 + v2 * v2)
    f["E"] = un * (rhoetot + pg)
    f["e"] = un * re_gd
    return f, {"u": un, "p": pg}


def face_flux(qm, qp, q, d, hc):
    cav = 0.5 * (q["c"] + sh(q["c"], d, -1))
    smallc = vmax(q["cs"], sh(q["cs"], d, -1))
    return riemann(qm, qp, d, cav, smallc, hc)


# ------------------------------------------------------------- transverse
def transdo(qn, f_l, g_l, f_r, g_r, cdtdx, trd, small_pres):
    """hydro_transdo for one side: qn corrected by the differences of the
    fluxes (f_l, g_l) and (f_r, g_r) of one transverse direction."""
    d = {k: cdtdx * (f_r[k] - f_l[k]) for k in ("r", "u0", "u1", "u2", "E")}
    dAup = g_r["p"] * g_r["u"] - g_l["p"] * g_l["u"]
    pav = 0.5 * (g_r["p"] + g_l["p"])
    dAu = g_r["u"] - g_l["u"]
    rr = qn["r"]
    u0, u1, u2 = qn["u0"], qn["u1"], qn["u2"]
    ekin = 0.5 * rr * (u0 * u0 + u1 * u1 + u2 * u2)
    ru0, ru1, ru2 = u0 * rr, u1 * rr, u2 * rr
    re = qn["re"] + ekin
    rrn, r0, r1, r2, ren = rr - d["r"], ru0 - d["u0"], ru1 - d["u1"], ru2 - d["u2"], re - d["E"]
    reset = (rrn < 0.0) if trd == 1 else torch.zeros_like(rr, dtype=torch.bool)
    rrn, r0, r1, r2, ren = (torch.where(reset, rr, rrn), torch.where(reset, ru0, r0),
                            torch.where(reset, ru1, r1), torch.where(reset, ru2, r2),
                            torch.where(reset, re, ren))
    out = {"r": rrn, "u0": r0 / rrn, "u1": r1 / rrn, "u2": r2 / rrn}
    out["re"] = ren - 0.5 * (r0 * r0 + r1 * r1 + r2 * r2) / rrn
    pnew = vmax(qn["p"] - cdtdx * (dAup + pav * dAu * (GAMMA - 1.0)), torch.zeros_like(rr) + small_pres)
    out["p"] = torch.where(reset, qn["p"], pnew)
    return out


def transdd(qn, fa_l, ga_l, fa_r, ga_r, fb_l, gb_l, fb_r, gb_r, ca, cb, trd, small_pres):
    """hydro_transdd for one side, zero sources."""
    d = {k: ca * (fa_r[k] - fa_l[k]) + cb * (fb_r[k] - fb_l[k]) for k in ("r", "u0", "u1", "u2", "E")}
    pxnew = ca * ((ga_r["p"] * ga_r["u"] - ga_l["p"] * ga_l["u"]) +
                  0.5 * (ga_r["p"] + ga_l["p"]) * (ga_r["u"] - ga_l["u"]) * (GAMMA - 1.0))
    pynew = cb * ((gb_r["p"] * gb_r["u"] - gb_l["p"] * gb_l["u"]) +
                  0.5 * (gb_r["p"] + gb_l["p"]) * (gb_r["u"] - gb_l["u"]) * (GAMMA - 1.0))
    rr = qn["r"]
    rrn = rr - d["r"]
    ru0, ru1, ru2 = rr * qn["u0"], rr * qn["u1"], rr * qn["u2"]
    ekin = 0.5 * (ru0 * ru0 + ru1 * ru1 + ru2 * ru2) / rr
    re = qn["re"] + ekin
    r0, r1, r2, ren = ru0 - d["u0"], ru1 - d["u1"], ru2 - d["u2"], re - d["E"]
    reset = (rrn < 0.0) if trd == 1 else torch.zeros_like(rr, dtype=torch.bool)
    rrn, r0, r1, r2, ren = (torch.where(reset, rr, rrn), torch.where(reset, ru0, r0),
                            torch.where(reset, ru1, r1), torch.where(reset, ru2, r2),
                            torch.where(reset, re, ren))
    out = {"r": rrn, "u0": r0 / rrn, "u1": r1 / rrn, "u2": r2 / rrn}
    out["re"] = ren - 0.5 * (r0 * r0 + r1 * r1 + r2 * r2) / rrn
    pn = torch.where(reset, qn["p"], qn["p"] - pxnew - pynew)
    out["p"] = vmax(pn, torch.zeros_like(rr) + small_pres)
    return out


def shift_dict(x, d, o):
    return {k: sh(v, d, o) for k, v in x.items()}


# ------------------------------------------------------------- the step
def level_step(S, dt, dx, hc):
    """One step of an n^3 periodic box: S is (6, n, n, n) [rho, rho u, rho v,
    rho w, rho E, rho e] in (z, y, x) order; returns S_new, same shape."""
    n = S.shape[-1]
    N = n + 2 * NG
    idx = torch.arange(-NG, n + NG) % n
    Sb = S[:, idx][:, :, idx][:, :, :, idx]
    Sb = reset_e(Sb, hc["eta2"])  # clean_state(S_old) on the source cells
    Sb = reset_e(Sb, hc["eta2"])  # clean_state(Sborder)
    q = prim(Sb, hc["small_num"])
    flat = flattening(q, N) if hc["use_flattening"] == 1 else torch.ones_like(q["p"])
    qc = {"r": q["r"], "u0": q["u"][0], "u1": q["u"][1], "u2": q["u"][2], "p": q["p"], "re": q["re"]}
    qp, qm, F1 = {}, {}, {}
    for d in range(3):
        qp[d], qm[d] = trace_ppm(q, flat, d, dt / dx[d], hc["small_dens"], hc["small_pres"])
        F1[d] = face_flux(qm[d], qp[d], q, d, hc)
    # transdo + Riemann: F2[(d, t)] = d-flux from states corrected by t
    F2 = {}
    for d in range(3):
        for t in range(3):
            if t == d:
                continue
            cdt = 1.0 / 3.0 * dt / dx[t]
            f, g = F1[t]
            sp = transdo(qp[d], f, g, shift_dict(f, t, 1), shift_dict(g, t, 1), cdt, hc["trd"], hc["small_pres"])
            fm, gm = shift_dict(f, d, -1), shift_dict(g, d, -1)
            sm = transdo(qm[d], fm, gm, shift_dict(fm, t, 1), shift_dict(gm, t, 1), cdt, hc["trd"], hc["small_pres"])
            F2[(d, t)] = face_flux(sm, sp, q, d, hc)
    # nodal divu (8-point)
    def avg4(v, a, b):
        return v + sh(v, a, -1) + sh(v, b, -1) + sh(sh(v, a, -1), b, -1)
    ux = avg4(q["u"][0] - sh(q["u"][0], 0, -1), 2, 1)
    vy = avg4(q["u"][1] - sh(q["u"][1], 1, -1), 2, 0)
    wz = avg4(q["u"][2] - sh(q["u"][2], 2, -1), 0, 1)
    divu = 0.25 * ux / dx[0] + 0.25 * vy / dx[1] + 0.25 * wz / dx[2]
    # transdd + final Riemann + artificial viscosity
    hdt = 0.5 * dt
    FL = {}
    for d in range(3):
        t0, t1 = [t for t in range(3) if t != d]
        ca, cb = hdt / dx[t0], hdt / dx[t1]
        fa, ga = F2[(t0, t1)]
        fb, gb = F2[(t1, t0)]

        def side(qn, o):
            fal, gal = shift_dict(fa, d, o), shift_dict(ga, d, o)
            fbl, gbl = shift_dict(fb, d, o), shift_dict(gb, d, o)
            return transdd(qn, fal, gal, shift_dict(fal, t0, 1), shift_dict(gal, t0, 1),
                           fbl, gbl, shift_dict(fbl, t1, 1), shift_dict(gbl, t1, 1),
                           ca, cb, hc["trd"], hc["small_pres"])

        sp = side(qp[d], 0)
        sm = side(qm[d], -1)
        f, g = face_flux(sm, sp, q, d, hc)
        xs = 0.25 * (divu + sh(divu, t0, 1) + sh(divu, t1, 1) + sh(sh(divu, t0, 1), t1, 1))
        div = hc["difmag"] * vmin(torch.zeros_like(xs), xs)
        area = dx[t0] * dx[t1]
        fu = [f["r"], f["u0"], f["u1"], f["u2"], f["E"], f["e"]]
        fu = [(fu[c] + dx[d] * div * (Sb[c] - sh(Sb[c], d, -1))) * area for c in range(6)]
        FL[d] = (fu, g)
    # update: hydro_update + p div u + Saxpy (valid cells)
    vol = dx[0] * dx[1] * dx[2]
    volinv = 1.0 / vol
    a = [dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1]]
    pd = 0.0
    for d in range(3):
        g = FL[d][1]
        pd = pd + (sh(g["p"], d, 1) + g["p"]) * (sh(g["u"], d, 1) * a[d] - g["u"] * a[d])
    pdivu = 0.5 * pd / vol
    out = []
    for c in range(6):
        F = [FL[d][0][c] for d in range(3)]
        ds = F[0] - sh(F[0], 0, 1)
        ds = ds + F[1] - sh(F[1], 1, 1)
        ds = ds + F[2] - sh(F[2], 2, 1)
        ds = ds * volinv
        if c == UEINT:
            ds = ds - pdivu
        out.append(Sb[c] + dt * ds)
    Sn = torch.stack(out)[:, NG:NG + n, NG:NG + n, NG:NG + n]
    Sn = reset_e(Sn, hc["eta2"])
    Sn = reset_e(Sn, hc["eta2"])
    return Sn


def default_consts(ppm_type=1, use_flattening=1, trd=1):
    return {"small": 1.e-8, "small_dens": 1.e-7, "small_pres": 1.e-200, "smallu": 1.e-12,
            "small_num": 1.e-8, "difmag": 0.1, "eta2": 1.e-4, "use_flattening": use_flattening,
            "trd": trd}
