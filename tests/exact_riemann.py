"""Exact solution of the 1-D Riemann problem for an ideal gas (Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4), used
to check the oracle's Sod tube against the physics, independently of the
reference code."""
from __future__ import annotations

import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        f = (p - pk) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1.0 - (p - pk) / (2.0 * (B + p)))
    else:  # rarefaction
        f = 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g=1.4):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = 0.5 * (pl + pr)
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-15 * p:
            break
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, rl, ul, pl, rr, ur, pr, g=1.4):
    """Density, velocity, pressure at positions x (array) and time t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((3, len(x)))
    for n, xx in enumerate(x):
        s = (xx - x0) / t
        if s <= us:  # left of contact
            if ps > pl:
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    out[:, n] = (rl, ul, pl)
                else:
                    r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                    out[:, n] = (r, us, ps)
            else:
                shl = ul - cl
                cml = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cml
                if s <= shl:
                    out[:, n] = (rl, ul, pl)
                elif s >= stl:
                    out[:, n] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[:, n] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    out[:, n] = (rr, ur, pr)
                else:
                    r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                    out[:, n] = (r, us, ps)
            else:
                shr = ur + cr
                cmr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + cmr
                if s >= shr:
                    out[:, n] = (rr, ur, pr)
                elif s <= str_:
                    out[:, n] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[:, n] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
