"""Exact solution of the 1-D Euler Riemann problem for an ideal gas.

Newton iteration on the star pressure with the shock (Rankine-Hugoniot) and
rarefaction (isentropic) branches of the pressure function, then sampling of
the self-similar solution (E. F. Toro, Riemann Solvers and Numerical Methods
for Fluid Dynamics, 3rd ed., ch. 4.2-4.5).  Test infrastructure only.
"""
from __future__ import annotations

import math

import numpy as np


def _f(p, rho, pk, ck, gamma):
    if p > pk:  # shock
        a = 2.0 / ((gamma + 1.0) * rho)
        b = (gamma - 1.0) / (gamma + 1.0) * pk
        q = math.sqrt(a / (p + b))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (b + p))
    # rarefaction
    pr = p / pk
    return (2.0 * ck / (gamma - 1.0) * (pr ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0),
            1.0 / (rho * ck) * pr ** (-(gamma + 1.0) / (2.0 * gamma)))


def star_state(left, right, gamma=1.4):
    rl, ul, pl = left
    rr, ur, pr = right
    cl = math.sqrt(gamma * pl / rl)
    cr = math.sqrt(gamma * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, gamma)
        fr, dr = _f(p, rr, pr, cr, gamma)
        dp = (fl + fr + (ur - ul)) / (dl + dr)
        p_new = max(1e-12, p - dp)
        if abs(p_new - p) < 1e-14 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, gamma)
    fr, _ = _f(p, rr, pr, cr, gamma)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, gamma=1.4):
    """Density, velocity and pressure at positions x and time t > 0."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(left, right, gamma)
    cl = math.sqrt(gamma * pl / rl)
    cr = math.sqrt(gamma * pr / rr)
    g1 = (gamma - 1.0) / (gamma + 1.0)
    out = np.zeros((3, len(x)))
    for i, xi in enumerate(np.asarray(x, dtype=float)):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:
                rs = rl * (ps / pl + g1) / (g1 * ps / pl + 1.0)
                sl = ul - cl * math.sqrt((gamma + 1) / (2 * gamma) * ps / pl
                                         + (gamma - 1) / (2 * gamma))
                w = (rl, ul, pl) if s <= sl else (rs, us, ps)
            else:
                rs = rl * (ps / pl) ** (1.0 / gamma)
                cs = cl * (ps / pl) ** ((gamma - 1) / (2 * gamma))
                if s <= ul - cl:
                    w = (rl, ul, pl)
                elif s >= us - cs:
                    w = (rs, us, ps)
                else:
                    u = 2.0 / (gamma + 1) * (cl + (gamma - 1) / 2 * ul + s)
                    c = 2.0 / (gamma + 1) * (cl + (gamma - 1) / 2 * (ul - s))
                    w = (rl * (c / cl) ** (2 / (gamma - 1)), u,
                         pl * (c / cl) ** (2 * gamma / (gamma - 1)))
        else:
            if ps > pr:
                rs = rr * (ps / pr + g1) / (g1 * ps / pr + 1.0)
                sr = ur + cr * math.sqrt((gamma + 1) / (2 * gamma) * ps / pr
                                         + (gamma - 1) / (2 * gamma))
                w = (rr, ur, pr) if s >= sr else (rs, us, ps)
            else:
                rs = rr * (ps / pr) ** (1.0 / gamma)
                cs = cr * (ps / pr) ** ((gamma - 1) / (2 * gamma))
                if s >= ur + cr:
                    w = (rr, ur, pr)
                elif s <= us + cs:
                    w = (rs, us, ps)
                else:
                    u = 2.0 / (gamma + 1) * (-cr + (gamma - 1) / 2 * ur + s)
                    c = 2.0 / (gamma + 1) * (cr - (gamma - 1) / 2 * (ur - s))
                    w = (rr * (c / cr) ** (2 / (gamma - 1)), u,
                         pr * (c / cr) ** (2 * gamma / (gamma - 1)))
        out[:, i] = w
    return out
