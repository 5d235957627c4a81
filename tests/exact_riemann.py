"""Exact Riemann solver for the 1-D ideal-gas Euler equations (Toro, *Riemann
Solvers and Numerical Methods for Fluid Dynamics*, 3rd ed., ch. 4: Newton
iteration for p* on eqs. 4.5-4.7, then self-similar sampling, sec. 4.5).

Written for the Sod gate of crates/cfd_solver/tests/sod_exact_convergence.rs:67-188,
whose own exact solver is the reference's independent check; this one is
independent of both the oracle and the GPU code.
"""
import math


def _fk(p, rho_k, p_k, a_k, g):
    if p > p_k:  # shock branch
        A = 2.0 / ((g + 1.0) * rho_k)
        B = (g - 1.0) / (g + 1.0) * p_k
        q = math.sqrt(A / (p + B))
        f = (p - p_k) * q
        df = q * (1.0 - 0.5 * (p - p_k) / (B + p))
    else:        # rarefaction branch
        f = 2.0 * a_k / (g - 1.0) * ((p / p_k) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho_k * a_k) * (p / p_k) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(g, rl, ul, pl, rr, ur, pr):
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(200):
        fl, dfl = _fk(p, rl, pl, al, g)
        fr, dfr = _fk(p, rr, pr, ar, g)
        dp = (fl + fr + (ur - ul)) / (dfl + dfr)
        pn = max(1e-12, p - dp)
        if abs(pn - p) <= 1e-14 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, al, g)
    fr, _ = _fk(p, rr, pr, ar, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample_density(g, rl, ul, pl, rr, ur, pr, ps, us, s):
    """Density of the exact solution at similarity coordinate s = x/t."""
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    if s <= us:  # left of the contact
        if ps > pl:  # left shock
            sl = ul - al * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
            if s <= sl:
                return rl
            return rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
        shl = ul - al
        als = al * (ps / pl) ** ((g - 1) / (2 * g))
        stl = us - als
        if s <= shl:
            return rl
        if s >= stl:
            return rl * (ps / pl) ** (1 / g)
        return rl * (2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)) ** (2 / (g - 1))
    if ps > pr:  # right shock
        sr = ur + ar * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
        if s >= sr:
            return rr
        return rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
    shr = ur + ar
    ars = ar * (ps / pr) ** ((g - 1) / (2 * g))
    str_ = us + ars
    if s >= shr:
        return rr
    if s <= str_:
        return rr * (ps / pr) ** (1 / g)
    return rr * (2 / (g + 1) - (g - 1) / ((g + 1) * ar) * (ur - s)) ** (2 / (g - 1))
