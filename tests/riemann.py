"""Exact solution of the 1-D Riemann problem, for the Sod checks."""
import numpy as np


def exact_riemann(x, t, left, right, gamma=1.4, x0=0.5):
    """Exact solution of the 1-D Riemann problem (Toro, Riemann Solvers and
    Numerical Methods for Fluid Dynamics, ch. 4): density at points x."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    lo, hi = 1e-8, 10 * max(pl, pr)
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if f(pm, rl, pl, cl) + f(pm, rr, pr, cr) + ur - ul > 0:
            hi = pm
        else:
            lo = pm
    p = 0.5 * (lo + hi)
    um = 0.5 * (ul + ur) + 0.5 * (f(p, rr, pr, cr) - f(p, rl, pl, cl))
    g = gamma
    out = np.empty_like(x)
    for k, xx in enumerate(x):
        s = (xx - x0) / t
        if s <= um:  # left of contact
            if p > pl:
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g))
                out[k] = rl if s < sl else rl * ((p / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pl + 1))
            else:
                cml = cl * (p / pl) ** ((g - 1) / (2 * g))
                if s < ul - cl:
                    out[k] = rl
                elif s > um - cml:
                    out[k] = rl * (p / pl) ** (1 / g)
                else:
                    out[k] = rl * (2 / (g + 1) + (g - 1) / ((g + 1) * cl) * (ul - s)) ** (2 / (g - 1))
        else:
            if p > pr:
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
                out[k] = rr if s > sr else rr * ((p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1))
            else:
                cmr = cr * (p / pr) ** ((g - 1) / (2 * g))
                if s > ur + cr:
                    out[k] = rr
                elif s < um + cmr:
                    out[k] = rr * (p / pr) ** (1 / g)
                else:
                    out[k] = rr * (2 / (g + 1) - (g - 1) / ((g + 1) * cr) * (ur - s)) ** (2 / (g - 1))
    return out
