"""ORACLE (test infrastructure only): exact solution of the 1-D Riemann
problem for an ideal gas, following the published method of E. F. Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4 (pressure
function, Newton iteration for p*, sampling of the wave pattern).

It plays the part of the reference's Fortran exact solver
(cases/shockTube/riemann.f, Fryxell) whose output at t = 0.006341272 ships as
cases/shockTube/output; tests check this module against that file.
"""
import numpy as np


def _f(p, rho, pk, ck, g):
    """Toro eq. 4.6-4.7: pressure function and derivative for one side."""
    if p > pk:                                   # shock
        A = 2. / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        s = np.sqrt(A / (p + B))
        return (p - pk) * s, s * (1 - 0.5 * (p - pk) / (B + p))
    pr = p / pk                                  # rarefaction
    f = 2 * ck / (g - 1) * (pr ** ((g - 1) / (2 * g)) - 1)
    return f, 1. / (rho * ck) * pr ** (-(g + 1) / (2 * g))


def star_state(rhoL, uL, pL, rhoR, uR, pR, g=1.4, tol=1e-14):
    cL, cR = np.sqrt(g * pL / rhoL), np.sqrt(g * pR / rhoR)
    p = max(tol, 0.5 * (pL + pR))                # arithmetic-mean guess (Toro 4.46)
    for _ in range(100):
        fL, dL = _f(p, rhoL, pL, cL, g)
        fR, dR = _f(p, rhoR, pR, cR, g)
        pn = p - (fL + fR + uR - uL) / (dL + dR)
        pn = max(pn, tol)
        if abs(pn - p) / (0.5 * (pn + p)) < tol:
            p = pn
            break
        p = pn
    fL, _ = _f(p, rhoL, pL, cL, g)
    fR, _ = _f(p, rhoR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(x, t, x0, left, right, g=1.4):
    """(rho, u, p) at positions x at time t for an initial jump at x0;
    left/right = (rho, u, p)."""
    rhoL, uL, pL = left
    rhoR, uR, pR = right
    ps, us = star_state(rhoL, uL, pL, rhoR, uR, pR, g)
    cL, cR = np.sqrt(g * pL / rhoL), np.sqrt(g * pR / rhoR)
    out = np.zeros((len(x), 3))
    for i, xi in enumerate(np.asarray(x, float)):
        S = (xi - x0) / t
        if S <= us:                              # left of contact
            if ps > pL:                          # left shock
                SL = uL - cL * np.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if S <= SL:
                    out[i] = (rhoL, uL, pL)
                else:
                    r = rhoL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
                    out[i] = (r, us, ps)
            else:                                # left rarefaction
                SHL = uL - cL
                cs = cL * (ps / pL) ** ((g - 1) / (2 * g))
                STL = us - cs
                if S <= SHL:
                    out[i] = (rhoL, uL, pL)
                elif S >= STL:
                    out[i] = (rhoL * (ps / pL) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - S))
                    r = rhoL * (c / cL) ** (2 / (g - 1))
                    out[i] = (r, 2 / (g + 1) * (cL + (g - 1) / 2 * uL + S), pL * (c / cL) ** (2 * g / (g - 1)))
        else:                                    # right of contact
            if ps > pR:                          # right shock
                SR = uR + cR * np.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if S >= SR:
                    out[i] = (rhoR, uR, pR)
                else:
                    r = rhoR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
                    out[i] = (r, us, ps)
            else:                                # right rarefaction
                SHR = uR + cR
                cs = cR * (ps / pR) ** ((g - 1) / (2 * g))
                STR = us + cs
                if S >= SHR:
                    out[i] = (rhoR, uR, pR)
                elif S <= STR:
                    out[i] = (rhoR * (ps / pR) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - S))
                    r = rhoR * (c / cR) ** (2 / (g - 1))
                    out[i] = (r, 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + S), pR * (c / cR) ** (2 * g / (g - 1)))
    return out
