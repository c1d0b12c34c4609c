"""The exact Riemann solution that the Sod runs are checked against, pinned to
the reference's own exact-solution output (cases/shockTube/output: Fryxell's
riemann.f, t = 0.006341272, rho/p/u at 500 cell centres) and to closed-form
properties of the wave pattern."""
import os

import numpy as np

from oracle import exact_riemann

HERE = os.path.dirname(__file__)
T_OUT = 0.006341272
LEFT = (1.0, 0.0, 1e5)
RIGHT = (0.125, 0.0, 1e4)


def golden():
    return np.loadtxt(os.path.join(HERE, 'golden', 'shocktube_exact_output.txt'), skiprows=2)


def test_against_reference_output():
    d = golden()
    sol = exact_riemann.sample(d[:, 1], T_OUT, 0.0, LEFT, RIGHT)
    # the golden file prints 8 significant digits
    for k, col in ((0, 2), (1, 4), (2, 3)):
        ref = d[:, col]
        got = sol[:, k]
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 2e-6 * scale, (k, np.abs(got - ref).max() / scale)


def test_star_state_rankine_hugoniot():
    """p*, u* satisfy the right shock's jump conditions (mass and momentum)
    and the left rarefaction's isentrope/Riemann invariant."""
    g = 1.4
    ps, us = exact_riemann.star_state(*LEFT, *RIGHT, g=g)
    rL, uL, pL = LEFT
    rR, uR, pR = RIGHT
    # right shock: rho*, shock speed S from mass, momentum balance
    rs = rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
    S = (rs * us - rR * uR) / (rs - rR)
    assert abs((rs * (us - S) ** 2 + ps) - (rR * (uR - S) ** 2 + pR)) < 1e-8 * pL
    # left rarefaction: u + 2c/(g-1) constant, p/rho^g constant
    cL = np.sqrt(g * pL / rL)
    rsl = rL * (ps / pL) ** (1 / g)
    cs = np.sqrt(g * ps / rsl)
    assert abs((uL + 2 * cL / (g - 1)) - (us + 2 * cs / (g - 1))) < 1e-9 * cL


def test_symmetric_problem():
    """Mirror symmetry: swapping the states and negating x and u mirrors the
    solution."""
    x = np.linspace(-4, 4, 41)
    a = exact_riemann.sample(x, 0.004, 0.0, LEFT, RIGHT)
    b = exact_riemann.sample(-x, 0.004, 0.0, (RIGHT[0], -RIGHT[1], RIGHT[2]), (LEFT[0], -LEFT[1], LEFT[2]))
    assert np.allclose(a[:, 0], b[:, 0], rtol=1e-12)
    assert np.allclose(a[:, 2], b[:, 2], rtol=1e-12)
    assert np.allclose(a[:, 1], -b[:, 1], rtol=1e-12, atol=1e-12)


def test_trivial_problem():
    x = np.linspace(-1, 1, 9)
    s = exact_riemann.sample(x, 0.01, 0.0, (1.2, 3.0, 2e5), (1.2, 3.0, 2e5))
    assert np.allclose(s, [[1.2, 3.0, 2e5]] * 9, rtol=1e-12)
