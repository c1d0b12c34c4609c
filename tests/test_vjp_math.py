"""CPU checks of the hand-written reverse arithmetic (adjoint_math.hpp)
through a host build of the same headers: dot-product tests against central
finite differences of the forward functions, on random states, for every
face kind, both Riemann solvers, constant and Sutherland viscosity, with and
without the limiter."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, 'tests', 'shim', 'math_shim.cpp')
OUT = os.path.join(ROOT, '_build', 'math_shim.so')


@pytest.fixture(scope='module')
def shim():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.check_call(['g++', '-O1', '-std=c++17', '-shared', '-fPIC', '-ffp-contract=off',
                           '-I', os.path.join(ROOT, 'adfvm_amd', 'csrc'), SRC, '-o', OUT])
    lib = C.CDLL(OUT)
    return lib


def P(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def rand_face(rng, kind):
    """States near (U ~ 30 m/s, T ~ 300 K, p ~ 1e5 Pa) with gradients."""
    def w():
        return np.array([30 + 5 * rng.normal(), 5 * rng.normal(), 5 * rng.normal(),
                         300 + 10 * rng.normal(), 1e5 * (1 + 0.05 * rng.normal())])
    g = lambda: rng.normal(size=15) * np.array([300] * 9 + [500] * 3 + [1e5] * 3)
    n = rng.normal(size=3); n /= np.linalg.norm(n)
    D = n + 0.2 * rng.normal(size=3); D /= np.linalg.norm(D)
    geo = np.concatenate([[1e-2], n, [0.4], [0.02], D, [0.15, 0.18], 0.01 * rng.normal(size=6)])
    wN = w()
    if kind == 1 and rng.random() < 0.5:
        wN[:3] = [0.2, -0.1, 0.05]   # near-stagnant ghost
    return w(), g(), wN, g(), geo


PHYS = {
    'roe_const': [1.4, 1004.5, 0.7, 1.8e-5, 0, 0, 1e-30, 0, 0, 0, 0],
    'lf_sutherland': [1.4, 1004.5, 0.7, 0, 1.4792e-6, 116., 1e-30, 1, 1, 1, 0],
    'roe_limited': [1.4, 1004.5, 0.7, 1.8e-5, 0, 0, 1e-30, 0, 0, 0, 1],
}


@pytest.mark.parametrize('phys', sorted(PHYS))
@pytest.mark.parametrize('kind', [0, 1, 2])
def test_face_flux_vjp_dot_product(shim, phys, kind):
    rng = np.random.default_rng(10 * kind + len(phys))
    p, pp = P(PHYS[phys])
    for trial in range(20):
        wP, gP, wN, gN, geo = rand_face(rng, kind)
        if kind != 0:
            gN = gP   # boundary/characteristic faces use the owner's gradient for the ghost
        x = np.concatenate([wP, gP, wN, gN])

        def F(xx):
            a = [P(xx[0:5]), P(xx[5:20]), P(xx[20:25]), P(xx[25:40] if kind == 0 else xx[5:20])]
            out, po = P(np.zeros(5))
            lam, pl = P(np.zeros(1))
            _, pg = P(geo)
            shim.shim_face_flux(pp, kind, a[0][1], a[1][1], a[2][1], a[3][1], pg, po, pl)
            return out.copy()
        Fb = rng.normal(size=5)
        outs = [np.zeros(5), np.zeros(15), np.zeros(5), np.zeros(15)]
        ptrs = [P(o) for o in outs]
        _, pg = P(geo)
        a = [P(wP), P(gP), P(wN), P(gN)]
        _, pFb = P(Fb)
        gNb_ptr = ptrs[1][1] if kind != 0 else ptrs[3][1]
        shim.shim_face_flux_vjp(pp, kind, a[0][1], a[1][1], a[2][1], a[3][1], pg, pFb, ptrs[0][1], ptrs[1][1],
                                ptrs[2][1], gNb_ptr)
        xb = np.concatenate([ptrs[0][0], ptrs[1][0], ptrs[2][0], ptrs[3][0]])
        v = rng.normal(size=40) * np.abs(x) * 1e-3 + 1e-3
        if kind != 0:
            v[25:40] = 0
        h = 1e-4
        fd = (8 * (F(x + h * v) - F(x - h * v)) - (F(x + 2 * h * v) - F(x - 2 * h * v))) / (12 * h)
        lhs = Fb @ fd
        rhs = xb @ v
        assert abs(lhs - rhs) <= 1e-6 * (abs(lhs) + np.abs(Fb).sum() * np.abs(fd).max() * 1e-3 + 1e-12), \
            (trial, lhs, rhs)


def test_primitive_vjp_dot_product(shim):
    rng = np.random.default_rng(0)
    p, pp = P(PHYS['roe_const'])
    for _ in range(20):
        q = np.array([1.2, 30., 4., -2., 2.6e5]) * (1 + 0.1 * rng.normal(size=5))

        def W(qq):
            w, pw = P(np.zeros(5))
            _, pq = P(qq)
            shim.shim_primitive(pp, pq, pw)
            return w.copy()
        wb = rng.normal(size=5)
        qb, pqb = P(np.zeros(5))
        _, pq = P(q)
        _, pwb = P(wb)
        shim.shim_primitive_vjp(pp, pq, pwb, pqb)
        v = rng.normal(size=5) * q * 1e-3
        h = 1e-3
        fd = (8 * (W(q + h * v) - W(q - h * v)) - (W(q + 2 * h * v) - W(q - 2 * h * v))) / (12 * h)
        assert abs(wb @ fd - qb @ v) <= 1e-7 * abs(wb @ fd) + 1e-12
