"""ORACLE (test infrastructure only): CPU float64 restatement of adFVM's
adjoint-viscosity stabilisation -- the viscous map of apps/adjoint.py:127-140:
computeAdjointViscosity (postpro.py:491-678) and viscositySolver
(postpro.py:680-719) with the GPU heat-equation solve of
matop_cuda.cpp:108-235.

Only tests/ may import this module, and only as the checker.

How each function is pinned (tests/test_oracle.py):
  gradients        exact for linear U, p on a Cartesian box whose fixedValue
                   ghosts hold the face values; divU = trace(gradU)
  matrix           closed forms: the uniform-state matrix (only the gradrho
                   term of postpro.py:553 survives), the linear-field
                   velocity block, and every entry coupling gradp or gradc
                   into rows/columns 0 and 4 on states where only gradp
                   (linear p, uniform T) or only gradc (uniform p, c linear)
                   is non-zero, with the largest eigenvalue of each such
                   matrix in closed form; symmetry
  max_eigenvalue   numpy.linalg.eigvalsh (LAPACK), the routine the
                   reference calls (dsyev/ssyev, scaling.cpp:107-127)
  heat_coefficients / jacobi
                   the converged sweeps against a direct sparse solve of
                   the same system (scipy.sparse.linalg.spsolve); constants
                   preserved; 'uniform' gives DT = scaling in closed form
Boundary faces carry no flux in the heat equation: the internal-neighbour
table (cellNeighboursMatOp, cmesh.cpp:250-263, -1 at every boundary face,
cyclic included) is the one whose -1 test the solver makes; the reference
binds mesh.cellNeighbours there, whose boundary entries index ghost cells
past the solver's internal-cell arrays.
"""
import numpy as np
import torch

from .fvm_oracle import CYCLIC_TYPES

UREF, TREF, PREF = 33., 300., 1e5            # density.py:57-59


def _np(x):
    return x.detach().numpy() if torch.is_tensor(x) else np.asarray(x)


def primitive_with_ghosts(oc, u, ghostInit=None):
    """(5, nIC) conservative state -> U (nC,3), T, p (nC,1) with ghost rows
    filled by the boundary conditions (density.py:416-430)."""
    x = torch.as_tensor(u, dtype=torch.float64)
    U, T, p = oc.primitive(x[0:1].T, x[1:4].T, x[4:5].T)
    U, T, p = oc.ghosts(U, T, p, ghostInit)
    return _np(U), _np(T), _np(p)


def viscosity_gradients(oc, U, T, p):
    """gradU (nIC,3,3), divU (nIC,), gradp (nIC,3), gradc (nIC,3):
    postpro.py:495-540 -- central face values on internal and coupled faces,
    the ghost's value on other boundary faces; op.grad / op.div, owner side
    only on boundary faces."""
    m = oc.mesh
    g, R = oc.gamma, oc.R
    nIC, nIF = oc.nIC, oc.nIF
    owner, nb = m.owner.astype(np.int64), m.neighbour.astype(np.int64)
    w = m.weights.reshape(-1, 1)
    A = m.areas.reshape(-1, 1)
    n = m.normals
    V = m.volumes.reshape(-1)
    c = np.sqrt(g * T * R)
    UF = U[owner] * w + U[nb] * (1 - w)
    pF = p[owner] * w + p[nb] * (1 - w)
    cF = c[nb] * (1 - w) + c[owner] * w
    for pt in oc.patches:
        if pt['n'] == 0 or m.boundary[pt['name']]['type'] in CYCLIC_TYPES:
            continue
        s = slice(pt['start'], pt['start'] + pt['n'])
        UF[s] = U[nb[s]]
        pF[s] = p[nb[s]]
        cF[s] = np.sqrt(g * T[nb[s]] * R)
    gradU = np.zeros((nIC, 3, 3))
    divU = np.zeros(nIC)
    gradp = np.zeros((nIC, 3))
    gradc = np.zeros((nIC, 3))
    for f in range(len(owner)):
        for side, sgn in ((owner[f], 1.0), (nb[f] if f < nIF else -1, -1.0)):
            if side < 0:
                continue
            k = sgn * A[f, 0] / V[side]
            gradU[side] += k * np.outer(UF[f], n[f])
            divU[side] += k * UF[f] @ n[f]
            gradp[side] += k * pF[f, 0] * n[f]
            gradc[side] += k * cF[f, 0] * n[f]
    return gradU, divU, gradp, gradc


def matrix(oc, vtype, U, T, p, gradU, divU, gradp, gradc):
    """MS = (M + M^T)/2 per cell (nIC,5,5), postpro.py:542-621."""
    g = oc.gamma
    sg, g1 = np.sqrt(g), g - 1
    sg1 = np.sqrt(g1)
    nIC = len(divU)
    Z = np.zeros(nIC)
    p = p[:nIC, 0]
    T = T[:nIC, 0]
    rho = p / (oc.R * T)
    c = np.sqrt(g * p / rho)
    gradrho = g * (gradp - (c * p)[:, None]) / (c * c)[:, None]
    gU = gradU
    if vtype == 'abarbanel':
        b = c / sg
        a = sg1 * c / sg
        gb = gradc / sg
        ga = gradc * sg1 / sg
        M1 = [[divU, gb[:, 0], gb[:, 1], gb[:, 2], Z],
              [gb[:, 0], divU, Z, Z, ga[:, 0]],
              [gb[:, 1], Z, divU, Z, ga[:, 1]],
              [gb[:, 2], Z, Z, divU, ga[:, 2]],
              [Z, ga[:, 0], ga[:, 1], ga[:, 2], divU]]
        t1 = (b / rho)[:, None] * gradrho
        t2 = (a / (2 * p))[:, None] * gradp
        t3 = 2 * ga / g1
        M2 = [[Z, t1[:, 0], t1[:, 1], t1[:, 2], sg1 * divU / 2],
              [Z, gU[:, 0, 0], gU[:, 0, 1], gU[:, 0, 2], t2[:, 0]],
              [Z, gU[:, 1, 0], gU[:, 1, 1], gU[:, 1, 2], t2[:, 1]],
              [Z, gU[:, 2, 0], gU[:, 2, 1], gU[:, 2, 2], t2[:, 2]],
              [Z, t3[:, 0], t3[:, 1], t3[:, 2], g1 * divU / 2]]
    elif vtype == 'turkel':
        M1 = [[divU, gradc[:, 0], gradc[:, 1], gradc[:, 2], Z],
              [gradc[:, 0], divU, Z, Z, Z],
              [gradc[:, 1], Z, divU, Z, Z],
              [gradc[:, 2], Z, Z, divU, Z],
              [Z, Z, Z, Z, divU]]
        t1 = gradp / (rho * c)[:, None]
        t2 = g1 * gradp / (2 * rho * c)[:, None]
        t3 = gradp * PREF / (2 * g * p * rho * UREF)[:, None]
        t4 = (gradp - (c * c)[:, None] * gradrho) * UREF / PREF
        M2 = [[g1 * divU / 2, t1[:, 0], t1[:, 1], t1[:, 2], divU * PREF / (2 * rho * c * UREF)],
              [t2[:, 0], gU[:, 0, 0], gU[:, 0, 1], gU[:, 0, 2], t3[:, 0]],
              [t2[:, 1], gU[:, 1, 0], gU[:, 1, 1], gU[:, 1, 2], t3[:, 1]],
              [t2[:, 2], gU[:, 2, 0], gU[:, 2, 1], gU[:, 2, 2], t3[:, 2]],
              [Z, t4[:, 0], t4[:, 1], t4[:, 2], g1 * divU / 2]]
    else:
        raise ValueError(vtype)
    M = np.stack([np.stack(r, -1) for r in M1], -2) / 2 - np.stack([np.stack(r, -1) for r in M2], -2)
    return (M + np.swapaxes(M, -1, -2)) / 2


def max_eigenvalue(MS):
    """Largest eigenvalue per cell (the last of LAPACK's ascending list,
    scaling.cpp:121-125)."""
    return np.linalg.eigvalsh(MS)[:, -1]


def scaled_viscosity(oc, vtype, u, scaling, ghostInit=None):
    """M_2norm * scaling / sqrt(sum M_2norm^2 V / sum V) per cell,
    postpro.py:623-652."""
    V = oc.mesh.volumes.reshape(-1)
    if vtype == 'uniform':
        e = np.ones(oc.nIC)
    else:
        U, T, p = primitive_with_ghosts(oc, u, ghostInit)
        gU, dU, gp, gc = viscosity_gradients(oc, U, T, p)
        e = max_eigenvalue(matrix(oc, vtype, U, T, p, gU, dU, gp, gc))
    N = (e * e * V / V.sum()).sum()
    return e * scaling / np.sqrt(N)


def heat_coefficients(oc, M, dt):
    """Per cell slot nd = -(A DT/delta) dt/V for internal neighbours (DT the
    central face value of M), cd = -sum nd (matop_cuda.cpp:187-205).
    Returns nd (nIC,6), cd (nIC,), neighbour (nIC,6) with -1 where no flux."""
    m = oc.mesh
    nIC = oc.nIC
    cf = m.cellFaces.astype(np.int64)
    owner, nbr = m.owner.astype(np.int64), m.neighbour.astype(np.int64)
    w = m.weights.reshape(-1)
    A, d = m.areas.reshape(-1), m.deltas.reshape(-1)
    V = m.volumes.reshape(-1)
    own = owner[cf] == np.arange(nIC)[:, None]
    other = np.where(own, nbr[cf], owner[cf])
    internal = cf < oc.nIF
    nb = np.where(internal, other, -1)
    DT = M[owner[cf]] * w[cf] + np.where(internal, M[np.clip(nbr[cf], 0, nIC - 1)], 0) * (1 - w[cf])
    nd = np.where(internal, -(A[cf] * DT / d[cf]) * dt / V[:, None], 0.)
    return nd, -nd.sum(1), nb


def jacobi(nd, cd, nb, b, iterations):
    """`iterations` Jacobi sweeps phi' = (b - sum_j nd_j phi[nb_j])/(cd + 1)
    from phi = b, per column of b (nIC,k) (matop_cuda.cpp:108-124,217-228)."""
    phi = b.copy()
    valid = nb >= 0
    idx = np.where(valid, nb, 0)
    for _ in range(iterations):
        lap = np.zeros_like(b)
        for j in range(6):
            lap = lap + np.where(valid[:, j:j + 1], nd[:, j:j + 1] * phi[idx[:, j]], 0.)
        phi = (b - lap) / (cd + 1)[:, None]
    return phi


def adjoint_viscosity(oc, vtype, u, ua, dt, scaling, iterations, ghostInit=None):
    """The viscous map's effect on the adjoint state (5, nIC) (apps/adjoint.py
    viscositySolver, postpro.py:680-719): returns (ua', M_2norm)."""
    M = scaled_viscosity(oc, vtype, u, scaling, ghostInit)
    nd, cd, nb = heat_coefficients(oc, M, dt)
    V = oc.mesh.volumes.reshape(-1)
    b = (np.asarray(ua, np.float64) / V).T
    phi = jacobi(nd, cd, nb, b, iterations)
    return (phi * V[:, None]).T, M
