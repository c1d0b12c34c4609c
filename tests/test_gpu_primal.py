"""HIP primal path against the oracle (same mesh, boundary spec, physics
and state).  Tolerances: FP64 kernels agree with the FP64 oracle to round-off
(relative L2 <= 1e-10 on gradients, residual and step); FP32 kernels are
checked against the same FP64 oracle with the FP32 tolerances stated per
test."""
import numpy as np
import pytest

from tests import casefactory as cf

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_fields(case, u):
    oc = case.oracle()
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    rho, rhoU, rhoE = t(u[0:1].T), t(u[1:4].T), t(u[4:5].T)
    gi = [t(a) for a in cf.ghost_init(case)]
    (d0, d1, d2, dtc), (U, T, p, gU, gT, gp) = oc.equation(rho, rhoU, rhoE, ghostInit=gi, with_fields=True)
    grad = np.concatenate([gU.reshape(-1, 9).T.numpy(), gT.reshape(-1, 3).T.numpy(), gp.reshape(-1, 3).T.numpy()])
    R = np.concatenate([d0.T.numpy(), d1.T.numpy(), d2.T.numpy()])
    prim = np.concatenate([U.T.numpy(), T.T.numpy(), p.T.numpy()])
    return oc, grad, R, dtc[:, 0].numpy(), prim


SHEAR = ((1., 0.3, 0.2), (0., 1., 0.25), (0., 0., 1.))

CASES = {
    'tgv6': lambda: (lambda c: (c[0], cf.tgv_state(*c)))(cf.tgv(6, seed=3)),
    'sod': lambda: (lambda c: (c, cf.sod_state(c)))(cf.sod(40)),
    'mixed_upt': lambda: (lambda c: (c, cf.smooth_state(c, 1)))(cf.mixed_box(1, 'CBC_UPT')),
    'mixed_totalpt_lf': lambda: (lambda c: (c, cf.smooth_state(c, 2)))(cf.mixed_box(2, 'CBC_TOTAL_PT', 'lf', 'lf', True)),
    'mixed_calc': lambda: (lambda c: (c, cf.smooth_state(c, 3)))(cf.mixed_box(3, 'calculated')),
    # drag on the no-slip wall 'top' (T zeroGradient there: the ghost T
    # follows the owner), Sutherland viscosity
    'mixed_drag': lambda: (lambda c: (c, cf.smooth_state(c, 4)))(cf.mixed_box(4, 'CBC_UPT', sutherland=True,
                                                                              objective=('drag', 'top'))),
    # a sheared (non-orthogonal) mesh: centre-to-centre directions differ
    # from the face normals, so the snGrad correction is active
    'mixed_skew': lambda: (lambda c: (c, cf.smooth_state(c, 6)))(cf.mixed_box(6, 'CBC_UPT', sutherland=True,
                                                                              objective=('kinetic_energy',),
                                                                              shear=SHEAR)),
}


@pytest.mark.parametrize('name', sorted(CASES))
def test_gradients_residual_fp64(name):
    case, u = CASES[name]()
    oc, grad_o, R_o, dtc_o, prim_o = oracle_fields(case, u)
    ctx = case.context(np.float64)
    ud = torch.as_tensor(u, device='cuda').contiguous()
    g = ctx.gradients(ud).cpu().numpy()
    prim = ctx.primitive_ghosts(ud).cpu().numpy()
    dtc = torch.empty(ctx.nIC, dtype=torch.float64, device='cuda')
    R = ctx.residual(ud, dtc_cell=dtc).cpu().numpy()
    e_prim, e_g, e_R, e_dtc = rel_l2(prim, prim_o), rel_l2(g, grad_o), rel_l2(R, R_o), rel_l2(dtc.cpu().numpy(), dtc_o)
    print('%s fp64 prim %.2e grad %.2e R %.2e dtc %.2e' % (name, e_prim, e_g, e_R, e_dtc))
    assert e_prim < 1e-13
    assert e_g < 1e-10
    for i in range(5):
        assert rel_l2(R[i], R_o[i]) < 1e-9 or np.abs(R[i] - R_o[i]).max() < 1e-9 * np.abs(R_o).max(), i
    assert e_dtc < 1e-12


@pytest.mark.parametrize('name', sorted(CASES))
def test_step_fp64_and_fp32(name):
    case, u = CASES[name]()
    oc = case.oracle()
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    gi = [t(a) for a in cf.ghost_init(case)]
    dt = 1e-6 if name.startswith('mixed') else 2e-5
    r, rU, rE, dtc_o, obj_o = oc.step(t(u[0:1].T), t(u[1:4].T), t(u[4:5].T), dt, ghostInit=gi)
    u_o = np.concatenate([r.T.numpy(), rU.T.numpy(), rE.T.numpy()])
    for prec, tol in ((np.float64, 1e-12), (np.float32, 2e-6)):
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        ctx.primal_step(ud, dt, dtc_out=dtc, obj_out=obj)
        un = ud.cpu().numpy().astype(np.float64)
        # per conserved variable: rho, rhoU (the 3-vector), rhoE
        errs = [rel_l2(un[s], u_o[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        e_dtc = abs(dtc.item() - dtc_o.item()) / dtc_o.item()
        print('%s %s step rel-L2 rho/rhoU/rhoE %s dtc %.2e obj %s vs %s' % (name, np.dtype(prec).name,
              ' '.join('%.2e' % e for e in errs), e_dtc, obj.item(), float(obj_o)))
        for k in range(3):
            assert errs[k] < tol, (k, errs[k])
        assert e_dtc < (1e-12 if prec == np.float64 else 1e-5)
        if case.objective:
            assert abs(obj.item() - float(obj_o)) <= (1e-12 if prec == np.float64 else 1e-5) * abs(float(obj_o))


def test_grad_cell_polynomial_like_reference():
    """tests/test_op.py:20-53: gradCell of U = xy + x^2 + y^2 + x on a
    uniform cyclic box, ghost values taken at the ghost centres; exact for a
    quadratic (max-relative 1e-9)."""
    from adfvm_amd import meshgen
    case = cf.CaseDef(meshgen.periodic_box((20, 20, 4), (1., 1., 0.1)), {}, 0.)
    m = case.mesh
    X, Y = m.cellCentres[:, 0], m.cellCentres[:, 1]
    Uc = X * Y + X ** 2 + Y ** 2 + X
    ref = np.stack([Y + 2 * X + 1, X + 2 * Y, 0 * X])[:, :m.nInternalCells]
    ctx = case.context(np.float64)
    g = ctx.grad_cell(torch.as_tensor(Uc[None, :], device='cuda').contiguous()).cpu().numpy()
    assert np.abs(g - ref).max() / np.abs(ref).max() < 1e-9


def test_interp_second_order_like_reference():
    """tests/test_interp.py:19-48: secondOrder of 2x + x^3 y^2 with Gauss
    gradients; max-relative error 1e-5 against the exact face values."""
    from adfvm_amd import meshgen
    case = cf.CaseDef(meshgen.periodic_box((100, 100, 10), (1., 1., 0.1)), {}, 0.)   # the convection case's grid
    m = case.mesh
    X, Y = m.cellCentres[:, 0], m.cellCentres[:, 1]
    Uc = 2 * X + X ** 3 * Y ** 2
    ctx = case.context(np.float64)
    phi = torch.as_tensor(Uc[None, :], device='cuda').contiguous()
    g = ctx.grad_cell(phi).cpu().numpy()                     # internal rows
    # ghost gradient rows: the default (cyclic) boundary copies the paired owner
    oc = case.oracle()
    gfull = g[:, oc.grad_src.numpy()]
    out = ctx.interp_second_order(phi, torch.as_tensor(np.ascontiguousarray(gfull), device='cuda'), 0).cpu().numpy()
    Xf, Yf = m.faceCentres[:, 0], m.faceCentres[:, 1]
    Ur = 2 * Xf + Xf ** 3 * Yf ** 2
    assert np.abs(out[0] - Ur).max() / np.abs(Ur).max() < 1e-5


def test_pack_unpack_roundtrip():
    case, params = cf.tgv(4)
    u = cf.tgv_state(case, params)
    ctx = case.context(np.float32)
    rho = torch.as_tensor(u[0:1].T.astype(np.float32).copy(), device='cuda')
    rhoU = torch.as_tensor(u[1:4].T.astype(np.float32).copy(), device='cuda')
    rhoE = torch.as_tensor(u[4:5].T.astype(np.float32).copy(), device='cuda')
    s = ctx.pack(rho, rhoU, rhoE)
    assert np.array_equal(s.cpu().numpy(), u.astype(np.float32))
    r2, ru2, re2 = ctx.unpack(s)
    assert torch.equal(r2, rho) and torch.equal(ru2, rhoU) and torch.equal(re2, rhoE)


def test_dot_and_energy_fp64_accumulation():
    case, params = cf.tgv(5)
    u = cf.tgv_state(case, params)
    ctx = case.context(np.float32)
    a = torch.as_tensor(u.astype(np.float32), device='cuda').contiguous()
    b = torch.ones_like(a) * 1e-3
    d = ctx.dot(a, b).item()
    ref = float((u.astype(np.float32).astype(np.float64) * np.float64(np.float32(1e-3))).sum())
    assert abs(d - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize('name', ['mixed_drag', 'mixed_skew', 'tgv6'])
def test_adjoint_energy_against_oracle(name):
    """adfvm_adjoint_energy (computeEnergy, cmesh.cpp:342-425) against the
    oracle's restatement on a random adjoint field: FP64 relative error
    <= 1e-12; the FP32 kernel (FP64 arithmetic on FP32 inputs and FP32 cell
    volumes) against the oracle on the same FP32-rounded state and adjoint
    <= 1e-6 (the volumes' rounding)."""
    case, u = CASES[name]()
    oc = case.oracle()
    rng = np.random.default_rng(5)
    ua = rng.normal(size=u.shape) * case.mesh.volumes.reshape(1, -1)
    for prec in (np.float64, np.float32):
        uu, aa = u.astype(prec).astype(np.float64), ua.astype(prec).astype(np.float64)
        t = lambda a: torch.as_tensor(a, dtype=torch.float64)
        want = float(oc.adjoint_energy(t(aa[0]), t(aa[1:4].T), t(aa[4]), t(uu[0:1].T), t(uu[1:4].T), t(uu[4:5].T)))
        ctx = case.context(prec)
        got = ctx.adjoint_energy(torch.as_tensor(ua.astype(prec), device='cuda').contiguous(),
                                 torch.as_tensor(u.astype(prec), device='cuda').contiguous()).item()
        print('%s %s energy %.15e vs %.15e' % (name, np.dtype(prec).name, got, want))
        assert abs(got - want) <= (1e-12 if prec == np.float64 else 1e-6) * abs(want)


def test_sod_1000_cells_200_steps_against_oracle():
    """The CPU-runnable shock tube: 1000 cells on [0, 1], left p = 1e5 Pa,
    T = 348.4 K, right p = 1e4 Pa, T = 278.7 K, inviscid Roe, second-order
    reconstruction (limited: unlimited, the first step's face pressure goes
    negative), SSP-RK3, 200 steps at CFL 0.5 under the reference's rule
    dt = min(2 CFL/dtc, 1.2 dt) (solver.py:359-366).  The HIP steps set the
    dt sequence; the oracle replays it.  FP64 relative L2 per conserved
    variable <= 1e-12; FP32 against the same FP64 oracle <= 1e-4 (limiter
    branches at near-ties can differ between precisions)."""
    case = cf.sod(1000)
    u0 = cf.sod_state(case)
    nsteps, cfl = 200, 0.5
    results = {}
    for prec in (np.float64, np.float32):
        ctx = case.context(prec)
        ud = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
        dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
        dt, dts = 1e-8, []
        for _ in range(nsteps):
            ctx.primal_step(ud, dt, dtc_out=dtc)
            dts.append(dt)
            dt = min(2 * cfl / dtc.item(), dt * 1.2)
        results[prec] = (ud.cpu().numpy().astype(np.float64), dts)
        assert np.isfinite(results[prec][0]).all()
    oc = case.oracle()
    x = torch.as_tensor(u0)
    for dt in results[np.float64][1]:
        r, rU, rE, _, _ = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt)
        x = torch.cat([r.T, rU.T, rE.T])
    u_o = x.numpy()
    print('sod 1000: t_end %.6e' % sum(results[np.float64][1]))
    for prec, tol in ((np.float64, 1e-12), (np.float32, 1e-4)):
        un = results[prec][0]
        errs = [rel_l2(un[s], u_o[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('sod 1000 %s 200 steps rel-L2 rho/rhoU/rhoE %s' % (np.dtype(prec).name,
              ' '.join('%.2e' % e for e in errs)))
        assert max(errs) < tol, errs


def test_tgv128_against_fp64_oracle_on_gpu():
    """The Taylor-Green case at its full size: 128^3 periodic box stored as
    an unstructured face list (seeded random permutation, then RCM), Mach
    0.1, Re 1600, 20 steps at CFL 0.5.  The FP64 oracle runs its torch
    restatement on the same GPU as the checker; the HIP kernels in FP64
    agree to round-off (<= 1e-12) and in FP32 within the stated 1e-5
    relative L2 per conserved variable."""
    case, params = cf.tgv(128, seed=0)
    u0 = cf.tgv_state(case, params)
    c0 = np.sqrt(cf.GAMMA * cf.R * params['T0'])
    dt = 0.5 * (2 * np.pi / 128) / (params['U0'] + c0)
    nsteps = 20
    oc = case.oracle().to('cuda')
    x = torch.as_tensor(u0, device='cuda')
    with torch.no_grad():
        for _ in range(nsteps):
            r, rU, rE, _, _ = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt)
            x = torch.cat([r.T, rU.T, rE.T])
    u_o = x.cpu().numpy()
    del oc, x
    torch.cuda.empty_cache()
    for prec, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        ctx = case.context(prec)
        u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
        for _ in range(nsteps):
            ctx.primal_step(u, dt)
        un = u.cpu().numpy().astype(np.float64)
        errs = [rel_l2(un[s], u_o[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('tgv128 %s %d steps vs oracle rel-L2 rho/rhoU/rhoE %s' % (np.dtype(prec).name, nsteps,
              ' '.join('%.2e' % e for e in errs)), flush=True)
        assert max(errs) < tol, errs


def test_tgv128_viscous_dissipation_closed_form():
    """A size-independent property at the full 128^3 size: the viscous part
    of the HIP residual's kinetic-energy rate (the residual with the case's
    viscosity minus the one without) equals -6 pi^3 mu U0^2 up to the
    second-order discretisation error, which the oracle shows at 3.8e-3 on
    24^3 and 2.1e-3 on 32^3, so about 1.3e-4 here; FP64 is held to 3e-4
    (measured 1.34e-4).  In FP32 the difference of two residuals also
    carries their rounding (measured 1.6e-4); held to 1e-3."""
    case, params = cf.tgv(128, seed=0)
    u0 = cf.tgv_state(case, params)
    V = case.mesh.volumes[:, 0]
    rho, U = u0[0], u0[1:4] / u0[0]
    inviscid = cf.CaseDef(case.md, {}, 0.0)
    inviscid._mesh = case.mesh
    exact = -6 * np.pi ** 3 * case.mu0 * params['U0'] ** 2
    for prec, tol in ((np.float64, 3e-4), (np.float32, 1e-3)):
        rates = []
        for c in (case, inviscid):
            ctx = c.context(prec)
            R = ctx.residual(torch.as_tensor(u0.astype(prec), device='cuda').contiguous())
            R = R.double().cpu().numpy()
            rates.append(float(((U * -R[1:4]).sum(0) + 0.5 * (U ** 2).sum(0) * R[0]) @ V))
            del ctx
        err = (rates[0] - rates[1]) / exact - 1
        print('TGV 128^3 %s viscous kinetic-energy rate: relative error %.2e' % (np.dtype(prec).name, err))
        assert abs(err) < tol


def test_periodic_conservation_fp32_and_fp64():
    """On a periodic box every face flux is added to one cell and taken from
    another, so sum(V*u) changes only by rounding.  TGV 16^3 over 300
    steps: the FP32 drift per conserved variable (relative to sum V|u|) stays
    at the level of unbiased rounding (measured in the 1e-9 range; held to
    3e-7), far below the 9e-6 a stage pair whose coefficients do not sum to
    one in FP32 (SSP-RK3's 1/3 and 2/3 both round up) would add over 300
    steps; FP64 is held to 1e-13."""
    case, params = cf.tgv(16, seed=5)
    u0 = cf.tgv_state(case, params)
    V = case.mesh.volumes[:, 0]
    c0 = np.sqrt(cf.GAMMA * cf.R * params['T0'])
    dt = 0.5 * (2 * np.pi / 16) / (params['U0'] + c0)
    for prec, tol in ((np.float32, 3e-7), (np.float64, 1e-13)):
        ctx = case.context(prec)
        u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
        a = u.double().cpu().numpy()
        for _ in range(300):
            ctx.primal_step(u, dt)
        b = u.double().cpu().numpy()
        drift = []
        for rows in (slice(0, 1), slice(1, 4), slice(4, 5)):
            d = np.abs(((b[rows] - a[rows]) * V).sum(1)).max() / (np.abs(a[rows]) * V).sum()
            drift.append(float(d))
        print('TGV 16^3 %s 300 steps: relative change of sum V*u rho/rhoU/rhoE %s'
              % (np.dtype(prec).name, ' '.join('%.2e' % d for d in drift)), flush=True)
        assert max(drift) <= tol, drift
