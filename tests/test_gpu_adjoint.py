"""HIP adjoint path against the oracle's autograd (the exact discrete
adjoint of the oracle's step) and against central finite differences of the
HIP primal itself.  FP64 tolerances are round-off level; FP32 tolerances are
stated per test."""
import numpy as np
import pytest

from tests import casefactory as cf
from tests.test_gpu_primal import CASES, rel_l2

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def oracle_residual_vjp(case, u, rbar):
    oc = case.oracle()
    x = torch.tensor(u, dtype=torch.float64, requires_grad=True)
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    d0, d1, d2, _ = oc.equation(x[0:1].T, x[1:4].T, x[4:5].T, ghostInit=gi)
    R = torch.cat([d0.T, d1.T, d2.T])
    (g,) = torch.autograd.grad(R, x, torch.as_tensor(rbar))
    return g.numpy()


def oracle_step_vjp(case, u, dt, ua, obja, source=None):
    oc = case.oracle()
    x = torch.tensor(u, dtype=torch.float64, requires_grad=True)
    S = None
    if source is not None:
        Sx = torch.tensor(source, dtype=torch.float64, requires_grad=True)
        S = [Sx[0:1].T, Sx[1:4].T, Sx[4:5].T]
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    r, rU, rE, dtc, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt, source=S, ghostInit=gi)
    un = torch.cat([r.T, rU.T, rE.T])
    L = (un * torch.as_tensor(ua)).sum() + obja * obj
    inputs = [x] + ([Sx] if source is not None else [])
    grads = torch.autograd.grad(L, inputs)
    return [g.numpy() for g in grads]


def scaled_rbar(case, u, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True) * 1e-3


def adjoint_case(name):
    """The primal cases, with the shock tube's exactly uniform (U = 0) halves
    made generic by 1% relative noise and a random O(1 m/s) velocity: at
    (near-)ties the limiter's clamp branch is decided by rounding in near-zero
    gradients, and a uniform velocity puts the Roe entropy fix's
    |Un_L - Un_R| on its kink at 0 -- subgradient choices no two evaluation
    orders (or precisions) share.  1% keeps the clamp margins far above FP32
    rounding while both branches still occur."""
    case, u = CASES[name]()
    if name == 'sod':
        rng = np.random.default_rng(5)
        u = u * (1 + 1e-2 * rng.normal(size=u.shape))
        u[1] += u[0] * rng.normal(size=u.shape[1])
    return case, u


# FP32 tolerances: the oracle itself evaluated by torch in float32 differs
# from its float64 autograd by 7.5e-4 (rho row) on the tgv6 residual VJP; the
# HIP FP32 kernels are held to that level.
@pytest.mark.parametrize('name', sorted(CASES))
def test_residual_vjp_matches_autograd(name):
    case, u = adjoint_case(name)
    rbar = scaled_rbar(case, u, 1)
    for prec, tol in ((np.float64, 1e-10), (np.float32, 1e-3)):
        # reference at this precision's rounded inputs (the limiter's branch
        # choice depends on the input bits)
        r = lambda a: a.astype(prec).astype(np.float64)
        ref = oracle_residual_vjp(case, r(u), r(rbar))
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        rb = torch.as_tensor(rbar.astype(prec), device='cuda').contiguous()
        got = ctx.residual_vjp(ud, rb).cpu().numpy().astype(np.float64)
        errs = [rel_l2(got[s], ref[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('%s %s residual VJP rel-L2 rho/rhoU/rhoE %s' % (name, np.dtype(prec).name, ' '.join('%.2e' % e for e in errs)))
        assert max(errs) < tol


@pytest.mark.parametrize('name', sorted(CASES))
def test_adjoint_step_matches_autograd(name):
    case, u = adjoint_case(name)
    rng = np.random.default_rng(7)
    ua = scaled_rbar(case, u, 2) * 1e3
    src = rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True) * 10
    dt = 1e-6 if name.startswith('mixed') else 2e-5
    obja = 1.0 if case.objective else 0.0
    for prec, tol in ((np.float64, 1e-11), (np.float32, 2e-4)):
        r = lambda a: a.astype(prec).astype(np.float64)
        ref_u, ref_s = oracle_step_vjp(case, r(u), dt, r(ua), obja, source=r(src))
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        uad = torch.as_tensor(ua.astype(prec), device='cuda').contiguous()
        sd = torch.as_tensor(src.astype(prec), device='cuda').contiguous()
        sg = torch.zeros_like(ud)
        ctx.adjoint_step(ud, dt, uad, obja=obja, source=sd, source_grad=sg)
        got = uad.cpu().numpy().astype(np.float64)
        errs = [rel_l2(got[s], ref_u[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        es = rel_l2(sg.cpu().numpy(), ref_s)
        print('%s %s adjoint step rel-L2 rho/rhoU/rhoE %s source %.2e' % (name, np.dtype(prec).name,
              ' '.join('%.2e' % e for e in errs), es))
        assert max(errs) < tol and es < tol
        # the primal state is untouched
        assert np.array_equal(ud.cpu().numpy(), u.astype(prec))


def test_multistep_gradient_vs_central_fd_fp64():
    """d/d eps of sum_n J(u_n^(1)) over 6 steps along a random initial-velocity
    perturbation: adjoint (6 reverse steps) vs central FD of the HIP primal,
    FP64, relative difference <= 1e-6.

    The base state is the Taylor-Green field plus seeded velocity noise: the
    exact TGV field has w = 0, which puts every z-face on the kink of
    |UnL - UnR| in the Roe entropy-fix width (riemann.py:50), where the
    objective has different one-sided derivatives and no FD can agree with an
    adjoint."""
    case, params = cf.tgv(6, seed=2)
    u0 = cf.tgv_state(case, params)
    u0[1:4] += u0[0] * np.random.default_rng(11).normal(size=u0[1:4].shape) * 0.3
    ctx = case.context(np.float64)
    nsteps, dt = 6, 4e-4
    rng = np.random.default_rng(3)
    dU = rng.normal(size=(3, ctx.nIC))
    rho0 = u0[0]
    v = np.zeros_like(u0)
    v[1:4] = rho0 * dU                             # perturb rhoU by rho*dU
    v[4] = (u0[1:4] * dU).sum(0)                   # first-order change of rhoE at fixed rho, p

    def J(u):
        ud = torch.as_tensor(u, device='cuda').contiguous()
        tot = 0.0
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        for _ in range(nsteps):
            ctx.primal_step(ud, dt, obj_out=obj)
            tot += obj.item()
        return tot

    # forward storing states, then reverse
    ud = torch.as_tensor(u0, device='cuda').contiguous()
    states = []
    for _ in range(nsteps):
        states.append(ud.clone())
        ctx.primal_step(ud, dt)
    ua = torch.zeros_like(ud)
    for k in reversed(range(nsteps)):
        ctx.adjoint_step(states[k], dt, ua, obja=1.0)
    adj = float((ua.cpu().numpy() * v).sum())
    eps = 1e-5
    fd = (J(u0 + eps * v) - J(u0 - eps * v)) / (2 * eps)
    print('multistep: adjoint %.12e fd %.12e rel %.2e' % (adj, fd, abs(adj - fd) / abs(fd)))
    assert abs(adj - fd) <= 1e-6 * abs(fd)


def test_adjoint_deterministic_bitwise():
    case, params = cf.tgv(6, seed=4)
    u = cf.tgv_state(case, params)
    ctx = case.context(np.float32)
    ud = torch.as_tensor(u.astype(np.float32), device='cuda').contiguous()
    outs = []
    for _ in range(3):
        ua = torch.ones_like(ud) * 1e-3
        ctx.adjoint_step(ud, 2e-5, ua, obja=1.0)
        outs.append(ua.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


@pytest.mark.parametrize('name', ['tgv6', 'mixed_upt'])
def test_tape_equals_recompute_bitwise(name):
    """An adjoint step fed the forward step's tape gives the same bits as one
    that recomputes its stages."""
    case, u = adjoint_case(name)
    for prec in (np.float32, np.float64):
        ctx = case.context(prec)
        dt = 1e-6
        u0 = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        tape = ctx.new_tapes(1)[0]
        u1 = u0.clone()
        ctx.primal_step(u1, dt, tape=tape)
        u2 = u0.clone()
        ctx.primal_step(u2, dt)
        assert torch.equal(u1, u2)
        ua0 = torch.as_tensor(scaled_rbar(case, u, 3).astype(prec), device='cuda').contiguous()
        a, b = ua0.clone(), ua0.clone()
        ctx.adjoint_step(u0, dt, a, obja=1.0, tape=tape)
        ctx.adjoint_step(u0, dt, b, obja=1.0)
        assert torch.equal(a, b)


BC_CASES = cf.bc_param_cases()


def oracle_bc_vjp(case, u, dt, ua, obja, params, rnd=lambda a: a):
    """d(<ua, step(u)> + obja J)/d(value) for each boundary-value parameter,
    by the oracle's autograd; rnd rounds the values as the device holds them."""
    bcs, ts = cf.bc_value_tensors(case, params)
    with torch.no_grad():
        for t in ts.values():
            t.copy_(torch.as_tensor(rnd(t.numpy())))
    oc = case.oracle(bcs=bcs)
    x = torch.tensor(u, dtype=torch.float64)
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    r, rU, rE, dtc, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt, ghostInit=gi)
    L = (torch.cat([r.T, rU.T, rE.T]) * torch.as_tensor(ua)).sum() + obja * obj
    grads = torch.autograd.grad(L, [ts[p[:3]] for p in params])
    return {p[:3]: g.numpy() for p, g in zip(params, grads)}


@pytest.mark.parametrize('name', sorted(BC_CASES))
def test_bc_value_gradient_matches_autograd(name):
    """The boundary-value gradients an adjoint step adds to bc_grad
    (parameter ('BCs', field, patch, key)) against the oracle's autograd:
    fixedValue U and T, CBC_UPT U0/T0/p0, CBC_TOTAL_PT Tt/pt/direction, with
    the drag objective on a fixedValue wall."""
    case, u, params = BC_CASES[name]()
    ua = scaled_rbar(case, u, 4) * 1e3
    dt = 1e-4
    obja = 1.0 if case.objective else 0.0
    # FP32: the rounding of the FP32 step and its reverse (measured up to
    # 2.3e-4, the inlet temperature value of totalpt_normal)
    for prec, tol in ((np.float64, 1e-11), (np.float32, 5e-4)):
        r = lambda a: a.astype(prec).astype(np.float64)
        ref = oracle_bc_vjp(case, r(u), dt, r(ua), obja, params, rnd=r)
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        uad = torch.as_tensor(ua.astype(prec), device='cuda').contiguous()
        bg = ctx.bc_state()
        ctx.adjoint_step(ud, dt, uad, obja=obja, bc_grad=bg)
        bgn = bg.cpu().numpy().astype(np.float64)
        for field, patch, key, k in params:
            rows, faces = ctx.bc_param(field, patch, key)
            got = bgn[rows, faces].T
            err = rel_l2(got, ref[(field, patch, key)])
            print('%s %s %s/%s/%s rel-L2 %.2e' % (name, np.dtype(prec).name, field, patch, key, err))
            assert err < tol, (field, patch, key, err)
        # faces of rules without values stay zero
        for patch, info in ctx._bcinfo.items():
            if info['cbc'] == 0 and all(t != 'fixedValue' for t in info['types'].values()):
                assert not bgn[:, info['start']:info['start'] + info['n']].any(), patch


def test_bc_value_gradient_vs_central_fd_multistep():
    """d/d eps of sum_n J_n over 5 steps (drag on the no-slip wall) along
    random perturbations of the CBC_UPT outlet's U0 and the inlet's fixed
    U: the adjoint sweep's bc_grad against central differences of the HIP
    primal with the values replaced on the device (set_bc_value), FP64."""
    from adfvm_amd import adjoint
    case, u, params = BC_CASES['upt_drag']()
    ctx = case.context(np.float64)
    nsteps, dt = 5, 1e-4
    u0 = torch.as_tensor(u, device='cuda').contiguous()
    bg = ctx.bc_state()
    adjoint.sweep(ctx, u0, [dt] * nsteps, schedule='store_all', bc_grad=bg)
    rng = np.random.default_rng(21)
    for field, patch, key, k in (('p', 'outlet', 'U0', 3), ('U', 'inlet', 'value', 3), ('p', 'outlet', 'p0', 1)):
        n = case.mesh.boundary[patch]['nFaces']
        base = np.broadcast_to(np.asarray(case.bcs[patch][field][key], np.float64).reshape(-1, k), (n, k)).copy()
        dv = rng.normal(size=(n, k)) * (np.abs(base).mean() + 1.0)
        rows, faces = ctx.bc_param(field, patch, key)
        adj = float((bg[rows, faces].T.cpu().numpy() * dv).sum())

        def J(s, eps):
            ctx.set_bc_value(field, patch, key, base + s * eps * dv)
            ud = u0.clone()
            obj = torch.zeros(1, dtype=torch.float64, device='cuda')
            tot = 0.0
            for _ in range(nsteps):
                ctx.primal_step(ud, dt, obj_out=obj)
                tot += obj.item()
            return tot
        eps = 1e-4
        fd = (8 * (J(1, eps) - J(-1, eps)) - (J(2, eps) - J(-2, eps))) / (12 * eps)
        ctx.set_bc_value(field, patch, key, base)
        rel = abs(adj - fd) / abs(fd)
        print('bc fd %s/%s/%s: adjoint %.12e fd %.12e rel %.2e' % (field, patch, key, adj, fd, rel))
        assert rel <= 1e-6, (field, patch, key, rel)


@pytest.mark.parametrize('name', ['mixed_upt', 'mixed_totalpt_lf', 'mixed_calc', 'mixed_drag'])
def test_adjoint_taylor_remainder_second_order(name):
    """Taylor remainder of one FP64 step's functional
    J(u) = <w, step(u)> + obja * objective(u):
    r(eps) = |J(u + eps v) - J(u) - eps <grad J, v>| with grad J from the
    adjoint step must fall as eps^2 (ratio 4 per halving); a gradient that
    is wrong in any direction leaves an O(eps) remainder (ratio 2).  The
    perturbation scale keeps r far above FP64 round-off and below the
    distance to the entropy fix's and limiter's branch points on these
    generic states (every ghost rule, Roe and LF, the drag objective).  The
    Taylor-Green adjoint is left to the central-difference tests: its face
    jumps |UnL - UnR| are so small that an eps on one side of every kink of
    the entropy fix leaves remainders at FP64 round-off."""
    case, u = adjoint_case(name)
    ctx = case.context(np.float64)
    rng = np.random.default_rng(21)
    scale = np.abs(u).mean(1, keepdims=True)
    v = rng.normal(size=u.shape) * scale
    w = rng.normal(size=u.shape) * scale * 1e-3
    dt = 1e-6 if name.startswith('mixed') else 2e-5
    obja = 1.0 if case.objective else 0.0

    def J(x):
        xd = torch.as_tensor(x, device='cuda').contiguous()
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        ctx.primal_step(xd, dt, obj_out=obj)
        return float((xd.cpu().numpy() * w).sum()) + obja * obj.item()

    ud = torch.as_tensor(u, device='cuda').contiguous()
    ua = torch.as_tensor(w, device='cuda').contiguous()
    ctx.adjoint_step(ud, dt, ua, obja=obja)
    dJ = float((ua.cpu().numpy() * v).sum())
    J0 = J(u)
    eps = [1e-4 * 2.0 ** -k for k in range(4)]
    r = [abs(J(u + e * v) - J0 - e * dJ) for e in eps]
    ratios = [r[k] / r[k + 1] for k in range(3)]
    print('%s Taylor remainders %s, ratios %s' % (name, ' '.join('%.3e' % x for x in r),
                                                  ' '.join('%.2f' % x for x in ratios)))
    assert all(3.5 < q < 4.5 for q in ratios)


def _degenerate_cases():
    """Edge cases of the index tables: periodic boxes of 1 cell (all six
    faces cyclic, the cell its own partner: six extra record planes) and 2
    cells per direction (each neighbour reached once through an internal
    face and once through a cyclic one); a mixed box with an empty (zero-face)
    patch listed first."""
    from collections import OrderedDict
    from adfvm_amd import meshgen
    out = {}
    for n in (1, 2):
        md = meshgen.periodic_box(n)
        case = cf.CaseDef(md, {}, 1e-3, objective=('kinetic_energy',))
        rng = np.random.default_rng(n)
        nIC = n ** 3
        U = rng.normal(size=(nIC, 3)) * 30
        T = 300 + rng.random((nIC, 1))
        p = 1e5 + 100 * rng.random((nIC, 1))
        out['periodic%d' % n] = (case, cf.state_from_prims(U, T, p))
    base = cf.mixed_box(1, 'CBC_UPT')
    md = base.md
    nF = len(md.owner)
    b = OrderedDict([('aaa_empty', dict(type='patch', startFace=nF, nFaces=0))])
    b.update(md.boundary)
    md2 = meshgen.MeshData(md.points, md.faces, md.owner, md.neighbour, b)
    bcs = dict(base.bcs)
    bcs['aaa_empty'] = {'U': {'type': 'fixedValue', 'value': np.zeros((0, 3))},
                        'T': {'type': 'zeroGradient'}, 'p': {'type': 'zeroGradient'}}
    case = cf.CaseDef(md2, bcs, base.mu0)
    out['mixed_empty_patch'] = (case, cf.smooth_state(case, 1))
    # internal faces in a random order, owners no longer non-decreasing: the
    # face adjoint cannot sum a cell's owner sides in one block and writes a
    # record per face side (the path of meshes not in owner order)
    for name, c0, u0 in (('unsorted_mixed', base, cf.smooth_state(base, 2)),
                         ('unsorted_periodic',) + out['periodic2']):
        m0 = c0.md
        nIF = len(m0.neighbour)
        order = np.random.default_rng(4).permutation(nIF)
        idx = np.concatenate([order, np.arange(nIF, len(m0.owner))])
        md3 = meshgen.MeshData(m0.points, m0.faces[idx], m0.owner[idx], m0.neighbour[order], m0.boundary,
                               m0.nCells)
        assert np.any(np.diff(md3.owner[:nIF]) < 0)
        c3 = cf.CaseDef(md3, c0.bcs, c0.mu0, objective=c0.objective)
        out[name] = (c3, u0)
    return out


@pytest.mark.parametrize('name', ['periodic1', 'periodic2', 'mixed_empty_patch', 'unsorted_mixed',
                                  'unsorted_periodic'])
def test_degenerate_meshes_step_and_adjoint(name):
    from tests.test_gpu_primal import rel_l2 as rl2
    case, u = _degenerate_cases()[name]
    dt = 1e-6
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    oc = case.oracle()
    gi = [t(a) for a in cf.ghost_init(case)]
    r, rU, rE, _, _ = oc.step(t(u[0:1].T), t(u[1:4].T), t(u[4:5].T), dt, ghostInit=gi)
    u_o = np.concatenate([r.T.numpy(), rU.T.numpy(), rE.T.numpy()])
    ua = scaled_rbar(case, u, 8) * 1e3
    obja = 1.0 if case.objective else 0.0
    (ref_u,) = oracle_step_vjp(case, u, dt, ua, obja)
    ctx = case.context(np.float64)
    ud = torch.as_tensor(u, device='cuda').contiguous()
    u1 = ud.clone()
    ctx.primal_step(u1, dt)
    e_step = rl2(u1.cpu().numpy(), u_o)
    uad = torch.as_tensor(ua, device='cuda').contiguous()
    ctx.adjoint_step(ud, dt, uad, obja=obja)
    e_adj = rl2(uad.cpu().numpy(), ref_u)
    print('%s: step rel-L2 %.2e, adjoint step rel-L2 %.2e' % (name, e_step, e_adj))
    assert e_step < 1e-12 and e_adj < 1e-11


def test_tgv128_adjoint_step_against_oracle_on_gpu():
    """One adjoint step at the full Taylor-Green size (128^3, random
    permutation + RCM, base state with seeded velocity noise off the
    |w| = 0 kink) against the oracle's autograd run by torch on the GPU:
    FP64 relative L2 <= 1e-11, FP32 <= 2e-4 per conserved variable."""
    case, params = cf.tgv(128, seed=1)
    u = cf.tgv_state(case, params)
    rng = np.random.default_rng(13)
    u[1:4] += u[0] * rng.normal(size=u[1:4].shape) * 0.3
    ua = rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True)
    c0 = np.sqrt(cf.GAMMA * cf.R * params['T0'])
    dt = 0.5 * (2 * np.pi / 128) / (params['U0'] + c0)
    oc = case.oracle().to('cuda')
    x = torch.tensor(u, dtype=torch.float64, device='cuda', requires_grad=True)
    r, rU, rE, _, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt)
    L = (torch.cat([r.T, rU.T, rE.T]) * torch.as_tensor(ua, device='cuda')).sum() + obj
    (g,) = torch.autograd.grad(L, x)
    ref = g.cpu().numpy()
    del oc, x, r, rU, rE, obj, L, g
    torch.cuda.empty_cache()
    for prec, tol in ((np.float64, 1e-11), (np.float32, 2e-4)):
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        uad = torch.as_tensor(ua.astype(prec), device='cuda').contiguous()
        ctx.adjoint_step(ud, dt, uad, obja=1.0)
        got = uad.cpu().numpy().astype(np.float64)
        errs = [rel_l2(got[s], ref[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('tgv128 %s adjoint step vs oracle autograd rel-L2 %s' % (np.dtype(prec).name,
              ' '.join('%.2e' % e for e in errs)), flush=True)
        assert max(errs) < tol, errs
