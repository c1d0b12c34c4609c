"""The objective evaluated on each step's start state u_n
(adfvm_set_objective_state with ADFVM_OBJSTATE_START) beside the reference's
hook on the first intermediate stage (density.py:412-413): the primal
objective and the adjoint step against the oracle's autograd of the same
definition, the multi-step gradient against central finite differences of
the HIP primal, the caller-cotangent entry on the same stage, and the
default left as the reference's."""
import numpy as np
import pytest

from tests import casefactory as cf
from tests.test_gpu_adjoint import adjoint_case, scaled_rbar
from tests.test_gpu_primal import rel_l2

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def oracle_start_vjp(case, u, dt, ua, obja):
    """(d(<ua, step(u)> + obja J(u))/du, J(u), dtc) by the oracle, with the
    objective on the start state."""
    oc = case.oracle()
    x = torch.tensor(u, dtype=torch.float64, requires_grad=True)
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    r, rU, rE, dtc, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt, ghostInit=gi, obj_on_start=True)
    un = torch.cat([r.T, rU.T, rE.T])
    L = (un * torch.as_tensor(ua)).sum() + obja * obj
    (g,) = torch.autograd.grad(L, x)
    return g.numpy(), float(obj.detach()), float(dtc.detach())


def test_default_is_the_reference_hook():
    case, params = cf.tgv(4, seed=1)
    ctx = case.context(np.float64)
    assert ctx.objective_state == 'stage1'
    ctx.objective_state = 'start'
    assert ctx.objective_state == 'start'
    with pytest.raises(ValueError):
        ctx.objective_state = 'stage2'


@pytest.mark.parametrize('name', ['tgv6', 'mixed_drag', 'mixed_skew'])
def test_start_state_step_matches_oracle(name):
    """FP64 round-off level (FP32 at the adjoint step's FP32 tolerance of
    test_gpu_adjoint.py): obj_out is J(u_n), dtc stays on the first
    intermediate stage, the adjoint adds dJ/du at u_n."""
    case, u = adjoint_case(name)
    assert case.objective
    dt = 1e-6 if name.startswith('mixed') else 2e-5
    ua = scaled_rbar(case, u, 2) * 1e3
    for prec, tol, otol in ((np.float64, 1e-11, 1e-12), (np.float32, 2e-4, 1e-5)):
        r = lambda a: a.astype(prec).astype(np.float64)
        ref_g, ref_obj, ref_dtc = oracle_start_vjp(case, r(u), dt, r(ua), 1.0)
        ctx = case.context(prec)
        ctx.objective_state = 'start'
        u0 = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        ud = u0.clone()
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
        ctx.primal_step(ud, dt, dtc_out=dtc, obj_out=obj)
        # the same objective kernel on u_n, launched on its own
        obj_n = torch.zeros(1, dtype=torch.float64, device='cuda')
        ctx.objective(u0, out=obj_n)
        assert obj.item() == obj_n.item()
        eo = abs(obj.item() - ref_obj) / abs(ref_obj)
        ed = abs(dtc.item() - ref_dtc) / abs(ref_dtc)
        uad = torch.as_tensor(ua.astype(prec), device='cuda').contiguous()
        ctx.adjoint_step(u0, dt, uad, obja=1.0)
        got = uad.cpu().numpy().astype(np.float64)
        errs = [rel_l2(got[s], ref_g[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('%s %s start-state objective %.2e dtc %.2e adjoint rel-L2 %s' % (
            name, np.dtype(prec).name, eo, ed, ' '.join('%.2e' % e for e in errs)))
        assert eo < otol and ed < otol and max(errs) < tol


def test_start_state_multistep_gradient_vs_central_fd_fp64():
    """d/d eps of sum_{n<N} J(u_n) over 6 steps along a random
    initial-velocity perturbation: adjoint vs central FD of the HIP primal,
    FP64, relative difference <= 1e-6 (the base state as in
    test_gpu_adjoint.py's stage-1 test: velocity noise keeps the Roe
    entropy-fix width off its kink)."""
    case, params = cf.tgv(6, seed=2)
    u0 = cf.tgv_state(case, params)
    u0[1:4] += u0[0] * np.random.default_rng(11).normal(size=u0[1:4].shape) * 0.3
    ctx = case.context(np.float64)
    ctx.objective_state = 'start'
    nsteps, dt = 6, 4e-4
    rng = np.random.default_rng(3)
    dU = rng.normal(size=(3, ctx.nIC))
    v = np.zeros_like(u0)
    v[1:4] = u0[0] * dU
    v[4] = (u0[1:4] * dU).sum(0)

    def J(u):
        ud = torch.as_tensor(u, device='cuda').contiguous()
        tot = 0.0
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        for _ in range(nsteps):
            ctx.primal_step(ud, dt, obj_out=obj)
            tot += obj.item()
        return tot

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
    print('start-state multistep: adjoint %.12e fd %.12e rel %.2e' % (adj, fd, abs(adj - fd) / abs(fd)))
    assert abs(adj - fd) <= 1e-6 * abs(fd)


def test_start_state_cotangent_equals_builtin_kinetic_energy():
    """A caller's objective cotangent added on the start state (a context
    without a built-in objective) gives the built-in kinetic-energy adjoint
    of the same definition, FP64 to 1e-12."""
    case_ke, params = cf.tgv(6, seed=2)
    case_none, _ = cf.tgv(6, seed=2, objective=None)
    u = cf.tgv_state(case_ke, params)
    u[1:4] += u[0] * np.random.default_rng(4).normal(size=u[1:4].shape) * 0.3
    dt = 4e-4
    ua0 = scaled_rbar(case_ke, u, 5) * 1e3
    # dJ/du of J = sum_c 0.5 |rhoU|^2 / rho V_c at u_n, by torch in FP64
    V = torch.as_tensor(case_ke.mesh.volumes[:u.shape[1]].reshape(-1), dtype=torch.float64)
    x = torch.tensor(u, dtype=torch.float64, requires_grad=True)
    J = (0.5 * (x[1:4] ** 2).sum(0) / x[0] * V).sum()
    (dJ,) = torch.autograd.grad(J, x)
    outs = []
    for case, grad in ((case_ke, None), (case_none, dJ)):
        ctx = case.context(np.float64)
        ctx.objective_state = 'start'
        ud = torch.as_tensor(u, device='cuda').contiguous()
        uad = torch.as_tensor(ua0, device='cuda').contiguous()
        og = None if grad is None else grad.to('cuda').contiguous()
        ctx.adjoint_step(ud, dt, uad, obja=1.0, obj_grad=og)
        outs.append(uad.cpu().numpy())
    errs = [rel_l2(outs[1][s], outs[0][s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
    print('cotangent vs built-in start-state kinetic energy rel-L2 %s' % ' '.join('%.2e' % e for e in errs))
    assert max(errs) < 1e-12


def test_switching_back_restores_the_reference_hook_bitwise():
    case, u = adjoint_case('tgv6')
    dt = 2e-5
    for prec in (np.float32, np.float64):
        ref = case.context(prec)
        ctx = case.context(prec)
        ctx.objective_state = 'start'
        ctx.objective_state = 'stage1'
        res = []
        for c in (ref, ctx):
            ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
            obj = torch.zeros(1, dtype=torch.float64, device='cuda')
            u1 = ud.clone()
            c.primal_step(u1, dt, obj_out=obj)
            ua = torch.as_tensor(scaled_rbar(case, u, 6).astype(prec), device='cuda').contiguous()
            c.adjoint_step(ud, dt, ua, obja=1.0)
            res.append((obj.item(), ua.cpu().numpy()))
        assert res[0][0] == res[1][0] and np.array_equal(res[0][1], res[1][1])


def test_graph_sweep_follows_the_objective_state():
    """A GraphSweep captured with the stage-1 objective, replayed after the
    context switched to the start-state objective, recaptures: its result
    equals an eager sweep in the new state bitwise."""
    from adfvm_amd.graphs import GraphSweep
    case, params = cf.tgv(8, seed=5)
    u = cf.tgv_state(case, params)
    u[1:4] += u[0] * np.random.default_rng(8).normal(size=u[1:4].shape) * 0.3
    ctx = case.context(np.float64)
    u0 = torch.as_tensor(u, device='cuda').contiguous()
    dt, interval, n = 2e-4, 2, 4
    sweep = GraphSweep(ctx, dt, interval, obja=1.0)
    J1, _ = sweep.run(u0, n)
    ctx.objective_state = 'start'
    J2, ua2 = sweep.run(u0, n)
    ua2 = ua2.clone()
    eager = GraphSweep(ctx, dt, interval, obja=1.0, graphs=False)
    J3, ua3 = eager.run(u0, n)
    print('graph sweep objective: stage1 %.15e start %.15e eager start %.15e' % (J1[0], J2[0], J3[0]))
    assert J1[0] != J2[0]
    assert J2[0] == J3[0] and torch.equal(ua2, ua3)
