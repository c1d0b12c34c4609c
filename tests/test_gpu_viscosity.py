"""HIP adjoint-viscosity stabilisation against the oracle
(oracle/viscosity_oracle.py): the scaled viscosity per cell and the smoothed
adjoint state, for the uniform, abarbanel and turkel symmetrisers on
periodic and mixed-boundary cases.  FP64 relative L2 <= 1e-10 (the largest
eigenvalue by in-register Jacobi rotations against LAPACK); FP32 <= 1e-4."""
import numpy as np
import pytest

from tests import casefactory as cf
from tests.test_gpu_primal import CASES, rel_l2
from tests.test_gpu_adjoint import scaled_rbar

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

VTYPES = {'uniform': 0, 'abarbanel': 1, 'turkel': 2}


@pytest.mark.parametrize('name', ['tgv6', 'mixed_upt', 'mixed_totalpt_lf', 'mixed_drag'])
@pytest.mark.parametrize('vtype', sorted(VTYPES))
def test_adjoint_viscosity_matches_oracle(name, vtype):
    from oracle import viscosity_oracle as vo
    case, u = CASES[name]()
    ua = scaled_rbar(case, u, 5) * 1e3
    # viscosity scaled so that dt*DT/h^2 ~ 0.1 on the smallest cells: the
    # implicit step then changes the state well above rounding
    h = case.mesh.volumes.min() ** (1. / 3)
    dt, iters = 2e-5, 25
    scaling = 0.1 * h * h / dt
    oc = case.oracle()
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    ref_ua, ref_M = vo.adjoint_viscosity(oc, vtype, u, ua, dt, scaling, iters, ghostInit=gi)
    for prec, tol in ((np.float64, 1e-10), (np.float32, 1e-4)):
        r = lambda a: a.astype(prec).astype(np.float64)
        if prec == np.float32:
            ref_ua, ref_M = vo.adjoint_viscosity(oc, vtype, r(u), r(ua), dt, scaling, iters, ghostInit=gi)
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        uad = torch.as_tensor(ua.astype(prec), device='cuda').contiguous()
        m2 = torch.zeros(ctx.nIC, dtype=ctx.tdtype, device='cuda')
        ctx.adjoint_viscosity(ud, uad, dt, vtype, scaling, iterations=iters, m2norm_out=m2)
        eM = rel_l2(m2.cpu().numpy(), ref_M)
        got = uad.cpu().numpy().astype(np.float64)
        errs = [rel_l2(got[s], ref_ua[s]) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('%s %s %s viscosity rel-L2 %.2e, smoothed adjoint rho/rhoU/rhoE %s' % (
            name, vtype, np.dtype(prec).name, eM, ' '.join('%.2e' % e for e in errs)))
        assert eM < tol and max(errs) < tol, (eM, errs)
        # the smoothing changed the state (not an identity)
        assert rel_l2(got, ua) > 1e-3


def test_adjoint_viscosity_bitwise_repeatable_and_no_iterations():
    case, u = CASES['mixed_upt']()
    ctx = case.context(np.float32)
    ud = torch.as_tensor(u.astype(np.float32), device='cuda').contiguous()
    ua = torch.as_tensor((scaled_rbar(case, u, 6) * 1e3).astype(np.float32), device='cuda').contiguous()
    outs = []
    for _ in range(2):
        a = ua.clone()
        ctx.adjoint_viscosity(ud, a, 1e-5, 'abarbanel', 1e-3, iterations=10)
        outs.append(a.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    a = ua.clone()
    ctx.adjoint_viscosity(ud, a, 1e-5, 'turkel', 1e-3, iterations=0)
    assert np.allclose(a.cpu().numpy(), ua.cpu().numpy(), rtol=1e-6)


@pytest.mark.parametrize('kind', ['p', 'c'])
@pytest.mark.parametrize('vtype', ['abarbanel', 'turkel'])
def test_adjoint_viscosity_pressure_and_sound_speed_states(kind, vtype):
    """The states on which the oracle's gradp and gradc couplings are pinned
    entry by entry (only gradp, or only gradc, non-zero; U = 0, fixedValue
    ghosts): HIP's scaled viscosity per cell against the oracle's."""
    from oracle import viscosity_oracle as vo
    from tests.test_oracle import _visc_state_box
    _, oc, u, _, _, _, _, (md, bcs) = _visc_state_box(kind)
    case = cf.CaseDef(md, bcs, 1e-5)
    scaling = 2.5
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    ref = vo.scaled_viscosity(oc, vtype, u, scaling, ghostInit=gi)
    for prec, tol in ((np.float64, 1e-10), (np.float32, 1e-4)):
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        uad = torch.zeros_like(ud)
        m2 = torch.zeros(ctx.nIC, dtype=ctx.tdtype, device='cuda')
        ctx.adjoint_viscosity(ud, uad, 1e-5, vtype, scaling, iterations=0, m2norm_out=m2)
        err = rel_l2(m2.cpu().numpy(), ref)
        print('%s-only %s %s: scaled viscosity rel-L2 %.2e' % (kind, vtype, np.dtype(prec).name, err))
        assert err < tol, err
