"""The reverse sweep with each segment's re-run on a second stream alongside
the next segment's adjoint steps (GraphSweep overlap) against the same sweep
in sequence: objective, adjoint state and boundary-value gradient bitwise
equal, eager, with per-segment graphs and as one whole-sweep graph, in FP32
and FP64, replayed twice."""
import numpy as np
import pytest

from tests import casefactory as cf

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _cases():
    case, params = cf.tgv(10, seed=6)
    u = cf.tgv_state(case, params)
    u[1:4] += u[0] * np.random.default_rng(9).normal(size=u[1:4].shape) * 0.3
    yield 'tgv10', case, u, 2e-4, False
    c2 = cf.mixed_box(4, 'CBC_UPT', sutherland=True, objective=('drag', 'top'))
    yield 'mixed_drag', c2, cf.smooth_state(c2, 4), 1e-6, True


@pytest.mark.parametrize('mode', ['eager', 'segment_graphs', 'whole_graph'])
def test_overlapped_reverse_sweep_is_bitwise_the_sequential_one(mode):
    from adfvm_amd.graphs import GraphSweep
    graphs, whole = mode != 'eager', mode == 'whole_graph'
    interval, n = 2, 8                               # four segments: both buffer sets, twice
    for name, case, u, dt, bcg in _cases():
        for prec in (np.float32, np.float64):
            ctx = case.context(prec)
            u0 = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
            out = {}
            for overlap in (False, True):
                sw = GraphSweep(ctx, dt, interval, obja=1.0, graphs=graphs, whole=whole, overlap=overlap,
                                bc_grad=bcg)
                assert sw.overlap == overlap
                for rep in range(2):                 # capture (or eager) run, then a replay
                    J, ua = sw.run(u0, n)
                    res = (J.copy(), ua.clone(), None if not bcg else sw.bc_grad.clone())
                    if rep == 1:
                        assert np.array_equal(res[0], first[0]) and torch.equal(res[1], first[1])
                    first = res
                out[overlap] = first
            (J0, ua0, b0), (J1, ua1, b1) = out[False], out[True]
            print('%s %s %s: J %.15e overlapped %.15e' % (name, np.dtype(prec).name, mode, J0[0], J1[0]))
            assert np.array_equal(J0, J1)
            assert torch.equal(ua0, ua1)
            if bcg:
                assert torch.equal(b0, b1)
