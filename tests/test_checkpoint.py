"""Checkpoint schedules (adfvm_amd.checkpoint) on a CPU toy system: every
schedule must hand the reverse sweep the exact start state of each step, in
reverse order, and the adjoint must equal the store-everything result;
revolve must meet Griewank & Walther's minimal forward-step count."""
import numpy as np
import pytest

from adfvm_amd import checkpoint


class ToyStepper(object):
    """u_{n+1} = u_n + h_n sin(u_n) elementwise; J = sum_n <c, u_n>
    (objective on each step's start state).  reverse() applies the exact
    transpose, so the final adjoint is dJ/du_0."""

    def __init__(self, n, dim=4, seed=0, tapes=True):
        rng = np.random.default_rng(seed)
        self.h = rng.uniform(0.05, 0.2, n)
        self.c = rng.standard_normal(dim)
        self.ua = np.zeros(dim)
        self.forwards = 0
        self.reversed = []
        self.tapes = tapes
        self.truth = None

    def clone(self, u):
        return u.copy()

    def forward(self, u, n, keep):
        self.forwards += 1
        tok = (n, np.sin(u).copy(), np.cos(u).copy()) if (keep and self.tapes) else None
        u += self.h[n] * np.sin(u)
        return tok

    def reverse(self, un, n, tok):
        if self.truth is not None:
            assert np.array_equal(un, self.truth[n]), n
        if tok is not None:
            assert tok[0] == n and np.array_equal(tok[2], np.cos(un))
        self.reversed.append(n)
        self.ua = self.ua * (1 + self.h[n] * np.cos(un)) + self.c


def trajectory(st, u0, n):
    u = u0.copy()
    out = [u.copy()]
    for k in range(n):
        u = u + st.h[k] * np.sin(u)
        out.append(u.copy())
    return out


def reference_adjoint(st, traj, n):
    ua = np.zeros_like(traj[0])
    for k in range(n - 1, -1, -1):
        ua = ua * (1 + st.h[k] * np.cos(traj[k])) + st.c
    return ua


def run(schedule, n, **kw):
    st = ToyStepper(n)
    u0 = np.random.default_rng(1).standard_normal(4)
    st.truth = trajectory(st, u0, n)
    u0c = u0.copy()
    if schedule == 'store_all':
        checkpoint.store_all(st, u0, 0, n)
    elif schedule == 'uniform':
        checkpoint.uniform(st, u0, 0, n, kw['interval'])
    else:
        checkpoint.revolve(st, u0, 0, n, kw['snapshots'])
    assert np.array_equal(u0, u0c)                      # u0 untouched
    assert st.reversed == list(range(n - 1, -1, -1))
    ref = reference_adjoint(st, st.truth, n)
    assert np.allclose(st.ua, ref, rtol=1e-13, atol=1e-13)
    # a sanity check of the toy adjoint itself: finite differences of J
    return st


def test_toy_adjoint_is_the_gradient():
    n = 7
    st = ToyStepper(n)
    u0 = np.random.default_rng(1).standard_normal(4)
    J = lambda u: sum(st.c @ x for x in trajectory(st, u, n)[:n])
    g = reference_adjoint(st, trajectory(st, u0, n), n)
    eps = 1e-6
    fd = np.array([(J(u0 + eps * e) - J(u0 - eps * e)) / (2 * eps) for e in np.eye(4)])
    assert np.allclose(g, fd, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize('n', [1, 2, 5, 17])
def test_store_all(n):
    st = run('store_all', n)
    assert st.forwards == max(n - 1, 0)


@pytest.mark.parametrize('n,interval', [(20, 10), (20, 20), (23, 5), (5, 1), (400, 20)])
def test_uniform(n, interval):
    st = run('uniform', n, interval=interval)
    starts = len(range(0, n, interval))
    # one sweep to the last segment start, then each segment re-run
    assert st.forwards == (starts - 1) * interval + (n - starts)


@pytest.mark.parametrize('n', [1, 2, 3, 10, 37, 100, 400])
@pytest.mark.parametrize('snapshots', [0, 1, 2, 3, 5, 10])
def test_revolve_minimal(n, snapshots):
    if snapshots == 0 and n > 100:
        pytest.skip('quadratic schedule')
    st = run('revolve', n, snapshots=snapshots)
    assert st.forwards == checkpoint.min_forward_steps(n, snapshots), (st.forwards,)


def test_beta_closed_forms():
    # one checkpoint (initial state only): t repetitions reverse t+1 steps
    assert [checkpoint.beta(1, t) for t in range(5)] == [1, 2, 3, 4, 5]
    # Pascal recursion
    for c in range(1, 6):
        for t in range(1, 6):
            assert checkpoint.beta(c, t) == checkpoint.beta(c - 1, t) + checkpoint.beta(c, t - 1)
    # no snapshots: n(n-1)/2 forward steps
    assert checkpoint.min_forward_steps(10, 0) == 45
    # enough snapshots to store everything: n-1 forward steps
    assert checkpoint.min_forward_steps(10, 9) == 9
