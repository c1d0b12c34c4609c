"""The C++/OpenMP FP64 CPU baseline (oracle/cpu_step.cpp, timed by bench.py
as the CPU path) against the oracle on a small periodic case: one step to
FP64 round-off, one adjoint step against the oracle's autograd.  CPU only."""
import os

import numpy as np
import pytest
import torch

from oracle import cpu_baseline
from tests import casefactory as cf


@pytest.fixture(scope='module')
def built():
    if not os.path.exists(cpu_baseline.LIB) or os.path.getmtime(cpu_baseline.LIB) < os.path.getmtime(cpu_baseline.SRC):
        cpu_baseline.build()
    return cpu_baseline.load()


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_cpu_step_and_adjoint_match_the_oracle(built):
    case, params = cf.tgv(6, seed=4)
    u = cf.tgv_state(case, params)
    rng = np.random.default_rng(2)
    u = u * (1 + 1e-3 * rng.normal(size=u.shape))           # off the exact TGV state (Roe's |.| kinks)
    dt = 2e-5
    cs = cpu_baseline.CpuStep(case.mesh, case.mu0)
    un = np.ascontiguousarray(u)
    x = torch.tensor(u, dtype=torch.float64, requires_grad=True)
    oc = case.oracle()
    r, rU, rE, dtc, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt)
    out = torch.cat([r.T, rU.T, rE.T])
    got = un.copy()
    J = cs.step(got, dt)
    for s in (slice(0, 1), slice(1, 4), slice(4, 5)):
        assert rel_l2(got[s], out.detach().numpy()[s]) < 1e-13
    assert abs(J - float(obj.detach())) <= 1e-13 * abs(float(obj.detach()))
    w = rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True) * 1e-3
    (g,) = torch.autograd.grad((out * torch.as_tensor(w)).sum() + 0.7 * obj, x)
    ua = np.ascontiguousarray(w.copy())
    cs.adjoint_step(un, dt, ua, 0.7)
    ref = g.numpy()
    for s in (slice(0, 1), slice(1, 4), slice(4, 5)):
        assert rel_l2(ua[s], ref[s]) < 1e-11, s
    assert cpu_baseline.CpuStep.threads() >= 1
