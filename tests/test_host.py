"""Host layer (no GPU): boundary specs, viscosity models, objectives, solver
configuration."""
import numpy as np
import pytest

from adfvm_amd import BCs, objectives, meshgen, _lib
from adfvm_amd.density import RCF, mu_model
from adfvm_amd.mesh import Mesh


def test_parse_value():
    assert np.array_equal(BCs.parse_value('uniform 300', 4, 1), np.full((4, 1), 300.))
    assert np.array_equal(BCs.parse_value('uniform (1 -2e-3 3)', 2, 3), [[1, -2e-3, 3]] * 2)
    v = BCs.parse_value('nonuniform List<vector> 2\n(\n(1 2 3)\n(4 5 6.5e1)\n)\n', 2, 3)
    assert np.array_equal(v, [[1, 2, 3], [4, 5, 65]])
    v = BCs.parse_value('nonuniform List<scalar> 3(1 -2 3e-1)', 3, 1)
    assert np.array_equal(v.ravel(), [1, -2, 0.3])
    assert np.array_equal(BCs.parse_value([1., 2., 3.], 2, 3), [[1, 2, 3]] * 2)
    with pytest.raises(ValueError):
        BCs.parse_value(np.zeros(5), 2, 3)
    with pytest.raises(ValueError):
        BCs.parse_value('table 3', 2, 1)


def test_validate():
    mesh = Mesh.from_data(meshgen.box((3, 3, 3), patches={'x0': ('inlet', 'patch'), 'x1': ('outlet', 'patch'),
                                                           'y0': ('walls', 'patch'), 'y1': ('walls', 'patch'),
                                                           'z0': ('sides', 'empty'), 'z1': ('sides', 'empty')}))
    bnd = {'U': {'inlet': {'type': 'fixedValue', 'value': 'uniform (10 0 0)'},
                 'walls': {'type': 'fixedValue', 'value': 'uniform (0 0 0)'},
                 'outlet': {'type': 'zeroGradient'}},
           'T': {'inlet': {'type': 'fixedValue', 'value': 'uniform 300'},
                 'walls': {'type': 'zeroGradient'}, 'outlet': {'type': 'zeroGradient'}},
           'p': {'inlet': {'type': 'zeroGradient'}, 'walls': {'type': 'zeroGradient'},
                 'outlet': {'type': 'CBC_UPT', 'U0': 'uniform (10 0 0)', 'T0': 'uniform 300', 'p0': 'uniform 1e5'}}}
    out = BCs.validate(mesh, bnd)
    assert out['inlet']['U']['value'].shape == (mesh.boundary['inlet']['nFaces'], 3)
    assert out['sides']['U']['type'] == 'empty'          # the mesh's default
    assert out['outlet']['p']['p0'].shape == (mesh.boundary['outlet']['nFaces'], 1)
    bad = {'U': {'inlet': {'type': 'fancyInlet'}}}
    with pytest.raises(ValueError):
        BCs.validate(mesh, bad)
    missing = {'p': {'outlet': {'type': 'CBC_TOTAL_PT', 'Tt': 300.}}}
    with pytest.raises(ValueError, match="needs 'pt'"):
        BCs.validate(mesh, missing)
    cyclic_on_patch = {'U': {'inlet': {'type': 'cyclic'}}}
    with pytest.raises(ValueError):
        BCs.validate(mesh, cyclic_on_patch)


def test_mu_model():
    assert mu_model(2.5e-5) == (_lib.MU_CONSTANT, 2.5e-5, 0., 0.)
    assert mu_model(lambda T: T * 0.) == (_lib.MU_CONSTANT, 0., 0., 0.)
    m, mu0, A, S = mu_model(lambda T: 1.4792e-06 * T ** 1.5 / (T + 116.))
    assert m == _lib.MU_SUTHERLAND and mu0 == 0.
    assert abs(A - 1.4792e-06) < 1e-15 and abs(S - 116.) < 1e-7
    assert mu_model(None)[0] == _lib.MU_SUTHERLAND
    with pytest.raises(ValueError):
        mu_model(lambda T: 1e-5 * T)


def test_objectives():
    assert objectives.resolve(None) is None
    assert objectives.resolve('kinetic_energy').code == _lib.OBJ_KINETIC_ENERGY
    d = objectives.resolve(('drag', 'cylinder'))
    assert d.code == _lib.OBJ_DRAG and d.patch == 'cylinder'
    with pytest.raises(ValueError):
        objectives.resolve(lambda fields, solver: 0)


def test_solver_config():
    md = meshgen.periodic_box(4)
    s = RCF(md, mu=0., CFL=0.6, objective='kinetic_energy')
    assert s.CFL == 0.6 and s.gamma == 1.4 and s.timeIntegrator == 'SSPRK'
    assert RCF.defaultConfig['CFL'] == 1.2          # class default not mutated
    assert abs(s.Cv - 1004.5 / 1.4) < 1e-12
    with pytest.raises(KeyError):
        RCF(md, notAnOption=1)
    with pytest.raises(ValueError):
        RCF(md, riemannSolver='hllc')
    ph = s.physics()
    assert ph['riemann'] == _lib.RIEMANN_ROE and ph['integrator'] == _lib.RK_SSPRK3
    assert ph['objective'] == _lib.OBJ_KINETIC_ENERGY


def test_primitive_conservative_round_trip():
    s = RCF(meshgen.periodic_box(2), mu=0.)
    rng = np.random.default_rng(1)
    U = rng.standard_normal((8, 3)) * 50
    T = 300 + rng.random((8, 1)) * 10
    p = 1e5 + rng.random((8, 1)) * 1e3
    U2, T2, p2 = s.primitive(*s.conservative(U, T, p))
    assert np.allclose(U2, U, rtol=1e-13, atol=1e-12)
    assert np.allclose(T2, T, rtol=1e-13) and np.allclose(p2, p, rtol=1e-13)
