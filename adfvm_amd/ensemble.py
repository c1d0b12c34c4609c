"""Batched ensemble: many perturbed instances of uneven mesh sizes advanced
primal plus adjoint as one mesh.

The instances are concatenated into one disjoint mesh
(`meshgen.concatenate`): every kernel launch covers all of them, so small
instances share the GPU instead of each under-filling it, and each instance's
arithmetic is exactly what it would be alone (own cell order, own slot
tables, own geometry).  The context evaluates the objective per instance
(`Context.set_segments`).  The reverse sweep is the uniform checkpoint
schedule in HIP graphs (`graphs.GraphSweep`): by default the whole sweep
is one graph launch (or one graph per checkpoint segment).  Instances are
independent, so the
adjoint of the summed objective gives every instance its own gradient.

    ens = TaylorGreenEnsemble(n_instances=64, sizes=(32, 48, 64, 96), seed=0)
    J, ua = ens.run(n_steps=200, interval=20)
"""
import numpy as np

from . import meshgen, config
from .cases import GAMMA, R, conservative
from .mesh import Mesh


class TaylorGreenEnsemble(object):
    """n_instances Taylor-Green boxes (periodic [0, 2pi]^3, Mach 0.1,
    rho0 = 1.2, T0 = 300 K, Re = 1600), sizes drawn by seed from `sizes`,
    each with a seeded random cell permutation + RCM, velocity scaled per
    cell by (1 + perturbation * U(-1, 1)).  One fixed dt for all: CFL 0.5 on
    the finest instance."""

    def __init__(self, n_instances=64, sizes=(32, 48, 64, 96), seed=0, perturbation=0.01, precision=np.float32,
                 mach=0.1, rho0=1.2, T0=300., Re=1600.):
        from .context import Context, _need_torch
        from . import _lib
        torch = _need_torch()
        self.torch = torch
        rng = np.random.default_rng(seed)
        self.sizes = [int(s) for s in rng.choice(np.asarray(sizes), n_instances)]
        mds = [meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=seed * 100003 + k)
               for k, n in enumerate(self.sizes)]
        md, self.offsets = meshgen.concatenate(mds, ['i%03d_' % k for k in range(n_instances)])
        del mds
        self.mesh = Mesh.from_data(md)
        m = self.mesh
        nIC = m.nInternalCells
        c0 = np.sqrt(GAMMA * R * T0)
        self.U0 = U0 = mach * c0
        x, y, z = [m.cellCentres[:nIC, i:i + 1] for i in range(3)]
        U = np.concatenate([U0 * np.sin(x) * np.cos(y) * np.cos(z), -U0 * np.cos(x) * np.sin(y) * np.cos(z),
                            0 * x], 1)
        U *= 1 + perturbation * rng.uniform(-1, 1, (nIC, 1))
        p = rho0 * R * T0 + rho0 * U0 ** 2 / 16 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
        T = T0 + 0 * x
        rho, rhoU, rhoE = conservative(U, T, p)
        u0 = np.concatenate([rho.T, rhoU.T, rhoE.T])
        self.mu = rho0 * U0 / Re
        self.dt = 0.5 * (2 * np.pi / max(self.sizes)) / (U0 + c0)
        ph = dict(gamma=GAMMA, Cp=1004.5, Pr=0.7, mu=(_lib.MU_CONSTANT, self.mu, 0., 0.),
                  riemann=_lib.RIEMANN_ROE, boundaryRiemann=_lib.RIEMANN_ROE, small=config.small_for(precision),
                  integrator=_lib.RK_SSPRK3, objective=_lib.OBJ_KINETIC_ENERGY)
        self.ctx = Context(m, {}, ph, precision)
        self.ctx.set_segments(self.offsets)
        self.nSeg = len(self.offsets) - 1
        self.u0 = torch.as_tensor(u0, dtype=self.ctx.tdtype, device='cuda').contiguous()
        self._sweeps = {}

    @property
    def n_cells(self):
        return self.mesh.nInternalCells

    def run(self, n_steps=200, interval=20, graphs=True, tapes=True, whole=None, overlap=None):
        """Primal + adjoint of J_k = sum over steps of instance k's kinetic
        energy (on each step's stage-1 state).  Returns (J (nSeg,) FP64
        host array, ua: d(sum_k J_k)/du0 on the device, i.e. each instance's
        own gradient).  With graphs, the whole sweep is one captured HIP
        graph, a single launch per run (whole=False: one graph per
        checkpoint segment, replayed per segment).  overlap: a segment's
        re-run alongside the next segment's adjoint steps (graphs.GraphSweep;
        None: when a second set of segment buffers fits)."""
        from .graphs import GraphSweep
        whole = graphs if whole is None else whole
        key = (interval, graphs, tapes, whole, overlap)
        if key not in self._sweeps:                 # captured once, replayed by later runs
            self._sweeps[key] = GraphSweep(self.ctx, self.dt, interval, 1.0, graphs, tapes, whole=whole,
                                           overlap=overlap)
        self.last_sweep = self._sweeps[key]
        return self.last_sweep.run(self.u0, n_steps)
