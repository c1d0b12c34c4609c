"""ORACLE (test infrastructure only): CPU float64 restatement of adFVM's
compressible finite-volume step, written with torch tensor ops so that
torch.autograd supplies the exact discrete adjoint (the role adpy's reverse
mode plays for the reference, apps/adjoint.py:101).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker; the product runs HIP kernels.

Each function cites the reference lines it restates.  Face->cell
accumulation runs over each cell's cellFaces slots in order (owned faces in
face order, then neighbour faces), a fixed order; the reference leaves the
order of its scatter-add (`Tensor.collate`) to adpy, so bitwise agreement with
the reference's sums is not claimed.

How each function is pinned (tests/test_oracle.py unless named):
  primitive/conservative   round trip and p = rho R T
  ghosts                   fixedValue/zeroGradient copies, slip (zero normal
                           velocity), CBC_TOTAL_PT total-temperature and
                           isentropic relations
  grad_cell                exact for linear fields; zero for constants on
                           an O-grid
  second_order             exact face values for linear fields, both sides;
                           limiter bounds and identity inside the bounds
  euler_roe / euler_lf     consistency, conservation F(L,R,n) = -F(R,L,-n),
                           rotational invariance, Roe upwinding of a contact,
                           LF dissipation >= Roe
  viscous_flux             zero for uniform fields; linear-T heat flux and
                           Couette shear in closed form
  accumulate / equation    sum_c V R = 0 on a periodic box; dtc by hand
  step                     conservation, constant-source update, SSP-RK3
                           third order, autograd vs central differences
                           (w.r.t. the state and w.r.t. boundary values), and
                           the shock tube against the exact solution
                           (tests/golden/shocktube_exact_output.txt, the
                           reference's cases/shockTube/output)
  objective_value          drag and kinetic energy in closed form; step()'s
                           objective on the stage-1 or (obj_on_start) the
                           input state, the latter the closed form again
  face_gradient            its defining property on a sheared (non-orthogonal)
                           mesh: the component along D is the difference
                           quotient, the part normal to D is unchanged
  viscous_flux             (also) the Newtonian stress with Stokes' hypothesis
                           mu (G + G^T - 2/3 tr(G) I) n for a linear velocity
                           of arbitrary gradient G (dilatation included) on a
                           sheared mesh; the 4/3 mu a normal stress of a pure
                           dilatation
  stabilise                immaterial: the residual and its autograd VJP are
                           unchanged (<= 1e-15) under the forms that keep
                           stabilise(0) > 0, x + SMALL (= x + sign(x) SMALL
                           with sign(0) = 1) and max(x, SMALL), on TGV, Sod,
                           mixed-boundary and uniform states
  symmetriser /            A0 = dU/dv, the inverse of the Jacobian of the
  adjoint_energy           entropy variables v(U) (autograd); w = e_1 gives
                           sum rho V; a uniform state in closed form
Parity unpinned: the bitwise order of the reference's face->cell sums
(adpy's Tensor.collate), which this restatement does not claim.

Inputs: an oracle mesh (oracle.mesh_oracle.build) and a case spec:
  bcs[patch][field] = {'type': 'cyclic'|'zeroGradient'|'empty'|'inletOutlet'|
                       'symmetryPlane'|'slip'|'fixedValue'|'calculated'|
                       'CBC_UPT'|'CBC_TOTAL_PT', 'value': array, ...}
  (missing fields default to the mesh's default boundary, mesh.py:702-711)
"""
import numpy as np
import torch

DT = torch.float64
# the reference's differentiable mesh arrays, in order (Mesh.gradFields, mesh.py:28-32)
MESH_GRAD_FIELDS = ('areas', 'volumesL', 'volumesR', 'weights', 'deltas', 'normals', 'deltasUnit',
                    'linearWeights', 'quadraticWeights', 'volumes')
CYCLIC_TYPES = ('cyclic', 'slidingPeriodic1D')
COPY_TYPES = ('zeroGradient', 'empty', 'inletOutlet')          # BCs.py:97-104,135-137
SYMMETRY_TYPES = ('symmetryPlane', 'slip')                       # BCs.py:120-135
DEFAULT_TYPES = ('cyclic', 'slidingPeriodic1D', 'processor', 'processorCyclic', 'symmetryPlane', 'empty')


def T_(a, dtype=DT):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype)


def bvalue(v, n, k):
    """A patch value given uniform (k numbers) or per face (n x k).  A torch
    tensor is used as it is (broadcast if uniform), so that autograd of the
    step reaches boundary-condition values."""
    if isinstance(v, torch.Tensor):
        t = v.to(DT)
        return (t.reshape(1, k).expand(n, k) if t.numel() == k else t.reshape(n, k)) * 1
    a = np.asarray(v, np.float64)
    if a.size == k:
        a = np.tile(a.reshape(1, k), (n, 1))
    return T_(a.reshape(n, k))


class Case(object):
    """Physics constants and boundary spec; defaults from RCF.defaultConfig
    (adFVM/density.py:22-39)."""

    def __init__(self, mesh, bcs=None, gamma=1.4, Cp=1004.5, Pr=0.7, mu=None,
                 riemann='roe', boundaryRiemann='roe', small=1e-30, objective=None,
                 integrator='SSPRK', limiter=False):
        self.mesh = mesh
        self.gamma, self.Cp, self.Pr = gamma, Cp, Pr
        self.Cv = Cp / gamma                      # density.py:44
        self.R = Cp - self.Cv                     # density.py:45
        self.mu = mu if mu is not None else (lambda T: 1.4792e-06 * T ** 1.5 / (T + 116.))
        self.riemann, self.boundaryRiemann = riemann, boundaryRiemann
        self.small = small
        self.objective = objective
        self.integrator = integrator
        self.limiter = limiter
        m = mesh
        self.nIC, self.nIF, self.nF, self.nC = m.nInternalCells, m.nInternalFaces, m.nFaces, m.nCells
        self.t = {k: T_(getattr(m, k)) for k in ['areas', 'normals', 'weights', 'deltas', 'deltasUnit',
                                                    'linearWeights', 'quadraticWeights', 'volumes']}
        self.t['volumes'] = self.t['volumes'].reshape(-1)
        self.t['areas'] = self.t['areas'].reshape(-1)
        self.t['weights'] = self.t['weights'].reshape(-1)
        self.t['deltas'] = self.t['deltas'].reshape(-1)
        self.t['quadraticWeights'] = self.t['quadraticWeights'].reshape(-1, 2, 3)
        self.owner = torch.as_tensor(m.owner.astype(np.int64))
        self.neighbour = torch.as_tensor(m.neighbour.astype(np.int64))
        self.cellFaces = torch.as_tensor(m.cellFaces.astype(np.int64))
        self.cellOwner = torch.as_tensor(m.cellOwner.astype(np.int64))
        self.cellNeighbours = torch.as_tensor(m.cellNeighbours.astype(np.int64))
        # per-face volumes of the divergence (mesh.py:108-109): owner side of
        # every face, neighbour side of internal faces; separate arrays of the
        # reference's step (inputs 'volumesL', 'volumesR', mesh.py:28-32)
        self.t['volumesL'] = self.t['volumes'][self.owner]
        self.t['volumesR'] = self.t['volumes'][self.neighbour[:self.nIF]]
        # patches: sorted (alphabetical) local patches, mesh.py:100-101
        self.patches = []
        bcs = bcs or {}
        for name in sorted(m.boundary.keys()):
            b = m.boundary[name]
            spec = dict(bcs.get(name, {}))
            for f in ('U', 'T', 'p'):
                if f not in spec:
                    spec[f] = {'type': b['type'] if b['type'] in DEFAULT_TYPES else 'zeroGradient'}
            ptype = b['type']
            cbc = spec['p']['type'] if spec['p']['type'] in ('CBC_UPT', 'CBC_TOTAL_PT') else None
            if cbc is not None:
                ptype = 'characteristic'              # BCs.py:155,167
            self.patches.append(dict(name=name, start=b['startFace'], n=b['nFaces'],
                                     cstart=b['startFace'] - self.nIF + self.nIC, type=ptype,
                                     nbr=m.boundary[b['neighbourPatch']]['startFace'] if b['type'] in CYCLIC_TYPES else None,
                                     spec=spec, cbc=cbc))
        # ghost gradient source (default boundaries of the gradient fields: copy
        # owner, or the paired owner on cyclic patches), density.py:376-385
        gsrc = torch.arange(self.nC)
        for p in self.patches:
            if p['n'] == 0:
                continue
            if m.boundary[p['name']]['type'] in CYCLIC_TYPES:
                gsrc[p['cstart']:p['cstart'] + p['n']] = self.owner[p['nbr']:p['nbr'] + p['n']]
            else:
                gsrc[p['cstart']:p['cstart'] + p['n']] = self.owner[p['start']:p['start'] + p['n']]
        self.grad_src = gsrc

    def to(self, device):
        """Move the case's tensors to `device` (torch runs the same
        restatement there; inputs must be on that device too).  Returns
        self."""
        self.t = {k: v.to(device) for k, v in self.t.items()}
        for k in ('owner', 'neighbour', 'cellFaces', 'cellOwner', 'cellNeighbours', 'grad_src'):
            setattr(self, k, getattr(self, k).to(device))
        return self

    # --------------------------------------------------------- primitives --
    def primitive(self, rho, rhoU, rhoE):
        """adFVM/density.py:162-171."""
        U = rhoU / rho
        E = rhoE / rho
        e = E - 0.5 * (U * U).sum(-1, keepdim=True)
        p = (self.gamma - 1) * rho * e
        T = e * (1. / self.Cv)
        return U, T, p

    def conservative(self, U, T, p):
        """adFVM/density.py:173-181."""
        e = self.Cv * T
        rho = p / (e * (self.gamma - 1))
        rhoE = rho * (e + 0.5 * (U * U).sum(-1, keepdim=True))
        rhoU = U * rho
        return rho, rhoU, rhoE

    # ------------------------------------------------------------ ghosts --
    def ghosts(self, U, T, p, ghostInit=None):
        """Fill ghost rows: per field, patches in sorted order (field.py:148-156,
        BCs.py:86-133), then characteristic patches (density.py:333-337,
        BCs.py:152-184).  U,T,p: internal (nIC,k); returns (nC,k)."""
        nG = self.nC - self.nIC
        if ghostInit is None:
            ghostInit = [torch.zeros(nG, 3, dtype=U.dtype, device=U.device),
                         torch.zeros(nG, 1, dtype=U.dtype, device=U.device),
                         torch.zeros(nG, 1, dtype=U.dtype, device=U.device)]
        fields = {'U': torch.cat([U, ghostInit[0]]), 'T': torch.cat([T, ghostInit[1]]),
                  'p': torch.cat([p, ghostInit[2]])}
        n_all = self.t['normals']
        for name in ('U', 'T', 'p'):
            phi = fields[name]
            parts = [phi[:self.nIC]]
            cur = self.nIC
            gh = phi[self.nIC:].clone()
            rows = {}
            for pt in self.patches:
                if pt['n'] == 0:
                    continue
                s, n, cs = pt['start'], pt['n'], pt['cstart']
                kind = pt['spec'][name]['type']
                own = self.owner[s:s + n]
                if kind in CYCLIC_TYPES:
                    val = phi[self.owner[pt['nbr']:pt['nbr'] + n]]
                elif kind in COPY_TYPES:
                    val = phi[own]
                elif kind in SYMMETRY_TYPES:
                    val = phi[own]
                    if name == 'U':
                        nn = n_all[s:s + n]
                        val = val - (val * nn).sum(-1, keepdim=True) * nn
                elif kind == 'fixedValue':
                    val = bvalue(pt['spec'][name]['value'], n, 3 if name == 'U' else 1).to(phi) * 1
                elif kind in ('calculated', 'CBC_UPT', 'CBC_TOTAL_PT'):
                    continue
                else:
                    raise NotImplementedError(kind)
                rows[cs - self.nIC] = (n, val)
            pieces, g = [], 0
            for off in sorted(rows):
                if off > g:
                    pieces.append(gh[g:off])
                pieces.append(rows[off][1])
                g = off + rows[off][0]
            if g < nG:
                pieces.append(gh[g:])
            fields[name] = torch.cat(parts + pieces)
        # characteristic BCs
        for pt in self.patches:
            if pt['cbc'] is None or pt['n'] == 0:
                continue
            s, n, cs = pt['start'], pt['n'], pt['cstart']
            sp = pt['spec']['p']
            if pt['cbc'] == 'CBC_UPT':
                Ub = bvalue(sp['U0'], n, 3).to(fields['U']) * 1
                Tb = bvalue(sp['T0'], n, 1).to(fields['U']) * 1
                pb = bvalue(sp['p0'], n, 1).to(fields['U']) * 1
            else:
                own = self.owner[s:s + n]
                Uo, To = fields['U'][own], fields['T'][own]
                d = bvalue(sp['direction'], n, 3).to(Uo) if 'direction' in sp else n_all[s:s + n]
                Tt = bvalue(sp['Tt'], n, 1).to(Uo)
                ptot = bvalue(sp['pt'], n, 1).to(Uo)
                Un = (Uo * d).sum(-1, keepdim=True)
                Ub = Un * d
                Tb = Tt - 0.5 * Un * Un / self.Cp
                pb = ptot * (To / Tt) ** (self.gamma / (self.gamma - 1))
            for name, val in (('U', Ub), ('T', Tb), ('p', pb)):
                phi = fields[name]
                fields[name] = torch.cat([phi[:cs], val.to(phi.dtype), phi[cs + n:]])
        return fields['U'], fields['T'], fields['p']

    # -------------------------------------------------------- operators --
    def grad_cell(self, phi):
        """Gauss gradient over the 6 faces, adFVM/op.py:45-63.  phi (nC,k) ->
        (nIC,k,3)."""
        S = self.t['areas'][self.cellFaces]                 # (nIC,6)
        N = self.t['normals'][self.cellFaces]               # (nIC,6,3)
        w = self.t['weights'][self.cellFaces]
        O = self.cellOwner.to(phi.dtype)
        N = 2 * N * O[..., None] - N
        w = w + O - 2 * w * O
        phiP = phi[self.cellNeighbours]                     # (nIC,6,k)
        phic = phi[:self.nIC]
        g = 0
        for i in range(6):
            phiF = phic * (-w[:, i:i + 1] + 1) + phiP[:, i] * w[:, i:i + 1]
            g = g + phiF[:, :, None] * (S[:, i, None, None] * N[:, i, None, :])
        return g / self.t['volumes'][:, None, None]

    def central(self, phi):
        """adFVM/interp.py:15-18 (face values on all faces)."""
        f = self.t['weights'][:, None]
        return phi[self.owner] * f + phi[self.neighbour] * (-f + 1)

    def second_order(self, phi, grad, swap, faces=None):
        """adFVM/interp.py:20-28: phi_C + (phi_D-phi_C)*lw + qw.grad(phi_C)."""
        faces = slice(None) if faces is None else faces
        p, n = self.owner[faces], self.neighbour[faces]
        gsrc = self.grad_src
        if swap:
            p, n = n, p
        phiC, phiD = phi[p], phi[n]
        lw = self.t['linearWeights'][faces][:, swap:swap + 1]
        qw = self.t['quadraticWeights'][faces][:, swap]        # (nf,3)
        gC = grad[gsrc[p]]                                       # (nf,k,3)
        out = phiC + (phiD - phiC) * lw + (gC * qw[:, None, :]).sum(-1)
        if self.limiter:
            # optional clamp between the two cell values (not in the reference);
            # torch.where so that ties pass the adjoint to one selected value
            lo = torch.where(phiC < phiD, phiC, phiD)
            hi = torch.where(phiC < phiD, phiD, phiC)
            out = torch.where(out < lo, lo, torch.where(out > hi, hi, out))
        return out

    # --------------------------------------------------- Riemann solvers --
    def stabilise(self, x):
        """eps.stabilise(SMALL), riemann.py:52; taken as x + sign(x)*SMALL
        (x >= 0 here).  adpy's definition is absent; every form that keeps
        stabilise(0) > 0 gives the same residual and VJP to <= 1e-15
        (test_stabilise_form_is_immaterial)."""
        return torch.where(x < 0, x - self.small, x + self.small)

    def euler_roe(self, pL, pR, TL, TR, UL, UR, rhoL, rhoR, rhoUL, rhoUR, rhoEL, rhoER, n):
        """adFVM/riemann.py:21-70."""
        g = self.gamma
        dot = lambda a, b: (a * b).sum(-1, keepdim=True)
        rhoUnL, rhoUnR = rhoL * dot(UL, n), rhoR * dot(UR, n)
        hL = g * pL / ((g - 1) * rhoL) + 0.5 * dot(UL, UL)
        hR = g * pR / ((g - 1) * rhoR) + 0.5 * dot(UR, UR)
        Frho = 0.5 * (rhoUnL + rhoUnR)
        FrhoU = 0.5 * (rhoUnL * UL + rhoUnR * UR + (pL + pR) * n)
        FrhoE = 0.5 * (rhoUnL * hL + rhoUnR * hR)
        sL, sR = rhoL.sqrt(), rhoR.sqrt()
        div = sL + sR
        UF = (UL * sL + UR * sR) / div
        hF = (hL * sL + hR * sR) / div
        qF = 0.5 * dot(UF, UF)
        a2F = (g - 1) * (hF - qF)
        aF = a2F.sqrt()
        UnF = dot(UF, n)
        drho = rhoR - rhoL
        drhoU = rhoR * UR - rhoL * UL
        drhoE = (hR * rhoR - pR) - (hL * rhoL - pL)
        lam1, lam2, lam3 = UnF.abs(), (UnF + aF).abs(), (UnF - aF).abs()
        eps = 0.5 * (rhoUnL / rhoL - rhoUnR / rhoR).abs()
        eps = eps + 0.5 * ((g * pL / rhoL).sqrt() - (g * pR / rhoR).sqrt()).abs()
        eps = self.stabilise(eps)
        fix = lambda lam: torch.where(lam < 2. * eps, .25 * lam * lam / eps + eps, lam)
        lam1, lam2, lam3 = fix(lam1), fix(lam2), fix(lam3)
        abv1 = 0.5 * (lam2 + lam3)
        abv2 = 0.5 * (lam2 - lam3)
        abv3 = abv1 - lam1
        abv4 = (g - 1) * (qF * drho - dot(UF, drhoU) + drhoE)
        abv5 = UnF * drho - dot(drhoU, n)
        abv6 = abv3 * abv4 / a2F - abv2 * abv5 / aF
        abv7 = abv3 * abv5 - abv2 * abv4 / aF
        Frho = Frho - 0.5 * (lam1 * drho + abv6)
        FrhoU = FrhoU - 0.5 * (lam1 * drhoU + UF * abv6 - abv7 * n)
        FrhoE = FrhoE - 0.5 * (lam1 * drhoE + hF * abv6 - UnF * abv7)
        return Frho, FrhoU, FrhoE

    def euler_lf(self, pL, pR, TL, TR, UL, UR, rhoL, rhoR, rhoUL, rhoUR, rhoEL, rhoER, n):
        """adFVM/riemann.py:6-18."""
        g = self.gamma
        UnL, UnR = (UL * n).sum(-1, keepdim=True), (UR * n).sum(-1, keepdim=True)
        cL = (g * pL / rhoL).sqrt()
        cR = (g * pR / rhoR).sqrt()
        aF = torch.maximum(UnL.abs() + cL, UnR.abs() + cR)
        Frho = 0.5 * (rhoL * UnL + rhoR * UnR) - 0.5 * aF * (rhoR - rhoL)
        FrhoU = 0.5 * (rhoUL * UnL + rhoUR * UnR + (pL + pR) * n) - 0.5 * aF * (rhoUR - rhoUL)
        FrhoE = 0.5 * ((rhoEL + pL) * UnL + (rhoER + pR) * UnR) - 0.5 * aF * (rhoER - rhoEL)
        return Frho, FrhoU, FrhoE

    def riemann_fn(self, name):
        return self.euler_roe if name == 'roe' else self.euler_lf

    # ------------------------------------------------------------ fluxes --
    def face_gradient(self, gradF, phiL, phiR, faces):
        """Non-orthogonal face gradient of density.py:197-200: the averaged
        gradient gradF (nf,k,3) with its component along the unit
        centre-to-centre direction D replaced by the difference quotient
        snGrad = (phiR - phiL)/delta (op.snGrad, op.py:66-67):
            gradF + snGrad D - (gradF . D) D."""
        D = self.t['deltasUnit'][faces][:, None, :]
        dl = self.t['deltas'][faces][:, None]
        sn = (phiR - phiL) / dl                              # (nf,k)
        return gradF + sn[:, :, None] * D - (gradF * D).sum(-1)[:, :, None] * D

    def viscous_flux(self, TL, TR, UL, UR, TF, UF, gradTF, gradUF, faces):
        """adFVM/density.py:192-222 with op.snGrad (op.py:66-67)."""
        mu = self.mu(TF)
        kappa = mu * (self.Cp / self.Pr)                     # density.py:46
        N = self.t['normals'][faces]
        gT = self.face_gradient(gradTF, TL, TR, faces)[:, 0]     # (nf,3)
        gU = self.face_gradient(gradUF, UL, UR, faces)           # (nf,3,3)
        qF = kappa * (gT * N).sum(-1, keepdim=True)
        tmp2 = ((gU + gU.transpose(1, 2)) * N[:, None, :]).sum(-1)
        tmp3 = (gU[:, 0, 0] + gU[:, 1, 1] + gU[:, 2, 2])[:, None]
        sigma = mu * (tmp2 - 2. / 3 * tmp3 * N)
        return -sigma, -(qF + (sigma * UF).sum(-1, keepdim=True))

    def face_flux(self, U, T, p, gU, gT, gp, faces, characteristic=False, boundary=False):
        """RCF.flux (density.py:253-305) or RCF.boundaryFlux (:307-331) on a
        face range.  Returns (Frho, FrhoU, FrhoE, maxaF) per face."""
        P, N = self.owner[faces], self.neighbour[faces]
        n = self.t['normals'][faces]
        gs = self.grad_src
        TL, TR = T[P], T[N]
        UL, UR = U[P], U[N]
        if boundary:
            pR = p[N]
            gUR, gTR = gU[gs[N]], gT[gs[N]]
            rhoR, rhoUR, rhoER = self.conservative(UR, TR, pR)
            Un = (UR * n).sum(-1, keepdim=True)
            Frho = rhoR * Un
            FrhoU = rhoUR * Un + pR * n
            FrhoE = (rhoER + pR) * Un
            vU, vE = self.viscous_flux(TL, TR, UL, UR, TR, UR, gTR, gUR, faces)
            aF = (self.Cp * TR * (self.gamma - 1)).sqrt()
            maxaF = (UR * n).sum(-1, keepdim=True).abs() + aF
            return Frho, FrhoU + vU, FrhoE + vE, maxaF
        ULF = self.second_order(U, gU, 0, faces)
        TLF = self.second_order(T, gT, 0, faces)
        pLF = self.second_order(p, gp, 0, faces)
        rhoLF, rhoULF, rhoELF = self.conservative(ULF, TLF, pLF)
        if characteristic:
            URF, TRF, pRF = UR, TR, p[N]
            rhoRF, rhoURF, rhoERF = self.conservative(URF, TRF, pRF)
            Frho, FrhoU, FrhoE = self.riemann_fn(self.boundaryRiemann)(
                pLF, pRF, TLF, TRF, ULF, URF, rhoLF, rhoRF, rhoULF, rhoURF, rhoELF, rhoERF, n)
            TF, UF = TR, UR
            gTF, gUF = gT[gs[N]], gU[gs[N]]
        else:
            URF = self.second_order(U, gU, 1, faces)
            TRF = self.second_order(T, gT, 1, faces)
            pRF = self.second_order(p, gp, 1, faces)
            rhoRF, rhoURF, rhoERF = self.conservative(URF, TRF, pRF)
            Frho, FrhoU, FrhoE = self.riemann_fn(self.riemann)(
                pLF, pRF, TLF, TRF, ULF, URF, rhoLF, rhoRF, rhoULF, rhoURF, rhoELF, rhoERF, n)
            UF = 0.5 * (ULF + URF)
            TF = 0.5 * (TLF + TRF)
            gTF = 0.5 * (gT[gs[P]] + gT[gs[N]])
            gUF = 0.5 * (gU[gs[P]] + gU[gs[N]])
        vU, vE = self.viscous_flux(TL, TR, UL, UR, TF, UF, gTF, gUF, faces)
        aF = (self.Cp * TF * (self.gamma - 1)).sqrt()
        maxaF = (UF * n).sum(-1, keepdim=True).abs() + aF
        return Frho, FrhoU + vU, FrhoE + vE, maxaF

    def all_face_fluxes(self, U, T, p, gU, gT, gp):
        """Per-face flux for every face: internal faces (density.py:390),
        then each sorted patch by its type (:391-400)."""
        nF = self.nF
        out = [None] * 4
        parts = [(slice(0, self.nIF), self.face_flux(U, T, p, gU, gT, gp, slice(0, self.nIF)))]
        for pt in self.patches:
            if pt['n'] == 0:
                continue
            sl = slice(pt['start'], pt['start'] + pt['n'])
            if pt['type'] in CYCLIC_TYPES:
                r = self.face_flux(U, T, p, gU, gT, gp, sl)
            elif pt['type'] == 'characteristic':
                r = self.face_flux(U, T, p, gU, gT, gp, sl, characteristic=True)
            else:
                r = self.face_flux(U, T, p, gU, gT, gp, sl, boundary=True)
            parts.append((sl, r))
        parts.sort(key=lambda x: x[0].start)
        for i in range(4):
            out[i] = torch.cat([r[i] for _, r in parts])
        return out

    def accumulate(self, faceVals, absolute=False):
        """div / absDiv (adFVM/op.py:12-29): owner += F*(A/volumesL),
        neighbour += -F*(A/volumesR) (or +, absolute), summed over the
        cell's cellFaces slots in order.  faceVals (nF,k) -> (nIC,k)."""
        A = self.t['areas']
        VL, VR = self.t['volumesL'], self.t['volumesR']
        own = self.cellOwner.bool()
        acc = 0
        for j in range(6):
            f = self.cellFaces[:, j]
            fr = f.clamp(max=self.nIF - 1)         # neighbour slots are internal faces
            wP = A[f] / VL[f]
            wN = A[f] / VR[fr] if self.nIF > 0 else torch.zeros_like(wP)
            wgt = torch.where(own[:, j], wP, wN if absolute else -wN)
            acc = acc + faceVals[f] * wgt[:, None]
        return acc

    def mesh_leaves(self):
        """The step's geometry inputs as autograd leaves, in the order and
        shapes of the reference's Mesh.gradFields (mesh.py:28-32), the
        parameter 'mesh' of apps/adjoint.py:107-110: areas (nF,1),
        volumesL (nF,1), volumesR (nIF,1), weights (nF,1), deltas (nF,1),
        normals (nF,3), deltasUnit (nF,3), linearWeights (nF,2),
        quadraticWeights (nF,2,3), volumes (nIC,1).  The case then computes
        with views of these leaves, each used where the reference uses its
        array (op.py, interp.py, density.py, BCs.py)."""
        shapes = {'areas': (self.nF, 1), 'volumesL': (self.nF, 1), 'volumesR': (self.nIF, 1),
                  'weights': (self.nF, 1), 'deltas': (self.nF, 1), 'normals': (self.nF, 3),
                  'deltasUnit': (self.nF, 3), 'linearWeights': (self.nF, 2),
                  'quadraticWeights': (self.nF, 2, 3), 'volumes': (self.nIC, 1)}
        leaves = []
        for k in MESH_GRAD_FIELDS:
            x = self.t[k].detach().clone().reshape(shapes[k]).requires_grad_(True)
            leaves.append(x)
            self.t[k] = x.reshape(self.t[k].shape)
        return leaves

    def equation(self, rho, rhoU, rhoE, ghostInit=None, with_fields=False):
        """RCF.equation (density.py:339-414): (drho, drhoU, drhoE, dtc_max)."""
        U, T, p = self.primitive(rho, rhoU, rhoE)
        U, T, p = self.ghosts(U, T, p, ghostInit)
        gU = self.grad_cell(U)
        gT = self.grad_cell(T)
        gp = self.grad_cell(p)
        # gradient ghost rows: copies of owner / paired-owner rows (grad_src)
        Fr, FrU, FrE, lam = self.all_face_fluxes(U, T, p, gU, gT, gp)
        drho = self.accumulate(Fr)
        drhoU = self.accumulate(FrU)
        drhoE = self.accumulate(FrE)
        dtc = self.accumulate(lam, absolute=True)
        out = (drho, drhoU, drhoE, dtc)
        if with_fields:
            return out, (U, T, p, gU, gT, gp)
        return out

    # ---------------------------------------------------------- objectives --
    def objective_value(self, rho, rhoU, rhoE, U, T, p):
        if self.objective is None:
            return torch.zeros((), dtype=rho.dtype, device=rho.device)
        kind = self.objective[0]
        if kind == 'kinetic_energy':
            Uc = rhoU / rho
            return (0.5 * rho[:, 0] * (Uc * Uc).sum(-1) * self.t['volumes']).sum()
        if kind == 'drag':
            pt = [q for q in self.patches if q['name'] == self.objective[1]][0]
            s, n = pt['start'], pt['n']
            faces = slice(s, s + n)
            N, P = self.neighbour[faces], self.owner[faces]
            U0, U0i = U[N][:, 0:1], U[P][:, 0:1]
            p0, T0 = p[N], T[N]
            nx = self.t['normals'][faces][:, 0:1]
            mungUx = self.mu(T0) * (U0 - U0i) / self.t['deltas'][faces][:, None]
            return ((p0 * nx - mungUx) * self.t['areas'][faces][:, None]).sum()
        raise NotImplementedError(kind)

    def symmetriser(self, rho, rhoU, rhoE):
        """The 5x5 matrix A0 per cell of computeEnergy, cmesh.cpp:384-418
        (nIC,5,5), in conservative order rho, rhoU, rhoE: rows
        [rho, rho u, rhoE], [rho u, rho u u^T + p I, rho H u],
        [rhoE, rho H u^T, rho H^2 - gamma p^2/(rho (gamma-1))]."""
        g = self.gamma
        r = rho.reshape(-1)
        u = rhoU.reshape(-1, 3) / r[:, None]
        E = rhoE.reshape(-1)
        q2 = (u * u).sum(-1)
        p = (E - r * q2 / 2) * (g - 1)
        H = g * p / (r * (g - 1)) + q2 / 2
        n = r.shape[0]
        A = torch.zeros(n, 5, 5, dtype=r.dtype)
        A[:, 0, 0] = r
        A[:, 0, 1:4] = A[:, 1:4, 0] = r[:, None] * u
        A[:, 0, 4] = A[:, 4, 0] = E
        A[:, 1:4, 1:4] = r[:, None, None] * u[:, :, None] * u[:, None, :] + p[:, None, None] * torch.eye(3, dtype=r.dtype)
        A[:, 1:4, 4] = A[:, 4, 1:4] = (r * H)[:, None] * u
        A[:, 4, 4] = r * H * H - g * p * p / (r * (g - 1))
        return A

    def adjoint_energy(self, rhoa, rhoUa, rhoEa, rho, rhoU, rhoE):
        """computeEnergy, cmesh.cpp:342-425: sum_c w^T A0 w V_c with
        w = (rhoa, rhoUa, rhoEa)/V_c, the adjoint fields kept times the
        cell volume (apps/adjoint.py:185-189)."""
        V = self.t['volumes']
        w = torch.cat([rhoa.reshape(-1, 1), rhoUa.reshape(-1, 3), rhoEa.reshape(-1, 1)], 1) / V[:, None]
        A = self.symmetriser(rho, rhoU, rhoE)
        return ((w[:, :, None] * A * w[:, None, :]).sum((1, 2)) * V).sum()

    # --------------------------------------------------------------- step --
    def coefficients(self):
        if self.integrator == 'SSPRK':                          # timestep.py:17-25
            alpha = [[1, 0, 0], [3. / 4, 1. / 4, 0], [1. / 3, 0, 2. / 3]]
            beta = [1, 1. / 4, 2. / 3]
        else:                                                   # timestep.py:9-13
            alpha, beta = [[1.]], [1.]
        return alpha, beta

    def step(self, rho, rhoU, rhoE, dt, source=None, ghostInit=None, obj_on_start=False):
        """One time step (timestep.py:27-55): returns (rho', rhoU', rhoE',
        dtc, obj) with dtc and obj taken on the stage-1 state
        (density.py:412-413).  obj_on_start: the objective on the step's
        start state instead (the library's ADFVM_OBJSTATE_START; no
        counterpart in the reference)."""
        alpha, beta = self.coefficients()
        fields = [[rho, rhoU, rhoE]]
        dtc = obj = None
        dtc_stage = 1 if len(beta) > 1 else 0
        obj_stage = 0 if obj_on_start else dtc_stage
        for i in range(len(beta)):
            (d0, d1, d2, dtcc) = self.equation(*fields[i], ghostInit=ghostInit)
            if i == dtc_stage:
                dtc = dtcc.max()
            if i == obj_stage:
                U, T, p = self.primitive(*fields[i])
                U, T, p = self.ghosts(U, T, p, ghostInit)
                obj = self.objective_value(*fields[i], U, T, p)
            L = [d0, d1, d2]
            S = source if source is not None else [0., 0., 0.]
            new = []
            for k in range(3):
                acc = 0
                for j in range(i + 1):
                    if alpha[i][j] != 0:
                        acc = acc + alpha[i][j] * fields[j][k]
                acc = acc + -beta[i] * (L[k] - S[k]) * dt
                new.append(acc)
            fields.append(new)
        return fields[-1][0], fields[-1][1], fields[-1][2], dtc, obj


def to_state(rho, rhoU, rhoE):
    """SoA state u[5][n] (numpy) from reference-layout arrays."""
    return np.concatenate([np.asarray(rho).reshape(1, -1), np.asarray(rhoU).T, np.asarray(rhoE).reshape(1, -1)])


def from_state(u):
    u = np.asarray(u)
    return u[0:1].T.copy(), u[1:4].T.copy(), u[4:5].T.copy()
