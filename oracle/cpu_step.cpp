// ORACLE-SIDE CPU BASELINE (test and measurement infrastructure only): a
// C++/OpenMP FP64 implementation of one explicit SSP-RK3 step of adFVM's
// compressible solver and of its discrete adjoint, on periodic (all-cyclic)
// hex meshes, for timing the CPU path beside the GPU in bench.py's
// cpu_baseline leg and for nothing else.  Only tests/ and that bench leg may
// load it; the product never does.
//
// It stands in for the reference's CPU path (adpy-generated C++ with OpenMP,
// adFVM/config.py:20,68,73-117), which cannot be built here (adpy is an
// empty submodule).  The step follows RCF.equation (adFVM/density.py:339-414)
// and timestep.SSPRK (timestep.py:17-55): Gauss cell gradients (op.py:45-63),
// second-order face values (interp.py:20-28), Roe or Lax-Friedrichs fluxes
// (riemann.py:6-70) with the viscous flux (density.py:192-222), the
// divergence over each cell's cellFaces slots (op.py:12-29) and the stage
// update; cyclic faces are computed from their owner's side only
// (density.py:395-396).  The kinetic-energy objective is taken on the
// stage-1 state (density.py:412-413).  The adjoint step recomputes the
// stages from u_n (the reference's reverse sweep re-runs the primal,
// apps/adjoint.py:211-217) and reverses them.  The per-face arithmetic is
// the host build of the same header functions the kernels use
// (adfvm_amd/csrc/fvm_math.hpp, adjoint_math.hpp); the CPU code around them
// (loops, layout, reductions) is this file's own.
//
// Layout: state u[5][nIC] (rows rho, rhoUx, rhoUy, rhoUz, rhoE), like the
// library's.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "adjoint_math.hpp"

using namespace adfvm;

namespace {

struct Stage { double a0, ai, beta; };
const Stage kStages[3] = {{0.0, 1.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};

struct CpuCase {
    long nIC = 0, nIF = 0, nF = 0;
    Phys<double> ph;
    std::vector<FaceGeo<double>> geo;     // per face
    std::vector<int> owner, other;         // per face: owner; neighbour or the paired face's owner
    std::vector<int> slotFace, slotOther;  // [nIC][6]
    std::vector<unsigned char> slotOwn, slotCyc;
    std::vector<int> slotPair;             // [nIC][6]: paired face of a cyclic slot
    std::vector<double> slotSN, slotW;     // [nIC][6][3] outward S n, [nIC][6] other side's weight
    std::vector<double> vol;
    // work
    std::vector<double> U[3], G[3], fb, recP, recN, cadj, acc0, tmp;
};

inline void prim_of(const CpuCase& c, const double* u, long i, double* w) {
    double q[5];
    for (int k = 0; k < 5; k++) q[k] = u[k * c.nIC + i];
    primitive(c.ph, q, w);
}

// Gauss gradients: grad[c*15 + ...] (U 9, T 3, p 3)
void gradients(const CpuCase& c, const double* u, double* grad) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < c.nIC; i++) {
        double wc[5], g[15] = {0};
        prim_of(c, u, i, wc);
        for (int j = 0; j < 6; j++) {
            const long s = i * 6 + j;
            double wo[5];
            prim_of(c, u, c.slotOther[s], wo);
            const double w = c.slotW[s];
            const double* SN = &c.slotSN[3 * s];
            for (int k = 0; k < 5; k++) {
                const double phiF = wc[k] * (1.0 - w) + wo[k] * w;
                for (int d = 0; d < 3; d++) g[3 * k + d] += phiF * SN[d];
            }
        }
        for (int k = 0; k < 15; k++) grad[i * 15 + k] = g[k] / c.vol[i];
    }
}

// face fluxes times area: fb[f*5 + k]
void fluxes(CpuCase& c, const double* u, const double* grad) {
#pragma omp parallel for schedule(static)
    for (long f = 0; f < c.nF; f++) {
        const int P = c.owner[f], N = c.other[f];
        double wP[5], wN[5], F[5], lam;
        prim_of(c, u, P, wP);
        prim_of(c, u, N, wN);
        face_flux(c.ph, FACE_INTERIOR, wP, &grad[15L * P], wN, &grad[15L * N], c.geo[f], F, &lam);
        for (int k = 0; k < 5; k++) c.fb[5 * f + k] = F[k] * c.geo[f].A;
    }
}

// stage update out = a0*u0 + ai*ui - beta*L*dt, L = (1/V) sum_slots +-F A
void stage_update(const CpuCase& c, const double* ui, const double* u0, const Stage& st, double dt, double* out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < c.nIC; i++) {
        double L[5] = {0};
        for (int j = 0; j < 6; j++) {
            const long s = i * 6 + j;
            const double* F = &c.fb[5L * c.slotFace[s]];
            for (int k = 0; k < 5; k++) L[k] += c.slotOwn[s] ? F[k] : -F[k];
        }
        for (int k = 0; k < 5; k++) {
            double a = st.ai * ui[k * c.nIC + i];
            if (st.a0 != 0.0) a += st.a0 * u0[k * c.nIC + i];
            out[k * c.nIC + i] = a - st.beta * (L[k] / c.vol[i]) * dt;
        }
    }
}

double kinetic_energy(const CpuCase& c, const double* u) {
    double e = 0;
#pragma omp parallel for schedule(static) reduction(+ : e)
    for (long i = 0; i < c.nIC; i++) {
        const double r = u[i], a = u[c.nIC + i] / r, b = u[2 * c.nIC + i] / r, d = u[3 * c.nIC + i] / r;
        e += 0.5 * r * (a * a + b * b + d * d) * c.vol[i];
    }
    return e;
}

// the three stage states (U[0] = u_n) and gradients of the step from u_n
void forward_stages(CpuCase& c, const double* un, double dt, double* uout, double* obj) {
    std::memcpy(c.U[0].data(), un, sizeof(double) * 5 * c.nIC);
    for (int i = 0; i < 3; i++) {
        if (i == 1 && obj) *obj = kinetic_energy(c, c.U[1].data());
        gradients(c, c.U[i].data(), c.G[i].data());
        fluxes(c, c.U[i].data(), c.G[i].data());
        double* out = i < 2 ? c.U[i + 1].data() : uout;
        stage_update(c, c.U[i].data(), c.U[0].data(), kStages[i], dt, out);
    }
}

// reverse of stage i: ua holds the adjoint of the stage's output; on return
// the adjoint of its input state U[i] (u0's share accumulated into acc0).
void reverse_stage(CpuCase& c, int i, double dt, double* ua, double obja, bool first) {
    const Stage st = kStages[i];
    const double* u = c.U[i].data();
    const double* g = c.G[i].data();
    const double rscale = -st.beta * dt;
    const long nIC = c.nIC;
    // face VJPs, one record of 20 (gradient 15, primitive 5) per face side
#pragma omp parallel for schedule(static)
    for (long f = 0; f < c.nF; f++) {
        const int P = c.owner[f], N = c.other[f];
        double wP[5], wN[5], Fb[5];
        prim_of(c, u, P, wP);
        prim_of(c, u, N, wN);
        const double A = c.geo[f].A;
        for (int k = 0; k < 5; k++) {
            Fb[k] = rscale * ua[k * nIC + P] * (A / c.vol[P]);
            if (f < c.nIF) Fb[k] += rscale * ua[k * nIC + N] * (-(A / c.vol[N]));
        }
        double* sP = &c.recP[20 * f];
        double* sN = &c.recN[20 * f];
        for (int k = 0; k < 20; k++) { sP[k] = 0; sN[k] = 0; }
        face_flux_vjp(c.ph, FACE_INTERIOR, wP, &g[15L * P], wN, &g[15L * N], c.geo[f], Fb, sP + 15, sP, sN + 15,
                      sN);
    }
    // each cell's own side of its faces, and on a cyclic slot the
    // neighbour side of the paired face
#pragma omp parallel for schedule(static)
    for (long i2 = 0; i2 < nIC; i2++) {
        double a[20] = {0};
        for (int j = 0; j < 6; j++) {
            const long s = i2 * 6 + j;
            const double* r = c.slotOwn[s] ? &c.recP[20L * c.slotFace[s]] : &c.recN[20L * c.slotFace[s]];
            for (int k = 0; k < 20; k++) a[k] += r[k];
            if (c.slotCyc[s]) {
                const double* r2 = &c.recN[20L * c.slotPair[s]];
                for (int k = 0; k < 20; k++) a[k] += r2[k];
            }
        }
        for (int k = 0; k < 20; k++) c.cadj[20 * i2 + k] = a[k];
    }
    // transpose of the Gauss gradient, primitive map, stage combination
    // (into a separate array: the loop reads neighbours' ua)
    std::vector<double>& tmp = c.tmp;
#pragma omp parallel for schedule(static)
    for (long i2 = 0; i2 < nIC; i2++) {
        double q[5], wc[5], wb[5];
        for (int k = 0; k < 5; k++) q[k] = u[k * nIC + i2];
        primitive(c.ph, q, wc);
        const double* own = &c.cadj[20 * i2];
        for (int k = 0; k < 5; k++) wb[k] = own[15 + k];
        const double Vc = c.vol[i2];
        for (int j = 0; j < 6; j++) {
            const long s = i2 * 6 + j;
            const double* SN = &c.slotSN[3 * s];
            const double w = c.slotW[s];
            double d[5];
            for (int k = 0; k < 5; k++) d[k] = (SN[0] * own[3 * k] + SN[1] * own[3 * k + 1] + SN[2] * own[3 * k + 2]) / Vc;
            for (int k = 0; k < 5; k++) wb[k] += (1.0 - w) * d[k];
            // the other side's gradient through this face: its outward S n
            // and its weight of this cell's value
            const int o = c.slotOther[s];
            double K[3];
            if (!c.slotCyc[s]) {
                const double kk = -(1.0 - w) / c.vol[o];
                for (int dd = 0; dd < 3; dd++) K[dd] = kk * SN[dd];
            } else {
                const long gf = c.slotPair[s];
                const double coef = 1.0 - c.geo[gf].w;
                for (int dd = 0; dd < 3; dd++) K[dd] = coef * (c.geo[gf].A * c.geo[gf].n[dd]) / c.vol[o];
            }
            const double* go = &c.cadj[20L * o];
            for (int k = 0; k < 5; k++) wb[k] += K[0] * go[3 * k] + K[1] * go[3 * k + 1] + K[2] * go[3 * k + 2];
        }
        double qb[5] = {0};
        primitive_vjp(c.ph, q, wb, qb);
        if (i == 1 && obja != 0.0) {
            const double r = q[0], U0 = q[1] / r, U1 = q[2] / r, U2 = q[3] / r;
            qb[0] += obja * (-0.5 * (U0 * U0 + U1 * U1 + U2 * U2) * Vc);
            qb[1] += obja * U0 * Vc; qb[2] += obja * U1 * Vc; qb[3] += obja * U2 * Vc;
        }
        for (int k = 0; k < 5; k++) {
            const long x = k * nIC + i2;
            const double un = ua[x];
            if (i > 0) c.acc0[x] = (first ? 0.0 : c.acc0[x]) + st.a0 * un;
            double r = st.ai * un + qb[k];
            if (i == 0) r += c.acc0[x];
            tmp[x] = r;
        }
    }
    std::memcpy(ua, tmp.data(), sizeof(double) * 5 * nIC);
}

}  // namespace

extern "C" {

// Periodic hex mesh with the reference's arrays (adfvm_mesh_build output):
// owner, neighbour (nF, ghost-extended), cellFaces / cellOwner (nIC x 6),
// pair (nF: the paired face of a cyclic boundary face, -1 for internal
// faces; every boundary face must be cyclic), FP64 geometry.
void* adfvm_cpu_create(long nIC, long nIF, long nF, const int* owner, const int* neighbour, const int* cellFaces,
                       const int* cellOwner, const int* pair, const double* areas, const double* normals,
                       const double* weights, const double* deltas, const double* deltasUnit,
                       const double* linearWeights, const double* quadraticWeights, const double* volumes,
                       double gamma, double Cp, double Pr, double mu0, double small, int riemann) {
    CpuCase* c = new CpuCase();
    c->nIC = nIC; c->nIF = nIF; c->nF = nF;
    Phys<double>& ph = c->ph;
    ph.gamma = gamma; ph.Cp = Cp; ph.Cv = Cp / gamma; ph.Pr = Pr; ph.invCv = 1.0 / ph.Cv; ph.small = small;
    ph.mu0 = mu0; ph.muA = 0; ph.muTs = 0; ph.muModel = 0; ph.riemann = riemann; ph.boundaryRiemann = riemann;
    ph.limit = 0;
    c->geo.resize(nF);
    c->owner.assign(owner, owner + nF);
    c->other.resize(nF);
    for (long f = 0; f < nF; f++) {
        FaceGeo<double>& g = c->geo[f];
        g.A = areas[f];
        for (int k = 0; k < 3; k++) { g.n[k] = normals[3 * f + k]; g.D[k] = deltasUnit[3 * f + k]; }
        g.w = weights[f]; g.delta = deltas[f];
        g.lw0 = linearWeights[2 * f]; g.lw1 = linearWeights[2 * f + 1];
        for (int k = 0; k < 3; k++) { g.qw0[k] = quadraticWeights[6 * f + k]; g.qw1[k] = quadraticWeights[6 * f + 3 + k]; }
        if (f < nIF) c->other[f] = neighbour[f];
        else if (pair[f] >= 0) c->other[f] = owner[pair[f]];
        else { delete c; return nullptr; }
    }
    c->slotFace.resize(6 * nIC); c->slotOther.resize(6 * nIC); c->slotOwn.resize(6 * nIC);
    c->slotCyc.resize(6 * nIC); c->slotPair.resize(6 * nIC); c->slotSN.resize(18 * nIC); c->slotW.resize(6 * nIC);
    for (long i = 0; i < nIC; i++)
        for (int j = 0; j < 6; j++) {
            const long s = i * 6 + j;
            const int f = cellFaces[s];
            const bool own = cellOwner[s] != 0;
            c->slotFace[s] = f;
            c->slotOwn[s] = own;
            c->slotCyc[s] = f >= nIF;
            c->slotPair[s] = f >= nIF ? pair[f] : -1;
            c->slotOther[s] = f >= nIF ? owner[pair[f]] : (own ? neighbour[f] : owner[f]);
            const double sg = own ? 1.0 : -1.0;
            for (int k = 0; k < 3; k++) c->slotSN[3 * s + k] = sg * (areas[f] * normals[3 * f + k]);
            c->slotW[s] = own ? 1.0 - weights[f] : weights[f];
        }
    c->vol.assign(volumes, volumes + nIC);
    for (int i = 0; i < 3; i++) { c->U[i].resize(5 * nIC); c->G[i].resize(15 * nIC); }
    c->fb.resize(5 * nF); c->recP.resize(20 * nF); c->recN.resize(20 * nF); c->cadj.resize(20 * nIC);
    c->acc0.resize(5 * nIC);
    c->tmp.resize(5 * nIC);
    return c;
}

void adfvm_cpu_destroy(void* h) { delete (CpuCase*)h; }

int adfvm_cpu_threads(void) { return omp_get_max_threads(); }

// u <- step(u); obj_out (nullable): kinetic energy on the stage-1 state
int adfvm_cpu_step(void* h, double* u, double dt, double* obj_out) {
    CpuCase* c = (CpuCase*)h;
    if (!c || !u) return 1;
    forward_stages(*c, u, dt, u, obj_out);
    return 0;
}

// ua <- adjoint of the step from u_n with the kinetic-energy objective
// weighted by obja (apps/adjoint.py:268-291 with dtca = 0)
int adfvm_cpu_adjoint_step(void* h, const double* un, double dt, double* ua, double obja) {
    CpuCase* c = (CpuCase*)h;
    if (!c || !un || !ua) return 1;
    std::vector<double> scratch(5 * c->nIC);
    forward_stages(*c, un, dt, scratch.data(), nullptr);
    for (int i = 2; i >= 0; i--) reverse_stage(*c, i, dt, ua, obja, i == 2);
    return 0;
}

}  // extern "C"
