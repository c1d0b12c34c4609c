// Test shim: exposes the per-face arithmetic of adfvm_amd/csrc/fvm_math.hpp
// and adjoint_math.hpp to Python through C, compiled for the host with g++.
// Used only by tests to check the hand-written reverse code on the CPU.
#include "adjoint_math.hpp"

using namespace adfvm;

static Phys<double> make(const double* p) {
    // p: gamma, Cp, Pr, mu0, muA, muTs, small, muModel, riemann, boundaryRiemann, limit
    Phys<double> ph;
    ph.gamma = p[0]; ph.Cp = p[1]; ph.Cv = p[1] / p[0]; ph.Pr = p[2]; ph.invCv = 1.0 / ph.Cv;
    ph.mu0 = p[3]; ph.muA = p[4]; ph.muTs = p[5]; ph.small = p[6];
    ph.muModel = (int)p[7]; ph.riemann = (int)p[8]; ph.boundaryRiemann = (int)p[9]; ph.limit = (int)p[10];
    return ph;
}

static FaceGeo<double> geo(const double* g) {
    // g: A, n[3], w, delta, D[3], lw0, lw1, qw0[3], qw1[3]
    FaceGeo<double> f;
    f.A = g[0]; for (int i = 0; i < 3; i++) f.n[i] = g[1 + i];
    f.w = g[4]; f.delta = g[5]; for (int i = 0; i < 3; i++) f.D[i] = g[6 + i];
    f.lw0 = g[9]; f.lw1 = g[10];
    for (int i = 0; i < 3; i++) { f.qw0[i] = g[11 + i]; f.qw1[i] = g[14 + i]; }
    return f;
}

extern "C" void shim_face_flux(const double* p, int kind, const double* wP, const double* gP, const double* wN,
                               const double* gN, const double* g, double* F, double* lam) {
    face_flux(make(p), kind, wP, gP, wN, gN, geo(g), F, lam);
}

extern "C" void shim_face_flux_vjp(const double* p, int kind, const double* wP, const double* gP, const double* wN,
                                   const double* gN, const double* g, const double* Fb, double* wPb, double* gPb,
                                   double* wNb, double* gNb) {
    face_flux_vjp(make(p), kind, wP, gP, wN, gN, geo(g), Fb, wPb, gPb, wNb, gNb);
}

extern "C" void shim_primitive(const double* p, const double* q, double* w) { primitive(make(p), q, w); }
extern "C" void shim_primitive_vjp(const double* p, const double* q, const double* wb, double* qb) {
    primitive_vjp(make(p), q, wb, qb);
}
