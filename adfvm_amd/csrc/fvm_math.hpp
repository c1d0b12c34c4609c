// Per-face and per-cell arithmetic of adFVM's compressible FV step, written
// once for device kernels (hipcc) and for a host test shim (g++).
//
// Layouts:  primitive w[5] = {Ux, Uy, Uz, T, p};
//           conservative q[5] = {rho, rhoUx, rhoUy, rhoUz, rhoE};
//           gradients g[15] = dU_i/dx_k at 3*i+k (9), dT/dx_k at 9+k, dp/dx_k at 12+k.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ADFVM_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define ADFVM_HD inline
#endif

namespace adfvm {

template <typename T> ADFVM_HD T fsqrt(T x) { return sqrt(x); }
template <typename T> ADFVM_HD T fabs_(T x) { return x < T(0) ? -x : x; }
template <typename T> ADFVM_HD T fmax_(T a, T b) { return a > b ? a : b; }

template <typename T>
struct Phys {
    T gamma, Cp, Cv, Pr, invCv, small;
    T mu0, muA, muTs;
    int muModel;           // 0 constant, 1 Sutherland
    int riemann;           // 0 Roe, 1 Lax-Friedrichs (internal + coupled faces)
    int boundaryRiemann;   // characteristic faces
    int limit;             // 1: clamp reconstructed values between the two cell values
};

// viscosity: constant, or Sutherland muA*T^1.5/(T+Ts) (adFVM/density.py:27)
template <typename T> ADFVM_HD T viscosity(const Phys<T>& ph, T Tf) {
    if (ph.muModel == 0) return ph.mu0;
    return ph.muA * (Tf * fsqrt(Tf)) / (Tf + ph.muTs);
}

// RCF.primitive, adFVM/density.py:162-171
template <typename T> ADFVM_HD void primitive(const Phys<T>& ph, const T* q, T* w) {
    T rho = q[0];
    T Ux = q[1] / rho, Uy = q[2] / rho, Uz = q[3] / rho;
    T E = q[4] / rho;
    T e = E - T(0.5) * (Ux * Ux + Uy * Uy + Uz * Uz);
    w[0] = Ux; w[1] = Uy; w[2] = Uz;
    w[4] = (ph.gamma - T(1)) * rho * e;
    w[3] = e * ph.invCv;
}

// RCF.conservative, adFVM/density.py:173-181
template <typename T> ADFVM_HD void conservative(const Phys<T>& ph, const T* w, T* q) {
    T e = ph.Cv * w[3];
    T rho = w[4] / (e * (ph.gamma - T(1)));
    T U2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    q[0] = rho;
    q[4] = rho * (e + T(0.5) * U2);
    q[1] = w[0] * rho; q[2] = w[1] * rho; q[3] = w[2] * rho;
}

// interp.secondOrder for one side, adFVM/interp.py:20-28:
//   phi_C + (phi_D - phi_C)*lw + qw . grad(phi_C)
// With limit set (an option the reference does not have; its secondOrder is
// unlimited), each value is clamped to [min(phi_C,phi_D), max(phi_C,phi_D)].
template <typename T>
ADFVM_HD void reconstruct(const T* wC, const T* gC, const T* wD, T lw, const T* qw, T* out, int limit) {
    for (int i = 0; i < 5; i++) {
        const T* g = gC + 3 * i;    // rows 0..2 U, 3 T (9..11), 4 p (12..14)
        T v = wC[i] + (wD[i] - wC[i]) * lw + (qw[0] * g[0] + qw[1] * g[1] + qw[2] * g[2]);
        if (limit) {
            T lo = wC[i] < wD[i] ? wC[i] : wD[i];
            T hi = wC[i] < wD[i] ? wD[i] : wC[i];
            v = v < lo ? lo : (v > hi ? hi : v);
        }
        out[i] = v;
    }
}

// riemann.eulerRoe, adFVM/riemann.py:21-70.  stabilise(eps) = eps + SMALL
// for eps >= 0 (adpy's definition is not in the reference).
template <typename T>
ADFVM_HD void roe_flux(const Phys<T>& ph, const T* wL, const T* qL, const T* wR, const T* qR,
                       const T* n, T* F) {
    const T g = ph.gamma, gm1 = ph.gamma - T(1);
    const T *UL = wL, *UR = wR;
    T pL = wL[4], pR = wR[4], rhoL = qL[0], rhoR = qR[0];
    T UnL = UL[0] * n[0] + UL[1] * n[1] + UL[2] * n[2];
    T UnR = UR[0] * n[0] + UR[1] * n[1] + UR[2] * n[2];
    T rhoUnL = rhoL * UnL, rhoUnR = rhoR * UnR;
    T hL = g * pL / (gm1 * rhoL) + T(0.5) * (UL[0] * UL[0] + UL[1] * UL[1] + UL[2] * UL[2]);
    T hR = g * pR / (gm1 * rhoR) + T(0.5) * (UR[0] * UR[0] + UR[1] * UR[1] + UR[2] * UR[2]);
    T Fr = T(0.5) * (rhoUnL + rhoUnR);
    T FU[3];
    for (int i = 0; i < 3; i++) FU[i] = T(0.5) * (rhoUnL * UL[i] + rhoUnR * UR[i] + (pL + pR) * n[i]);
    T FE = T(0.5) * (rhoUnL * hL + rhoUnR * hR);

    T sL = fsqrt(rhoL), sR = fsqrt(rhoR);
    T dv = sL + sR;
    T UF[3];
    for (int i = 0; i < 3; i++) UF[i] = (UL[i] * sL + UR[i] * sR) / dv;
    T hF = (hL * sL + hR * sR) / dv;
    T qF = T(0.5) * (UF[0] * UF[0] + UF[1] * UF[1] + UF[2] * UF[2]);
    T a2F = gm1 * (hF - qF);
    T aF = fsqrt(a2F);
    T UnF = UF[0] * n[0] + UF[1] * n[1] + UF[2] * n[2];

    T drho = rhoR - rhoL;
    T drhoU[3];
    for (int i = 0; i < 3; i++) drhoU[i] = rhoR * UR[i] - rhoL * UL[i];
    T drhoE = (hR * rhoR - pR) - (hL * rhoL - pL);

    T lam1 = fabs_(UnF), lam2 = fabs_(UnF + aF), lam3 = fabs_(UnF - aF);
    T eps = T(0.5) * fabs_(rhoUnL / rhoL - rhoUnR / rhoR);
    eps += T(0.5) * fabs_(fsqrt(g * pL / rhoL) - fsqrt(g * pR / rhoR));
    eps = eps < T(0) ? eps - ph.small : eps + ph.small;
    T e2 = T(2) * eps;
    if (lam1 < e2) lam1 = T(0.25) * lam1 * lam1 / eps + eps;
    if (lam2 < e2) lam2 = T(0.25) * lam2 * lam2 / eps + eps;
    if (lam3 < e2) lam3 = T(0.25) * lam3 * lam3 / eps + eps;

    T abv1 = T(0.5) * (lam2 + lam3);
    T abv2 = T(0.5) * (lam2 - lam3);
    T abv3 = abv1 - lam1;
    T abv4 = gm1 * (qF * drho - (UF[0] * drhoU[0] + UF[1] * drhoU[1] + UF[2] * drhoU[2]) + drhoE);
    T abv5 = UnF * drho - (drhoU[0] * n[0] + drhoU[1] * n[1] + drhoU[2] * n[2]);
    T abv6 = abv3 * abv4 / a2F - abv2 * abv5 / aF;
    T abv7 = abv3 * abv5 - abv2 * abv4 / aF;

    F[0] = Fr - T(0.5) * (lam1 * drho + abv6);
    for (int i = 0; i < 3; i++) F[1 + i] = FU[i] - T(0.5) * (lam1 * drhoU[i] + UF[i] * abv6 - abv7 * n[i]);
    F[4] = FE - T(0.5) * (lam1 * drhoE + hF * abv6 - UnF * abv7);
}

// riemann.eulerLaxFriedrichs, adFVM/riemann.py:6-18
template <typename T>
ADFVM_HD void lf_flux(const Phys<T>& ph, const T* wL, const T* qL, const T* wR, const T* qR,
                      const T* n, T* F) {
    const T g = ph.gamma;
    T pL = wL[4], pR = wR[4], rhoL = qL[0], rhoR = qR[0];
    T UnL = wL[0] * n[0] + wL[1] * n[1] + wL[2] * n[2];
    T UnR = wR[0] * n[0] + wR[1] * n[1] + wR[2] * n[2];
    T cL = fsqrt(g * pL / rhoL), cR = fsqrt(g * pR / rhoR);
    T aL = fabs_(UnL) + cL, aR = fabs_(UnR) + cR;
    T aF = fmax_(aL, aR);
    F[0] = T(0.5) * (rhoL * UnL + rhoR * UnR) - T(0.5) * aF * (rhoR - rhoL);
    for (int i = 0; i < 3; i++)
        F[1 + i] = T(0.5) * (qL[1 + i] * UnL + qR[1 + i] * UnR + (pL + pR) * n[i]) - T(0.5) * aF * (qR[1 + i] - qL[1 + i]);
    F[4] = T(0.5) * ((qL[4] + pL) * UnL + (qR[4] + pR) * UnR) - T(0.5) * aF * (qR[4] - qL[4]);
}

// RCF.viscousFlux, adFVM/density.py:192-222 (snGrad, op.py:66-67), added to F.
//   TL,UL / TR,UR: cell values either side; TF, UF: face values;
//   gTF[3], gUF[9]: face gradients before the snGrad correction.
template <typename T>
ADFVM_HD void viscous_flux(const Phys<T>& ph, T TL, T TR, const T* UL, const T* UR, T TF,
                           const T* UF, const T* gTF0, const T* gUF0, const T* n, T delta,
                           const T* D, T* F) {
    T mu = viscosity(ph, TF);
    T kappa = mu * (ph.Cp / ph.Pr);
    T snT = (TR - TL) / delta;
    T gTD = gTF0[0] * D[0] + gTF0[1] * D[1] + gTF0[2] * D[2];
    T gT[3];
    for (int k = 0; k < 3; k++) gT[k] = gTF0[k] + snT * D[k] - gTD * D[k];
    T gU[9];
    for (int i = 0; i < 3; i++) {
        T snU = (UR[i] - UL[i]) / delta;
        T gD = gUF0[3 * i] * D[0] + gUF0[3 * i + 1] * D[1] + gUF0[3 * i + 2] * D[2];
        for (int k = 0; k < 3; k++) gU[3 * i + k] = gUF0[3 * i + k] + snU * D[k] - gD * D[k];
    }
    T qF = kappa * (gT[0] * n[0] + gT[1] * n[1] + gT[2] * n[2]);
    T tmp3 = gU[0] + gU[4] + gU[8];
    T sigma[3];
    for (int i = 0; i < 3; i++) {
        T t2 = (gU[3 * i] + gU[i]) * n[0] + (gU[3 * i + 1] + gU[3 + i]) * n[1] + (gU[3 * i + 2] + gU[6 + i]) * n[2];
        sigma[i] = mu * (t2 - T(2.0 / 3.0) * tmp3 * n[i]);
    }
    F[1] -= sigma[0]; F[2] -= sigma[1]; F[3] -= sigma[2];
    F[4] -= qF + (sigma[0] * UF[0] + sigma[1] * UF[1] + sigma[2] * UF[2]);
}

// face kinds
enum { FACE_INTERIOR = 0, FACE_CHARACTERISTIC = 1, FACE_BOUNDARY = 2 };

// Face geometry for one face (gathered from the SoA tables).
template <typename T> struct FaceGeo {
    T A, n[3], w, delta, D[3], lw0, lw1, qw0[3], qw1[3];
};

// One face's flux F[5] (not yet multiplied by the area) and the spectral
// radius term lam = |UF.n| + sqrt(Cp TF (gamma-1)):
//   FACE_INTERIOR        RCF.flux, internal or coupled face (density.py:253-305)
//   FACE_CHARACTERISTIC  RCF.flux with characteristic=True (density.py:266-274)
//   FACE_BOUNDARY        RCF.boundaryFlux (density.py:307-331)
// P = owner side, N = neighbour (or ghost) side; for ghosts gN is the owner's
// gradient (default gradient boundary conditions, density.py:376-385).
template <typename T>
ADFVM_HD void face_flux(const Phys<T>& ph, int kind, const T* wP, const T* gP, const T* wN,
                        const T* gN, const FaceGeo<T>& f, T* F, T* lam) {
    T TF, UF[3], gTF[3], gUF[9];
    if (kind == FACE_BOUNDARY) {
        T qR[5];
        conservative(ph, wN, qR);
        T Un = wN[0] * f.n[0] + wN[1] * f.n[1] + wN[2] * f.n[2];
        F[0] = qR[0] * Un;
        for (int i = 0; i < 3; i++) F[1 + i] = qR[1 + i] * Un + wN[4] * f.n[i];
        F[4] = (qR[4] + wN[4]) * Un;
        TF = wN[3];
        for (int i = 0; i < 3; i++) UF[i] = wN[i];
        for (int k = 0; k < 3; k++) gTF[k] = gN[9 + k];
        for (int k = 0; k < 9; k++) gUF[k] = gN[k];
    } else {
        T wLF[5], qLF[5], wRF[5], qRF[5];
        reconstruct(wP, gP, wN, f.lw0, f.qw0, wLF, ph.limit);
        conservative(ph, wLF, qLF);
        if (kind == FACE_CHARACTERISTIC) {
            for (int i = 0; i < 5; i++) wRF[i] = wN[i];
            conservative(ph, wRF, qRF);
            if (ph.boundaryRiemann == 0) roe_flux(ph, wLF, qLF, wRF, qRF, f.n, F);
            else lf_flux(ph, wLF, qLF, wRF, qRF, f.n, F);
            TF = wN[3];
            for (int i = 0; i < 3; i++) UF[i] = wN[i];
            for (int k = 0; k < 3; k++) gTF[k] = gN[9 + k];
            for (int k = 0; k < 9; k++) gUF[k] = gN[k];
        } else {
            reconstruct(wN, gN, wP, f.lw1, f.qw1, wRF, ph.limit);
            conservative(ph, wRF, qRF);
            if (ph.riemann == 0) roe_flux(ph, wLF, qLF, wRF, qRF, f.n, F);
            else lf_flux(ph, wLF, qLF, wRF, qRF, f.n, F);
            for (int i = 0; i < 3; i++) UF[i] = T(0.5) * (wLF[i] + wRF[i]);
            TF = T(0.5) * (wLF[3] + wRF[3]);
            for (int k = 0; k < 3; k++) gTF[k] = T(0.5) * (gP[9 + k] + gN[9 + k]);
            for (int k = 0; k < 9; k++) gUF[k] = T(0.5) * (gP[k] + gN[k]);
        }
    }
    viscous_flux(ph, wP[3], wN[3], wP, wN, TF, UF, gTF, gUF, f.n, f.delta, f.D, F);
    T aF = fsqrt(ph.Cp * TF * (ph.gamma - T(1)));
    *lam = fabs_(UF[0] * f.n[0] + UF[1] * f.n[1] + UF[2] * f.n[2]) + aF;
}

}  // namespace adfvm
