// Reverse-mode (vector-Jacobian) arithmetic of fvm_math.hpp, derived by hand
// line by line from the forward code.  Every *_vjp function ADDS the adjoint
// of its inputs into the output arrays (so outputs may alias).
//
// Conventions match torch autograd, which the oracle uses: d|x|/dx = sign(x)
// with sign(0) = 0; a switch passes the adjoint to the branch it selected;
// max(a, b) with a == b passes half to each.
#pragma once
#include "fvm_math.hpp"

namespace adfvm {

template <typename T> ADFVM_HD T sgn(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

// d mu / dT
template <typename T> ADFVM_HD T viscosity_dT(const Phys<T>& ph, T Tf) {
    if (ph.muModel == 0) return T(0);
    T s = fsqrt(Tf);
    T d = Tf + ph.muTs;
    // mu = A*(T*sqrt(T))/(T+Ts)
    return ph.muA * (T(1.5) * s * d - Tf * s) / (d * d);
}

// primitive(q) -> w ; qb += (dw/dq)^T wb.
// The chain rule through e = E - |U|^2/2 gives for rho the terms
//   pb*(g-1)*e - eb*E/rho - sum_i Ub_i U_i/rho,  eb = Tb/Cv + pb*(g-1)*rho,
// whose p parts cancel down to pb*(g-1)|U|^2/2 (e ~ 1e3 |U|^2 at low Mach);
// they are summed in that cancelled form, which FP32 evaluates accurately:
//   rhob = (g-1)/2 |U|^2 pb + (|U|^2/2 - e) Tb/(Cv rho) - (U.wbU)/rho
template <typename T> ADFVM_HD void primitive_vjp(const Phys<T>& ph, const T* q, const T* wb, T* qb) {
    T rho = q[0];
    T U[3] = {q[1] / rho, q[2] / rho, q[3] / rho};
    T E = q[4] / rho;
    T U2 = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
    T e = E - T(0.5) * U2;
    T gm1 = ph.gamma - T(1);
    T Tb = wb[3] * ph.invCv;                 // adjoint of e through T
    T pb = wb[4];
    // eb = Tb + pb*gm1*rho ; qb4 = eb/rho
    qb[4] += Tb / rho + pb * gm1;
    for (int i = 0; i < 3; i++) qb[1 + i] += wb[i] / rho - (Tb / rho + pb * gm1) * U[i];
    T UwU = U[0] * wb[0] + U[1] * wb[1] + U[2] * wb[2];
    qb[0] += T(0.5) * gm1 * U2 * pb + (T(0.5) * U2 - e) * Tb / rho - UwU / rho;
}

// conservative(w) -> q ; wb += (dq/dw)^T qb
template <typename T> ADFVM_HD void conservative_vjp(const Phys<T>& ph, const T* w, const T* qb, T* wb) {
    T gm1 = ph.gamma - T(1);
    T e = ph.Cv * w[3];
    T rho = w[4] / (e * gm1);
    T U2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    T rhob = qb[0] + qb[4] * (e + T(0.5) * U2);
    T eb = qb[4] * rho;
    for (int i = 0; i < 3; i++) {
        rhob += qb[1 + i] * w[i];
        wb[i] += qb[1 + i] * rho + qb[4] * rho * w[i];
    }
    wb[4] += rhob / (e * gm1);
    eb += -rhob * rho / e;
    wb[3] += eb * ph.Cv;
}

// reconstruct (one side): out = clamp(wC + (wD-wC)*lw + qw.g)
template <typename T>
ADFVM_HD void reconstruct_vjp(const T* wC, const T* gC, const T* wD, T lw, const T* qw, const T* ob,
                              int limit, T* wCb, T* gCb, T* wDb) {
    for (int i = 0; i < 5; i++) {
        T b = ob[i];
        if (limit) {
            const T* g = gC + 3 * i;
            T v = wC[i] + (wD[i] - wC[i]) * lw + (qw[0] * g[0] + qw[1] * g[1] + qw[2] * g[2]);
            T lo = wC[i] < wD[i] ? wC[i] : wD[i];
            T hi = wC[i] < wD[i] ? wD[i] : wC[i];
            // out = v < lo ? lo : (v > hi ? hi : v), lo/hi selected as in
            // reconstruct(); the adjoint follows the selected value
            if (v < lo) {
                if (wC[i] < wD[i]) wCb[i] += b; else wDb[i] += b;
                continue;
            }
            if (v > hi) {
                if (wC[i] < wD[i]) wDb[i] += b; else wCb[i] += b;
                continue;
            }
        }
        wCb[i] += b * (T(1) - lw);
        wDb[i] += b * lw;
        for (int k = 0; k < 3; k++) gCb[3 * i + k] += b * qw[k];
    }
}

// Roe flux reverse.  Inputs used: U (w[0..2]), p (w[4]), rho (q[0]).
template <typename T>
ADFVM_HD void roe_flux_vjp(const Phys<T>& ph, const T* wL, const T* qL, const T* wR, const T* qR, const T* n,
                           const T* Fb, T* wLb, T* qLb, T* wRb, T* qRb) {
    const T g = ph.gamma, gm1 = ph.gamma - T(1);
    const T *UL = wL, *UR = wR;
    T pL = wL[4], pR = wR[4], rhoL = qL[0], rhoR = qR[0];
    // ---- forward recompute
    T UnL = UL[0] * n[0] + UL[1] * n[1] + UL[2] * n[2];
    T UnR = UR[0] * n[0] + UR[1] * n[1] + UR[2] * n[2];
    T rhoUnL = rhoL * UnL, rhoUnR = rhoR * UnR;
    T hL = g * pL / (gm1 * rhoL) + T(0.5) * (UL[0] * UL[0] + UL[1] * UL[1] + UL[2] * UL[2]);
    T hR = g * pR / (gm1 * rhoR) + T(0.5) * (UR[0] * UR[0] + UR[1] * UR[1] + UR[2] * UR[2]);
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
    T cL = fsqrt(g * pL / rhoL), cR = fsqrt(g * pR / rhoR);
    T ea = rhoUnL / rhoL - rhoUnR / rhoR;
    T eb_ = cL - cR;
    T eps = T(0.5) * fabs_(ea);
    eps += T(0.5) * fabs_(eb_);
    eps = eps < T(0) ? eps - ph.small : eps + ph.small;
    T e2 = T(2) * eps;
    bool f1 = lam1 < e2, f2 = lam2 < e2, f3 = lam3 < e2;
    T l1 = f1 ? T(0.25) * lam1 * lam1 / eps + eps : lam1;
    T l2 = f2 ? T(0.25) * lam2 * lam2 / eps + eps : lam2;
    T l3 = f3 ? T(0.25) * lam3 * lam3 / eps + eps : lam3;
    T abv1 = T(0.5) * (l2 + l3);
    T abv2 = T(0.5) * (l2 - l3);
    T abv3 = abv1 - l1;
    T abv4 = gm1 * (qF * drho - (UF[0] * drhoU[0] + UF[1] * drhoU[1] + UF[2] * drhoU[2]) + drhoE);
    T abv5 = UnF * drho - (drhoU[0] * n[0] + drhoU[1] * n[1] + drhoU[2] * n[2]);
    T abv6 = abv3 * abv4 / a2F - abv2 * abv5 / aF;
    T abv7 = abv3 * abv5 - abv2 * abv4 / aF;
    (void)abv1;
    // ---- reverse
    T l1b = 0, drhob = 0, abv6b = 0, abv7b = 0, drhoEb = 0, hFb = 0, UnFb = 0;
    T drhoUb[3] = {0, 0, 0}, UFb[3] = {0, 0, 0}, FU0b[3];
    T Fr0b = Fb[0];
    l1b += -T(0.5) * Fb[0] * drho;
    drhob += -T(0.5) * Fb[0] * l1;
    abv6b += -T(0.5) * Fb[0];
    for (int i = 0; i < 3; i++) {
        T b = Fb[1 + i];
        FU0b[i] = b;
        l1b += -T(0.5) * b * drhoU[i];
        drhoUb[i] += -T(0.5) * b * l1;
        UFb[i] += -T(0.5) * b * abv6;
        abv6b += -T(0.5) * b * UF[i];
        abv7b += T(0.5) * b * n[i];
    }
    T FE0b = Fb[4];
    l1b += -T(0.5) * Fb[4] * drhoE;
    drhoEb += -T(0.5) * Fb[4] * l1;
    hFb += -T(0.5) * Fb[4] * abv6;
    abv6b += -T(0.5) * Fb[4] * hF;
    UnFb += T(0.5) * Fb[4] * abv7;
    abv7b += T(0.5) * Fb[4] * UnF;
    // abv7 = abv3*abv5 - abv2*abv4/aF
    T abv3b = abv7b * abv5, abv5b = abv7b * abv3;
    T abv2b = -abv7b * abv4 / aF, abv4b = -abv7b * abv2 / aF;
    T aFb = abv7b * abv2 * abv4 / (aF * aF);
    // abv6 = abv3*abv4/a2F - abv2*abv5/aF
    abv3b += abv6b * abv4 / a2F;
    abv4b += abv6b * abv3 / a2F;
    T a2Fb = -abv6b * abv3 * abv4 / (a2F * a2F);
    abv2b += -abv6b * abv5 / aF;
    abv5b += -abv6b * abv2 / aF;
    aFb += abv6b * abv2 * abv5 / (aF * aF);
    // abv5 = UnF*drho - drhoU.n
    UnFb += abv5b * drho;
    drhob += abv5b * UnF;
    for (int i = 0; i < 3; i++) drhoUb[i] += -abv5b * n[i];
    // abv4 = gm1*(qF*drho - UF.drhoU + drhoE)
    T t4 = gm1 * abv4b;
    T qFb = t4 * drho;
    drhob += t4 * qF;
    for (int i = 0; i < 3; i++) {
        UFb[i] += -t4 * drhoU[i];
        drhoUb[i] += -t4 * UF[i];
    }
    drhoEb += t4;
    // abv3 = abv1 - l1 ; abv1, abv2
    T abv1b = abv3b;
    l1b += -abv3b;
    T l2b = T(0.5) * abv1b + T(0.5) * abv2b;
    T l3b = T(0.5) * abv1b - T(0.5) * abv2b;
    // entropy fix
    T epsb = 0, lam1b, lam2b, lam3b;
    if (f1) { lam1b = l1b * T(0.5) * lam1 / eps; epsb += l1b * (T(1) - T(0.25) * lam1 * lam1 / (eps * eps)); } else lam1b = l1b;
    if (f2) { lam2b = l2b * T(0.5) * lam2 / eps; epsb += l2b * (T(1) - T(0.25) * lam2 * lam2 / (eps * eps)); } else lam2b = l2b;
    if (f3) { lam3b = l3b * T(0.5) * lam3 / eps; epsb += l3b * (T(1) - T(0.25) * lam3 * lam3 / (eps * eps)); } else lam3b = l3b;
    // eps = 0.5|ea| + 0.5|eb| (+ small)
    T eab = T(0.5) * sgn(ea) * epsb, ebb = T(0.5) * sgn(eb_) * epsb;
    T rhoUnLb = eab / rhoL, rhoUnRb = -eab / rhoR;
    T rhoLb = -eab * rhoUnL / (rhoL * rhoL), rhoRb = eab * rhoUnR / (rhoR * rhoR);
    T cLb = ebb, cRb = -ebb;
    T pLb = cLb * g / (T(2) * cL * rhoL), pRb = cRb * g / (T(2) * cR * rhoR);
    rhoLb += -cLb * g * pL / (T(2) * cL * rhoL * rhoL);
    rhoRb += -cRb * g * pR / (T(2) * cR * rhoR * rhoR);
    // lam
    T s1 = sgn(UnF), s2 = sgn(UnF + aF), s3 = sgn(UnF - aF);
    UnFb += lam1b * s1 + lam2b * s2 + lam3b * s3;
    aFb += lam2b * s2 - lam3b * s3;
    // drhoE
    T hRb = drhoEb * rhoR, hLb = -drhoEb * rhoL;
    rhoRb += drhoEb * hR;
    rhoLb += -drhoEb * hL;
    pRb += -drhoEb;
    pLb += drhoEb;
    T ULb[3] = {0, 0, 0}, URb[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        rhoRb += drhoUb[i] * UR[i];
        URb[i] += drhoUb[i] * rhoR;
        rhoLb += -drhoUb[i] * UL[i];
        ULb[i] += -drhoUb[i] * rhoL;
    }
    rhoRb += drhob;
    rhoLb += -drhob;
    for (int i = 0; i < 3; i++) UFb[i] += UnFb * n[i];
    a2Fb += aFb * T(0.5) / aF;
    hFb += gm1 * a2Fb;
    qFb += -gm1 * a2Fb;
    for (int i = 0; i < 3; i++) UFb[i] += qFb * UF[i];
    hLb += hFb * sL / dv;
    hRb += hFb * sR / dv;
    T sLb = hFb * hL / dv, sRb = hFb * hR / dv;
    T dvb = -hFb * hF / dv;
    for (int i = 0; i < 3; i++) {
        ULb[i] += UFb[i] * sL / dv;
        URb[i] += UFb[i] * sR / dv;
        sLb += UFb[i] * UL[i] / dv;
        sRb += UFb[i] * UR[i] / dv;
        dvb += -UFb[i] * UF[i] / dv;
    }
    sLb += dvb;
    sRb += dvb;
    rhoLb += sLb * T(0.5) / sL;
    rhoRb += sRb * T(0.5) / sR;
    // FE0 = 0.5*(rhoUnL*hL + rhoUnR*hR)
    rhoUnLb += T(0.5) * FE0b * hL;
    hLb += T(0.5) * FE0b * rhoUnL;
    rhoUnRb += T(0.5) * FE0b * hR;
    hRb += T(0.5) * FE0b * rhoUnR;
    // FU0
    for (int i = 0; i < 3; i++) {
        rhoUnLb += T(0.5) * FU0b[i] * UL[i];
        ULb[i] += T(0.5) * FU0b[i] * rhoUnL;
        rhoUnRb += T(0.5) * FU0b[i] * UR[i];
        URb[i] += T(0.5) * FU0b[i] * rhoUnR;
        pLb += T(0.5) * FU0b[i] * n[i];
        pRb += T(0.5) * FU0b[i] * n[i];
    }
    rhoUnLb += T(0.5) * Fr0b;
    rhoUnRb += T(0.5) * Fr0b;
    // h = g p/(gm1 rho) + 0.5 U.U
    pLb += hLb * g / (gm1 * rhoL);
    rhoLb += -hLb * g * pL / (gm1 * rhoL * rhoL);
    pRb += hRb * g / (gm1 * rhoR);
    rhoRb += -hRb * g * pR / (gm1 * rhoR * rhoR);
    for (int i = 0; i < 3; i++) { ULb[i] += hLb * UL[i]; URb[i] += hRb * UR[i]; }
    // rhoUn = rho*Un
    rhoLb += rhoUnLb * UnL;
    rhoRb += rhoUnRb * UnR;
    T UnLb = rhoUnLb * rhoL, UnRb = rhoUnRb * rhoR;
    for (int i = 0; i < 3; i++) { ULb[i] += UnLb * n[i]; URb[i] += UnRb * n[i]; }
    for (int i = 0; i < 3; i++) { wLb[i] += ULb[i]; wRb[i] += URb[i]; }
    wLb[4] += pLb; wRb[4] += pRb;
    qLb[0] += rhoLb; qRb[0] += rhoRb;
}

// Lax-Friedrichs flux reverse.  Inputs used: U, p (w) and rho, rhoU, rhoE (q).
template <typename T>
ADFVM_HD void lf_flux_vjp(const Phys<T>& ph, const T* wL, const T* qL, const T* wR, const T* qR, const T* n,
                          const T* Fb, T* wLb, T* qLb, T* wRb, T* qRb) {
    const T g = ph.gamma;
    T pL = wL[4], pR = wR[4], rhoL = qL[0], rhoR = qR[0];
    T UnL = wL[0] * n[0] + wL[1] * n[1] + wL[2] * n[2];
    T UnR = wR[0] * n[0] + wR[1] * n[1] + wR[2] * n[2];
    T cL = fsqrt(g * pL / rhoL), cR = fsqrt(g * pR / rhoR);
    T aL = fabs_(UnL) + cL, aR = fabs_(UnR) + cR;
    T aF = aL > aR ? aL : aR;
    T aFb = -T(0.5) * Fb[0] * (rhoR - rhoL) - T(0.5) * Fb[4] * (qR[4] - qL[4]);
    qLb[0] += T(0.5) * Fb[0] * UnL + T(0.5) * Fb[0] * aF;
    qRb[0] += T(0.5) * Fb[0] * UnR - T(0.5) * Fb[0] * aF;
    T UnLb = T(0.5) * Fb[0] * rhoL + T(0.5) * Fb[4] * (qL[4] + pL);
    T UnRb = T(0.5) * Fb[0] * rhoR + T(0.5) * Fb[4] * (qR[4] + pR);
    for (int i = 0; i < 3; i++) {
        T b = Fb[1 + i];
        aFb += -T(0.5) * b * (qR[1 + i] - qL[1 + i]);
        UnLb += T(0.5) * b * qL[1 + i];
        UnRb += T(0.5) * b * qR[1 + i];
        qLb[1 + i] += T(0.5) * b * UnL + T(0.5) * b * aF;
        qRb[1 + i] += T(0.5) * b * UnR - T(0.5) * b * aF;
        wLb[4] += T(0.5) * b * n[i];
        wRb[4] += T(0.5) * b * n[i];
    }
    wLb[4] += T(0.5) * Fb[4] * UnL;
    wRb[4] += T(0.5) * Fb[4] * UnR;
    qLb[4] += T(0.5) * Fb[4] * UnL + T(0.5) * Fb[4] * aF;
    qRb[4] += T(0.5) * Fb[4] * UnR - T(0.5) * Fb[4] * aF;
    T aLb, aRb;
    if (aL > aR) { aLb = aFb; aRb = 0; }
    else if (aR > aL) { aLb = 0; aRb = aFb; }
    else { aLb = T(0.5) * aFb; aRb = T(0.5) * aFb; }
    UnLb += aLb * sgn(UnL);
    UnRb += aRb * sgn(UnR);
    T cLb = aLb, cRb = aRb;
    wLb[4] += cLb * g / (T(2) * cL * rhoL);
    qLb[0] += -cLb * g * pL / (T(2) * cL * rhoL * rhoL);
    wRb[4] += cRb * g / (T(2) * cR * rhoR);
    qRb[0] += -cRb * g * pR / (T(2) * cR * rhoR * rhoR);
    for (int i = 0; i < 3; i++) { wLb[i] += UnLb * n[i]; wRb[i] += UnRb * n[i]; }
}

// viscous_flux reverse: adjoint of (TL, TR, UL, UR, TF, UF, gTF0, gUF0).
template <typename T>
ADFVM_HD void viscous_flux_vjp(const Phys<T>& ph, T TL, T TR, const T* UL, const T* UR, T TF, const T* UF,
                               const T* gTF0, const T* gUF0, const T* n, T delta, const T* D, const T* Fb,
                               T* TLb, T* TRb, T* ULb, T* URb, T* TFb, T* UFb, T* gTF0b, T* gUF0b) {
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
    T gTn = gT[0] * n[0] + gT[1] * n[1] + gT[2] * n[2];
    T tmp3 = gU[0] + gU[4] + gU[8];
    T t2[3], sigma[3];
    for (int i = 0; i < 3; i++) {
        t2[i] = (gU[3 * i] + gU[i]) * n[0] + (gU[3 * i + 1] + gU[3 + i]) * n[1] + (gU[3 * i + 2] + gU[6 + i]) * n[2];
        sigma[i] = mu * (t2[i] - T(2.0 / 3.0) * tmp3 * n[i]);
    }
    // reverse
    T sigb[3];
    for (int i = 0; i < 3; i++) {
        sigb[i] = -Fb[1 + i] - Fb[4] * UF[i];
        UFb[i] += -Fb[4] * sigma[i];
    }
    T qFb = -Fb[4];
    T mub = 0, tmp3b = 0, t2b[3];
    for (int i = 0; i < 3; i++) {
        mub += sigb[i] * (t2[i] - T(2.0 / 3.0) * tmp3 * n[i]);
        t2b[i] = sigb[i] * mu;
        tmp3b += -T(2.0 / 3.0) * sigb[i] * mu * n[i];
    }
    T gUb[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) gUb[3 * a + b] = t2b[a] * n[b] + t2b[b] * n[a];
    gUb[0] += tmp3b; gUb[4] += tmp3b; gUb[8] += tmp3b;
    T kappab = qFb * gTn;
    T gTb[3];
    for (int k = 0; k < 3; k++) gTb[k] = qFb * kappa * n[k];
    for (int i = 0; i < 3; i++) {
        T snUb = 0, gDb = 0;
        for (int k = 0; k < 3; k++) {
            gUF0b[3 * i + k] += gUb[3 * i + k];
            snUb += gUb[3 * i + k] * D[k];
            gDb += -gUb[3 * i + k] * D[k];
        }
        for (int k = 0; k < 3; k++) gUF0b[3 * i + k] += gDb * D[k];
        URb[i] += snUb / delta;
        ULb[i] += -snUb / delta;
    }
    T snTb = 0, gTDb = 0;
    for (int k = 0; k < 3; k++) {
        gTF0b[k] += gTb[k];
        snTb += gTb[k] * D[k];
        gTDb += -gTb[k] * D[k];
    }
    for (int k = 0; k < 3; k++) gTF0b[k] += gTDb * D[k];
    *TRb += snTb / delta;
    *TLb += -snTb / delta;
    mub += kappab * (ph.Cp / ph.Pr);
    *TFb += mub * viscosity_dT(ph, TF);
}

template <typename T>
ADFVM_HD void riemann_vjp(const Phys<T>& ph, int which, const T* wL, const T* qL, const T* wR, const T* qR,
                          const T* n, const T* Fb, T* wLb, T* qLb, T* wRb, T* qRb) {
    if (which == 0) roe_flux_vjp(ph, wL, qL, wR, qR, n, Fb, wLb, qLb, wRb, qRb);
    else lf_flux_vjp(ph, wL, qL, wR, qR, n, Fb, wLb, qLb, wRb, qRb);
}

// Reverse of face_flux with respect to the F output (the lam output's
// adjoint is zero in the adjoint solver, apps/adjoint.py:268: dtca = 0).
// Adds into wPb[5], gPb[15], wNb[5], gNb[15] (gNb may alias gPb).
template <typename T>
ADFVM_HD void face_flux_vjp(const Phys<T>& ph, int kind, const T* wP, const T* gP, const T* wN, const T* gN,
                            const FaceGeo<T>& f, const T* Fb, T* wPb, T* gPb, T* wNb, T* gNb) {
    T TF, UF[3], gTF[3], gUF[9];
    T wLF[5], qLF[5], wRF[5], qRF[5];
    if (kind == FACE_BOUNDARY) {
        TF = wN[3];
        for (int i = 0; i < 3; i++) UF[i] = wN[i];
        for (int k = 0; k < 3; k++) gTF[k] = gN[9 + k];
        for (int k = 0; k < 9; k++) gUF[k] = gN[k];
    } else {
        reconstruct(wP, gP, wN, f.lw0, f.qw0, wLF, ph.limit);
        conservative(ph, wLF, qLF);
        if (kind == FACE_CHARACTERISTIC) {
            for (int i = 0; i < 5; i++) wRF[i] = wN[i];
            TF = wN[3];
            for (int i = 0; i < 3; i++) UF[i] = wN[i];
            for (int k = 0; k < 3; k++) gTF[k] = gN[9 + k];
            for (int k = 0; k < 9; k++) gUF[k] = gN[k];
        } else {
            reconstruct(wN, gN, wP, f.lw1, f.qw1, wRF, ph.limit);
            for (int i = 0; i < 3; i++) UF[i] = T(0.5) * (wLF[i] + wRF[i]);
            TF = T(0.5) * (wLF[3] + wRF[3]);
            for (int k = 0; k < 3; k++) gTF[k] = T(0.5) * (gP[9 + k] + gN[9 + k]);
            for (int k = 0; k < 9; k++) gUF[k] = T(0.5) * (gP[k] + gN[k]);
        }
        conservative(ph, wRF, qRF);
    }
    // viscous part
    T TFb = 0, UFb[3] = {0, 0, 0}, gTFb[3] = {0, 0, 0}, gUFb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    viscous_flux_vjp(ph, wP[3], wN[3], wP, wN, TF, UF, gTF, gUF, f.n, f.delta, f.D, Fb,
                     &wPb[3], &wNb[3], wPb, wNb, &TFb, UFb, gTFb, gUFb);
    if (kind == FACE_BOUNDARY) {
        wNb[3] += TFb;
        for (int i = 0; i < 3; i++) wNb[i] += UFb[i];
        for (int k = 0; k < 3; k++) gNb[9 + k] += gTFb[k];
        for (int k = 0; k < 9; k++) gNb[k] += gUFb[k];
        // physical flux of the ghost state
        T qR[5];
        conservative(ph, wN, qR);
        T Un = wN[0] * f.n[0] + wN[1] * f.n[1] + wN[2] * f.n[2];
        T Unb = Fb[0] * qR[0] + Fb[4] * (qR[4] + wN[4]);
        T qRb[5] = {Fb[0] * Un, 0, 0, 0, Fb[4] * Un};
        T pb = Fb[4] * Un;
        for (int i = 0; i < 3; i++) {
            Unb += Fb[1 + i] * qR[1 + i];
            qRb[1 + i] = Fb[1 + i] * Un;
            pb += Fb[1 + i] * f.n[i];
        }
        wNb[4] += pb;
        for (int i = 0; i < 3; i++) wNb[i] += Unb * f.n[i];
        conservative_vjp(ph, wN, qRb, wNb);
        return;
    }
    T wLFb[5] = {0, 0, 0, 0, 0}, wRFb[5] = {0, 0, 0, 0, 0};
    T qLFb[5] = {0, 0, 0, 0, 0}, qRFb[5] = {0, 0, 0, 0, 0};
    if (kind == FACE_CHARACTERISTIC) {
        wNb[3] += TFb;
        for (int i = 0; i < 3; i++) wNb[i] += UFb[i];
        for (int k = 0; k < 3; k++) gNb[9 + k] += gTFb[k];
        for (int k = 0; k < 9; k++) gNb[k] += gUFb[k];
        riemann_vjp(ph, ph.boundaryRiemann, wLF, qLF, wRF, qRF, f.n, Fb, wLFb, qLFb, wRFb, qRFb);
    } else {
        for (int i = 0; i < 3; i++) { wLFb[i] += T(0.5) * UFb[i]; wRFb[i] += T(0.5) * UFb[i]; }
        wLFb[3] += T(0.5) * TFb; wRFb[3] += T(0.5) * TFb;
        for (int k = 0; k < 3; k++) { gPb[9 + k] += T(0.5) * gTFb[k]; gNb[9 + k] += T(0.5) * gTFb[k]; }
        for (int k = 0; k < 9; k++) { gPb[k] += T(0.5) * gUFb[k]; gNb[k] += T(0.5) * gUFb[k]; }
        riemann_vjp(ph, ph.riemann, wLF, qLF, wRF, qRF, f.n, Fb, wLFb, qLFb, wRFb, qRFb);
    }
    conservative_vjp(ph, wLF, qLFb, wLFb);
    conservative_vjp(ph, wRF, qRFb, wRFb);
    reconstruct_vjp(wP, gP, wN, f.lw0, f.qw0, wLFb, ph.limit, wPb, gPb, wNb);
    if (kind == FACE_CHARACTERISTIC) {
        for (int i = 0; i < 5; i++) wNb[i] += wRFb[i];
    } else {
        reconstruct_vjp(wN, gN, wP, f.lw1, f.qw1, wRFb, ph.limit, wNb, gNb, wPb);
    }
}

}  // namespace adfvm
