// Forward-mode dual numbers: a value and one tangent.  The face arithmetic of
// fvm_math.hpp is written once over its scalar type, so evaluating it on
// Dual<S> with the tangent seeded on one input gives that input's exact
// partial derivative alongside the value (the mesh-geometry derivatives of
// the face flux use it: one evaluation per geometry component).
#pragma once
#include "fvm_math.hpp"

namespace adfvm {

template <typename S> struct Dual {
    S v, d;
    ADFVM_HD Dual() : v(0), d(0) {}
    ADFVM_HD Dual(double x) : v(S(x)), d(0) {}
    ADFVM_HD Dual(float x) : v(S(x)), d(0) {}
    ADFVM_HD Dual(int x) : v(S(x)), d(0) {}
    ADFVM_HD Dual(S value, S tangent, int) : v(value), d(tangent) {}
    ADFVM_HD Dual& operator+=(const Dual& b) { v += b.v; d += b.d; return *this; }
    ADFVM_HD Dual& operator-=(const Dual& b) { v -= b.v; d -= b.d; return *this; }
    ADFVM_HD Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
    ADFVM_HD Dual& operator/=(const Dual& b) { *this = *this / b; return *this; }
    ADFVM_HD friend Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d, 0); }
    ADFVM_HD friend Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d, 0); }
    ADFVM_HD friend Dual operator-(const Dual& a) { return Dual(-a.v, -a.d, 0); }
    ADFVM_HD friend Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d, 0); }
    ADFVM_HD friend Dual operator/(const Dual& a, const Dual& b) {
        const S q = a.v / b.v;
        return Dual(q, (a.d - q * b.d) / b.v, 0);
    }
    ADFVM_HD friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    ADFVM_HD friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    ADFVM_HD friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    ADFVM_HD friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    ADFVM_HD friend Dual sqrt(const Dual& a) {
        const S r = sqrt(a.v);
        return Dual(r, a.d / (S(2) * r), 0);
    }
    // a^b; the log term only when the exponent carries a tangent (a may be
    // negative where it does not)
    ADFVM_HD friend Dual pow(const Dual& a, const Dual& b) {
        const S r = pow(a.v, b.v);
        S t = b.v * pow(a.v, b.v - S(1)) * a.d;
        if (b.d != S(0)) t += r * log(a.v) * b.d;
        return Dual(r, t, 0);
    }
};

// |x| with the derivative sign(x) * dx, sign(0) = 0: the convention of the
// hand-written adjoints (adjoint_math.hpp) and of torch, so that at a kink
// (a velocity component exactly 0) the geometry derivatives take the same
// one-sided choice as the state derivatives
template <typename S> ADFVM_HD Dual<S> fabs_(Dual<S> x) {
    const S s = x.v > S(0) ? S(1) : (x.v < S(0) ? S(-1) : S(0));
    return Dual<S>(x.v < S(0) ? -x.v : x.v, s * x.d, 0);
}

// max(a, b) passing half of the tangent to each at a tie (the adjoints'
// and torch's convention)
template <typename S> ADFVM_HD Dual<S> fmax_(Dual<S> a, Dual<S> b) {
    if (a.v > b.v) return a;
    if (b.v > a.v) return b;
    return Dual<S>(a.v, S(0.5) * (a.d + b.d), 0);
}

// The physical constants as duals with zero tangents.
template <typename S> ADFVM_HD Phys<Dual<S>> phys_dual(const Phys<S>& p) {
    Phys<Dual<S>> q;
    q.gamma = p.gamma; q.Cp = p.Cp; q.Cv = p.Cv; q.Pr = p.Pr; q.invCv = p.invCv; q.small = p.small;
    q.mu0 = p.mu0; q.muA = p.muA; q.muTs = p.muTs;
    q.muModel = p.muModel; q.riemann = p.riemann; q.boundaryRiemann = p.boundaryRiemann; q.limit = p.limit;
    return q;
}

}  // namespace adfvm
