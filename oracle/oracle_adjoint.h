// ORACLE — test infrastructure only. Not part of the product; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Reference adjoint of the K17 collision by forward-mode automatic differentiation: the
// oracle's own templated k17_collide (oracle_k17.h, the restatement of
// src/lbm/collision/K17CompressibleNavierStokes.h:87-591) is evaluated on dual numbers
// v + e dv, once per input direction, which gives every column of the node Jacobian exactly
// (to FP64 round-off). No derivative formula is written by hand here, so the GPU's
// hand-written reverse mode is checked against an independent method.
#pragma once

#include <cmath>

#include "oracle_k17.h"

namespace orc {

struct Dual {
    double v, d;
    Dual(double x = 0.0) : v(x), d(0.0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
    Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
    Dual& operator*=(const Dual& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
};
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(const Dual& a, const Dual& b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
// |x| with the derivative sign(x); at x = 0 the limiter term x |x| has derivative 0 either way
inline Dual abs(const Dual& a) { return a.v < 0.0 ? -a : a; }

// (d post / d f)^T g for one node; ub (optional): adjoint of the half-step velocity shift.
// sb (optional): adjoint of the velocity scale p.vscale.
inline void k17_vjp_dual(const double* fin, const double* g, const K17Param<double>& p, double* fb, double* ub,
                         double* sb = nullptr)
{
    K17Param<Dual> q{p.omega, p.limP, p.limM, p.limD, p.fx, p.fy, p.fz, p.vscale};
    for (int col = 0; col < 31; ++col) {
        if (col >= 27 && col < 30 && !ub) continue;
        if (col == 30 && !sb) continue;
        Dual f[27];
        for (int e = 0; e < 27; ++e) f[e] = Dual(fin[e], col == e ? 1.0 : 0.0);
        K17Param<Dual> qq = q;
        if (col == 27) qq.fx = Dual(p.fx, 1.0);
        if (col == 28) qq.fy = Dual(p.fy, 1.0);
        if (col == 29) qq.fz = Dual(p.fz, 1.0);
        if (col == 30) qq.vscale = Dual(p.vscale, 1.0);
        k17_collide(f, qq, static_cast<Macro<Dual>*>(nullptr));
        double s = 0.0;
        for (int e = 0; e < 27; ++e) s += f[e].d * g[e];
        if (col < 27) fb[col] = s;
        else if (col < 30) ub[col - 27] = s;
        else *sb = s;
    }
}

} // namespace orc
