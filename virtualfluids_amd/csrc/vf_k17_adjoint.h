// Hand-written reverse mode (vector-Jacobian product) of the FP32 K17 collision of vf_k17.h.
//
// For one node with pre-collision shifted populations f and an adjoint g of its
// post-collision populations, k17_vjp returns fb = (d post / d f)^T g, and the adjoint of
// the half-step velocity shift (fx, fy, fz) for force gradients. The collision is the chain
//
//   f -> (drho, j, u) -> forward chimera z, y, x (central moments m_abc)
//     -> relaxation and cumulant algebra (the inputs N_abc of the backward chimera)
//     -> backward chimera x, y, z -> post-collision populations,
//
// and the reverse walks it backwards: the backward chimera lines are linear in their inputs
// for a fixed velocity (3x3 transposes plus a velocity term), the moment algebra is
// differentiated term by term, the limiter x -> x N / (|x| + L) has the derivative
// N L / (|x| + L)^2, and the forward chimera lines and the macroscopic sums close the chain.
// Every forward intermediate the reverse needs is recomputed from f (nothing is stored per
// node between steps). Arrays inside this file are indexed idx(a, b, c) = a + 3 b + 9 c: for
// populations (a, b, c) is the velocity plus one per axis, for moments the orders per axis.
#pragma once

#include "vf_k17.h"

namespace vfa {
namespace adj {

VF_HD constexpr int idx(int a, int b, int c) { return a + 3 * b + 9 * c; }

// D3Q27 direction of idx(i, j, k): velocity (i-1, j-1, k-1)
VF_HD constexpr int dir_of_idx(int i)
{
    // (cx, cy, cz) = (i%3-1, (i/3)%3-1, i/9-1)
    return i == idx(1, 1, 1) ? d000
         : i == idx(2, 1, 1) ? dP00 : i == idx(0, 1, 1) ? dM00
         : i == idx(1, 2, 1) ? d0P0 : i == idx(1, 0, 1) ? d0M0
         : i == idx(1, 1, 2) ? d00P : i == idx(1, 1, 0) ? d00M
         : i == idx(2, 2, 1) ? dPP0 : i == idx(0, 0, 1) ? dMM0
         : i == idx(2, 0, 1) ? dPM0 : i == idx(0, 2, 1) ? dMP0
         : i == idx(2, 1, 2) ? dP0P : i == idx(0, 1, 0) ? dM0M
         : i == idx(2, 1, 0) ? dP0M : i == idx(0, 1, 2) ? dM0P
         : i == idx(1, 2, 2) ? d0PP : i == idx(1, 0, 0) ? d0MM
         : i == idx(1, 2, 0) ? d0PM : i == idx(1, 0, 2) ? d0MP
         : i == idx(2, 2, 2) ? dPPP : i == idx(0, 2, 2) ? dMPP
         : i == idx(2, 0, 2) ? dPMP : i == idx(0, 0, 2) ? dMMP
         : i == idx(2, 2, 0) ? dPPM : i == idx(0, 2, 0) ? dMPM
         : i == idx(2, 0, 0) ? dPMM : dMMM;
}

// Chimera weights K of a line: x lines (b, c) -> v(b) v(c); y lines (a, c) -> w(a) v(c);
// z lines (a, b) -> w(a) w(b), with v = (1, 0, 1/3) and w = (1/6, 2/3, 1/6).
VF_HD constexpr float kv(int i) { return i == 1 ? 0.0f : (i == 0 ? 1.0f : 1.0f / 3.0f); }
VF_HD constexpr float kw(int i) { return i == 1 ? 2.0f / 3.0f : 1.0f / 6.0f; }

// forward chimera line: (a, b, c) -> (s, d - u (s+K), e - 2 u d + (s+K) u^2)
VF_HD void line_fwd(float& a, float& b, float& c, float u, float K)
{
    const float e = a + c, d = c - a, s = e + b, sK = s + K;
    a = s;
    b = d - u * sK;
    c = e - 2.0f * u * d + sK * u * u;
}

// its reverse: (a, b, c) are the line's inputs, (A, B, C) the adjoints of its outputs;
// returns the adjoints of the inputs in (A, B, C) and adds the velocity term to ub
VF_HD void line_fwd_rev(float a, float b, float c, float u, float K, float& A, float& B, float& C, float& ub)
{
    const float e = a + c, d = c - a, s = e + b, sK = s + K;
    ub += -sK * B + (2.0f * u * sK - 2.0f * d) * C;
    const float sb = A - u * B + u * u * C;
    const float db = B - 2.0f * u * C;
    const float eb = C + sb;
    A = eb - db;
    B = sb;
    C = eb + db;
}

// backward chimera line (the inverse transform): lo = a0 (a+K) + b0 b + c/2,
// hi = a2 (a+K) + b2 b + c/2, mid = a - lo - hi
VF_HD void line_bwd(float& a, float& b, float& c, float u, float K)
{
    const float a0 = 0.5f * (u * u - u), b0 = u - 0.5f, a2 = 0.5f * (u * u + u), b2 = u + 0.5f;
    const float aK = a + K;
    const float lo = a0 * aK + b0 * b + 0.5f * c;
    const float hi = a2 * aK + b2 * b + 0.5f * c;
    const float mid = (a - lo) - hi;
    a = lo;
    b = mid;
    c = hi;
}

VF_HD void line_bwd_rev(float a, float b, float u, float K, float& A, float& B, float& C, float& ub)
{
    const float a0 = 0.5f * (u * u - u), b0 = u - 0.5f, a2 = 0.5f * (u * u + u), b2 = u + 0.5f;
    const float aK = a + K;
    const float L = A, M = B, H = C;
    ub += (L - M) * (b0 * aK + b) + (H - M) * (b2 * aK + b);
    A = a0 * L + a2 * H + (1.0f - a0 - a2) * M;
    B = b0 * L + b2 * H - (b0 + b2) * M;
    C = 0.5f * (L + H) - M;
}

// ---- forward state --------------------------------------------------------------------------
struct Macro {
    float drho, R, jx, jy, jz, vx, vy, vz;
};

VF_HD void macro_of(const float (&p)[27], const K17Rates& r, float vscale, Macro& m)
{
    float drho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float a = p[idx(i, j, 0)], b = p[idx(i, j, 1)], c = p[idx(i, j, 2)];
            const float s = (a + c) + b;
            drho += s;
            jx += float(i - 1) * s;
            jy += float(j - 1) * s;
            jz += c - a;
        }
    m.drho = drho;
    m.R = fast_rcp(drho + 1.0f);
    m.jx = jx;
    m.jy = jy;
    m.jz = jz;
    const float rs = m.R * vscale;
    m.vx = jx * rs + r.fx;
    m.vy = jy * rs + r.fy;
    m.vz = jz * rs + r.fz;
}

// z pass (all 27 outputs are used downstream)
VF_HD void pass_z(float (&p)[27], float vz)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) line_fwd(p[idx(i, j, 0)], p[idx(i, j, 1)], p[idx(i, j, 2)], vz, kw(i) * kw(j));
}
VF_HD void pass_y(float (&p)[27], float vy)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) line_fwd(p[idx(i, 0, k)], p[idx(i, 1, k)], p[idx(i, 2, k)], vy, kw(i) * kv(k));
}
VF_HD void pass_x(float (&p)[27], float vx)
{
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) line_fwd(p[idx(0, j, k)], p[idx(1, j, k)], p[idx(2, j, k)], vx, kv(j) * kv(k));
}

// Moments of the cumulant stage that the relaxation reads, and everything derived from them.
struct Mid {
    float m000, m100, m010, m001, m110, m101, m011, m200, m020, m002, m111;
    float m210, m012, m201, m021, m120, m102;
    float mxxMyy0, mxxMzz0, dxux, dyuy, dzuz, pTrace, mxxMyy, mxxMzz;
    float n200, n020, n002, n110, n101, n011, n111;
    float c[6], inv[6], invD;
    float n210, n012, n201, n021, n120, n102;
    float q110, q101, q011, n211, n121, n112, n220, n202, n022;
    float Q122, Q212, Q221, n122, n212, n221, t1, t2, n222;
};

VF_HD void mid_forward(const float (&m)[27], const Macro& mc, const K17Rates& r, Mid& s, float (&N)[27])
{
    const float R = mc.R, drho = mc.drho;
    const float vx2 = mc.vx * mc.vx, vy2 = mc.vy * mc.vy, vz2 = mc.vz * mc.vz;
    s.m000 = m[idx(0, 0, 0)];
    s.m100 = m[idx(1, 0, 0)];
    s.m010 = m[idx(0, 1, 0)];
    s.m001 = m[idx(0, 0, 1)];
    s.m110 = m[idx(1, 1, 0)];
    s.m101 = m[idx(1, 0, 1)];
    s.m011 = m[idx(0, 1, 1)];
    s.m200 = m[idx(2, 0, 0)];
    s.m020 = m[idx(0, 2, 0)];
    s.m002 = m[idx(0, 0, 2)];
    s.m111 = m[idx(1, 1, 1)];
    s.m210 = m[idx(2, 1, 0)];
    s.m012 = m[idx(0, 1, 2)];
    s.m201 = m[idx(2, 0, 1)];
    s.m021 = m[idx(0, 2, 1)];
    s.m120 = m[idx(1, 2, 0)];
    s.m102 = m[idx(1, 0, 2)];
    // second order
    const float mxxPyyPzz = s.m200 + s.m020 + s.m002;
    s.mxxMyy0 = s.m200 - s.m020;
    s.mxxMzz0 = s.m200 - s.m002;
    s.dxux = r.mHalfOmega * (s.mxxMyy0 + s.mxxMzz0) + 0.5f * (s.m000 - mxxPyyPzz);
    s.dyuy = r.threeHalfOmega * s.mxxMyy0 + s.dxux;
    s.dzuz = r.threeHalfOmega * s.mxxMzz0 + s.dxux;
    const float Px = vx2 * s.dxux, Py = vy2 * s.dyuy, Pz = vz2 * s.dzuz;
    s.pTrace = s.m000 - 1.5f * (Px + Py + Pz);
    s.mxxMyy = r.oneMinusOmega * s.mxxMyy0 - r.galileo * (Px - Py);
    s.mxxMzz = r.oneMinusOmega * s.mxxMzz0 - r.galileo * (Px - Pz);
    s.n110 = r.oneMinusOmega * s.m110;
    s.n101 = r.oneMinusOmega * s.m101;
    s.n011 = r.oneMinusOmega * s.m011;
    s.n200 = (1.0f / 3.0f) * (s.mxxMyy + s.mxxMzz + s.pTrace);
    s.n020 = (1.0f / 3.0f) * (-2.0f * s.mxxMyy + s.mxxMzz + s.pTrace);
    s.n002 = (1.0f / 3.0f) * (s.mxxMyy - 2.0f * s.mxxMzz + s.pTrace);
    // third order, limited
    s.invD = fast_rcp(__builtin_fabsf(s.m111) + r.limD);
    s.n111 = s.m111 * r.limNumD * s.invD;
    s.c[0] = s.m210 + s.m012;
    s.c[1] = s.m210 - s.m012;
    s.c[2] = s.m201 + s.m021;
    s.c[3] = s.m201 - s.m021;
    s.c[4] = s.m120 + s.m102;
    s.c[5] = s.m120 - s.m102;
    float cl[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float L = (i & 1) ? r.limM : r.limP, Nn = (i & 1) ? r.limNumM : r.limNumP;
        s.inv[i] = fast_rcp(__builtin_fabsf(s.c[i]) + L);
        cl[i] = s.c[i] * Nn * s.inv[i];
    }
    s.n210 = 0.5f * (cl[0] + cl[1]);
    s.n012 = 0.5f * (cl[0] - cl[1]);
    s.n201 = 0.5f * (cl[2] + cl[3]);
    s.n021 = 0.5f * (cl[2] - cl[3]);
    s.n120 = 0.5f * (cl[4] + cl[5]);
    s.n102 = 0.5f * (cl[4] - cl[5]);
    // fourth order
    s.q110 = s.n110 * R;
    s.q101 = s.n101 * R;
    s.q011 = s.n011 * R;
    s.n211 = r.B4 * s.m011 + (1.0f / 3.0f) * ((3.0f * s.n200 + 1.0f) * s.q011 + 6.0f * s.n110 * s.q101);
    s.n121 = r.B4 * s.m101 + (1.0f / 3.0f) * ((3.0f * s.n020 + 1.0f) * s.q101 + 6.0f * s.n110 * s.q011);
    s.n112 = r.B4 * s.m110 + (1.0f / 3.0f) * ((3.0f * s.n002 + 1.0f) * s.q110 + 6.0f * s.n101 * s.q011);
    const float rhoTerm = drho * R * (1.0f / 9.0f);
    s.n220 = r.A4 * (s.dxux + s.dyuy) +
             ((s.n200 * s.n020 + 2.0f * s.n110 * s.n110) + (1.0f / 3.0f) * (s.n200 + s.n020)) * R - rhoTerm;
    s.n202 = r.A4 * (s.dxux + s.dzuz) +
             ((s.n200 * s.n002 + 2.0f * s.n101 * s.n101) + (1.0f / 3.0f) * (s.n200 + s.n002)) * R - rhoTerm;
    s.n022 = r.A4 * (s.dyuy + s.dzuz) +
             ((s.n002 * s.n020 + 2.0f * s.n011 * s.n011) + (1.0f / 3.0f) * (s.n002 + s.n020)) * R - rhoTerm;
    // fifth order
    const float P122 = s.n002 * s.n120 + s.n020 * s.n102 + 4.0f * s.n011 * s.n111 + 2.0f * (s.n101 * s.n021 + s.n110 * s.n012);
    const float P212 = s.n002 * s.n210 + s.n200 * s.n012 + 4.0f * s.n101 * s.n111 + 2.0f * (s.n011 * s.n201 + s.n110 * s.n102);
    const float P221 = s.n200 * s.n021 + s.n020 * s.n201 + 4.0f * s.n110 * s.n111 + 2.0f * (s.n101 * s.n120 + s.n011 * s.n210);
    s.Q122 = (1.0f / 3.0f) * (3.0f * P122 + s.n120 + s.n102);
    s.Q212 = (1.0f / 3.0f) * (3.0f * P212 + s.n210 + s.n012);
    s.Q221 = (1.0f / 3.0f) * (3.0f * P221 + s.n021 + s.n201);
    s.n122 = s.Q122 * R;
    s.n212 = s.Q212 * R;
    s.n221 = s.Q221 * R;
    // sixth order
    const float sq110 = s.n110 * s.n110, sq101 = s.n101 * s.n101, sq011 = s.n011 * s.n011;
    const float tr2 = s.n200 + s.n020 + s.n002;
    s.t1 = -4.0f * s.n111 * s.n111 - (s.n200 * s.n022 + s.n020 * s.n202 + s.n002 * s.n220) -
           4.0f * (s.n011 * s.n211 + s.n101 * s.n121 + s.n110 * s.n112) -
           2.0f * (s.n120 * s.n102 + s.n210 * s.n012 + s.n201 * s.n021) - (1.0f / 3.0f) * (s.n022 + s.n202 + s.n220) -
           (1.0f / 9.0f) * tr2;
    s.t2 = 4.0f * (sq101 * s.n020 + sq011 * s.n200 + sq110 * s.n002) + 2.0f * (s.n200 * s.n020 * s.n002) +
           16.0f * s.n110 * s.n101 * s.n011 +
           (2.0f / 3.0f) * (2.0f * (sq101 + sq011 + sq110) + (s.n002 * s.n020 + s.n002 * s.n200 + s.n020 * s.n200) +
                            (1.0f / 3.0f) * tr2) +
           (1.0f / 27.0f) * (drho * drho - drho);
    s.n222 = -(s.t1 + s.t2 * R) * R;

    N[idx(0, 0, 0)] = s.m000;
    N[idx(1, 0, 0)] = -s.m100;
    N[idx(0, 1, 0)] = -s.m010;
    N[idx(0, 0, 1)] = -s.m001;
    N[idx(2, 0, 0)] = s.n200;
    N[idx(0, 2, 0)] = s.n020;
    N[idx(0, 0, 2)] = s.n002;
    N[idx(1, 1, 0)] = s.n110;
    N[idx(1, 0, 1)] = s.n101;
    N[idx(0, 1, 1)] = s.n011;
    N[idx(1, 1, 1)] = s.n111;
    N[idx(2, 1, 0)] = s.n210;
    N[idx(0, 1, 2)] = s.n012;
    N[idx(2, 0, 1)] = s.n201;
    N[idx(0, 2, 1)] = s.n021;
    N[idx(1, 2, 0)] = s.n120;
    N[idx(1, 0, 2)] = s.n102;
    N[idx(2, 1, 1)] = s.n211;
    N[idx(1, 2, 1)] = s.n121;
    N[idx(1, 1, 2)] = s.n112;
    N[idx(2, 2, 0)] = s.n220;
    N[idx(2, 0, 2)] = s.n202;
    N[idx(0, 2, 2)] = s.n022;
    N[idx(1, 2, 2)] = s.n122;
    N[idx(2, 1, 2)] = s.n212;
    N[idx(2, 2, 1)] = s.n221;
    N[idx(2, 2, 2)] = s.n222;
}

// Adjoint of mid_forward: B holds the adjoints of N on entry; on exit mb holds the adjoints of
// the moments m (idx order, zero where unused), and drho, R and u^2 adjoints are added.
VF_HD void mid_reverse(const Mid& s, const Macro& mc, const K17Rates& r, const float (&B)[27], float (&mb)[27],
                       float& drhob, float& Rb, float& vx2b, float& vy2b, float& vz2b)
{
    const float R = mc.R, drho = mc.drho;
    const float vx2 = mc.vx * mc.vx, vy2 = mc.vy * mc.vy, vz2 = mc.vz * mc.vz;
    float b200 = B[idx(2, 0, 0)], b020 = B[idx(0, 2, 0)], b002 = B[idx(0, 0, 2)];
    float b110 = B[idx(1, 1, 0)], b101 = B[idx(1, 0, 1)], b011 = B[idx(0, 1, 1)], b111 = B[idx(1, 1, 1)];
    float b210 = B[idx(2, 1, 0)], b012 = B[idx(0, 1, 2)], b201 = B[idx(2, 0, 1)], b021 = B[idx(0, 2, 1)];
    float b120 = B[idx(1, 2, 0)], b102 = B[idx(1, 0, 2)];
    float b211 = B[idx(2, 1, 1)], b121 = B[idx(1, 2, 1)], b112 = B[idx(1, 1, 2)];
    float b220 = B[idx(2, 2, 0)], b202 = B[idx(2, 0, 2)], b022 = B[idx(0, 2, 2)];
    const float b122 = B[idx(1, 2, 2)], b212 = B[idx(2, 1, 2)], b221 = B[idx(2, 2, 1)], b222 = B[idx(2, 2, 2)];
#pragma unroll
    for (int i = 0; i < 27; ++i) mb[i] = 0.0f;
    mb[idx(0, 0, 0)] = B[idx(0, 0, 0)];
    mb[idx(1, 0, 0)] = -B[idx(1, 0, 0)];
    mb[idx(0, 1, 0)] = -B[idx(0, 1, 0)];
    mb[idx(0, 0, 1)] = -B[idx(0, 0, 1)];

    // sixth order: n222 = -(t1 + t2 R) R
    {
        const float t1b = -R * b222;
        const float t2b = -R * R * b222;
        Rb += -(s.t1 + 2.0f * s.t2 * R) * b222;
        // t1
        b111 += -8.0f * s.n111 * t1b;
        b200 += (-s.n022 - 1.0f / 9.0f) * t1b;
        b020 += (-s.n202 - 1.0f / 9.0f) * t1b;
        b002 += (-s.n220 - 1.0f / 9.0f) * t1b;
        b022 += (-s.n200 - 1.0f / 3.0f) * t1b;
        b202 += (-s.n020 - 1.0f / 3.0f) * t1b;
        b220 += (-s.n002 - 1.0f / 3.0f) * t1b;
        b011 += -4.0f * s.n211 * t1b;
        b211 += -4.0f * s.n011 * t1b;
        b101 += -4.0f * s.n121 * t1b;
        b121 += -4.0f * s.n101 * t1b;
        b110 += -4.0f * s.n112 * t1b;
        b112 += -4.0f * s.n110 * t1b;
        b120 += -2.0f * s.n102 * t1b;
        b102 += -2.0f * s.n120 * t1b;
        b210 += -2.0f * s.n012 * t1b;
        b012 += -2.0f * s.n210 * t1b;
        b201 += -2.0f * s.n021 * t1b;
        b021 += -2.0f * s.n201 * t1b;
        // t2
        const float sq110 = s.n110 * s.n110, sq101 = s.n101 * s.n101, sq011 = s.n011 * s.n011;
        b200 += (4.0f * sq011 + 2.0f * s.n020 * s.n002 + (2.0f / 3.0f) * (s.n002 + s.n020) + 2.0f / 9.0f) * t2b;
        b020 += (4.0f * sq101 + 2.0f * s.n200 * s.n002 + (2.0f / 3.0f) * (s.n002 + s.n200) + 2.0f / 9.0f) * t2b;
        b002 += (4.0f * sq110 + 2.0f * s.n200 * s.n020 + (2.0f / 3.0f) * (s.n020 + s.n200) + 2.0f / 9.0f) * t2b;
        b110 += (8.0f * s.n110 * s.n002 + 16.0f * s.n101 * s.n011 + (8.0f / 3.0f) * s.n110) * t2b;
        b101 += (8.0f * s.n101 * s.n020 + 16.0f * s.n110 * s.n011 + (8.0f / 3.0f) * s.n101) * t2b;
        b011 += (8.0f * s.n011 * s.n200 + 16.0f * s.n110 * s.n101 + (8.0f / 3.0f) * s.n011) * t2b;
        drhob += (1.0f / 27.0f) * (2.0f * drho - 1.0f) * t2b;
    }
    // fifth order: n_abc = Q R, Q = P + (linear)/3
    {
        Rb += s.Q122 * b122 + s.Q212 * b212 + s.Q221 * b221;
        float Pb = R * b122;  // dQ/dP = 1
        b120 += Pb * (1.0f / 3.0f) + Pb * s.n002;
        b102 += Pb * (1.0f / 3.0f) + Pb * s.n020;
        b002 += Pb * s.n120;
        b020 += Pb * s.n102;
        b011 += 4.0f * Pb * s.n111;
        b111 += 4.0f * Pb * s.n011;
        b101 += 2.0f * Pb * s.n021;
        b021 += 2.0f * Pb * s.n101;
        b110 += 2.0f * Pb * s.n012;
        b012 += 2.0f * Pb * s.n110;
        Pb = R * b212;
        b210 += Pb * (1.0f / 3.0f) + Pb * s.n002;
        b012 += Pb * (1.0f / 3.0f) + Pb * s.n200;
        b002 += Pb * s.n210;
        b200 += Pb * s.n012;
        b101 += 4.0f * Pb * s.n111;
        b111 += 4.0f * Pb * s.n101;
        b011 += 2.0f * Pb * s.n201;
        b201 += 2.0f * Pb * s.n011;
        b110 += 2.0f * Pb * s.n102;
        b102 += 2.0f * Pb * s.n110;
        Pb = R * b221;
        b021 += Pb * (1.0f / 3.0f) + Pb * s.n200;
        b201 += Pb * (1.0f / 3.0f) + Pb * s.n020;
        b200 += Pb * s.n021;
        b020 += Pb * s.n201;
        b110 += 4.0f * Pb * s.n111;
        b111 += 4.0f * Pb * s.n110;
        b101 += 2.0f * Pb * s.n120;
        b120 += 2.0f * Pb * s.n101;
        b011 += 2.0f * Pb * s.n210;
        b210 += 2.0f * Pb * s.n011;
    }
    // fourth order
    float dxuxb = 0.0f, dyuyb = 0.0f, dzuzb = 0.0f;
    {
        dxuxb += r.A4 * (b220 + b202);
        dyuyb += r.A4 * (b220 + b022);
        dzuzb += r.A4 * (b202 + b022);
        Rb += ((s.n200 * s.n020 + 2.0f * s.n110 * s.n110) + (1.0f / 3.0f) * (s.n200 + s.n020)) * b220 +
              ((s.n200 * s.n002 + 2.0f * s.n101 * s.n101) + (1.0f / 3.0f) * (s.n200 + s.n002)) * b202 +
              ((s.n002 * s.n020 + 2.0f * s.n011 * s.n011) + (1.0f / 3.0f) * (s.n002 + s.n020)) * b022;
        float Sb = R * b220;
        b200 += Sb * (s.n020 + 1.0f / 3.0f);
        b020 += Sb * (s.n200 + 1.0f / 3.0f);
        b110 += Sb * 4.0f * s.n110;
        Sb = R * b202;
        b200 += Sb * (s.n002 + 1.0f / 3.0f);
        b002 += Sb * (s.n200 + 1.0f / 3.0f);
        b101 += Sb * 4.0f * s.n101;
        Sb = R * b022;
        b002 += Sb * (s.n020 + 1.0f / 3.0f);
        b020 += Sb * (s.n002 + 1.0f / 3.0f);
        b011 += Sb * 4.0f * s.n011;
        // rhoTerm = drho R / 9 enters all three with a minus sign
        const float rtb = -(b220 + b202 + b022) * (1.0f / 9.0f);
        drhob += R * rtb;
        Rb += drho * rtb;
        // n211, n121, n112
        float q110b = 0.0f, q101b = 0.0f, q011b = 0.0f;
        mb[idx(0, 1, 1)] += r.B4 * b211;
        b200 += s.q011 * b211;
        q011b += (1.0f / 3.0f) * (3.0f * s.n200 + 1.0f) * b211;
        b110 += 2.0f * s.q101 * b211;
        q101b += 2.0f * s.n110 * b211;
        mb[idx(1, 0, 1)] += r.B4 * b121;
        b020 += s.q101 * b121;
        q101b += (1.0f / 3.0f) * (3.0f * s.n020 + 1.0f) * b121;
        b110 += 2.0f * s.q011 * b121;
        q011b += 2.0f * s.n110 * b121;
        mb[idx(1, 1, 0)] += r.B4 * b112;
        b002 += s.q110 * b112;
        q110b += (1.0f / 3.0f) * (3.0f * s.n002 + 1.0f) * b112;
        b101 += 2.0f * s.q011 * b112;
        q011b += 2.0f * s.n101 * b112;
        b110 += R * q110b;
        b101 += R * q101b;
        b011 += R * q011b;
        Rb += s.n110 * q110b + s.n101 * q101b + s.n011 * q011b;
    }
    // third order: x -> x N / (|x| + L) has the derivative N L / (|x| + L)^2
    {
        mb[idx(1, 1, 1)] += b111 * r.limNumD * r.limD * s.invD * s.invD;
        float cb[6];
        cb[0] = 0.5f * (b210 + b012);
        cb[1] = 0.5f * (b210 - b012);
        cb[2] = 0.5f * (b201 + b021);
        cb[3] = 0.5f * (b201 - b021);
        cb[4] = 0.5f * (b120 + b102);
        cb[5] = 0.5f * (b120 - b102);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const float L = (i & 1) ? r.limM : r.limP, Nn = (i & 1) ? r.limNumM : r.limNumP;
            cb[i] *= Nn * L * s.inv[i] * s.inv[i];
        }
        mb[idx(2, 1, 0)] += cb[0] + cb[1];
        mb[idx(0, 1, 2)] += cb[0] - cb[1];
        mb[idx(2, 0, 1)] += cb[2] + cb[3];
        mb[idx(0, 2, 1)] += cb[2] - cb[3];
        mb[idx(1, 2, 0)] += cb[4] + cb[5];
        mb[idx(1, 0, 2)] += cb[4] - cb[5];
    }
    // second order
    {
        mb[idx(1, 1, 0)] += r.oneMinusOmega * b110;
        mb[idx(1, 0, 1)] += r.oneMinusOmega * b101;
        mb[idx(0, 1, 1)] += r.oneMinusOmega * b011;
        const float yyb = (b200 - 2.0f * b020 + b002) * (1.0f / 3.0f);  // adjoint of mxxMyy (relaxed)
        const float zzb = (b200 + b020 - 2.0f * b002) * (1.0f / 3.0f);  // adjoint of mxxMzz (relaxed)
        const float ptb = (b200 + b020 + b002) * (1.0f / 3.0f);         // adjoint of pTrace
        float yy0b = r.oneMinusOmega * yyb, zz0b = r.oneMinusOmega * zzb;
        const float Pxb = -r.galileo * (yyb + zzb) - 1.5f * ptb;
        const float Pyb = r.galileo * yyb - 1.5f * ptb;
        const float Pzb = r.galileo * zzb - 1.5f * ptb;
        mb[idx(0, 0, 0)] += ptb;
        vx2b += Pxb * s.dxux;
        vy2b += Pyb * s.dyuy;
        vz2b += Pzb * s.dzuz;
        dxuxb += Pxb * vx2;
        dyuyb += Pyb * vy2;
        dzuzb += Pzb * vz2;
        zz0b += r.threeHalfOmega * dzuzb;
        dxuxb += dzuzb;
        yy0b += r.threeHalfOmega * dyuyb;
        dxuxb += dyuyb;
        yy0b += r.mHalfOmega * dxuxb;
        zz0b += r.mHalfOmega * dxuxb;
        mb[idx(0, 0, 0)] += 0.5f * dxuxb;
        const float trb = -0.5f * dxuxb;  // adjoint of mxxPyyPzz
        mb[idx(2, 0, 0)] += trb + yy0b + zz0b;
        mb[idx(0, 2, 0)] += trb - yy0b;
        mb[idx(0, 0, 2)] += trb - zz0b;
    }
}

// backward chimera x, y, z on moments M (idx order), giving populations in idx order
VF_HD void backward_chimera(float (&M)[27], float vx, float vy, float vz)
{
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) line_bwd(M[idx(0, j, k)], M[idx(1, j, k)], M[idx(2, j, k)], vx, kv(j) * kv(k));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) line_bwd(M[idx(i, 0, k)], M[idx(i, 1, k)], M[idx(i, 2, k)], vy, kw(i) * kv(k));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) line_bwd(M[idx(i, j, 0)], M[idx(i, j, 1)], M[idx(i, j, 2)], vz, kw(i) * kw(j));
}

// its reverse: N the moments entering the passes; A holds the adjoints of the outputs on entry
// and the adjoints of N on return; the velocity terms are added to vxb, vyb, vzb
VF_HD void backward_chimera_rev(const float (&N)[27], float vx, float vy, float vz, float (&A)[27], float& vxb,
                                float& vyb, float& vzb)
{
    {
        float Y[27];  // input of the z pass = N after the x and y passes
#pragma unroll
        for (int i = 0; i < 27; ++i) Y[i] = N[i];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) line_bwd(Y[idx(0, j, k)], Y[idx(1, j, k)], Y[idx(2, j, k)], vx, kv(j) * kv(k));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) line_bwd(Y[idx(i, 0, k)], Y[idx(i, 1, k)], Y[idx(i, 2, k)], vy, kw(i) * kv(k));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                line_bwd_rev(Y[idx(i, j, 0)], Y[idx(i, j, 1)], vz, kw(i) * kw(j), A[idx(i, j, 0)], A[idx(i, j, 1)],
                             A[idx(i, j, 2)], vzb);
    }
    {
        float X[27];  // input of the y pass = N after the x pass
#pragma unroll
        for (int i = 0; i < 27; ++i) X[i] = N[i];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) line_bwd(X[idx(0, j, k)], X[idx(1, j, k)], X[idx(2, j, k)], vx, kv(j) * kv(k));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                line_bwd_rev(X[idx(i, 0, k)], X[idx(i, 1, k)], vy, kw(i) * kv(k), A[idx(i, 0, k)], A[idx(i, 1, k)],
                             A[idx(i, 2, k)], vyb);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            line_bwd_rev(N[idx(0, j, k)], N[idx(1, j, k)], vx, kv(j) * kv(k), A[idx(0, j, k)], A[idx(1, j, k)],
                         A[idx(2, j, k)], vxb);
}

} // namespace adj

// Reverse-mode of k17_collide for one node. f: shifted pre-collision populations (direction
// order); g: adjoint of the post-collision populations (direction order). fb receives the
// adjoint of f (direction order); ub (optional) receives the adjoint of the velocity shift
// (r.fx, r.fy, r.fz); sb (optional) the adjoint of the velocity scale vscale.
VF_HD void k17_vjp(const float (&f)[27], const float (&g)[27], const K17Rates& r, float (&fb)[27], float* ub = nullptr,
                   float vscale = 1.0f, float* sb = nullptr)
{
    using namespace adj;
    float p[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) p[i] = f[dir_of_idx(i)];
    Macro mc;
    macro_of(p, r, vscale, mc);

    // forward to the moments m, and the cumulant stage
    float m[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) m[i] = p[i];
    pass_z(m, mc.vz);
    pass_y(m, mc.vy);
    pass_x(m, mc.vx);
    Mid s;
    float N[27];
    mid_forward(m, mc, r, s, N);

    // ---- reverse of the backward chimera (x, y, z): A = adjoints -------------------------
    float vxb = 0.0f, vyb = 0.0f, vzb = 0.0f;
    float A[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) A[i] = g[dir_of_idx(i)];
    backward_chimera_rev(N, mc.vx, mc.vy, mc.vz, A, vxb, vyb, vzb);

    // ---- reverse of the cumulant stage ---------------------------------------------------
    float mb[27];
    float drhob = 0.0f, Rb = 0.0f, vx2b = 0.0f, vy2b = 0.0f, vz2b = 0.0f;
    mid_reverse(s, mc, r, A, mb, drhob, Rb, vx2b, vy2b, vz2b);
    vxb += 2.0f * mc.vx * vx2b;
    vyb += 2.0f * mc.vy * vy2b;
    vzb += 2.0f * mc.vz * vz2b;

    // ---- reverse of the forward chimera (x, y, z) -------------------------------------------
    // the x pass reads the y pass output, the y pass the z pass output
    float Zo[27], Yo[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) Zo[i] = p[i];
    pass_z(Zo, mc.vz);
#pragma unroll
    for (int i = 0; i < 27; ++i) Yo[i] = Zo[i];
    pass_y(Yo, mc.vy);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            line_fwd_rev(Yo[idx(0, j, k)], Yo[idx(1, j, k)], Yo[idx(2, j, k)], mc.vx, kv(j) * kv(k), mb[idx(0, j, k)],
                         mb[idx(1, j, k)], mb[idx(2, j, k)], vxb);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            line_fwd_rev(Zo[idx(i, 0, k)], Zo[idx(i, 1, k)], Zo[idx(i, 2, k)], mc.vy, kw(i) * kv(k), mb[idx(i, 0, k)],
                         mb[idx(i, 1, k)], mb[idx(i, 2, k)], vyb);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            line_fwd_rev(p[idx(i, j, 0)], p[idx(i, j, 1)], p[idx(i, j, 2)], mc.vz, kw(i) * kw(j), mb[idx(i, j, 0)],
                         mb[idx(i, j, 1)], mb[idx(i, j, 2)], vzb);

    // ---- reverse of the macroscopic values ----------------------------------------------
    if (ub) {
        ub[0] = vxb;
        ub[1] = vyb;
        ub[2] = vzb;
    }
    const float rs = mc.R * vscale;
    const float jxb = vxb * rs, jyb = vyb * rs, jzb = vzb * rs;
    const float vj = vxb * mc.jx + vyb * mc.jy + vzb * mc.jz;
    Rb += vscale * vj;
    if (sb) *sb = mc.R * vj;
    drhob += -Rb * mc.R * mc.R;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float sb = drhob + float(i - 1) * jxb + float(j - 1) * jyb;
            mb[idx(i, j, 0)] += sb - jzb;
            mb[idx(i, j, 1)] += sb;
            mb[idx(i, j, 2)] += sb + jzb;
        }
#pragma unroll
    for (int i = 0; i < 27; ++i) fb[dir_of_idx(i)] = mb[i];
}

} // namespace vfa
