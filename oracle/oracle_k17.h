// ORACLE — test infrastructure only. Not part of the product; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// CPU restatement of the D3Q27 cumulant K17 compressible collision on shifted
// ("well-conditioned") populations f_i - w_i, in the precision T (double or float).
//
// Follows, operation by operation so that the FP64 result agrees with the
// reference to the last bit:
//   moments        src/lbm/MacroscopicQuantities.h:55-133
//   chimera        src/lbm/ChimeraTransformation.h:58-147
//   collision      src/lbm/collision/K17CompressibleNavierStokes.h:87-591
// The chimera passes are written here as tables of (a,b,c) triples; the
// cumulant algebra keeps the reference's grouping of terms (Geier et al. 2017,
// JCP 10.1016/j.jcp.2017.05.040, Eqs. 20-56, 111-123).
#pragma once

#include <cmath>

#include "oracle_d3q27.h"

namespace orc {

template <class T> struct K17Param {
    T omega;
    T limP, limM, limD; // quadric limiters P, M, D (src/lbm/collision/K17CompressibleNavierStokes.h:177-179)
    T fx, fy, fz;       // half-step velocity shift (body force), already scaled by the caller
    T vscale = T(1);    // Brinkman factor 1 - porosity/2 on the velocity (no reference counterpart; 1: none)
};

template <class T> struct Macro {
    T drho, vx, vy, vz;
};

// ---- macroscopic moments (src/lbm/MacroscopicQuantities.h:55-64, 72-91) --------------------
template <class T> inline T density(const T* f)
{
    const T corners = ((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]));
    const T edges = ((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P])) +
                    ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]));
    const T faces = (f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0]) + (f[d00M] + f[d00P]);
    return ((corners + edges) + f[d000]) + faces;
}

template <class T> inline T momentum_x(const T* f)
{
    return (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
           (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
           (f[dP00] - f[dM00]);
}

template <class T> inline T momentum_y(const T* f)
{
    return (((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
           (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
           (f[d0P0] - f[d0M0]);
}

template <class T> inline T momentum_z(const T* f)
{
    return (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
           (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
           (f[d00P] - f[d00M]);
}

// getCompressibleMacroscopicValues (src/lbm/MacroscopicQuantities.h:116-133)
template <class T> inline void compressible_macro(const T* f, T& drho, T& inv_rho, T& vx, T& vy, T& vz)
{
    drho = density(f);
    vx = momentum_x(f);
    vy = momentum_y(f);
    vz = momentum_z(f);
    inv_rho = T(1) / (drho + T(1));
    vx *= inv_rho;
    vy *= inv_rho;
    vz *= inv_rho;
}

// getPressure (src/lbm/MacroscopicQuantities.h:139-150)
template <class T> inline T pressure(const T* f, T rho, T vx, T vy, T vz)
{
    const T faces = f[dP00] + f[dM00] + f[d0P0] + f[d0M0] + f[d00P] + f[d00M];
    const T edges = f[dPP0] + f[dMM0] + f[dPM0] + f[dMP0] + f[dP0P] + f[dM0M] + f[dP0M] + f[dM0P] + f[d0PP] +
                    f[d0MM] + f[d0PM] + f[d0MP];
    const T corners = f[dPPP] + f[dMMP] + f[dPMP] + f[dMPP] + f[dPPM] + f[dMMM] + f[dPMM] + f[dMPM];
    return (faces + T(2) * edges + T(3) * corners - rho - (vx * vx + vy * vy + vz * vz) * (T(1) + rho)) * T(0.5) +
           rho;
}

// ---- 1-D chimera transforms on one (a,b,c) line (src/lbm/ChimeraTransformation.h) ------------
// Kinv == 0 selects the K-free form (forwardChimera / backwardChimera).
template <class T> inline void chimera_fwd(T& a, T& b, T& c, T u, T u2, T Kinv, T K)
{
    if (Kinv == T(0)) { // src/lbm/ChimeraTransformation.h:94-102
        const T s = (a + c) + b;
        const T d = c - a;
        c = (c + a) + (u2 * s - T(2) * u * d);
        b = d - u * s;
        a = s;
        return;
    }
    // src/lbm/ChimeraTransformation.h:58-69
    const T e = a + c;
    const T d = c - a;
    T s = e + b;
    a = s;
    s *= Kinv;
    s += T(1);
    b = (d * Kinv - s * u) * K;
    c = ((e - T(2) * d * u) * Kinv + u2 * s) * K;
}

template <class T> inline void chimera_bwd(T& a, T& b, T& c, T u, T u2, T Kinv, T K)
{
    if (Kinv == T(0)) { // src/lbm/ChimeraTransformation.h:111-119
        const T lo = (c + a * (u2 - u)) * T(0.5) + b * (u - T(0.5));
        const T mid = ((a - c) - a * u2) - T(2) * b * u;
        c = (c + a * (u2 + u)) * T(0.5) + b * (u + T(0.5));
        b = mid;
        a = lo;
        return;
    }
    // src/lbm/ChimeraTransformation.h:77-86
    const T lo = (((c - b) * T(0.5) + b * u) * Kinv + (a * Kinv + T(1)) * (u2 - u) * T(0.5)) * K;
    const T mid = (((a - c) - T(2) * b * u) * Kinv + (a * Kinv + T(1)) * (-u2)) * K;
    c = (((c + b) * T(0.5) + b * u) * Kinv + (a * Kinv + T(1)) * (u2 + u) * T(0.5)) * K;
    a = lo;
    b = mid;
}

// One line of a chimera pass: three population slots plus the (1/K, K) pair.
// 1/K and K are kept as integer fractions so that T(num)/T(den) rounds exactly like
// the reference's own constants (e.g. 1.0f / 36.0f in the FP32 build).
struct ChimeraLine {
    int a, b, c;
    int kinvNum, kinvDen, kNum, kDen; // kinvNum == 0: K-free line
};

// Z pass: src/lbm/collision/K17CompressibleNavierStokes.h:190-198 (and 549-557 backwards)
static const ChimeraLine kPassZ[9] = {
    {dMMM, dMM0, dMMP, 36, 1, 1, 36}, {dM0M, dM00, dM0P, 9, 1, 1, 9}, {dMPM, dMP0, dMPP, 36, 1, 1, 36},
    {d0MM, d0M0, d0MP, 9, 1, 1, 9},   {d00M, d000, d00P, 9, 4, 4, 9}, {d0PM, d0P0, d0PP, 9, 1, 1, 9},
    {dPMM, dPM0, dPMP, 36, 1, 1, 36}, {dP0M, dP00, dP0P, 9, 1, 1, 9}, {dPPM, dPP0, dPPP, 36, 1, 1, 36},
};
// Y pass: src/lbm/collision/K17CompressibleNavierStokes.h:202-210 (and 537-545 backwards)
static const ChimeraLine kPassY[9] = {
    {dMMM, dM0M, dMPM, 6, 1, 1, 6},   {dMM0, dM00, dMP0, 0, 1, 0, 1},          {dMMP, dM0P, dMPP, 18, 1, 1, 18},
    {d0MM, d00M, d0PM, 3, 2, 2, 3}, {d0M0, d000, d0P0, 0, 1, 0, 1},       {d0MP, d00P, d0PP, 9, 2, 2, 9},
    {dPMM, dP0M, dPPM, 6, 1, 1, 6},   {dPM0, dP00, dPP0, 0, 1, 0, 1},          {dPMP, dP0P, dPPP, 18, 1, 1, 18},
};
// X pass, forward: src/lbm/collision/K17CompressibleNavierStokes.h:214-222.
// The last line uses (1/K, K) = (3, 1/9) in the reference; kept as is. Its outputs
// (m022's first and second x-moments, i.e. m122 and m222) only feed cumulants that
// K17 overwrites with O5 = O6 = 1, so the value does not reach the result.
static const ChimeraLine kPassXFwd[9] = {
    {dMMM, d0MM, dPMM, 1, 1, 1, 1},        {dM0M, d00M, dP0M, 0, 1, 0, 1},          {dMPM, d0PM, dPPM, 3, 1, 1, 3},
    {dMM0, d0M0, dPM0, 0, 1, 0, 1},        {dM00, d000, dP00, 0, 1, 0, 1},          {dMP0, d0P0, dPP0, 0, 1, 0, 1},
    {dMMP, d0MP, dPMP, 3, 1, 1, 3},   {dM0P, d00P, dP0P, 0, 1, 0, 1},          {dMPP, d0PP, dPPP, 3, 1, 1, 9},
};
// X pass, backward: src/lbm/collision/K17CompressibleNavierStokes.h:525-533
static const ChimeraLine kPassXBwd[9] = {
    {dMMM, d0MM, dPMM, 1, 1, 1, 1},        {dM0M, d00M, dP0M, 0, 1, 0, 1},          {dMPM, d0PM, dPPM, 3, 1, 1, 3},
    {dMM0, d0M0, dPM0, 0, 1, 0, 1},        {dM00, d000, dP00, 0, 1, 0, 1},          {dMP0, d0P0, dPP0, 0, 1, 0, 1},
    {dMMP, d0MP, dPMP, 3, 1, 1, 3},   {dM0P, d00P, dP0P, 0, 1, 0, 1},          {dMPP, d0PP, dPPP, 9, 1, 1, 9},
};

template <class T> inline void pass_fwd(T* f, const ChimeraLine* lines, T u, T u2)
{
    for (int i = 0; i < 9; ++i)
        chimera_fwd(f[lines[i].a], f[lines[i].b], f[lines[i].c], u, u2, T(lines[i].kinvNum) / T(lines[i].kinvDen), T(lines[i].kNum) / T(lines[i].kDen));
}

template <class T> inline void pass_bwd(T* f, const ChimeraLine* lines, T u, T u2)
{
    for (int i = 0; i < 9; ++i)
        chimera_bwd(f[lines[i].a], f[lines[i].b], f[lines[i].c], u, u2, T(lines[i].kinvNum) / T(lines[i].kinvDen), T(lines[i].kNum) / T(lines[i].kDen));
}

// ---- the collision itself -----------------------------------------------------------------
// After the forward passes the slot of direction (cx,cy,cz) holds the central moment
// m_{abc} with a = cx+1, b = cy+1, c = cz+1 (src/lbm/collision/K17CompressibleNavierStokes.h:121-147).
template <class T> void k17_collide(T* f, const K17Param<T>& p, Macro<T>* out)
{
    T drho, oneOverRho, vvx, vvy, vvz;
    compressible_macro(f, drho, oneOverRho, vvx, vvy, vvz);
    vvx = vvx * p.vscale + p.fx; // vscale = 1 leaves the value unchanged bit for bit
    vvy = vvy * p.vscale + p.fy;
    vvz = vvz * p.vscale + p.fz;
    const T vx2 = vvx * vvx, vy2 = vvy * vvy, vz2 = vvz * vvz;

    pass_fwd(f, kPassZ, vvz, vz2);
    pass_fwd(f, kPassY, vvy, vy2);
    pass_fwd(f, kPassXFwd, vvx, vx2);

    T &m000 = f[dMMM], &m100 = f[d0MM], &m200 = f[dPMM];
    T &m010 = f[dM0M], &m110 = f[d00M], &m210 = f[dP0M];
    T &m020 = f[dMPM], &m120 = f[d0PM], &m220 = f[dPPM];
    T &m001 = f[dMM0], &m101 = f[d0M0], &m201 = f[dPM0];
    T &m011 = f[dM00], &m111 = f[d000], &m211 = f[dP00];
    T &m021 = f[dMP0], &m121 = f[d0P0], &m221 = f[dPP0];
    T &m002 = f[dMMP], &m102 = f[d0MP], &m202 = f[dPMP];
    T &m012 = f[dM0P], &m112 = f[d00P], &m212 = f[dP0P];
    T &m022 = f[dMPP], &m122 = f[d0PP], &m222 = f[dPPP];

    const T omega = p.omega;
    const T c1 = T(1), c2 = T(2), c3 = T(3), c4 = T(4), c6 = T(6), c9 = T(9), c16 = T(16);
    const T c1o2 = T(1) / T(2), c1o3 = T(1) / T(3), c2o3 = T(2) / T(3), c1o9 = T(1) / T(9), c1o27 = T(1) / T(27);
    const T c3o2 = T(3) / T(2);

    // relaxation rates, simplified for a unit bulk rate (Eqs. 111-113, 115-116)
    const T OxxPyyPzz = c1;
    const T OxyyPxzz = T(8) * (-c2 + omega) * (c1 + c2 * omega) / (-T(8) - T(14) * omega + T(7) * omega * omega);
    const T OxyyMxzz = T(8) * (-c2 + omega) * (-T(7) + c4 * omega) / (T(56) - T(50) * omega + c9 * omega * omega);
    const T Oxyz = T(24) * (-c2 + omega) * (-c2 - T(7) * omega + c3 * omega * omega) /
                   (T(48) + T(152) * omega - T(130) * omega * omega + T(29) * omega * omega * omega);
    const T O4 = c1, O5 = c1, O6 = c1;
    const T factorA = (c4 + c2 * omega - c3 * omega * omega) / (c2 - T(7) * omega + T(5) * omega * omega);
    const T factorB = (c4 + T(28) * omega - T(14) * omega * omega) / (c6 - T(21) * omega + T(15) * omega * omega);

    // cumulants of order 4-6 from central moments (Eqs. 20-23)
    T cm211 = m211 - ((m200 + c1o3) * m011 + c2 * m110 * m101) * oneOverRho;
    T cm121 = m121 - ((m020 + c1o3) * m101 + c2 * m110 * m011) * oneOverRho;
    T cm112 = m112 - ((m002 + c1o3) * m110 + c2 * m101 * m011) * oneOverRho;
    T cm220 = m220 - (((m200 * m020 + c2 * m110 * m110) + c1o3 * (m200 + m020)) * oneOverRho - c1o9 * (drho * oneOverRho));
    T cm202 = m202 - (((m200 * m002 + c2 * m101 * m101) + c1o3 * (m200 + m002)) * oneOverRho - c1o9 * (drho * oneOverRho));
    T cm022 = m022 - (((m002 * m020 + c2 * m011 * m011) + c1o3 * (m002 + m020)) * oneOverRho - c1o9 * (drho * oneOverRho));
    T cm122 = m122 - ((m002 * m120 + m020 * m102 + c4 * m011 * m111 + c2 * (m101 * m021 + m110 * m012)) +
                      c1o3 * (m120 + m102)) * oneOverRho;
    T cm212 = m212 - ((m002 * m210 + m200 * m012 + c4 * m101 * m111 + c2 * (m011 * m201 + m110 * m102)) +
                      c1o3 * (m210 + m012)) * oneOverRho;
    T cm221 = m221 - ((m200 * m021 + m020 * m201 + c4 * m110 * m111 + c2 * (m101 * m120 + m011 * m210)) +
                      c1o3 * (m021 + m201)) * oneOverRho;
    T cm222 = m222 + ((-c4 * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
                       c4 * (m011 * m211 + m101 * m121 + m110 * m112) -
                       c2 * (m120 * m102 + m210 * m012 + m201 * m021)) * oneOverRho +
                      (c4 * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) +
                       c2 * (m200 * m020 * m002) + c16 * m110 * m101 * m011) * oneOverRho * oneOverRho -
                      c1o3 * (m022 + m202 + m220) * oneOverRho - c1o9 * (m200 + m020 + m002) * oneOverRho +
                      (c2 * (m101 * m101 + m011 * m011 + m110 * m110) +
                       (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * (m002 + m020 + m200)) *
                          oneOverRho * oneOverRho * c2o3 +
                      c1o27 * ((drho * drho - drho) * oneOverRho * oneOverRho));

    // 2nd/3rd order linear combinations
    T mxxPyyPzz = m200 + m020 + m002;
    T mxxMyy = m200 - m020;
    T mxxMzz = m200 - m002;
    T mxxyPyzz = m210 + m012, mxxyMyzz = m210 - m012;
    T mxxzPyyz = m201 + m021, mxxzMyyz = m201 - m021;
    T mxyyPxzz = m120 + m102, mxyyMxzz = m120 - m102;

    // velocity gradients times rho (Eqs. 27-32)
    const T Dxy = -c3 * omega * m110;
    const T Dxz = -c3 * omega * m101;
    const T Dyz = -c3 * omega * m011;
    const T dxux = c1o2 * (-omega) * (mxxMyy + mxxMzz) + c1o2 * OxxPyyPzz * (m000 - mxxPyyPzz);
    const T dyuy = dxux + omega * c3o2 * mxxMyy;
    const T dzuz = dxux + omega * c3o2 * mxxMzz;

    // relax 2nd order with Galilean correction (Eqs. 33-35)
    mxxPyyPzz += OxxPyyPzz * (m000 - mxxPyyPzz) - c3 * (c1 - c1o2 * OxxPyyPzz) * (vx2 * dxux + vy2 * dyuy + vz2 * dzuz);
    mxxMyy += omega * (-mxxMyy) - c3 * (c1 + c1o2 * (-omega)) * (vx2 * dxux - vy2 * dyuy);
    mxxMzz += omega * (-mxxMzz) - c3 * (c1 + c1o2 * (-omega)) * (vx2 * dxux - vz2 * dzuz);
    m011 += omega * (-m011);
    m101 += omega * (-m101);
    m110 += omega * (-m110);

    // limited relaxation of 3rd order (Eqs. 116-123)
    auto limited = [](T rate, T x, T lim) {
        using std::abs;
        return rate + (T(1) - rate) * abs(x) / (abs(x) + lim);
    };
    T w = limited(Oxyz, m111, p.limD);
    m111 += w * (-m111);
    w = limited(OxyyPxzz, mxxyPyzz, p.limP);
    mxxyPyzz += w * (-mxxyPyzz);
    w = limited(OxyyMxzz, mxxyMyzz, p.limM);
    mxxyMyzz += w * (-mxxyMyzz);
    w = limited(OxyyPxzz, mxxzPyyz, p.limP);
    mxxzPyyz += w * (-mxxzPyyz);
    w = limited(OxyyMxzz, mxxzMyyz, p.limM);
    mxxzMyyz += w * (-mxxzMyyz);
    w = limited(OxyyPxzz, mxyyPxzz, p.limP);
    mxyyPxzz += w * (-mxyyPxzz);
    w = limited(OxyyMxzz, mxyyMxzz, p.limM);
    mxyyMxzz += w * (-mxyyMxzz);

    // back to moments
    m200 = c1o3 * (mxxMyy + mxxMzz + mxxPyyPzz);
    m020 = c1o3 * (-c2 * mxxMyy + mxxMzz + mxxPyyPzz);
    m002 = c1o3 * (mxxMyy - c2 * mxxMzz + mxxPyyPzz);
    m210 = (mxxyMyzz + mxxyPyzz) * c1o2;
    m012 = (-mxxyMyzz + mxxyPyzz) * c1o2;
    m201 = (mxxzMyyz + mxxzPyyz) * c1o2;
    m021 = (-mxxzMyyz + mxxzPyyz) * c1o2;
    m120 = (mxyyMxzz + mxyyPxzz) * c1o2;
    m102 = (-mxyyMxzz + mxyyPxzz) * c1o2;

    // 4th order to the modified equilibrium (Eqs. 43-48); 5th and 6th to zero
    const T nuTerm = (c1 / omega - c1o2);
    cm022 = -O4 * nuTerm * (dyuy + dzuz) * c2o3 * factorA + (c1 - O4) * (cm022);
    cm202 = -O4 * nuTerm * (dxux + dzuz) * c2o3 * factorA + (c1 - O4) * (cm202);
    cm220 = -O4 * nuTerm * (dyuy + dxux) * c2o3 * factorA + (c1 - O4) * (cm220);
    cm112 = -O4 * nuTerm * Dxy * c1o3 * factorB + (c1 - O4) * (cm112);
    cm121 = -O4 * nuTerm * Dxz * c1o3 * factorB + (c1 - O4) * (cm121);
    cm211 = -O4 * nuTerm * Dyz * c1o3 * factorB + (c1 - O4) * (cm211);
    cm122 += O5 * (-cm122);
    cm212 += O5 * (-cm212);
    cm221 += O5 * (-cm221);
    cm222 += O6 * (-cm222);

    // cumulants back to central moments (Eqs. 53-56)
    m211 = cm211 + c1o3 * ((c3 * m200 + c1) * m011 + c6 * m110 * m101) * oneOverRho;
    m121 = cm121 + c1o3 * ((c3 * m020 + c1) * m101 + c6 * m110 * m011) * oneOverRho;
    m112 = cm112 + c1o3 * ((c3 * m002 + c1) * m110 + c6 * m101 * m011) * oneOverRho;
    m220 = cm220 + (((m200 * m020 + c2 * m110 * m110) * c9 + c3 * (m200 + m020)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m202 = cm202 + (((m200 * m002 + c2 * m101 * m101) * c9 + c3 * (m200 + m002)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m022 = cm022 + (((m002 * m020 + c2 * m011 * m011) * c9 + c3 * (m002 + m020)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m122 = cm122 + c1o3 * (c3 * (m002 * m120 + m020 * m102 + c4 * m011 * m111 + c2 * (m101 * m021 + m110 * m012)) +
                           (m120 + m102)) * oneOverRho;
    m212 = cm212 + c1o3 * (c3 * (m002 * m210 + m200 * m012 + c4 * m101 * m111 + c2 * (m011 * m201 + m110 * m102)) +
                           (m210 + m012)) * oneOverRho;
    m221 = cm221 + c1o3 * (c3 * (m200 * m021 + m020 * m201 + c4 * m110 * m111 + c2 * (m101 * m120 + m011 * m210)) +
                           (m021 + m201)) * oneOverRho;
    m222 = cm222 - ((-c4 * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
                     c4 * (m011 * m211 + m101 * m121 + m110 * m112) -
                     c2 * (m120 * m102 + m210 * m012 + m201 * m021)) * oneOverRho +
                    (c4 * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) +
                     c2 * (m200 * m020 * m002) + c16 * m110 * m101 * m011) * oneOverRho * oneOverRho -
                    c1o3 * (m022 + m202 + m220) * oneOverRho - c1o9 * (m200 + m020 + m002) * oneOverRho +
                    (c2 * (m101 * m101 + m011 * m011 + m110 * m110) +
                     (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * (m002 + m020 + m200)) *
                        oneOverRho * oneOverRho * c2o3 +
                    c1o27 * ((drho * drho - drho) * oneOverRho * oneOverRho));

    // first-order sign flip (src/lbm/collision/K17CompressibleNavierStokes.h:512-514)
    m100 = -m100;
    m010 = -m010;
    m001 = -m001;

    pass_bwd(f, kPassXBwd, vvx, vx2);
    pass_bwd(f, kPassY, vvy, vy2);
    pass_bwd(f, kPassZ, vvz, vz2);

    if (out) {
        out->drho = drho;
        out->vx = vvx;
        out->vy = vvy;
        out->vz = vvz;
    }
}

} // namespace orc
