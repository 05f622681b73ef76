// FP32 D3Q27 cumulant K17 collision for gfx950, on shifted populations f_i - w_i.
//
// Same mathematics as vf::lbm::runK17CompressibleNavierStokes<TurbulenceModel::None>
// (src/lbm/collision/K17CompressibleNavierStokes.h:87-591; Geier et al. 2017,
// JCP 10.1016/j.jcp.2017.05.040), rearranged for a memory-bound GPU kernel:
//   * every omega-only quantity (relaxation rates, factorA/B terms, limiter numerators)
//     is computed once per launch on the host (K17Rates);
//   * with O4 = O5 = O6 = 1 the pre-collision cumulants of order 4-6 are replaced, not
//     relaxed, so neither they nor the moments that only feed them (m220, m202, m022,
//     m211, m121, m112, m221, m212, m122, m222) are formed;
//   * the limited third-order relaxation  x <- x - w x,  w = O + (1-O)|x|/(|x|+L)
//     is written  x <- x * (1-O) L / (|x| + L)  (one reciprocal);
//   * chimera lines carry the lattice weight K additively (f + K is the unshifted
//     partial moment), which is the reference's K / 1/K form with K * (1/K) = 1.
// Results agree with the FP64 restatement to FP32 round-off; the tests state the bound.
#pragma once

#include "vf_common.h"

namespace vfa {

// Per-launch constants derived from omega on the host (make_k17_rates below).
struct K17Rates {
    float omega;         // w
    float oneMinusOmega; // 1 - w
    float mHalfOmega;    // -w/2
    float threeHalfOmega;// 3w/2
    float mThreeOmega;   // -3w
    float galileo;       // 3 (1 - w/2): Galilean correction factor of the deviatoric part
    float limNumP, limNumM, limNumD; // (1-O) L for the P, M and D limiter groups
    float limP, limM, limD;          // quadric limiters L
    float A4;            // -(1/w - 1/2) (2/3) factorA: diagonal 4th-order equilibrium per unit rho*du
    float B4;            // -(1/w - 1/2) (1/3) factorB * (-3w): off-diagonal 4th order per unit m_ab
    float fx, fy, fz;    // half-step velocity shift (forces * 1/2 * forceFactor)
};

// The collision helpers compile for the host as well (the CPU tests evaluate the same code).
#define VF_HD __host__ __device__ __forceinline__

VF_HD float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// forward chimera line: a,b,c -> zeroth, first, second central moment along the line
template <int KNUM, int KDEN>
VF_HD void fwd(float& a, float& b, float& c, float u, float u2, float m2u)
{
    const float e = a + c;
    const float d = c - a;
    const float s = e + b;
    const float sK = KNUM ? s + float(KNUM) / float(KDEN) : s;
    a = s;
    b = __builtin_fmaf(-u, sK, d);
    c = __builtin_fmaf(m2u, d, __builtin_fmaf(sK, u2, e));
}
// forward line when only the zeroth and first moments are used later
template <int KNUM, int KDEN>
VF_HD void fwd01(float& a, float& b, float c, float u)
{
    const float d = c - a;
    const float s = (a + c) + b;
    const float sK = KNUM ? s + float(KNUM) / float(KDEN) : s;
    a = s;
    b = __builtin_fmaf(-u, sK, d);
}
// Per-pass coefficients of the backward chimera line (u = velocity component):
//   lo  = a0 (a+K) + b0 b + c/2,   hi = a2 (a+K) + b2 b + c/2,   mid = a - lo - hi
// with a0 = (u^2-u)/2, b0 = u - 1/2, a2 = (u^2+u)/2, b2 = u + 1/2. The three populations
// of a line sum to its zeroth moment a, which fixes mid exactly.
struct BwdPass {
    float a0, b0, a2, b2;
};
VF_HD BwdPass bwd_pass(float u, float u2)
{
    BwdPass p;
    p.a0 = 0.5f * (u2 - u);
    p.b0 = u - 0.5f;
    p.a2 = 0.5f * (u2 + u);
    p.b2 = u + 0.5f;
    return p;
}
template <int KNUM, int KDEN>
VF_HD void bwd(float& a, float& b, float& c, const BwdPass& p)
{
    constexpr float K = float(KNUM) / float(KDEN);
    const float hc = 0.5f * c;
    const float lo = __builtin_fmaf(p.a0, a, __builtin_fmaf(p.b0, b, KNUM ? hc + K * p.a0 : hc));
    const float hi = __builtin_fmaf(p.a2, a, __builtin_fmaf(p.b2, b, KNUM ? hc + K * p.a2 : hc));
    const float mid = (a - lo) - hi;
    a = lo;
    b = mid;
    c = hi;
}

// Per-launch constants from the level's omega, limiters and force (double precision on the
// host; relaxation rates of src/lbm/collision/K17CompressibleNavierStokes.h:251-273).
inline K17Rates make_k17_rates(const vf_lbm_parameter& p)
{
    const double w = p.omega;
    const double OxyyPxzz = 8.0 * (-2.0 + w) * (1.0 + 2.0 * w) / (-8.0 - 14.0 * w + 7.0 * w * w);
    const double OxyyMxzz = 8.0 * (-2.0 + w) * (-7.0 + 4.0 * w) / (56.0 - 50.0 * w + 9.0 * w * w);
    const double Oxyz = 24.0 * (-2.0 + w) * (-2.0 - 7.0 * w + 3.0 * w * w) /
                        (48.0 + 152.0 * w - 130.0 * w * w + 29.0 * w * w * w);
    const double factorA = (4.0 + 2.0 * w - 3.0 * w * w) / (2.0 - 7.0 * w + 5.0 * w * w);
    const double factorB = (4.0 + 28.0 * w - 14.0 * w * w) / (6.0 - 21.0 * w + 15.0 * w * w);
    const double nu = 1.0 / w - 0.5;
    K17Rates r;
    r.omega = (float)w;
    r.oneMinusOmega = (float)(1.0 - w);
    r.mHalfOmega = (float)(-0.5 * w);
    r.threeHalfOmega = (float)(1.5 * w);
    r.mThreeOmega = (float)(-3.0 * w);
    r.galileo = (float)(3.0 * (1.0 - 0.5 * w));
    r.limP = p.quadricLimiters[0];
    r.limM = p.quadricLimiters[1];
    r.limD = p.quadricLimiters[2];
    r.limNumP = (float)((1.0 - OxyyPxzz) * (double)p.quadricLimiters[0]);
    r.limNumM = (float)((1.0 - OxyyMxzz) * (double)p.quadricLimiters[1]);
    r.limNumD = (float)((1.0 - Oxyz) * (double)p.quadricLimiters[2]);
    r.A4 = (float)(-nu * (2.0 / 3.0) * factorA);
    r.B4 = (float)(-nu * (1.0 / 3.0) * factorB * (-3.0 * w));
    r.fx = p.forces[0] * 0.5f * p.forceFactor;
    r.fy = p.forces[1] * 0.5f * p.forceFactor;
    r.fz = p.forces[2] * 0.5f * p.forceFactor;
    return r;
}

struct MacroOut {
    float drho, vx, vy, vz;
};

// f: 27 shifted pre-collision populations in D3Q27 order; replaced by post-collision values.
// vscale: the Brinkman factor 1 - porosity/2 that scales the collision velocity (1: none)
VF_HD void k17_collide(float (&f)[27], const K17Rates& r, MacroOut& mo, float vscale = 1.0f)
{
    float& fMMM = f[dMMM]; float& f0MM = f[d0MM]; float& fPMM = f[dPMM];
    float& fM0M = f[dM0M]; float& f00M = f[d00M]; float& fP0M = f[dP0M];
    float& fMPM = f[dMPM]; float& f0PM = f[d0PM]; float& fPPM = f[dPPM];
    float& fMM0 = f[dMM0]; float& f0M0 = f[d0M0]; float& fPM0 = f[dPM0];
    float& fM00 = f[dM00]; float& f000 = f[d000]; float& fP00 = f[dP00];
    float& fMP0 = f[dMP0]; float& f0P0 = f[d0P0]; float& fPP0 = f[dPP0];
    float& fMMP = f[dMMP]; float& f0MP = f[d0MP]; float& fPMP = f[dPMP];
    float& fM0P = f[dM0P]; float& f00P = f[d00P]; float& fP0P = f[dP0P];
    float& fMPP = f[dMPP]; float& f0PP = f[d0PP]; float& fPPP = f[dPPP];

    // ---- z pass, velocity-free half: per (x,y) line e = a+c, d = c-a, s = e+b ---------------
    // The nine line sums s carry the density and the x/y momenta, the nine differences d the
    // z momentum, so the macroscopic values come out of the first chimera pass.
    float eMM, dMM, sMM, eM0, dM0, sM0, eMP, dMP, sMP;
    float e0M, d0M, s0M, e00, d00, s00, e0P, d0P, s0P;
    float ePM, dPM, sPM, eP0, dP0, sP0, ePP, dPP, sPP;
    auto zsum = [](float a, float b, float c, float& e, float& d, float& s) {
        e = a + c;
        d = c - a;
        s = e + b;
    };
    zsum(fMMM, fMM0, fMMP, eMM, dMM, sMM);
    zsum(fM0M, fM00, fM0P, eM0, dM0, sM0);
    zsum(fMPM, fMP0, fMPP, eMP, dMP, sMP);
    zsum(f0MM, f0M0, f0MP, e0M, d0M, s0M);
    zsum(f00M, f000, f00P, e00, d00, s00);
    zsum(f0PM, f0P0, f0PP, e0P, d0P, s0P);
    zsum(fPMM, fPM0, fPMP, ePM, dPM, sPM);
    zsum(fP0M, fP00, fP0P, eP0, dP0, sP0);
    zsum(fPPM, fPP0, fPPP, ePP, dPP, sPP);
    const float colM = (sMM + sMP) + sM0, colP = (sPM + sPP) + sP0, col0 = (s0M + s0P) + s00;
    const float drho = (colM + colP) + col0;
    const float jx = colP - colM;
    const float jy = ((sMP + sPP) + s0P) - ((sMM + sPM) + s0M);
    const float jz = (((dMM + dMP) + (dPM + dPP)) + ((dM0 + dP0) + (d0M + d0P))) + d00;

    const float oneOverRho = fast_rcp(drho + 1.0f);
    const float rs = oneOverRho * vscale;
    const float vx = __builtin_fmaf(jx, rs, r.fx);
    const float vy = __builtin_fmaf(jy, rs, r.fy);
    const float vz = __builtin_fmaf(jz, rs, r.fz);
    const float vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;

    // ---- z pass, velocity half: b = d - (s+K) u, c = e - 2 u d + (s+K) u^2 ------------------
    {
        const float m2u = -2.0f * vz;
        auto zfin = [&](float& a, float& b, float& c, float e, float d, float sv, float K) {
            const float sK = sv + K;
            a = sv;
            b = __builtin_fmaf(-vz, sK, d);
            c = __builtin_fmaf(m2u, d, __builtin_fmaf(sK, vz2, e));
        };
        zfin(fMMM, fMM0, fMMP, eMM, dMM, sMM, 1.0f / 36.0f);
        zfin(fM0M, fM00, fM0P, eM0, dM0, sM0, 1.0f / 9.0f);
        zfin(fMPM, fMP0, fMPP, eMP, dMP, sMP, 1.0f / 36.0f);
        zfin(f0MM, f0M0, f0MP, e0M, d0M, s0M, 1.0f / 9.0f);
        zfin(f00M, f000, f00P, e00, d00, s00, 4.0f / 9.0f);
        zfin(f0PM, f0P0, f0PP, e0P, d0P, s0P, 1.0f / 9.0f);
        zfin(fPMM, fPM0, fPMP, ePM, dPM, sPM, 1.0f / 36.0f);
        zfin(fP0M, fP00, fP0P, eP0, dP0, sP0, 1.0f / 9.0f);
        zfin(fPPM, fPP0, fPPP, ePP, dPP, sPP, 1.0f / 36.0f);
    }
    {
        const float m2u = -2.0f * vy;
        fwd<1, 6>(fMMM, fM0M, fMPM, vy, vy2, m2u);
        fwd<0, 1>(fMM0, fM00, fMP0, vy, vy2, m2u);
        fwd01<1, 18>(fMMP, fM0P, fMPP, vy);
        fwd<2, 3>(f0MM, f00M, f0PM, vy, vy2, m2u);
        fwd<0, 1>(f0M0, f000, f0P0, vy, vy2, m2u);
        fwd01<2, 9>(f0MP, f00P, f0PP, vy);
        fwd<1, 6>(fPMM, fP0M, fPPM, vy, vy2, m2u);
        fwd<0, 1>(fPM0, fP00, fPP0, vy, vy2, m2u);
        fwd01<1, 18>(fPMP, fP0P, fPPP, vy);
    }
    {
        const float m2u = -2.0f * vx;
        fwd<1, 1>(fMMM, f0MM, fPMM, vx, vx2, m2u);  // m000 m100 m200
        fwd<0, 1>(fM0M, f00M, fP0M, vx, vx2, m2u);  // m010 m110 m210
        fwd01<1, 3>(fMPM, f0PM, fPPM, vx);          // m020 m120
        fwd<0, 1>(fMM0, f0M0, fPM0, vx, vx2, m2u);  // m001 m101 m201
        fwd01<0, 1>(fM00, f000, fP00, vx);          // m011 m111
        fMP0 = (fMP0 + fPP0) + f0P0;                // m021
        fwd01<1, 3>(fMMP, f0MP, fPMP, vx);          // m002 m102
        fM0P = (fM0P + fP0P) + f00P;                // m012
    }
    const float m000 = fMMM, m100 = f0MM, m200 = fPMM;
    const float m010 = fM0M, m110 = f00M, m210 = fP0M;
    const float m020 = fMPM, m120 = f0PM;
    const float m001 = fMM0, m101 = f0M0, m201 = fPM0;
    const float m011 = fM00, m111 = f000;
    const float m021 = fMP0;
    const float m002 = fMMP, m102 = f0MP;
    const float m012 = fM0P;

    // ---- second order: gradients (times rho) and relaxation with Galilean correction --------
    const float mxxPyyPzz = m200 + m020 + m002;
    float mxxMyy = m200 - m020;
    float mxxMzz = m200 - m002;
    const float dxux = __builtin_fmaf(r.mHalfOmega, mxxMyy + mxxMzz, 0.5f * (m000 - mxxPyyPzz));
    const float dyuy = __builtin_fmaf(r.threeHalfOmega, mxxMyy, dxux);
    const float dzuz = __builtin_fmaf(r.threeHalfOmega, mxxMzz, dxux);
    const float vx2dxux = vx2 * dxux;
    const float pTrace = m000 - 1.5f * (vx2dxux + __builtin_fmaf(vy2, dyuy, vz2 * dzuz));
    mxxMyy = __builtin_fmaf(r.oneMinusOmega, mxxMyy, -r.galileo * (vx2dxux - vy2 * dyuy));
    mxxMzz = __builtin_fmaf(r.oneMinusOmega, mxxMzz, -r.galileo * (vx2dxux - vz2 * dzuz));
    const float n110 = r.oneMinusOmega * m110;
    const float n101 = r.oneMinusOmega * m101;
    const float n011 = r.oneMinusOmega * m011;
    const float n200 = (1.0f / 3.0f) * (mxxMyy + mxxMzz + pTrace);
    const float n020 = (1.0f / 3.0f) * (__builtin_fmaf(-2.0f, mxxMyy, mxxMzz) + pTrace);
    const float n002 = (1.0f / 3.0f) * (__builtin_fmaf(-2.0f, mxxMzz, mxxMyy) + pTrace);

    // ---- third order with quadric limiters ------------------------------------------------
    const float n111 = m111 * r.limNumD * fast_rcp(__builtin_fabsf(m111) + r.limD);
    float c1 = m210 + m012, c2 = m210 - m012;
    float c3 = m201 + m021, c4 = m201 - m021;
    float c5 = m120 + m102, c6 = m120 - m102;
    c1 *= r.limNumP * fast_rcp(__builtin_fabsf(c1) + r.limP);
    c2 *= r.limNumM * fast_rcp(__builtin_fabsf(c2) + r.limM);
    c3 *= r.limNumP * fast_rcp(__builtin_fabsf(c3) + r.limP);
    c4 *= r.limNumM * fast_rcp(__builtin_fabsf(c4) + r.limM);
    c5 *= r.limNumP * fast_rcp(__builtin_fabsf(c5) + r.limP);
    c6 *= r.limNumM * fast_rcp(__builtin_fabsf(c6) + r.limM);
    const float n210 = 0.5f * (c1 + c2), n012 = 0.5f * (c1 - c2);
    const float n201 = 0.5f * (c3 + c4), n021 = 0.5f * (c3 - c4);
    const float n120 = 0.5f * (c5 + c6), n102 = 0.5f * (c5 - c6);

    // ---- fourth order: equilibrium cumulants turned back into central moments -----------
    const float q110 = n110 * oneOverRho, q101 = n101 * oneOverRho, q011 = n011 * oneOverRho;
    const float n211 = __builtin_fmaf(r.B4, m011, (1.0f / 3.0f) * __builtin_fmaf(__builtin_fmaf(3.0f, n200, 1.0f), q011, 6.0f * n110 * q101));
    const float n121 = __builtin_fmaf(r.B4, m101, (1.0f / 3.0f) * __builtin_fmaf(__builtin_fmaf(3.0f, n020, 1.0f), q101, 6.0f * n110 * q011));
    const float n112 = __builtin_fmaf(r.B4, m110, (1.0f / 3.0f) * __builtin_fmaf(__builtin_fmaf(3.0f, n002, 1.0f), q110, 6.0f * n101 * q011));
    const float rhoTerm = drho * oneOverRho * (1.0f / 9.0f);
    const float n220 = __builtin_fmaf(r.A4, dxux + dyuy,
                        __builtin_fmaf(__builtin_fmaf(n200, n020, 2.0f * n110 * n110), oneOverRho,
                                       (1.0f / 3.0f) * (n200 + n020) * oneOverRho) - rhoTerm);
    const float n202 = __builtin_fmaf(r.A4, dxux + dzuz,
                        __builtin_fmaf(__builtin_fmaf(n200, n002, 2.0f * n101 * n101), oneOverRho,
                                       (1.0f / 3.0f) * (n200 + n002) * oneOverRho) - rhoTerm);
    const float n022 = __builtin_fmaf(r.A4, dyuy + dzuz,
                        __builtin_fmaf(__builtin_fmaf(n002, n020, 2.0f * n011 * n011), oneOverRho,
                                       (1.0f / 3.0f) * (n002 + n020) * oneOverRho) - rhoTerm);

    // ---- fifth order (cumulants relaxed to zero) -----------------------------------------
    const float n122 = (__builtin_fmaf(3.0f, n002 * n120 + n020 * n102 + 4.0f * n011 * n111 + 2.0f * (n101 * n021 + n110 * n012),
                                       n120 + n102) * (1.0f / 3.0f)) * oneOverRho;
    const float n212 = (__builtin_fmaf(3.0f, n002 * n210 + n200 * n012 + 4.0f * n101 * n111 + 2.0f * (n011 * n201 + n110 * n102),
                                       n210 + n012) * (1.0f / 3.0f)) * oneOverRho;
    const float n221 = (__builtin_fmaf(3.0f, n200 * n021 + n020 * n201 + 4.0f * n110 * n111 + 2.0f * (n101 * n120 + n011 * n210),
                                       n021 + n201) * (1.0f / 3.0f)) * oneOverRho;

    // ---- sixth order (cumulant relaxed to zero) -------------------------------------------
    const float sq110 = n110 * n110, sq101 = n101 * n101, sq011 = n011 * n011;
    const float tr2 = n200 + n020 + n002;
    const float t1 = -4.0f * n111 * n111 - (n200 * n022 + n020 * n202 + n002 * n220) -
                     4.0f * (n011 * n211 + n101 * n121 + n110 * n112) -
                     2.0f * (n120 * n102 + n210 * n012 + n201 * n021) -
                     (1.0f / 3.0f) * (n022 + n202 + n220) - (1.0f / 9.0f) * tr2;
    const float t2 = 4.0f * (sq101 * n020 + sq011 * n200 + sq110 * n002) + 2.0f * (n200 * n020 * n002) +
                     16.0f * n110 * n101 * n011 +
                     (2.0f / 3.0f) * (2.0f * (sq101 + sq011 + sq110) + (n002 * n020 + n002 * n200 + n020 * n200) +
                                      (1.0f / 3.0f) * tr2) +
                     (1.0f / 27.0f) * (drho * drho - drho);
    const float n222 = -(t1 + t2 * oneOverRho) * oneOverRho;

    // ---- backward chimera: x, y, z ---------------------------------------------------------
    float M000 = m000, M100 = -m100, M200 = n200;
    float M010 = -m010, M110 = n110, M210 = n210;
    float M020 = n020, M120 = n120, M220 = n220;
    float M001 = -m001, M101 = n101, M201 = n201;
    float M011 = n011, M111 = n111, M211 = n211;
    float M021 = n021, M121 = n121, M221 = n221;
    float M002 = n002, M102 = n102, M202 = n202;
    float M012 = n012, M112 = n112, M212 = n212;
    float M022 = n022, M122 = n122, M222 = n222;
    {
        const BwdPass p = bwd_pass(vx, vx2);
        bwd<1, 1>(M000, M100, M200, p);
        bwd<0, 1>(M010, M110, M210, p);
        bwd<1, 3>(M020, M120, M220, p);
        bwd<0, 1>(M001, M101, M201, p);
        bwd<0, 1>(M011, M111, M211, p);
        bwd<0, 1>(M021, M121, M221, p);
        bwd<1, 3>(M002, M102, M202, p);
        bwd<0, 1>(M012, M112, M212, p);
        bwd<1, 9>(M022, M122, M222, p);
    }
    {
        const BwdPass p = bwd_pass(vy, vy2);
        bwd<1, 6>(M000, M010, M020, p);
        bwd<0, 1>(M001, M011, M021, p);
        bwd<1, 18>(M002, M012, M022, p);
        bwd<2, 3>(M100, M110, M120, p);
        bwd<0, 1>(M101, M111, M121, p);
        bwd<2, 9>(M102, M112, M122, p);
        bwd<1, 6>(M200, M210, M220, p);
        bwd<0, 1>(M201, M211, M221, p);
        bwd<1, 18>(M202, M212, M222, p);
    }
    {
        const BwdPass p = bwd_pass(vz, vz2);
        bwd<1, 36>(M000, M001, M002, p);
        bwd<1, 9>(M010, M011, M012, p);
        bwd<1, 36>(M020, M021, M022, p);
        bwd<1, 9>(M100, M101, M102, p);
        bwd<4, 9>(M110, M111, M112, p);
        bwd<1, 9>(M120, M121, M122, p);
        bwd<1, 36>(M200, M201, M202, p);
        bwd<1, 9>(M210, M211, M212, p);
        bwd<1, 36>(M220, M221, M222, p);
    }
    // moment m_abc is the population of direction (a-1, b-1, c-1)
    fMMM = M000; f0MM = M100; fPMM = M200;
    fM0M = M010; f00M = M110; fP0M = M210;
    fMPM = M020; f0PM = M120; fPPM = M220;
    fMM0 = M001; f0M0 = M101; fPM0 = M201;
    fM00 = M011; f000 = M111; fP00 = M211;
    fMP0 = M021; f0P0 = M121; fPP0 = M221;
    fMMP = M002; f0MP = M102; fPMP = M202;
    fM0P = M012; f00P = M112; fP0P = M212;
    fMPP = M022; f0PP = M122; fPPP = M222;

    mo.drho = drho;
    mo.vx = vx;
    mo.vy = vy;
    mo.vz = vz;
}

} // namespace vfa
