// Compressible grid-refinement interpolation of one interface cell, and its reverse mode.
//
//   arithmetic  src/lbm/interpolation/InterpolationCoefficients.h:63-432,
//               src/lbm/refinement/InterpolationCF.h:57-306, InterpolationFC.h:56-263
//   kernels     vf_scaling.hip (forward and reverse)
//
// The 8 source nodes of a cell sit at (sx, sy, sz)/2, n = ix + 2 iy + 4 iz (s = -1 for i = 0).
// Each gives drho, u and the non-equilibrium second moments; every coefficient of the
// interpolation polynomials is a short combination of the signed cell sums
// S_p(q) = sum_n p(n) q(n), p in {1, sx, sy, sz, sxsy, sxsz, sysz, sxsysz}; a rebuilt node's
// zeroth/first/second moments go through the K17 backward chimera to populations.
//
// The reverse (vector-Jacobian products) walks the same chain backwards: the rebuild through
// the transposed backward chimera lines (vf_k17_adjoint.h), the polynomial evaluation and the
// coefficients (both linear, apart from the LaplaceRho term and the density factor of the
// non-equilibrium scale), and the source moments (u = j / rho and products of u). The host
// compiles the same functions (the tests check them against dual numbers on the CPU).
#pragma once

#include "vf_k17_adjoint.h"

namespace vfa {

struct SrcMom {
    float drho, vx, vy, vz, kxy, kyz, kxz, kxxMyy, kxxMzz;
};

// The nine population sums a source node's moments come from
struct SrcSums {
    float drho, jx, jy, jz, pxy, pyz, pxz, pxxyy, pxxzz;
};

VF_HD SrcSums source_sums_of(const float (&f)[27])
{
    SrcSums s{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const float v = f[d];
        s.drho += v;
        if (cx(d)) s.jx += float(cx(d)) * v;
        if (cy(d)) s.jy += float(cy(d)) * v;
        if (cz(d)) s.jz += float(cz(d)) * v;
        if (cx(d) * cy(d)) s.pxy += float(cx(d) * cy(d)) * v;
        if (cy(d) * cz(d)) s.pyz += float(cy(d) * cz(d)) * v;
        if (cx(d) * cz(d)) s.pxz += float(cx(d) * cz(d)) * v;
        if (cx(d) * cx(d) - cy(d) * cy(d)) s.pxxyy += float(cx(d) * cx(d) - cy(d) * cy(d)) * v;
        if (cx(d) * cx(d) - cz(d) * cz(d)) s.pxxzz += float(cx(d) * cx(d) - cz(d) * cz(d)) * v;
    }
    return s;
}

// MomentsOnSourceNode::calculate (InterpolationCoefficients.h:86-144) from the node's sums
VF_HD SrcMom source_moments_of_sums(const SrcSums& s, float omega)
{
    const float ir = 1.0f / (s.drho + 1.0f);
    SrcMom m;
    m.drho = s.drho;
    m.vx = s.jx * ir;
    m.vy = s.jy * ir;
    m.vz = s.jz * ir;
    m.kxy = -3.0f * omega * (s.pxy * ir - m.vx * m.vy);
    m.kyz = -3.0f * omega * (s.pyz * ir - m.vy * m.vz);
    m.kxz = -3.0f * omega * (s.pxz * ir - m.vx * m.vz);
    m.kxxMyy = -1.5f * omega * (s.pxxyy * ir - (m.vx * m.vx - m.vy * m.vy));
    m.kxxMzz = -1.5f * omega * (s.pxxzz * ir - (m.vx * m.vx - m.vz * m.vz));
    return m;
}

// ... and from 27 populations
VF_HD SrcMom source_moments_of(const float (&f)[27], float omega)
{
    return source_moments_of_sums(source_sums_of(f), omega);
}

struct Coef {
    float a000, a100, a010, a001, a200, a020, a002, a110, a101, a011, a111;
    float b000, b100, b010, b001, b200, b020, b002, b110, b101, b011, b111;
    float c000, c100, c010, c001, c200, c020, c002, c110, c101, c011, c111;
    float d000, d100, d010, d001, d110, d101, d011, d111;
    float LaplaceRho;
};

VF_HD Coef coef_zero()
{
    Coef c;
    float* p = &c.a000;
    for (int i = 0; i < (int)(sizeof(Coef) / sizeof(float)); ++i) p[i] = 0.f;
    return c;
}

// signed cell sums of one quantity: [1, x, y, z, xy, xz, yz, xyz]
struct Sums {
    float s[8];
};
VF_HD float corner_sign(int n, int p)
{
    const float sx = (n & 1) ? 1.f : -1.f, sy = (n & 2) ? 1.f : -1.f, sz = (n & 4) ? 1.f : -1.f;
    switch (p) {
    case 1: return sx;
    case 2: return sy;
    case 3: return sz;
    case 4: return sx * sy;
    case 5: return sx * sz;
    case 6: return sy * sz;
    case 7: return sx * sy * sz;
    default: return 1.f;
    }
}
VF_HD Sums signed_sums(const float (&q)[8])
{
    Sums r;
#pragma unroll
    for (int p = 0; p < 8; ++p) r.s[p] = 0.f;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const float sx = (n & 1) ? 1.f : -1.f, sy = (n & 2) ? 1.f : -1.f, sz = (n & 4) ? 1.f : -1.f;
        r.s[0] += q[n];
        r.s[1] += sx * q[n];
        r.s[2] += sy * q[n];
        r.s[3] += sz * q[n];
        r.s[4] += sx * sy * q[n];
        r.s[5] += sx * sz * q[n];
        r.s[6] += sy * sz * q[n];
        r.s[7] += sx * sy * sz * q[n];
    }
    return r;
}

// the nine per-node quantities as sums; reverse: the adjoints of the sums back to the nodes
struct CellSums {
    Sums VX, VY, VZ, RH, KXY, KYZ, KXZ, KXXYY, KXXZZ;
};
VF_HD CellSums cell_sums(const SrcMom (&m)[8])
{
    CellSums S;
    float q[8];
#define VF_SUMS(name, field)                  \
    for (int n = 0; n < 8; ++n) q[n] = m[n].field; \
    S.name = signed_sums(q);
    VF_SUMS(VX, vx)
    VF_SUMS(VY, vy)
    VF_SUMS(VZ, vz)
    VF_SUMS(RH, drho)
    VF_SUMS(KXY, kxy)
    VF_SUMS(KYZ, kyz)
    VF_SUMS(KXZ, kxz)
    VF_SUMS(KXXYY, kxxMyy)
    VF_SUMS(KXXZZ, kxxMzz)
#undef VF_SUMS
    return S;
}

enum { P1 = 0, PX, PY, PZ, PXY, PXZ, PYZ, PXYZ };

VF_HD Coef coefficients(const SrcMom (&m)[8], float xo, float yo, float zo)
{
    const CellSums S = cell_sums(m);
    const Sums &VX = S.VX, &VY = S.VY, &VZ = S.VZ, &RH = S.RH, &KXY = S.KXY, &KYZ = S.KYZ, &KXZ = S.KXZ,
               &KXXYY = S.KXXYY, &KXXZZ = S.KXXZZ;
    Coef c;
    const float s64 = 1.f / 64.f, s16 = 1.f / 16.f, e8 = 0.125f;
    c.a000 = s64 * (2.f * (-KXY.s[PY] - KXZ.s[PZ] + VY.s[PXY] + VZ.s[PXZ]) + 8.f * VX.s[P1] - KXXYY.s[PX] - KXXZZ.s[PX]);
    c.b000 = s64 * (2.f * (KXXYY.s[PY] - KXY.s[PX] - KYZ.s[PZ] + VX.s[PXY] + VZ.s[PYZ]) + 8.f * VY.s[P1] - KXXZZ.s[PY]);
    c.c000 = s64 * (2.f * (KXXZZ.s[PZ] - KXZ.s[PX] - KYZ.s[PY] + VX.s[PXZ] + VY.s[PYZ]) + 8.f * VZ.s[P1] - KXXYY.s[PZ]);
    c.a100 = 0.25f * VX.s[PX];
    c.b100 = 0.25f * VY.s[PX];
    c.c100 = 0.25f * VZ.s[PX];
    c.a200 = s16 * (2.f * (VY.s[PXY] + VZ.s[PXZ]) + KXXYY.s[PX] + KXXZZ.s[PX]);
    c.b200 = e8 * (-2.f * VX.s[PXY] + KXY.s[PX]);
    c.c200 = e8 * (-2.f * VX.s[PXZ] + KXZ.s[PX]);
    c.a010 = 0.25f * VX.s[PY];
    c.b010 = 0.25f * VY.s[PY];
    c.c010 = 0.25f * VZ.s[PY];
    c.a020 = e8 * (-2.f * VY.s[PXY] + KXY.s[PY]);
    c.b020 = s16 * (2.f * (-KXXYY.s[PY] + VX.s[PXY] + VZ.s[PYZ]) + KXXZZ.s[PY]);
    c.c020 = e8 * (-2.f * VY.s[PYZ] + KYZ.s[PY]);
    c.a001 = 0.25f * VX.s[PZ];
    c.b001 = 0.25f * VY.s[PZ];
    c.c001 = 0.25f * VZ.s[PZ];
    c.a002 = e8 * (-2.f * VZ.s[PXZ] + KXZ.s[PZ]);
    c.b002 = e8 * (-2.f * VZ.s[PYZ] + KYZ.s[PZ]);
    c.c002 = s16 * (2.f * (-KXXZZ.s[PZ] + VX.s[PXZ] + VY.s[PYZ]) + KXXYY.s[PZ]);
    c.a110 = 0.5f * VX.s[PXY];
    c.b110 = 0.5f * VY.s[PXY];
    c.c110 = 0.5f * VZ.s[PXY];
    c.a101 = 0.5f * VX.s[PXZ];
    c.b101 = 0.5f * VY.s[PXZ];
    c.c101 = 0.5f * VZ.s[PXZ];
    c.a011 = 0.5f * VX.s[PYZ];
    c.b011 = 0.5f * VY.s[PYZ];
    c.c011 = 0.5f * VZ.s[PYZ];
    c.a111 = VX.s[PXYZ];
    c.b111 = VY.s[PXYZ];
    c.c111 = VZ.s[PXYZ];
    c.LaplaceRho = (xo != 0.f || yo != 0.f || zo != 0.f)
                       ? 0.f
                       : -3.f * (c.a100 * c.a100 + c.b010 * c.b010 + c.c001 * c.c001) -
                             6.f * (c.b100 * c.a010 + c.c100 * c.a001 + c.c010 * c.b001);
    c.d000 = e8 * RH.s[P1];
    c.d100 = 0.25f * RH.s[PX];
    c.d010 = 0.25f * RH.s[PY];
    c.d001 = 0.25f * RH.s[PZ];
    c.d110 = 0.5f * RH.s[PXY];
    c.d101 = 0.5f * RH.s[PXZ];
    c.d011 = 0.5f * RH.s[PYZ];
    c.d111 = RH.s[PXYZ];
    // expansion point shifted by the cell offset (refinement towards walls)
    const float xx = xo * xo, yy = yo * yo, zz = zo * zo;
    c.a000 += xo * c.a100 + yo * c.a010 + zo * c.a001 + xx * c.a200 + yy * c.a020 + zz * c.a002 + xo * yo * c.a110 +
              xo * zo * c.a101 + yo * zo * c.a011;
    c.a100 += 2.f * xo * c.a200 + yo * c.a110 + zo * c.a101;
    c.a010 += 2.f * yo * c.a020 + xo * c.a110 + zo * c.a011;
    c.a001 += 2.f * zo * c.a002 + xo * c.a101 + yo * c.a011;
    c.b000 += xo * c.b100 + yo * c.b010 + zo * c.b001 + xx * c.b200 + yy * c.b020 + zz * c.b002 + xo * yo * c.b110 +
              xo * zo * c.b101 + yo * zo * c.b011;
    c.b100 += 2.f * xo * c.b200 + yo * c.b110 + zo * c.b101;
    c.b010 += 2.f * yo * c.b020 + xo * c.b110 + zo * c.b011;
    c.b001 += 2.f * zo * c.b002 + xo * c.b101 + yo * c.b011;
    c.c000 += xo * c.c100 + yo * c.c010 + zo * c.c001 + xx * c.c200 + yy * c.c020 + zz * c.c002 + xo * yo * c.c110 +
              xo * zo * c.c101 + yo * zo * c.c011;
    c.c100 += 2.f * xo * c.c200 + yo * c.c110 + zo * c.c101;
    c.c010 += 2.f * yo * c.c020 + xo * c.c110 + zo * c.c011;
    c.c001 += 2.f * zo * c.c002 + xo * c.c101 + yo * c.c011;
    c.d000 += xo * c.d100 + yo * c.d010 + zo * c.d001 + xo * yo * c.d110 + xo * zo * c.d101 + yo * zo * c.d011;
    c.d100 += yo * c.d110 + zo * c.d101;
    c.d010 += xo * c.d110 + zo * c.d011;
    c.d001 += xo * c.d101 + yo * c.d011;
    return c;
}

// The nine inputs of a rebuilt node: density, velocity and the five second-order moments
// (normal-stress differences and off-diagonal moments, non-equilibrium included).
struct NodeMoments {
    float press, vx, vy, vz, mxxMyy, mxxMzz, m011, m101, m110;
};

// central moments of a rebuilt node before the backward chimera, idx(a, b, c) order
VF_HD void initial_moments(const NodeMoments& q, float (&M)[27])
{
    using adj::idx;
#pragma unroll
    for (int i = 0; i < 27; ++i) M[i] = 0.f;
    M[idx(0, 0, 0)] = q.press;
    M[idx(2, 0, 0)] = (1.f / 3.f) * (q.mxxMyy + q.mxxMzz + q.press);
    M[idx(0, 2, 0)] = (1.f / 3.f) * (-2.f * q.mxxMyy + q.mxxMzz + q.press);
    M[idx(0, 0, 2)] = (1.f / 3.f) * (q.mxxMyy - 2.f * q.mxxMzz + q.press);
    M[idx(1, 1, 0)] = q.m110;
    M[idx(1, 0, 1)] = q.m101;
    M[idx(0, 1, 1)] = q.m011;
    M[idx(2, 2, 0)] = q.press * (1.f / 9.f);
    M[idx(2, 0, 2)] = q.press * (1.f / 9.f);
    M[idx(0, 2, 2)] = q.press * (1.f / 9.f);
    M[idx(2, 2, 2)] = q.press * (1.f / 27.f);
}

// Central moments (zeroth, second order with non-equilibrium, fourth and sixth order at their
// rest values, everything else zero) -> shifted populations through the backward chimera
// passes of the collision (InterpolationCF.h:189-291).
VF_HD void rebuild_populations(float (&f)[27], const NodeMoments& q)
{
    using adj::idx;
    float M[27];
    initial_moments(q, M);
    {
        const BwdPass p = bwd_pass(q.vx, q.vx * q.vx);
        bwd<1, 1>(M[idx(0, 0, 0)], M[idx(1, 0, 0)], M[idx(2, 0, 0)], p);
        bwd<0, 1>(M[idx(0, 1, 0)], M[idx(1, 1, 0)], M[idx(2, 1, 0)], p);
        bwd<1, 3>(M[idx(0, 2, 0)], M[idx(1, 2, 0)], M[idx(2, 2, 0)], p);
        bwd<0, 1>(M[idx(0, 0, 1)], M[idx(1, 0, 1)], M[idx(2, 0, 1)], p);
        bwd<0, 1>(M[idx(0, 1, 1)], M[idx(1, 1, 1)], M[idx(2, 1, 1)], p);
        bwd<0, 1>(M[idx(0, 2, 1)], M[idx(1, 2, 1)], M[idx(2, 2, 1)], p);
        bwd<1, 3>(M[idx(0, 0, 2)], M[idx(1, 0, 2)], M[idx(2, 0, 2)], p);
        bwd<0, 1>(M[idx(0, 1, 2)], M[idx(1, 1, 2)], M[idx(2, 1, 2)], p);
        bwd<1, 9>(M[idx(0, 2, 2)], M[idx(1, 2, 2)], M[idx(2, 2, 2)], p);
    }
    {
        const BwdPass p = bwd_pass(q.vy, q.vy * q.vy);
        bwd<1, 6>(M[idx(0, 0, 0)], M[idx(0, 1, 0)], M[idx(0, 2, 0)], p);
        bwd<0, 1>(M[idx(0, 0, 1)], M[idx(0, 1, 1)], M[idx(0, 2, 1)], p);
        bwd<1, 18>(M[idx(0, 0, 2)], M[idx(0, 1, 2)], M[idx(0, 2, 2)], p);
        bwd<2, 3>(M[idx(1, 0, 0)], M[idx(1, 1, 0)], M[idx(1, 2, 0)], p);
        bwd<0, 1>(M[idx(1, 0, 1)], M[idx(1, 1, 1)], M[idx(1, 2, 1)], p);
        bwd<2, 9>(M[idx(1, 0, 2)], M[idx(1, 1, 2)], M[idx(1, 2, 2)], p);
        bwd<1, 6>(M[idx(2, 0, 0)], M[idx(2, 1, 0)], M[idx(2, 2, 0)], p);
        bwd<0, 1>(M[idx(2, 0, 1)], M[idx(2, 1, 1)], M[idx(2, 2, 1)], p);
        bwd<1, 18>(M[idx(2, 0, 2)], M[idx(2, 1, 2)], M[idx(2, 2, 2)], p);
    }
    {
        const BwdPass p = bwd_pass(q.vz, q.vz * q.vz);
        bwd<1, 36>(M[idx(0, 0, 0)], M[idx(0, 0, 1)], M[idx(0, 0, 2)], p);
        bwd<1, 9>(M[idx(0, 1, 0)], M[idx(0, 1, 1)], M[idx(0, 1, 2)], p);
        bwd<1, 36>(M[idx(0, 2, 0)], M[idx(0, 2, 1)], M[idx(0, 2, 2)], p);
        bwd<1, 9>(M[idx(1, 0, 0)], M[idx(1, 0, 1)], M[idx(1, 0, 2)], p);
        bwd<4, 9>(M[idx(1, 1, 0)], M[idx(1, 1, 1)], M[idx(1, 1, 2)], p);
        bwd<1, 9>(M[idx(1, 2, 0)], M[idx(1, 2, 1)], M[idx(1, 2, 2)], p);
        bwd<1, 36>(M[idx(2, 0, 0)], M[idx(2, 0, 1)], M[idx(2, 0, 2)], p);
        bwd<1, 9>(M[idx(2, 1, 0)], M[idx(2, 1, 1)], M[idx(2, 1, 2)], p);
        bwd<1, 36>(M[idx(2, 2, 0)], M[idx(2, 2, 1)], M[idx(2, 2, 2)], p);
    }
#pragma unroll
    for (int i = 0; i < 27; ++i) f[adj::dir_of_idx(i)] = M[i];
}

// interpolateCF: the node at (x, y, z) of the coarse cell (coarse units, centre 0)
VF_HD NodeMoments cf_node_moments(const Coef& c, float x, float y, float z, float eps, float omegaF)
{
    NodeMoments q;
    q.press = c.d000 + x * c.d100 + y * c.d010 + z * c.d001 + x * y * c.d110 + x * z * c.d101 + y * z * c.d011 +
              x * y * z * c.d111 + 3.f * x * x * c.LaplaceRho;
    q.vx = c.a000 + x * c.a100 + y * c.a010 + z * c.a001 + x * x * c.a200 + y * y * c.a020 + z * z * c.a002 +
           x * y * c.a110 + x * z * c.a101 + y * z * c.a011 + x * y * z * c.a111;
    q.vy = c.b000 + x * c.b100 + y * c.b010 + z * c.b001 + x * x * c.b200 + y * y * c.b020 + z * z * c.b002 +
           x * y * c.b110 + x * z * c.b101 + y * z * c.b011 + x * y * z * c.b111;
    q.vz = c.c000 + x * c.c100 + y * c.c010 + z * c.c001 + x * x * c.c200 + y * y * c.c020 + z * z * c.c002 +
           x * y * c.c110 + x * z * c.c101 + y * z * c.c011 + x * y * z * c.c111;
    const float scale = eps / omegaF * (1.f + q.press);
    const float dxMdy = c.a100 - c.b010 + 2.f * c.a200 * x - c.b110 * x + c.a110 * y - 2.f * c.b020 * y + c.a101 * z -
                        c.b011 * z - c.b111 * x * z + c.a111 * y * z;
    const float dxMdz = c.a100 - c.c001 + 2.f * c.a200 * x - c.c101 * x + c.a110 * y - c.c011 * y - c.c111 * x * y +
                        c.a101 * z - 2.f * c.c002 * z + c.a111 * y * z;
    const float syz = c.b001 + c.c010 + c.b101 * x + c.c110 * x + c.b011 * y + 2.f * c.c020 * y + c.b111 * x * y +
                      2.f * c.b002 * z + c.c011 * z + c.c111 * x * z;
    const float sxz = c.a001 + c.c100 + c.a101 * x + 2.f * c.c200 * x + c.a011 * y + c.c110 * y + c.a111 * x * y +
                      2.f * c.a002 * z + c.c101 * z + c.c111 * y * z;
    const float sxy = c.a010 + c.b100 + c.a110 * x + 2.f * c.b200 * x + 2.f * c.a020 * y + c.b110 * y + c.a011 * z +
                      c.b101 * z + c.a111 * x * z + c.b111 * y * z;
    q.mxxMyy = -(2.f / 3.f) * dxMdy * scale;
    q.mxxMzz = -(2.f / 3.f) * dxMdz * scale;
    q.m011 = -(1.f / 3.f) * syz * scale;
    q.m101 = -(1.f / 3.f) * sxz * scale;
    q.m110 = -(1.f / 3.f) * sxy * scale;
    return q;
}

// interpolateFC: the coarse node at the centre of the fine cell (eps 2)
VF_HD NodeMoments fc_node_moments(const Coef& c, float omegaC)
{
    NodeMoments q;
    q.press = c.d000 - 2.f * c.LaplaceRho * 0.125f;
    const float scale = 2.f / omegaC * (1.f + q.press);
    q.vx = c.a000;
    q.vy = c.b000;
    q.vz = c.c000;
    q.mxxMyy = -(2.f / 3.f) * (c.a100 - c.b010) * scale;
    q.mxxMzz = -(2.f / 3.f) * (c.a100 - c.c001) * scale;
    q.m011 = -(1.f / 3.f) * (c.b001 + c.c010) * scale;
    q.m101 = -(1.f / 3.f) * (c.a001 + c.c100) * scale;
    q.m110 = -(1.f / 3.f) * (c.a010 + c.b100) * scale;
    return q;
}

// position of fine child q of a CF cell, coarse units
VF_HD float cf_child_offset(int q, int axis) { return ((q >> axis) & 1) ? 0.25f : -0.25f; }

// ================================ reverse mode ==============================================

// (d rebuilt populations / d node moments)^T g, g by D3Q27 direction
VF_HD NodeMoments rebuild_vjp(const NodeMoments& q, const float (&g)[27])
{
    using adj::idx;
    float M[27];
    initial_moments(q, M);
    float A[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) A[i] = g[adj::dir_of_idx(i)];
    float vxb = 0.f, vyb = 0.f, vzb = 0.f;
    adj::backward_chimera_rev(M, q.vx, q.vy, q.vz, A, vxb, vyb, vzb);
    NodeMoments b;
    b.vx = vxb;
    b.vy = vyb;
    b.vz = vzb;
    const float t = 1.f / 3.f;
    b.press = A[idx(0, 0, 0)] + t * (A[idx(2, 0, 0)] + A[idx(0, 2, 0)] + A[idx(0, 0, 2)]) +
              (1.f / 9.f) * (A[idx(2, 2, 0)] + A[idx(2, 0, 2)] + A[idx(0, 2, 2)]) + (1.f / 27.f) * A[idx(2, 2, 2)];
    b.mxxMyy = t * A[idx(2, 0, 0)] - 2.f * t * A[idx(0, 2, 0)] + t * A[idx(0, 0, 2)];
    b.mxxMzz = t * A[idx(2, 0, 0)] + t * A[idx(0, 2, 0)] - 2.f * t * A[idx(0, 0, 2)];
    b.m110 = A[idx(1, 1, 0)];
    b.m101 = A[idx(1, 0, 1)];
    b.m011 = A[idx(0, 1, 1)];
    return b;
}

// adds (d node moments / d coefficients)^T qb to cb for CF child at (x, y, z)
VF_HD void cf_node_moments_vjp(const Coef& c, float x, float y, float z, float eps, float omegaF,
                               const NodeMoments& qb, Coef& cb)
{
    const float press = c.d000 + x * c.d100 + y * c.d010 + z * c.d001 + x * y * c.d110 + x * z * c.d101 +
                        y * z * c.d011 + x * y * z * c.d111 + 3.f * x * x * c.LaplaceRho;
    const float scale = eps / omegaF * (1.f + press);
    const float dxMdy = c.a100 - c.b010 + 2.f * c.a200 * x - c.b110 * x + c.a110 * y - 2.f * c.b020 * y + c.a101 * z -
                        c.b011 * z - c.b111 * x * z + c.a111 * y * z;
    const float dxMdz = c.a100 - c.c001 + 2.f * c.a200 * x - c.c101 * x + c.a110 * y - c.c011 * y - c.c111 * x * y +
                        c.a101 * z - 2.f * c.c002 * z + c.a111 * y * z;
    const float syz = c.b001 + c.c010 + c.b101 * x + c.c110 * x + c.b011 * y + 2.f * c.c020 * y + c.b111 * x * y +
                      2.f * c.b002 * z + c.c011 * z + c.c111 * x * z;
    const float sxz = c.a001 + c.c100 + c.a101 * x + 2.f * c.c200 * x + c.a011 * y + c.c110 * y + c.a111 * x * y +
                      2.f * c.a002 * z + c.c101 * z + c.c111 * y * z;
    const float sxy = c.a010 + c.b100 + c.a110 * x + 2.f * c.b200 * x + 2.f * c.a020 * y + c.b110 * y + c.a011 * z +
                      c.b101 * z + c.a111 * x * z + c.b111 * y * z;
    const float tt = 2.f / 3.f, t = 1.f / 3.f;
    const float dxMdyb = -tt * scale * qb.mxxMyy, dxMdzb = -tt * scale * qb.mxxMzz;
    const float syzb = -t * scale * qb.m011, sxzb = -t * scale * qb.m101, sxyb = -t * scale * qb.m110;
    const float scaleb = -tt * (dxMdy * qb.mxxMyy + dxMdz * qb.mxxMzz) - t * (syz * qb.m011 + sxz * qb.m101 + sxy * qb.m110);
    const float pb = qb.press + scaleb * eps / omegaF;
    // press
    cb.d000 += pb;
    cb.d100 += x * pb;
    cb.d010 += y * pb;
    cb.d001 += z * pb;
    cb.d110 += x * y * pb;
    cb.d101 += x * z * pb;
    cb.d011 += y * z * pb;
    cb.d111 += x * y * z * pb;
    cb.LaplaceRho += 3.f * x * x * pb;
    // velocity
    cb.a000 += qb.vx; cb.a100 += x * qb.vx; cb.a010 += y * qb.vx; cb.a001 += z * qb.vx;
    cb.a200 += x * x * qb.vx; cb.a020 += y * y * qb.vx; cb.a002 += z * z * qb.vx;
    cb.a110 += x * y * qb.vx; cb.a101 += x * z * qb.vx; cb.a011 += y * z * qb.vx; cb.a111 += x * y * z * qb.vx;
    cb.b000 += qb.vy; cb.b100 += x * qb.vy; cb.b010 += y * qb.vy; cb.b001 += z * qb.vy;
    cb.b200 += x * x * qb.vy; cb.b020 += y * y * qb.vy; cb.b002 += z * z * qb.vy;
    cb.b110 += x * y * qb.vy; cb.b101 += x * z * qb.vy; cb.b011 += y * z * qb.vy; cb.b111 += x * y * z * qb.vy;
    cb.c000 += qb.vz; cb.c100 += x * qb.vz; cb.c010 += y * qb.vz; cb.c001 += z * qb.vz;
    cb.c200 += x * x * qb.vz; cb.c020 += y * y * qb.vz; cb.c002 += z * z * qb.vz;
    cb.c110 += x * y * qb.vz; cb.c101 += x * z * qb.vz; cb.c011 += y * z * qb.vz; cb.c111 += x * y * z * qb.vz;
    // dxMdy
    cb.a100 += dxMdyb; cb.b010 -= dxMdyb; cb.a200 += 2.f * x * dxMdyb; cb.b110 -= x * dxMdyb; cb.a110 += y * dxMdyb;
    cb.b020 -= 2.f * y * dxMdyb; cb.a101 += z * dxMdyb; cb.b011 -= z * dxMdyb; cb.b111 -= x * z * dxMdyb;
    cb.a111 += y * z * dxMdyb;
    // dxMdz
    cb.a100 += dxMdzb; cb.c001 -= dxMdzb; cb.a200 += 2.f * x * dxMdzb; cb.c101 -= x * dxMdzb; cb.a110 += y * dxMdzb;
    cb.c011 -= y * dxMdzb; cb.c111 -= x * y * dxMdzb; cb.a101 += z * dxMdzb; cb.c002 -= 2.f * z * dxMdzb;
    cb.a111 += y * z * dxMdzb;
    // syz
    cb.b001 += syzb; cb.c010 += syzb; cb.b101 += x * syzb; cb.c110 += x * syzb; cb.b011 += y * syzb;
    cb.c020 += 2.f * y * syzb; cb.b111 += x * y * syzb; cb.b002 += 2.f * z * syzb; cb.c011 += z * syzb;
    cb.c111 += x * z * syzb;
    // sxz
    cb.a001 += sxzb; cb.c100 += sxzb; cb.a101 += x * sxzb; cb.c200 += 2.f * x * sxzb; cb.a011 += y * sxzb;
    cb.c110 += y * sxzb; cb.a111 += x * y * sxzb; cb.a002 += 2.f * z * sxzb; cb.c101 += z * sxzb;
    cb.c111 += y * z * sxzb;
    // sxy
    cb.a010 += sxyb; cb.b100 += sxyb; cb.a110 += x * sxyb; cb.b200 += 2.f * x * sxyb; cb.a020 += 2.f * y * sxyb;
    cb.b110 += y * sxyb; cb.a011 += z * sxyb; cb.b101 += z * sxyb; cb.a111 += x * z * sxyb; cb.b111 += y * z * sxyb;
}

// adds (d node moments / d coefficients)^T qb to cb for the FC node
VF_HD void fc_node_moments_vjp(const Coef& c, float omegaC, const NodeMoments& qb, Coef& cb)
{
    const float press = c.d000 - 2.f * c.LaplaceRho * 0.125f;
    const float scale = 2.f / omegaC * (1.f + press);
    const float tt = 2.f / 3.f, t = 1.f / 3.f;
    const float A = c.a100 - c.b010, B = c.a100 - c.c001, Cc = c.b001 + c.c010, D = c.a001 + c.c100,
                E = c.a010 + c.b100;
    const float Ab = -tt * scale * qb.mxxMyy, Bb = -tt * scale * qb.mxxMzz, Cb = -t * scale * qb.m011,
                Db = -t * scale * qb.m101, Eb = -t * scale * qb.m110;
    const float scaleb = -tt * (A * qb.mxxMyy + B * qb.mxxMzz) - t * (Cc * qb.m011 + D * qb.m101 + E * qb.m110);
    const float pb = qb.press + scaleb * 2.f / omegaC;
    cb.d000 += pb;
    cb.LaplaceRho += -0.25f * pb;
    cb.a000 += qb.vx;
    cb.b000 += qb.vy;
    cb.c000 += qb.vz;
    cb.a100 += Ab + Bb;
    cb.b010 -= Ab;
    cb.c001 -= Bb;
    cb.b001 += Cb;
    cb.c010 += Cb;
    cb.a001 += Db;
    cb.c100 += Db;
    cb.a010 += Eb;
    cb.b100 += Eb;
}

// (d coefficients / d signed sums)^T cb: the adjoints of the cell sums. The coefficients before
// the offset shift are recomputed where the LaplaceRho term needs them.
VF_HD CellSums coefficients_vjp_sums(const SrcMom (&m)[8], float xo, float yo, float zo, Coef cb)
{
    // ---- reverse of the offset shift (statements in reverse order) ----
    const float xx = xo * xo, yy = yo * yo, zz = zo * zo;
    cb.d101 += xo * cb.d001; cb.d011 += yo * cb.d001;
    cb.d110 += xo * cb.d010; cb.d011 += zo * cb.d010;
    cb.d110 += yo * cb.d100; cb.d101 += zo * cb.d100;
    cb.d100 += xo * cb.d000; cb.d010 += yo * cb.d000; cb.d001 += zo * cb.d000; cb.d110 += xo * yo * cb.d000;
    cb.d101 += xo * zo * cb.d000; cb.d011 += yo * zo * cb.d000;
#define VF_SHIFT_REV(P)                                                                                           \
    cb.P##002 += 2.f * zo * cb.P##001; cb.P##101 += xo * cb.P##001; cb.P##011 += yo * cb.P##001;                  \
    cb.P##020 += 2.f * yo * cb.P##010; cb.P##110 += xo * cb.P##010; cb.P##011 += zo * cb.P##010;                  \
    cb.P##200 += 2.f * xo * cb.P##100; cb.P##110 += yo * cb.P##100; cb.P##101 += zo * cb.P##100;                  \
    cb.P##100 += xo * cb.P##000; cb.P##010 += yo * cb.P##000; cb.P##001 += zo * cb.P##000;                        \
    cb.P##200 += xx * cb.P##000; cb.P##020 += yy * cb.P##000; cb.P##002 += zz * cb.P##000;                        \
    cb.P##110 += xo * yo * cb.P##000; cb.P##101 += xo * zo * cb.P##000; cb.P##011 += yo * zo * cb.P##000;
    VF_SHIFT_REV(c)
    VF_SHIFT_REV(b)
    VF_SHIFT_REV(a)
#undef VF_SHIFT_REV
    const CellSums S = cell_sums(m);
    // ---- LaplaceRho (only without offsets), from the unshifted first-order coefficients ----
    if (!(xo != 0.f || yo != 0.f || zo != 0.f)) {
        const float a100 = 0.25f * S.VX.s[PX], b010 = 0.25f * S.VY.s[PY], c001 = 0.25f * S.VZ.s[PZ];
        const float b100 = 0.25f * S.VY.s[PX], a010 = 0.25f * S.VX.s[PY], c100 = 0.25f * S.VZ.s[PX];
        const float a001 = 0.25f * S.VX.s[PZ], c010 = 0.25f * S.VZ.s[PY], b001 = 0.25f * S.VY.s[PZ];
        const float L = cb.LaplaceRho;
        cb.a100 += -6.f * a100 * L;
        cb.b010 += -6.f * b010 * L;
        cb.c001 += -6.f * c001 * L;
        cb.b100 += -6.f * a010 * L;
        cb.a010 += -6.f * b100 * L;
        cb.c100 += -6.f * a001 * L;
        cb.a001 += -6.f * c100 * L;
        cb.c010 += -6.f * b001 * L;
        cb.b001 += -6.f * c010 * L;
    }
    // ---- coefficient formulas -> adjoints of the signed sums ----
    CellSums Sb;
    {
        float* p = &Sb.VX.s[0];
        for (int i = 0; i < (int)(sizeof(CellSums) / sizeof(float)); ++i) p[i] = 0.f;
    }
    const float s64 = 1.f / 64.f, s16 = 1.f / 16.f, e8 = 0.125f;
    // a000, b000, c000
    Sb.KXY.s[PY] += -2.f * s64 * cb.a000; Sb.KXZ.s[PZ] += -2.f * s64 * cb.a000; Sb.VY.s[PXY] += 2.f * s64 * cb.a000;
    Sb.VZ.s[PXZ] += 2.f * s64 * cb.a000; Sb.VX.s[P1] += 8.f * s64 * cb.a000; Sb.KXXYY.s[PX] += -s64 * cb.a000;
    Sb.KXXZZ.s[PX] += -s64 * cb.a000;
    Sb.KXXYY.s[PY] += 2.f * s64 * cb.b000; Sb.KXY.s[PX] += -2.f * s64 * cb.b000; Sb.KYZ.s[PZ] += -2.f * s64 * cb.b000;
    Sb.VX.s[PXY] += 2.f * s64 * cb.b000; Sb.VZ.s[PYZ] += 2.f * s64 * cb.b000; Sb.VY.s[P1] += 8.f * s64 * cb.b000;
    Sb.KXXZZ.s[PY] += -s64 * cb.b000;
    Sb.KXXZZ.s[PZ] += 2.f * s64 * cb.c000; Sb.KXZ.s[PX] += -2.f * s64 * cb.c000; Sb.KYZ.s[PY] += -2.f * s64 * cb.c000;
    Sb.VX.s[PXZ] += 2.f * s64 * cb.c000; Sb.VY.s[PYZ] += 2.f * s64 * cb.c000; Sb.VZ.s[P1] += 8.f * s64 * cb.c000;
    Sb.KXXYY.s[PZ] += -s64 * cb.c000;
    // first order
    Sb.VX.s[PX] += 0.25f * cb.a100; Sb.VY.s[PX] += 0.25f * cb.b100; Sb.VZ.s[PX] += 0.25f * cb.c100;
    Sb.VX.s[PY] += 0.25f * cb.a010; Sb.VY.s[PY] += 0.25f * cb.b010; Sb.VZ.s[PY] += 0.25f * cb.c010;
    Sb.VX.s[PZ] += 0.25f * cb.a001; Sb.VY.s[PZ] += 0.25f * cb.b001; Sb.VZ.s[PZ] += 0.25f * cb.c001;
    // second order, x
    Sb.VY.s[PXY] += 2.f * s16 * cb.a200; Sb.VZ.s[PXZ] += 2.f * s16 * cb.a200; Sb.KXXYY.s[PX] += s16 * cb.a200;
    Sb.KXXZZ.s[PX] += s16 * cb.a200;
    Sb.VX.s[PXY] += -2.f * e8 * cb.b200; Sb.KXY.s[PX] += e8 * cb.b200;
    Sb.VX.s[PXZ] += -2.f * e8 * cb.c200; Sb.KXZ.s[PX] += e8 * cb.c200;
    // second order, y
    Sb.VY.s[PXY] += -2.f * e8 * cb.a020; Sb.KXY.s[PY] += e8 * cb.a020;
    Sb.KXXYY.s[PY] += -2.f * s16 * cb.b020; Sb.VX.s[PXY] += 2.f * s16 * cb.b020; Sb.VZ.s[PYZ] += 2.f * s16 * cb.b020;
    Sb.KXXZZ.s[PY] += s16 * cb.b020;
    Sb.VY.s[PYZ] += -2.f * e8 * cb.c020; Sb.KYZ.s[PY] += e8 * cb.c020;
    // second order, z
    Sb.VZ.s[PXZ] += -2.f * e8 * cb.a002; Sb.KXZ.s[PZ] += e8 * cb.a002;
    Sb.VZ.s[PYZ] += -2.f * e8 * cb.b002; Sb.KYZ.s[PZ] += e8 * cb.b002;
    Sb.KXXZZ.s[PZ] += -2.f * s16 * cb.c002; Sb.VX.s[PXZ] += 2.f * s16 * cb.c002; Sb.VY.s[PYZ] += 2.f * s16 * cb.c002;
    Sb.KXXYY.s[PZ] += s16 * cb.c002;
    // mixed and third order
    Sb.VX.s[PXY] += 0.5f * cb.a110; Sb.VY.s[PXY] += 0.5f * cb.b110; Sb.VZ.s[PXY] += 0.5f * cb.c110;
    Sb.VX.s[PXZ] += 0.5f * cb.a101; Sb.VY.s[PXZ] += 0.5f * cb.b101; Sb.VZ.s[PXZ] += 0.5f * cb.c101;
    Sb.VX.s[PYZ] += 0.5f * cb.a011; Sb.VY.s[PYZ] += 0.5f * cb.b011; Sb.VZ.s[PYZ] += 0.5f * cb.c011;
    Sb.VX.s[PXYZ] += cb.a111; Sb.VY.s[PXYZ] += cb.b111; Sb.VZ.s[PXYZ] += cb.c111;
    // density
    Sb.RH.s[P1] += e8 * cb.d000; Sb.RH.s[PX] += 0.25f * cb.d100; Sb.RH.s[PY] += 0.25f * cb.d010;
    Sb.RH.s[PZ] += 0.25f * cb.d001; Sb.RH.s[PXY] += 0.5f * cb.d110; Sb.RH.s[PXZ] += 0.5f * cb.d101;
    Sb.RH.s[PYZ] += 0.5f * cb.d011; Sb.RH.s[PXYZ] += cb.d111;
    return Sb;
}

// signed sums -> the adjoint of node n's moments (n may be a run-time value)
VF_HD SrcMom node_of_sums(const CellSums& Sb, int n)
{
    auto back = [&](const Sums& s) {
        float v = 0.f;
#pragma unroll
        for (int p = 0; p < 8; ++p) v += corner_sign(n, p) * s.s[p];
        return v;
    };
    SrcMom mb;
    mb.vx = back(Sb.VX);
    mb.vy = back(Sb.VY);
    mb.vz = back(Sb.VZ);
    mb.drho = back(Sb.RH);
    mb.kxy = back(Sb.KXY);
    mb.kyz = back(Sb.KYZ);
    mb.kxz = back(Sb.KXZ);
    mb.kxxMyy = back(Sb.KXXYY);
    mb.kxxMzz = back(Sb.KXXZZ);
    return mb;
}

// (d coefficients / d source moments)^T cb for the 8 nodes (mb overwritten)
VF_HD void coefficients_vjp(const SrcMom (&m)[8], float xo, float yo, float zo, Coef cb, SrcMom (&mb)[8])
{
    const CellSums Sb = coefficients_vjp_sums(m, xo, yo, zo, cb);
    for (int n = 0; n < 8; ++n) mb[n] = node_of_sums(Sb, n);
}

// (d source moments / d populations)^T mb, from the node's sums (the moments are linear in the
// populations through them, so the sums are all the reverse needs)
VF_HD void source_moments_vjp_sums(const SrcSums& s, float omega, const SrcMom& mb, float (&fb)[27])
{
    const float drho = s.drho, jx = s.jx, jy = s.jy, jz = s.jz, pxy = s.pxy, pyz = s.pyz, pxz = s.pxz;
    const float pxxyy = s.pxxyy, pxxzz = s.pxxzz;
    const float ir = 1.0f / (drho + 1.0f);
    const float vx = jx * ir, vy = jy * ir, vz = jz * ir;
    const float w3 = -3.0f * omega, w15 = -1.5f * omega;
    // k_ab = w (p_ab ir - u_a u_b)
    const float pxyb = w3 * ir * mb.kxy, pyzb = w3 * ir * mb.kyz, pxzb = w3 * ir * mb.kxz;
    const float pxxyyb = w15 * ir * mb.kxxMyy, pxxzzb = w15 * ir * mb.kxxMzz;
    const float vxb = mb.vx - w3 * (vy * mb.kxy + vz * mb.kxz) - w15 * 2.f * vx * (mb.kxxMyy + mb.kxxMzz);
    const float vyb = mb.vy - w3 * (vx * mb.kxy + vz * mb.kyz) + w15 * 2.f * vy * mb.kxxMyy;
    const float vzb = mb.vz - w3 * (vy * mb.kyz + vx * mb.kxz) + w15 * 2.f * vz * mb.kxxMzz;
    const float irb = w3 * (pxy * mb.kxy + pyz * mb.kyz + pxz * mb.kxz) + w15 * (pxxyy * mb.kxxMyy + pxxzz * mb.kxxMzz) +
                      jx * vxb + jy * vyb + jz * vzb;
    const float jxb = ir * vxb, jyb = ir * vyb, jzb = ir * vzb;
    const float drhob = mb.drho - ir * ir * irb;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        fb[d] = drhob + float(cx(d)) * jxb + float(cy(d)) * jyb + float(cz(d)) * jzb + float(cx(d) * cy(d)) * pxyb +
                float(cy(d) * cz(d)) * pyzb + float(cx(d) * cz(d)) * pxzb +
                float(cx(d) * cx(d) - cy(d) * cy(d)) * pxxyyb + float(cx(d) * cx(d) - cz(d) * cz(d)) * pxxzzb;
    }
}

VF_HD void source_moments_vjp(const float (&f)[27], float omega, const SrcMom& mb, float (&fb)[27])
{
    source_moments_vjp_sums(source_sums_of(f), omega, mb, fb);
}

// ---- whole cells: src[8][27] source populations, g the adjoints of the written populations ---

// CF: g[8][27] of the eight fine children -> srcb[8][27] of the eight coarse nodes
VF_HD void cf_cell_vjp(const float (&src)[8][27], float omegaC, float omegaF, float xo, float yo, float zo,
                       const float (&g)[8][27], float (&srcb)[8][27])
{
    SrcMom m[8];
    for (int n = 0; n < 8; ++n) m[n] = source_moments_of(src[n], omegaC);
    const Coef c = coefficients(m, xo, yo, zo);
    Coef cb = coef_zero();
    for (int q = 0; q < 8; ++q) {
        const float x = cf_child_offset(q, 0), y = cf_child_offset(q, 1), z = cf_child_offset(q, 2);
        const NodeMoments nm = cf_node_moments(c, x, y, z, 0.5f, omegaF);
        const NodeMoments nb = rebuild_vjp(nm, g[q]);
        cf_node_moments_vjp(c, x, y, z, 0.5f, omegaF, nb, cb);
    }
    SrcMom mb[8];
    coefficients_vjp(m, xo, yo, zo, cb, mb);
    for (int n = 0; n < 8; ++n) source_moments_vjp(src[n], omegaC, mb[n], srcb[n]);
}

// FC: g[27] of the coarse node -> srcb[8][27] of the eight fine nodes
VF_HD void fc_cell_vjp(const float (&src)[8][27], float omegaF, float omegaC, float xo, float yo, float zo,
                       const float (&g)[27], float (&srcb)[8][27])
{
    SrcMom m[8];
    for (int n = 0; n < 8; ++n) m[n] = source_moments_of(src[n], omegaF);
    const Coef c = coefficients(m, xo, yo, zo);
    Coef cb = coef_zero();
    const NodeMoments nm = fc_node_moments(c, omegaC);
    const NodeMoments nb = rebuild_vjp(nm, g);
    fc_node_moments_vjp(c, omegaC, nb, cb);
    SrcMom mb[8];
    coefficients_vjp(m, xo, yo, zo, cb, mb);
    for (int n = 0; n < 8; ++n) source_moments_vjp(src[n], omegaF, mb[n], srcb[n]);
}

} // namespace vfa
