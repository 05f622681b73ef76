// Grid-refinement interpolation kernels (compressible, K17 chimera moments).
//
//   coarse -> fine  scaleCoarseToFineCompressible_Device (src/gpu/core/GridScaling/scaleCoarseToFineCompressible.cu:305-528)
//   fine -> coarse  scaleFineToCoarseCompressible_Device (src/gpu/core/GridScaling/scaleFineToCoarseCompressible.cu:76-129)
//   arithmetic      src/lbm/interpolation/InterpolationCoefficients.h, src/lbm/refinement/Interpolation{CF,FC}.h
//
// One thread per interface cell. The 8 source nodes of a cell are the corners of the cell
// whose lowest corner is the listed index, n = ix + 2 iy + 4 iz (M = 0, P = 1); each source
// node's pre-collision state gives drho, u and the non-equilibrium second moments; every
// coefficient of the interpolation polynomials is a short combination of the signed sums
// S_p(q) = sum_n p(n) q(n), p in {1, sx, sy, sz, sxsy, sxsz, sysz, sxsysz}. CF writes the 8
// fine nodes at (+-1/4, +-1/4, +-1/4) of the coarse cell with the fine grid's even parity
// (the fine grid has taken two steps); FC writes the one coarse node at the centre of the
// fine cell with the coarse parity, reading the fine grid with the even parity.
#include "vf_k17.h"
#include "vf_lattice.h"

namespace vfa {

struct SrcMom {
    float drho, vx, vy, vz, kxy, kyz, kxz, kxxMyy, kxxMzz;
};

template <class S>
__device__ __forceinline__ SrcMom source_moments(const S* f, const Addr& a, uint32_t k, bool even, float omega)
{
    uint32_t n[8];
    cell_nodes(a, k, n);
    float drho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f, pxy = 0.f, pyz = 0.f, pxz = 0.f, pxxyy = 0.f, pxxzz = 0.f;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const float v = read_pre(f, a, n, d, even);
        drho += v;
        if (cx(d)) jx += float(cx(d)) * v;
        if (cy(d)) jy += float(cy(d)) * v;
        if (cz(d)) jz += float(cz(d)) * v;
        if (cx(d) * cy(d)) pxy += float(cx(d) * cy(d)) * v;
        if (cy(d) * cz(d)) pyz += float(cy(d) * cz(d)) * v;
        if (cx(d) * cz(d)) pxz += float(cx(d) * cz(d)) * v;
        if (cx(d) * cx(d) - cy(d) * cy(d)) pxxyy += float(cx(d) * cx(d) - cy(d) * cy(d)) * v;
        if (cx(d) * cx(d) - cz(d) * cz(d)) pxxzz += float(cx(d) * cx(d) - cz(d) * cz(d)) * v;
    }
    const float ir = 1.0f / (drho + 1.0f);
    SrcMom m;
    m.drho = drho;
    m.vx = jx * ir;
    m.vy = jy * ir;
    m.vz = jz * ir;
    m.kxy = -3.0f * omega * (pxy * ir - m.vx * m.vy);
    m.kyz = -3.0f * omega * (pyz * ir - m.vy * m.vz);
    m.kxz = -3.0f * omega * (pxz * ir - m.vx * m.vz);
    m.kxxMyy = -1.5f * omega * (pxxyy * ir - (m.vx * m.vx - m.vy * m.vy));
    m.kxxMzz = -1.5f * omega * (pxxzz * ir - (m.vx * m.vx - m.vz * m.vz));
    return m;
}

struct Coef {
    float a000, a100, a010, a001, a200, a020, a002, a110, a101, a011, a111;
    float b000, b100, b010, b001, b200, b020, b002, b110, b101, b011, b111;
    float c000, c100, c010, c001, c200, c020, c002, c110, c101, c011, c111;
    float d000, d100, d010, d001, d110, d101, d011, d111;
    float LaplaceRho;
};

// signed cell sums of one quantity: [1, x, y, z, xy, xz, yz, xyz]
struct Sums {
    float s[8];
};
__device__ __forceinline__ Sums signed_sums(const float (&q)[8])
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

__device__ Coef coefficients(const SrcMom (&m)[8], float xo, float yo, float zo)
{
    enum { P1 = 0, PX, PY, PZ, PXY, PXZ, PYZ, PXYZ };
    float q[8];
#define VF_SUMS(name, field)                              \
    _Pragma("unroll") for (int n = 0; n < 8; ++n) q[n] = m[n].field; \
    const Sums name = signed_sums(q);
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

// Central moments of a rebuilt node (zeroth, second order with non-equilibrium, fourth and
// sixth order at their rest values, everything else zero) -> shifted populations, through the
// same backward chimera passes as the collision (src/lbm/refinement/InterpolationCF.h:189-291).
__device__ void rebuild_populations(float (&f)[27], float press, float vx, float vy, float vz, float mxxMyy,
                                    float mxxMzz, float m011, float m101, float m110)
{
    float M000 = press, M100 = 0.f, M200 = (1.f / 3.f) * (mxxMyy + mxxMzz + press);
    float M010 = 0.f, M110 = m110, M210 = 0.f;
    float M020 = (1.f / 3.f) * (-2.f * mxxMyy + mxxMzz + press), M120 = 0.f, M220 = press * (1.f / 9.f);
    float M001 = 0.f, M101 = m101, M201 = 0.f;
    float M011 = m011, M111 = 0.f, M211 = 0.f;
    float M021 = 0.f, M121 = 0.f, M221 = 0.f;
    float M002 = (1.f / 3.f) * (mxxMyy - 2.f * mxxMzz + press), M102 = 0.f, M202 = press * (1.f / 9.f);
    float M012 = 0.f, M112 = 0.f, M212 = 0.f;
    float M022 = press * (1.f / 9.f), M122 = 0.f, M222 = press * (1.f / 27.f);
    {
        const BwdPass p = bwd_pass(vx, vx * vx);
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
        const BwdPass p = bwd_pass(vy, vy * vy);
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
        const BwdPass p = bwd_pass(vz, vz * vz);
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
    f[dMMM] = M000; f[d0MM] = M100; f[dPMM] = M200;
    f[dM0M] = M010; f[d00M] = M110; f[dP0M] = M210;
    f[dMPM] = M020; f[d0PM] = M120; f[dPPM] = M220;
    f[dMM0] = M001; f[d0M0] = M101; f[dPM0] = M201;
    f[dM00] = M011; f[d000] = M111; f[dP00] = M211;
    f[dMP0] = M021; f[d0P0] = M121; f[dPP0] = M221;
    f[dMMP] = M002; f[d0MP] = M102; f[dPMP] = M202;
    f[dM0P] = M012; f[d00P] = M112; f[dP0P] = M212;
    f[dMPP] = M022; f[d0PP] = M122; f[dPPP] = M222;
}

__device__ __forceinline__ uint32_t corner(const Addr& a, uint32_t base, int n)
{
    uint32_t k = base;
    if (n & 1) k = a.nbx ? a.nbx[k] : (k % a.nx + 1 == (uint32_t)a.nx ? k + 1 - a.nx : k + 1);
    if (n & 2) {
        if (a.nbx) k = a.nby[k];
        else {
            const uint32_t y = (k / a.nx) % a.ny;
            k = (y + 1 == (uint32_t)a.ny) ? k - (uint32_t)a.nx * (a.ny - 1) : k + a.nx;
        }
    }
    if (n & 4) {
        if (a.nbx) k = a.nbz[k];
        else {
            const uint32_t z = k / ((uint32_t)a.nx * a.ny);
            k = (z + 1 == (uint32_t)a.nz) ? k - (uint32_t)a.nx * a.ny * (a.nz - 1) : k + (uint32_t)a.nx * a.ny;
        }
    }
    return k;
}

struct ScaleArgs {
    const uint32_t* cellsC;
    const uint32_t* cellsF;
    uint32_t n;
    const float* ox;
    const float* oy;
    const float* oz;
    float omegaC, omegaF;
};

template <class S>
__global__ __launch_bounds__(64) void scale_cf_kernel(const S* __restrict__ fc, S* __restrict__ ff, Addr ac, Addr af,
                                                      ScaleArgs s, int evenC)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    SrcMom m[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) m[n] = source_moments(fc, ac, corner(ac, s.cellsC[i], n), evenC != 0, s.omegaC);
    const float xo = s.ox ? s.ox[i] : 0.f, yo = s.oy ? s.oy[i] : 0.f, zo = s.oz ? s.oz[i] : 0.f;
    const Coef c = coefficients(m, xo, yo, zo);
    const float eps = 0.5f;
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        const float x = (q & 1) ? 0.25f : -0.25f, y = (q & 2) ? 0.25f : -0.25f, z = (q & 4) ? 0.25f : -0.25f;
        const float press = c.d000 + x * c.d100 + y * c.d010 + z * c.d001 + x * y * c.d110 + x * z * c.d101 +
                            y * z * c.d011 + x * y * z * c.d111 + 3.f * x * x * c.LaplaceRho;
        const float vx = c.a000 + x * c.a100 + y * c.a010 + z * c.a001 + x * x * c.a200 + y * y * c.a020 +
                         z * z * c.a002 + x * y * c.a110 + x * z * c.a101 + y * z * c.a011 + x * y * z * c.a111;
        const float vy = c.b000 + x * c.b100 + y * c.b010 + z * c.b001 + x * x * c.b200 + y * y * c.b020 +
                         z * z * c.b002 + x * y * c.b110 + x * z * c.b101 + y * z * c.b011 + x * y * z * c.b111;
        const float vz = c.c000 + x * c.c100 + y * c.c010 + z * c.c001 + x * x * c.c200 + y * y * c.c020 +
                         z * z * c.c002 + x * y * c.c110 + x * z * c.c101 + y * z * c.c011 + x * y * z * c.c111;
        const float scale = eps / s.omegaF * (1.f + press);
        const float dxMdy = c.a100 - c.b010 + 2.f * c.a200 * x - c.b110 * x + c.a110 * y - 2.f * c.b020 * y +
                            c.a101 * z - c.b011 * z - c.b111 * x * z + c.a111 * y * z;
        const float dxMdz = c.a100 - c.c001 + 2.f * c.a200 * x - c.c101 * x + c.a110 * y - c.c011 * y -
                            c.c111 * x * y + c.a101 * z - 2.f * c.c002 * z + c.a111 * y * z;
        const float syz = c.b001 + c.c010 + c.b101 * x + c.c110 * x + c.b011 * y + 2.f * c.c020 * y +
                          c.b111 * x * y + 2.f * c.b002 * z + c.c011 * z + c.c111 * x * z;
        const float sxz = c.a001 + c.c100 + c.a101 * x + 2.f * c.c200 * x + c.a011 * y + c.c110 * y +
                          c.a111 * x * y + 2.f * c.a002 * z + c.c101 * z + c.c111 * y * z;
        const float sxy = c.a010 + c.b100 + c.a110 * x + 2.f * c.b200 * x + 2.f * c.a020 * y + c.b110 * y +
                          c.a011 * z + c.b101 * z + c.a111 * x * z + c.b111 * y * z;
        float f[27];
        rebuild_populations(f, press, vx, vy, vz, -(2.f / 3.f) * dxMdy * scale, -(2.f / 3.f) * dxMdz * scale,
                            -(1.f / 3.f) * syz * scale, -(1.f / 3.f) * sxz * scale, -(1.f / 3.f) * sxy * scale);
        uint32_t nn[8];
        cell_nodes(af, corner(af, s.cellsF[i], q), nn);
#pragma unroll
        for (int d = 0; d < 27; ++d) write_pre(ff, af, nn, d, true, f[d]);
    }
}

template <class S>
__global__ __launch_bounds__(64) void scale_fc_kernel(S* __restrict__ fc, const S* __restrict__ ff, Addr ac, Addr af,
                                                      ScaleArgs s, int evenC)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    SrcMom m[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) m[n] = source_moments(ff, af, corner(af, s.cellsF[i], n), true, s.omegaF);
    const float xo = s.ox ? s.ox[i] : 0.f, yo = s.oy ? s.oy[i] : 0.f, zo = s.oz ? s.oz[i] : 0.f;
    const Coef c = coefficients(m, xo, yo, zo);
    const float press = c.d000 - 2.f * c.LaplaceRho * 0.125f;
    const float scale = 2.f / s.omegaC * (1.f + press);
    float f[27];
    rebuild_populations(f, press, c.a000, c.b000, c.c000, -(2.f / 3.f) * (c.a100 - c.b010) * scale,
                        -(2.f / 3.f) * (c.a100 - c.c001) * scale, -(1.f / 3.f) * (c.b001 + c.c010) * scale,
                        -(1.f / 3.f) * (c.a001 + c.c100) * scale, -(1.f / 3.f) * (c.a010 + c.b100) * scale);
    uint32_t nn[8];
    cell_nodes(ac, s.cellsC[i], nn);
#pragma unroll
    for (int d = 0; d < 27; ++d) write_pre(fc, ac, nn, d, evenC != 0, f[d]);
}

static int check_level(const vf_lbm_parameter* p)
{
    if (!p || !p->distributions || p->numberOfNodes == 0 || p->numberOfNodes >= (1ull << 30)) return 0;
    if (!p->neighborX && (uint64_t)p->nx * p->ny * p->nz != p->numberOfNodes) return 0;
    return 1;
}

static int launch_scale(bool fine_to_coarse, const vf_lbm_parameter* c, const vf_lbm_parameter* f,
                        const vf_interpolation_cells* cells, const vf_interpolation_cell_neighbor* off, void* stream)
{
    if (!check_level(c) || !check_level(f) || !cells) {
        set_last_error("grid scaling needs two valid levels and the interface cells");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (cells->numberOfCells == 0) return VF_OK;
    if (!cells->coarseCellIndices || !cells->fineCellIndices) {
        set_last_error("interface cells need coarse and fine cell indices");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (c->storage != f->storage) {
        set_last_error("both levels must use the same population storage type");
        return VF_ERR_INVALID_ARGUMENT;
    }
    if (!(c->omega > 0.f && c->omega < 2.f && f->omega > 0.f && f->omega < 2.f)) {
        set_last_error("omega must lie in (0, 2) on both levels");
        return VF_ERR_INVALID_ARGUMENT;
    }
    ScaleArgs s{cells->coarseCellIndices, cells->fineCellIndices, cells->numberOfCells,
                off ? off->x : nullptr, off ? off->y : nullptr, off ? off->z : nullptr, c->omega, f->omega};
    const Addr ac = make_addr(*c), af = make_addr(*f);
    const dim3 grid((cells->numberOfCells + 63) / 64), block(64);
    hipStream_t st = (hipStream_t)stream;
#define VF_LAUNCH(S)                                                                                                 \
    if (fine_to_coarse)                                                                                              \
        hipLaunchKernelGGL(scale_fc_kernel<S>, grid, block, 0, st, (S*)c->distributions, (const S*)f->distributions, \
                           ac, af, s, c->isEvenTimestep);                                                            \
    else                                                                                                             \
        hipLaunchKernelGGL(scale_cf_kernel<S>, grid, block, 0, st, (const S*)c->distributions, (S*)f->distributions, \
                           ac, af, s, c->isEvenTimestep);
    switch (c->storage) {
    case VF_STORAGE_FP32: VF_LAUNCH(float) break;
    case VF_STORAGE_FP16: VF_LAUNCH(__half) break;
    case VF_STORAGE_BF16: VF_LAUNCH(__hip_bfloat16) break;
    default: set_last_error("unknown storage type"); return VF_ERR_INVALID_ARGUMENT;
    }
#undef VF_LAUNCH
    return check_launch(fine_to_coarse ? "scaleFineToCoarseCompressible_Device execution failed"
                                       : "scaleCoarseToFineCompressible_Device execution failed");
}

} // namespace vfa

using namespace vfa;

extern "C" int vf_ScaleCoarseToFineCompressible(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                const vf_interpolation_cells* coarseToFine,
                                                const vf_interpolation_cell_neighbor* offsets, void* stream)
{
    return launch_scale(false, coarse, fine, coarseToFine, offsets, stream);
}

extern "C" int vf_ScaleFineToCoarseCompressible(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                const vf_interpolation_cells* fineToCoarse,
                                                const vf_interpolation_cell_neighbor* offsets, void* stream)
{
    return launch_scale(true, coarse, fine, fineToCoarse, offsets, stream);
}
