// ORACLE — test infrastructure only. Not part of the product; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// CPU restatement of the compressible grid-refinement interpolation:
//   source-node moments     src/lbm/interpolation/InterpolationCoefficients.h:74-144
//   polynomial coefficients src/lbm/interpolation/InterpolationCoefficients.h:201-430
//   coarse -> fine node     src/lbm/refinement/InterpolationCF.h:57-306
//   fine -> coarse node     src/lbm/refinement/InterpolationFC.h:56-263
//
// The eight source nodes of a cell sit at (sx, sy, sz)/2 with s = -1 (M) or +1 (P).
// Every coefficient of the reference is a short combination of the signed cell sums
//   S_p(q) = sum_n p(n) q(n),  p in {1, sx, sy, sz, sxsy, sxsz, sysz, sxsysz},
// of the nodes' velocities, drho and non-equilibrium second moments; this file writes
// them that way. Rounding therefore differs from the reference's summation order at the
// 1e-16 level; tests compare against the reference build with a relative tolerance.
#pragma once

#include "oracle_k17.h"

namespace orc {

// node index n = ix + 2 iy + 4 iz, i = 0 for M and 1 for P
inline int node_sign(int n, int axis) { return ((n >> axis) & 1) ? 1 : -1; }

template <class T> struct SourceMoments {
    T drho, vx, vy, vz, kxy, kyz, kxz, kxxMyy, kxxMzz;
};

// MomentsOnSourceNode::calculate: macroscopic values and the non-equilibrium part of the
// second-order moments, scaled by -3 omega (off-diagonal) and -3/2 omega (normal differences).
template <class T> SourceMoments<T> source_moments(const T* f, T omega)
{
    SourceMoments<T> m;
    T inv_rho;
    compressible_macro(f, m.drho, inv_rho, m.vx, m.vy, m.vz);
    T pxy = 0, pyz = 0, pxz = 0, pxxyy = 0, pxxzz = 0;
    for (int d = 0; d < 27; ++d) {
        pxy += T(CX[d] * CY[d]) * f[d];
        pyz += T(CY[d] * CZ[d]) * f[d];
        pxz += T(CX[d] * CZ[d]) * f[d];
        pxxyy += T(CX[d] * CX[d] - CY[d] * CY[d]) * f[d];
        pxxzz += T(CX[d] * CX[d] - CZ[d] * CZ[d]) * f[d];
    }
    m.kxy = -T(3) * omega * (pxy * inv_rho - m.vx * m.vy);
    m.kyz = -T(3) * omega * (pyz * inv_rho - m.vy * m.vz);
    m.kxz = -T(3) * omega * (pxz * inv_rho - m.vx * m.vz);
    m.kxxMyy = -T(1.5) * omega * (pxxyy * inv_rho - (m.vx * m.vx - m.vy * m.vy));
    m.kxxMzz = -T(1.5) * omega * (pxxzz * inv_rho - (m.vx * m.vx - m.vz * m.vz));
    return m;
}

// Coefficients of the interpolation polynomials (InterpolationCoefficients):
//   u_x = a000 + a100 x + ... + a111 xyz,  u_y: b..., u_z: c..., drho: d...
template <class T> struct Coefficients {
    T a000, a100, a010, a001, a200, a020, a002, a110, a101, a011;
    T b000, b100, b010, b001, b200, b020, b002, b110, b101, b011;
    T c000, c100, c010, c001, c200, c020, c002, c110, c101, c011;
    T d000, d100, d010, d001, d110, d101, d011;
    T a111, b111, c111, d111;
    T LaplaceRho;
};

template <class T> Coefficients<T> coefficients(const SourceMoments<T> (&m)[8], T xoff, T yoff, T zoff)
{
    // signed sums over the cell
    enum { P1 = 0, PX, PY, PZ, PXY, PXZ, PYZ, PXYZ };
    auto S = [&](T SourceMoments<T>::*q, int pat) {
        T s = T(0);
        for (int n = 0; n < 8; ++n) {
            const int sx = node_sign(n, 0), sy = node_sign(n, 1), sz = node_sign(n, 2);
            int w = 1;
            switch (pat) {
            case PX: w = sx; break;
            case PY: w = sy; break;
            case PZ: w = sz; break;
            case PXY: w = sx * sy; break;
            case PXZ: w = sx * sz; break;
            case PYZ: w = sy * sz; break;
            case PXYZ: w = sx * sy * sz; break;
            default: break;
            }
            s += T(w) * (m[n].*q);
        }
        return s;
    };
    using M = SourceMoments<T>;
    const T h = T(0.5), q4 = T(0.25), e8 = T(1) / T(8), s16 = T(1) / T(16), s64 = T(1) / T(64), two = T(2);
    Coefficients<T> c;
    c.a000 = s64 * (two * (-S(&M::kxy, PY) - S(&M::kxz, PZ) + S(&M::vy, PXY) + S(&M::vz, PXZ)) + T(8) * S(&M::vx, P1) -
                    S(&M::kxxMyy, PX) - S(&M::kxxMzz, PX));
    c.b000 = s64 * (two * (S(&M::kxxMyy, PY) - S(&M::kxy, PX) - S(&M::kyz, PZ) + S(&M::vx, PXY) + S(&M::vz, PYZ)) +
                    T(8) * S(&M::vy, P1) - S(&M::kxxMzz, PY));
    c.c000 = s64 * (two * (S(&M::kxxMzz, PZ) - S(&M::kxz, PX) - S(&M::kyz, PY) + S(&M::vx, PXZ) + S(&M::vy, PYZ)) +
                    T(8) * S(&M::vz, P1) - S(&M::kxxMyy, PZ));
    c.a100 = q4 * S(&M::vx, PX);
    c.b100 = q4 * S(&M::vy, PX);
    c.c100 = q4 * S(&M::vz, PX);
    c.a200 = s16 * (two * (S(&M::vy, PXY) + S(&M::vz, PXZ)) + S(&M::kxxMyy, PX) + S(&M::kxxMzz, PX));
    c.b200 = e8 * (-two * S(&M::vx, PXY) + S(&M::kxy, PX));
    c.c200 = e8 * (-two * S(&M::vx, PXZ) + S(&M::kxz, PX));
    c.a010 = q4 * S(&M::vx, PY);
    c.b010 = q4 * S(&M::vy, PY);
    c.c010 = q4 * S(&M::vz, PY);
    c.a020 = e8 * (-two * S(&M::vy, PXY) + S(&M::kxy, PY));
    c.b020 = s16 * (two * (-S(&M::kxxMyy, PY) + S(&M::vx, PXY) + S(&M::vz, PYZ)) + S(&M::kxxMzz, PY));
    c.c020 = e8 * (-two * S(&M::vy, PYZ) + S(&M::kyz, PY));
    c.a001 = q4 * S(&M::vx, PZ);
    c.b001 = q4 * S(&M::vy, PZ);
    c.c001 = q4 * S(&M::vz, PZ);
    c.a002 = e8 * (-two * S(&M::vz, PXZ) + S(&M::kxz, PZ));
    c.b002 = e8 * (-two * S(&M::vz, PYZ) + S(&M::kyz, PZ));
    c.c002 = s16 * (two * (-S(&M::kxxMzz, PZ) + S(&M::vx, PXZ) + S(&M::vy, PYZ)) + S(&M::kxxMyy, PZ));
    c.a110 = h * S(&M::vx, PXY);
    c.b110 = h * S(&M::vy, PXY);
    c.c110 = h * S(&M::vz, PXY);
    c.a101 = h * S(&M::vx, PXZ);
    c.b101 = h * S(&M::vy, PXZ);
    c.c101 = h * S(&M::vz, PXZ);
    c.a011 = h * S(&M::vx, PYZ);
    c.b011 = h * S(&M::vy, PYZ);
    c.c011 = h * S(&M::vz, PYZ);
    c.a111 = S(&M::vx, PXYZ);
    c.b111 = S(&M::vy, PXYZ);
    c.c111 = S(&M::vz, PXYZ);
    // pressure curvature, only for a cell centred interpolation point
    c.LaplaceRho = (xoff != T(0) || yoff != T(0) || zoff != T(0))
                       ? T(0)
                       : -T(3) * (c.a100 * c.a100 + c.b010 * c.b010 + c.c001 * c.c001) -
                             T(6) * (c.b100 * c.a010 + c.c100 * c.a001 + c.c010 * c.b001);
    c.d000 = e8 * S(&M::drho, P1);
    c.d100 = q4 * S(&M::drho, PX);
    c.d010 = q4 * S(&M::drho, PY);
    c.d001 = q4 * S(&M::drho, PZ);
    c.d110 = h * S(&M::drho, PXY);
    c.d101 = h * S(&M::drho, PXZ);
    c.d011 = h * S(&M::drho, PYZ);
    c.d111 = S(&M::drho, PXYZ);

    // shift the expansion point by the offset (refinement towards walls)
    const T xx = xoff * xoff, yy = yoff * yoff, zz = zoff * zoff;
    c.a000 = c.a000 + xoff * c.a100 + yoff * c.a010 + zoff * c.a001 + xx * c.a200 + yy * c.a020 + zz * c.a002 +
             xoff * yoff * c.a110 + xoff * zoff * c.a101 + yoff * zoff * c.a011;
    c.a100 = c.a100 + two * xoff * c.a200 + yoff * c.a110 + zoff * c.a101;
    c.a010 = c.a010 + two * yoff * c.a020 + xoff * c.a110 + zoff * c.a011;
    c.a001 = c.a001 + two * zoff * c.a002 + xoff * c.a101 + yoff * c.a011;
    c.b000 = c.b000 + xoff * c.b100 + yoff * c.b010 + zoff * c.b001 + xx * c.b200 + yy * c.b020 + zz * c.b002 +
             xoff * yoff * c.b110 + xoff * zoff * c.b101 + yoff * zoff * c.b011;
    c.b100 = c.b100 + two * xoff * c.b200 + yoff * c.b110 + zoff * c.b101;
    c.b010 = c.b010 + two * yoff * c.b020 + xoff * c.b110 + zoff * c.b011;
    c.b001 = c.b001 + two * zoff * c.b002 + xoff * c.b101 + yoff * c.b011;
    c.c000 = c.c000 + xoff * c.c100 + yoff * c.c010 + zoff * c.c001 + xx * c.c200 + yy * c.c020 + zz * c.c002 +
             xoff * yoff * c.c110 + xoff * zoff * c.c101 + yoff * zoff * c.c011;
    c.c100 = c.c100 + two * xoff * c.c200 + yoff * c.c110 + zoff * c.c101;
    c.c010 = c.c010 + two * yoff * c.c020 + xoff * c.c110 + zoff * c.c011;
    c.c001 = c.c001 + two * zoff * c.c002 + xoff * c.c101 + yoff * c.c011;
    c.d000 = c.d000 + xoff * c.d100 + yoff * c.d010 + zoff * c.d001 + xoff * yoff * c.d110 + xoff * zoff * c.d101 +
             yoff * zoff * c.d011;
    c.d100 = c.d100 + yoff * c.d110 + zoff * c.d101;
    c.d010 = c.d010 + xoff * c.d110 + zoff * c.d011;
    c.d001 = c.d001 + xoff * c.d101 + yoff * c.d011;
    return c;
}

// Central moments of the rebuilt node -> shifted populations (the K17 backward passes).
template <class T>
void moments_to_populations(T* f, T press, T vx, T vy, T vz, T mxxMyy, T mxxMzz, T m011, T m101, T m110)
{
    for (int d = 0; d < 27; ++d) f[d] = T(0);
    // slot of m_abc: direction (a-1, b-1, c-1)
    f[dMMM] = press;                                                   // m000
    const T mxxPyyPzz = press;
    f[dPMM] = T(1) / T(3) * (mxxMyy + mxxMzz + mxxPyyPzz);             // m200
    f[dMPM] = T(1) / T(3) * (-T(2) * mxxMyy + mxxMzz + mxxPyyPzz);     // m020
    f[dMMP] = T(1) / T(3) * (mxxMyy - T(2) * mxxMzz + mxxPyyPzz);      // m002
    f[dM00] = m011;
    f[d0M0] = m101;
    f[d00M] = m110;
    f[dMPP] = press * (T(1) / T(9));                                   // m022
    f[dPMP] = press * (T(1) / T(9));                                   // m202
    f[dPPM] = press * (T(1) / T(9));                                   // m220
    f[dPPP] = press * (T(1) / T(27));                                  // m222
    pass_bwd(f, kPassXBwd, vx, vx * vx);
    pass_bwd(f, kPassY, vy, vy * vy);
    pass_bwd(f, kPassZ, vz, vz * vz);
}

// interpolateCF: the fine node at (x, y, z) (coarse units, cell centre at 0)
template <class T> void interpolate_cf(T* f, T omegaF, T eps, const Coefficients<T>& c, T x, T y, T z)
{
    const T press = c.d000 + x * c.d100 + y * c.d010 + z * c.d001 + x * y * c.d110 + x * z * c.d101 + y * z * c.d011 +
                    x * y * z * c.d111 + T(3) * x * x * c.LaplaceRho;
    const T vx = c.a000 + x * c.a100 + y * c.a010 + z * c.a001 + x * x * c.a200 + y * y * c.a020 + z * z * c.a002 +
                 x * y * c.a110 + x * z * c.a101 + y * z * c.a011 + x * y * z * c.a111;
    const T vy = c.b000 + x * c.b100 + y * c.b010 + z * c.b001 + x * x * c.b200 + y * y * c.b020 + z * z * c.b002 +
                 x * y * c.b110 + x * z * c.b101 + y * z * c.b011 + x * y * z * c.b111;
    const T vz = c.c000 + x * c.c100 + y * c.c010 + z * c.c001 + x * x * c.c200 + y * y * c.c020 + z * z * c.c002 +
                 x * y * c.c110 + x * z * c.c101 + y * z * c.c011 + x * y * z * c.c111;
    const T scale = eps / omegaF * (T(1) + press);
    // velocity gradients of the polynomial at (x, y, z), the non-equilibrium of the new node
    const T dxMdy = c.a100 - c.b010 + T(2) * c.a200 * x - c.b110 * x + c.a110 * y - T(2) * c.b020 * y + c.a101 * z -
                    c.b011 * z - c.b111 * x * z + c.a111 * y * z;
    const T dxMdz = c.a100 - c.c001 + T(2) * c.a200 * x - c.c101 * x + c.a110 * y - c.c011 * y - c.c111 * x * y +
                    c.a101 * z - T(2) * c.c002 * z + c.a111 * y * z;
    const T syz = c.b001 + c.c010 + c.b101 * x + c.c110 * x + c.b011 * y + T(2) * c.c020 * y + c.b111 * x * y +
                  T(2) * c.b002 * z + c.c011 * z + c.c111 * x * z;
    const T sxz = c.a001 + c.c100 + c.a101 * x + T(2) * c.c200 * x + c.a011 * y + c.c110 * y + c.a111 * x * y +
                  T(2) * c.a002 * z + c.c101 * z + c.c111 * y * z;
    const T sxy = c.a010 + c.b100 + c.a110 * x + T(2) * c.b200 * x + T(2) * c.a020 * y + c.b110 * y + c.a011 * z +
                  c.b101 * z + c.a111 * x * z + c.b111 * y * z;
    moments_to_populations(f, press, vx, vy, vz, -T(2) / T(3) * dxMdy * scale, -T(2) / T(3) * dxMdz * scale,
                           -T(1) / T(3) * syz * scale, -T(1) / T(3) * sxz * scale, -T(1) / T(3) * sxy * scale);
}

// interpolateFC: the coarse node at the cell centre
template <class T> void interpolate_fc(T* f, T eps, T omegaC, const Coefficients<T>& c)
{
    const T press = c.d000 - T(2) * c.LaplaceRho * (T(1) / T(8));
    const T scale = eps / omegaC * (T(1) + press);
    moments_to_populations(f, press, c.a000, c.b000, c.c000, -T(2) / T(3) * (c.a100 - c.b010) * scale,
                           -T(2) / T(3) * (c.a100 - c.c001) * scale, -T(1) / T(3) * (c.b001 + c.c010) * scale,
                           -T(1) / T(3) * (c.a001 + c.c100) * scale, -T(1) / T(3) * (c.a010 + c.b100) * scale);
}

} // namespace orc
