// ORACLE — test infrastructure only (builds oracle/_ref/, never shipped).
//
// A thin C-ABI harness around the reference's own header-only LBM arithmetic
// (src/lbm/**), compiled directly from the sources under the reference tree so
// that tests can check the CPU restatement in oracle_k17.h against the real code.
// Built twice: with -DVF_DOUBLE_ACCURACY (real = double) and without (real = float).
#include <cstring>

#include <lbm/ChimeraTransformation.h>
#include <lbm/MacroscopicQuantities.h>
#include <lbm/collision/K17CompressibleNavierStokes.h>
#include <lbm/interpolation/InterpolationCoefficients.h>
#include <lbm/refinement/InterpolationCF.h>
#include <lbm/refinement/InterpolationFC.h>

extern "C" int ref_real_size() { return (int)sizeof(real); }

// K17 collision on n independent nodes (node-major [n][27]); macro4 = drho, vx, vy, vz.
extern "C" void ref_k17_nodes(const real* fin, real* fout, long long n, real omega, const real* lim3,
                              const real* force3, real* macro4)
{
    real lim[3] = {lim3[0], lim3[1], lim3[2]};
    for (long long i = 0; i < n; ++i) {
        vf::lbm::CollisionParameter p;
        std::memcpy(p.distribution, fin + 27 * i, 27 * sizeof(real));
        p.omega = omega;
        p.quadricLimiter = lim;
        p.forceX = force3 ? force3[0] : real(0);
        p.forceY = force3 ? force3[1] : real(0);
        p.forceZ = force3 ? force3[2] : real(0);
        vf::lbm::MacroscopicValues mv;
        vf::lbm::TurbulentViscosity tv;
        vf::lbm::runK17CompressibleNavierStokes<vf::lbm::TurbulenceModel::None>(p, mv, tv);
        std::memcpy(fout + 27 * i, p.distribution, 27 * sizeof(real));
        if (macro4) {
            macro4[4 * i + 0] = mv.rho;
            macro4[4 * i + 1] = mv.vx;
            macro4[4 * i + 2] = mv.vy;
            macro4[4 * i + 3] = mv.vz;
        }
    }
}

// kind: 0 forwardInverseChimeraWithK, 1 backwardInverseChimeraWithK, 2 forwardChimera, 3 backwardChimera
extern "C" void ref_chimera(int kind, real* abc, real u, real u2, real Kinv, real K)
{
    if (kind == 0) vf::lbm::forwardInverseChimeraWithK(abc[0], abc[1], abc[2], u, u2, Kinv, K);
    if (kind == 1) vf::lbm::backwardInverseChimeraWithK(abc[0], abc[1], abc[2], u, u2, Kinv, K);
    if (kind == 2) vf::lbm::forwardChimera(abc[0], abc[1], abc[2], u, u2);
    if (kind == 3) vf::lbm::backwardChimera(abc[0], abc[1], abc[2], u, u2);
}

// out: drho, jx, jy, jz, 1/rho, ux, uy, uz, pressure(using drho, ux, uy, uz)
extern "C" void ref_moments(const real* f, real* out9)
{
    out9[0] = vf::lbm::getDensity(f);
    out9[1] = vf::lbm::getIncompressibleVelocityX1(f);
    out9[2] = vf::lbm::getIncompressibleVelocityX2(f);
    out9[3] = vf::lbm::getIncompressibleVelocityX3(f);
    real d, ir, ux, uy, uz;
    vf::lbm::getCompressibleMacroscopicValues(f, d, ir, ux, uy, uz);
    out9[4] = ir;
    out9[5] = ux;
    out9[6] = uy;
    out9[7] = uz;
    out9[8] = vf::lbm::getPressure(f, d, ux, uy, uz);
}

// One interpolation cell through the reference's MomentsOnSourceNodeSet, calculateCoefficients,
// interpolateCF / interpolateFC. f8: node-major [8][27] with n = ix + 2 iy + 4 iz (M = 0, P = 1).
// coeff42 follows the member order of InterpolationCoefficients.
extern "C" void ref_interp_cell(const real* f8, real omegaSrc, real xo, real yo, real zo, int mode, real omegaDst,
                                real eps, real x, real y, real z, real* coeff42, real* fout27)
{
    vf::lbm::MomentsOnSourceNodeSet set;
    set.calculateMMM(f8 + 27 * 0, omegaSrc);
    set.calculatePMM(f8 + 27 * 1, omegaSrc);
    set.calculateMPM(f8 + 27 * 2, omegaSrc);
    set.calculatePPM(f8 + 27 * 3, omegaSrc);
    set.calculateMMP(f8 + 27 * 4, omegaSrc);
    set.calculatePMP(f8 + 27 * 5, omegaSrc);
    set.calculateMPP(f8 + 27 * 6, omegaSrc);
    set.calculatePPP(f8 + 27 * 7, omegaSrc);
    vf::lbm::InterpolationCoefficients c;
    set.calculateCoefficients(c, xo, yo, zo);
    static_assert(sizeof(c) == 42 * sizeof(real), "coefficient layout");
    std::memcpy(coeff42, &c, sizeof(c));
    if (mode == 0)
        vf::lbm::interpolateCF(fout27, omegaDst, eps, c, x, y, z);
    else
        vf::lbm::interpolateFC(fout27, eps, omegaDst, c);
}
