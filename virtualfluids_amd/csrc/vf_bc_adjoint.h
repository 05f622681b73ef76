// Per-node reverse mode of the boundary-condition kernels (host and device).
//
// bc_post_vjp: for one BC node with post-collision values `post` and link distances q[d]
// (active when 0 <= q <= 1), G[r] is the adjoint of the value written for the incoming
// direction r = inv(d); returns the adjoint of the node's post-collision values and of the wall
// velocity. pressure_vjp: for one pressure-BC node, G[d] is the adjoint of the written value of
// direction d (zero where the face writes nothing); returns the adjoint of the neighbour's
// pre-collision values and of the prescribed density.
#pragma once

#include "vf_k17.h"

namespace vfa {

enum BcKind : int { kNoSlipBB = 0, kVelocityBB = 1, kNoSlipInterp = 2, kVelocityInterp = 3, kVelocityPressureInterp = 4 };

// One wave per workgroup for the BC kernels (one thread per BC node): a boundary has a few
// thousand to a few hundred thousand nodes, each a short latency-bound chain, and 64-thread
// workgroups spread them over every CU (4 096 inflow nodes are 64 workgroups, not 16).
constexpr int kBcBlock = 64;

template <int KIND>
VF_HD void bc_post_vjp(const float (&post)[27], const float (&q)[27], float omega, const float (&G)[27],
                       float (&fb)[27], float (&Ub)[3])
{
    constexpr bool interp = KIND >= kNoSlipInterp;
    constexpr bool velocity = KIND == kVelocityBB || KIND == kVelocityInterp || KIND == kVelocityPressureInterp;
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    if (interp) {
#pragma unroll
        for (int d = 0; d < 27; ++d) {
            drho += post[d];
            if (cx(d)) jx += float(cx(d)) * post[d];
            if (cy(d)) jy += float(cy(d)) * post[d];
            if (cz(d)) jz += float(cz(d)) * post[d];
        }
    }
    const float ir = 1.0f / (1.0f + drho);
    const float j2 = jx * jx + jy * jy + jz * jz;
    const float om1 = 1.0f / (1.0f - omega);
#pragma unroll
    for (int d = 0; d < 27; ++d) fb[d] = 0.0f;
    Ub[0] = Ub[1] = Ub[2] = 0.0f;
    float drhob = 0.0f, jxb = 0.0f, jyb = 0.0f, jzb = 0.0f;
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const float qd = q[d];
        if (!(qd >= 0.0f && qd <= 1.0f)) continue;
        const int r = inv(d);
        const float w = weight(d);
        const float g = G[r];
        float wallFactor = 6.0f * w; // d value / d (c.U) = -wallFactor
        if (!interp) {
            fb[d] += g;
        } else {
            const float A = 0.5f * (1.0f - qd) / (1.0f + qd), Bq = 1.0f / (1.0f + qd);
            fb[d] += g * (A + A * om1 + Bq * qd);
            fb[r] += g * (-A + A * om1 + Bq * qd);
            // feq = w (drho + (4.5 (c.j)^2 - 1.5 |j|^2) / (1 + drho))
            const float cj = float(cx(d)) * jx + float(cy(d)) * jy + float(cz(d)) * jz;
            const float Q = 4.5f * cj * cj - 1.5f * j2;
            const float feqb = g * (-2.0f * A * omega * om1);
            drhob += feqb * w * (1.0f - Q * ir * ir);
            const float t = feqb * w * ir;
            jxb += t * (9.0f * cj * float(cx(d)) - 3.0f * jx);
            jyb += t * (9.0f * cj * float(cy(d)) - 3.0f * jy);
            jzb += t * (9.0f * cj * float(cz(d)) - 3.0f * jz);
            if (KIND == kVelocityPressureInterp) drhob -= g * w;
            wallFactor *= Bq;
        }
        if (velocity) {
            Ub[0] -= wallFactor * float(cx(d)) * g;
            Ub[1] -= wallFactor * float(cy(d)) * g;
            Ub[2] -= wallFactor * float(cz(d)) * g;
        }
    }
    if (interp) {
#pragma unroll
        for (int e = 0; e < 27; ++e) fb[e] += drhob + float(cx(e)) * jxb + float(cy(e)) * jyb + float(cz(e)) * jzb;
    }
}

// v_d = fn_d + w_d (rhoBC - drho) (1 + E_d(u)),  E_d = 3 c.u + 4.5 (c.u)^2 - 1.5 |u|^2
VF_HD void pressure_vjp(const float (&fn)[27], float rhoBC, const float (&G)[27], float (&fnb)[27], float& rhob)
{
    float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        drho += fn[d];
        if (cx(d)) jx += float(cx(d)) * fn[d];
        if (cy(d)) jy += float(cy(d)) * fn[d];
        if (cz(d)) jz += float(cz(d)) * fn[d];
    }
    const float ir = 1.0f / (drho + 1.0f);
    const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    const float usq = ux * ux + uy * uy + uz * uz;
    float A1 = 0.0f, Vx = 0.0f, Vy = 0.0f, Vz = 0.0f;
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const float w = weight(d);
        const float cu = float(cx(d)) * ux + float(cy(d)) * uy + float(cz(d)) * uz;
        const float E = 3.0f * cu + 4.5f * cu * cu - 1.5f * usq;
        const float t = G[d] * w;
        A1 += t * (1.0f + E);
        Vx += t * (3.0f * float(cx(d)) + 9.0f * cu * float(cx(d)) - 3.0f * ux);
        Vy += t * (3.0f * float(cy(d)) + 9.0f * cu * float(cy(d)) - 3.0f * uy);
        Vz += t * (3.0f * float(cz(d)) + 9.0f * cu * float(cz(d)) - 3.0f * uz);
    }
    const float s = (rhoBC - drho) * ir;
    Vx *= s;
    Vy *= s;
    Vz *= s;
#pragma unroll
    for (int e = 0; e < 27; ++e)
        fnb[e] = G[e] - A1 + Vx * (float(cx(e)) - ux) + Vy * (float(cy(e)) - uy) + Vz * (float(cz(e)) - uz);
    rhob = A1;
}

} // namespace vfa
