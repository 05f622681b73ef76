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
#include "vf_lattice.h"

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

// ---- lane-per-direction forms of the BC kernels and their adjoints ----------------------------
// 32 lanes per BC node (lane d < 27 takes direction d; lanes 27..31 hold zeros but take part in
// every shuffle), 8 nodes per 256-thread workgroup, for lists whose consecutive nodes are not
// neighbours in x (faces normal to x: every node and direction is a cache line of its own). One
// thread per node leaves such a face of ten thousand nodes at a few waves per CU, each waiting on
// 27-deep chains of scattered loads; here the scattered loads of a node are one load per lane
// and 32 times as many waves are in flight. Lists along x rows (walls normal to y or z) keep one
// thread per node, whose loads of a direction are contiguous (vf_q_boundary_conditions::
// nodesAlongRows). The node sums (density, momentum, and the adjoint sums that reach all 27
// directions) are butterflies over the lane group.
constexpr int kLaneGroup = 32;
constexpr int kLaneBlock = 256;
constexpr int kLaneNodes = kLaneBlock / kLaneGroup;

constexpr uint32_t dir_bits(int axis, int sign)
{
    uint32_t m = 0;
    for (int d = 0; d < 27; ++d)
        if ((axis == 0 ? cx(d) : (axis == 1 ? cy(d) : cz(d))) == sign) m |= 1u << d;
    return m;
}

struct LaneDir {
    int d;       // direction of this lane, valid for d < 27
    bool valid;
    int cx, cy, cz;
    int r;       // inv(d)
    float w;     // weight(d), 0 for the idle lanes
};

__device__ __forceinline__ LaneDir lane_dir()
{
    LaneDir L;
    L.d = (int)(threadIdx.x & (kLaneGroup - 1));
    L.valid = L.d < 27;
    const uint32_t bit = L.valid ? 1u << L.d : 0u;
    L.cx = int((dir_bits(0, 1) & bit) != 0) - int((dir_bits(0, -1) & bit) != 0);
    L.cy = int((dir_bits(1, 1) & bit) != 0) - int((dir_bits(1, -1) & bit) != 0);
    L.cz = int((dir_bits(2, 1) & bit) != 0) - int((dir_bits(2, -1) & bit) != 0);
    L.r = L.d == 0 ? 0 : (L.d <= 18 ? ((L.d & 1) ? L.d + 1 : L.d - 1) : 45 - L.d);
    L.w = !L.valid ? 0.0f : (L.d == 0 ? 8.0f / 27.0f : (L.d <= 6 ? 2.0f / 27.0f : (L.d <= 18 ? 1.0f / 54.0f : 1.0f / 216.0f)));
    return L;
}

__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int m = kLaneGroup / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, kLaneGroup);
    return v;
}

// node k moved by +1 along the axes whose flag is set (the cell node that holds a direction)
__device__ __forceinline__ uint32_t moved_node(const Addr& a, uint32_t k, bool sx, bool sy, bool sz)
{
    if (a.nbx) {
        if (sx) k = a.nbx[k];
        if (sy) k = a.nby[k];
        if (sz) k = a.nbz[k];
        return k;
    }
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
    uint32_t x = k % nx, t = k / nx, y = t % ny, z = t / ny;
    if (sx) x = x + 1 == nx ? 0 : x + 1;
    if (sy) y = y + 1 == ny ? 0 : y + 1;
    if (sz) z = z + 1 == nz ? 0 : z + 1;
    return x + nx * (y + ny * z);
}

} // namespace vfa
